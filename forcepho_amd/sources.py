"""Sources and scenes: the map between a flat parameter vector and the
per-source ``source_struct_dtype`` records the kernel reads.

Mirrors forcepho/sources.py: ``Scene.set_all_source_params`` /
``get_all_source_params`` / ``get_proposal`` (:46-97) and ``Galaxy`` with its
parameter order [flux_1..flux_B, ra, dec, q, pa, sersic, rhalf]
(``SHAPE_COLS`` :527) and ``proposal()`` (:607-621).

The reference obtains the Sersic mixture amplitudes and their derivatives
from scipy splines of a fitted table (sources.py:623-681).  Here the amplitude
model is pluggable: ``LogNormalMixtureAmplitudes`` is a smooth analytic model
with exact derivatives (used by the synthetic scenes), and
``SplineMixtureAmplitudes`` wraps per-radius scipy RectBivariateSplines of a
table (ngrid x rgrid) exactly as the reference does.
"""

import numpy as np

from .proposal import source_struct_dtype

__all__ = ["Scene", "Galaxy", "LogNormalMixtureAmplitudes", "SplineMixtureAmplitudes",
           "DEFAULT_RADII"]

# The first nine radii of forcepho's default Sersic mixture grid,
# exp(arange(log 0.005, log 6, log 2)) arcsec (mixtures/sersic_mix_oned.py:331).
DEFAULT_RADII = 0.005 * 2.0 ** np.arange(9)


class LogNormalMixtureAmplitudes:
    """Mixture amplitudes A_i(n, rh) proportional to a Gaussian in ln(r_i)
    centred at ln(rh) - 0.15 (n - 1) with width 0.6 + 0.25 n, normalised to
    sum to one.  Derivatives are analytic:
        dA_i/dt = A_i (dl_i/dt - sum_j A_j dl_j/dt),  l_i = -z_i^2/2."""

    def __init__(self, radii=DEFAULT_RADII):
        self.radii = np.asarray(radii, dtype=np.float64)
        # a zero radius (point component) is placed at 1e-4 arcsec in ln r
        self._lnr = np.log(np.maximum(self.radii, 1e-4))

    def _z(self, n, rh):
        sig = 0.6 + 0.25 * n
        mu = np.log(rh) - 0.15 * (n - 1.0)
        return (self._lnr - mu) / sig, sig

    def amplitudes(self, n, rh):
        z, _ = self._z(n, rh)
        z2 = z * z
        w = np.exp(-0.5 * (z2 - z2.min()))  # shifted: no underflow of every weight
        return w / w.sum()

    def derivatives(self, n, rh):
        """(dA/dn, dA/drh)"""
        A = self.amplitudes(n, rh)
        z, sig = self._z(n, rh)
        dl_drh = z / (sig * rh)
        dl_dn = z / sig * (-0.15) + z * z * 0.25 / sig
        dA_dn = A * (dl_dn - np.sum(A * dl_dn))
        dA_drh = A * (dl_drh - np.sum(A * dl_drh))
        return dA_dn, dA_drh


class SplineMixtureAmplitudes:
    """Amplitudes from per-radius scipy splines of a fitted table, as
    forcepho/sources.py:623-681 builds them: RectBivariateSplines over the
    (nsersic, rh) grid (``spline_type="rect"``, the reference's default) or
    SmoothBivariateSplines of the flat table (``"smooth"``, :640-642)."""

    _file_cache = {}

    def __init__(self, nsersic_grid, rh_grid, amplitudes, radii, smoothing=None):
        from scipy.interpolate import RectBivariateSpline
        A = np.asarray(amplitudes, dtype=np.float64)  # [n_n, n_rh, n_radii]
        self.radii = np.asarray(radii, dtype=np.float64)
        self.splines = [RectBivariateSpline(nsersic_grid, rh_grid, A[:, :, i], s=smoothing)
                        for i in range(A.shape[-1])]
        self.sersic_range = (float(np.min(nsersic_grid)), float(np.max(nsersic_grid)))
        self.rh_range = (float(np.min(rh_grid)), float(np.max(rh_grid)))

    @classmethod
    def from_table(cls, nsersic, rh, amplitudes, radii, smoothing=None, spline_type="rect"):
        """Splines of a flat table of rows (nsersic[k], rh[k], amplitudes[k, :]).
        "rect": rows ordered sersic-major over the grid of unique values
        (forcepho/sources.py:633-638); "smooth": any scattered rows (:640-642)."""
        A = np.asarray(amplitudes)
        if spline_type == "rect":
            ngrid, rgrid = np.unique(nsersic), np.unique(rh)
            if A.shape[0] != len(ngrid) * len(rgrid):
                raise ValueError("table rows do not form a (nsersic, rh) grid")
            A = A.reshape(len(ngrid), len(rgrid), A.shape[-1])
            return cls(ngrid, rgrid, A, radii, smoothing=smoothing)
        if spline_type != "smooth":
            raise ValueError("spline_type must be 'rect' or 'smooth'")
        from scipy.interpolate import SmoothBivariateSpline
        obj = cls.__new__(cls)
        obj.radii = np.asarray(radii, dtype=np.float64)
        obj.splines = [SmoothBivariateSpline(nsersic, rh, A[:, i], s=smoothing)
                       for i in range(A.shape[-1])]
        obj.sersic_range = (float(np.min(nsersic)), float(np.max(nsersic)))
        obj.rh_range = (float(np.min(rh)), float(np.max(rh)))
        return obj

    @classmethod
    def from_file(cls, splinedata, smoothing=None, spline_type="rect"):
        """Splines of an HDF5 table with datasets nsersic, rh, amplitudes,
        radii (forcepho/sources.py:623-645); cached per file and options."""
        key = (str(splinedata), smoothing, spline_type)
        if key not in cls._file_cache:
            from .formats.hdf5 import File
            with File(splinedata) as f:
                n, r = f["nsersic"][:], f["rh"][:]
                A, radii = f["amplitudes"][:], f["radii"][:]
            cls._file_cache[key] = cls.from_table(n, r, A, radii, smoothing=smoothing,
                                                  spline_type=spline_type)
        return cls._file_cache[key]

    def amplitudes(self, n, rh):
        return np.squeeze(np.array([s(n, rh) for s in self.splines]))

    def derivatives(self, n, rh):
        dn = np.squeeze(np.array([s(n, rh, dx=1) for s in self.splines]))
        dr = np.squeeze(np.array([s(n, rh, dy=1) for s in self.splines]))
        return dn, dr


class Galaxy:
    """A Sersic galaxy: fluxes per band, ra/dec (degrees), q = sqrt(b/a), pa
    (radians), sersic index and half-light radius (arcsec)."""

    SHAPE_COLS = ["ra", "dec", "q", "pa", "sersic", "rhalf"]
    npos = 2
    nshape = 4

    def __init__(self, filters=("band",), amplitude_model=None, radii=None):
        self.filternames = list(filters)
        self.flux = np.zeros(len(self.filternames))
        self.amplitude_model = amplitude_model or LogNormalMixtureAmplitudes(
            DEFAULT_RADII if radii is None else radii)
        self.radii = self.amplitude_model.radii
        self.ra = self.dec = 0.0
        self.q, self.pa, self.sersic, self.rhalf = 1.0, 0.0, 2.0, 0.1
        self.fixed = False
        self.proposal_struct = np.zeros(1, dtype=source_struct_dtype)

    @property
    def n_bands(self):
        return len(self.filternames)

    @property
    def n_params(self):
        return self.n_bands + self.npos + self.nshape

    @property
    def n_gauss(self):
        return len(self.radii)

    @property
    def parameter_names(self):
        return np.array([f"{f}_{self.id}" for f in self.filternames] +
                        [f"{c}_{self.id}" for c in self.SHAPE_COLS])

    def set_params(self, theta):
        theta = np.asarray(theta, dtype=np.float64)
        assert len(theta) == self.n_params, "parameter vector length mismatch"
        nb = self.n_bands
        self.flux = theta[:nb].copy()
        self.ra, self.dec, self.q, self.pa, self.sersic, self.rhalf = theta[nb:nb + 6]

    def get_param_vector(self):
        return np.concatenate([self.flux, [self.ra, self.dec, self.q, self.pa,
                                           self.sersic, self.rhalf]])

    def from_catalog_row(self, row, extra_parameters=(), filternames=None):
        """Parameters from a catalog row with SHAPE_COLS and one column per
        filter name (forcepho/sources.py:458-468)."""
        for p in list(self.SHAPE_COLS) + list(extra_parameters):
            setattr(self, p, float(row[p]))
        if filternames is not None:
            self.filternames = list(filternames)
            self.flux = np.zeros(self.n_bands)
        for j, b in enumerate(self.filternames):
            self.flux[j] = row[b]

    @property
    def amplitudes(self):
        return self.amplitude_model.amplitudes(self.sersic, self.rhalf)

    @property
    def damplitude_dsersic(self):
        return self.amplitude_model.derivatives(self.sersic, self.rhalf)[0]

    @property
    def damplitude_drh(self):
        return self.amplitude_model.derivatives(self.sersic, self.rhalf)[1]

    def proposal(self):
        """One source_struct_dtype record (forcepho/sources.py:607-621)."""
        p = self.proposal_struct
        p["fluxes"][0, :self.n_bands] = self.flux
        p["ra"], p["dec"], p["q"], p["pa"] = self.ra, self.dec, self.q, self.pa
        p["nsersic"], p["rh"] = self.sersic, self.rhalf
        dn, dr = self.amplitude_model.derivatives(self.sersic, self.rhalf)
        p["mixture_amplitudes"][0, :self.n_gauss] = self.amplitudes
        p["damplitude_drh"][0, :self.n_gauss] = dr
        p["damplitude_dnsersic"][0, :self.n_gauss] = dn
        return p


class Scene:
    """An ordered collection of sources and the map to a flat parameter
    vector (forcepho/sources.py:28-215)."""

    def __init__(self, sources=(), catalog=None, **catkwargs):
        self.sources = list(sources)
        self.sky_zero = None  # set when coordinates are shifted to patch offsets
        if catalog is not None:
            self.sources = Scene.from_catalog(catalog, **catkwargs).sources
        for i, s in enumerate(self.sources):
            s.id = i

    @classmethod
    def from_catalog(cls, catalog, filternames=("band",), splinedata=None,
                     spline_smoothing=None, spline_type="rect", amplitude_model=None,
                     extra_parameters=()):
        """A Scene of Galaxies, one per catalog row (forcepho/sources.py:154-182).
        Mixture amplitudes come from ``amplitude_model``, else from
        ``splinedata`` (an HDF5 table path, or an amplitude model object),
        else the default log-normal model."""
        if amplitude_model is None and splinedata is not None and not isinstance(splinedata, str):
            amplitude_model = splinedata
        elif amplitude_model is None and splinedata:
            amplitude_model = SplineMixtureAmplitudes.from_file(splinedata, spline_smoothing,
                                                                spline_type)
        sources = []
        for row in np.atleast_1d(catalog):
            g = Galaxy(filters=filternames, amplitude_model=amplitude_model)
            g.from_catalog_row(row, extra_parameters=extra_parameters)
            sources.append(g)
        return cls(sources)

    def __len__(self):
        return len(self.sources)

    @property
    def active_sources(self):
        return [s for s in self.sources if not s.fixed]

    @property
    def n_active(self):
        return len(self.active_sources)

    @property
    def parameter_names(self):
        return np.concatenate([s.parameter_names for s in self.sources])

    def set_all_source_params(self, Theta):
        start = 0
        for source in self.active_sources:
            end = start + source.n_params
            source.set_params(Theta[start:end])
            start = end

    def get_all_source_params(self, active=True):
        srcs = self.active_sources if active else self.sources
        return np.concatenate([s.get_param_vector() for s in srcs])

    def get_proposal(self, active=True):
        srcs = self.active_sources if active else [s for s in self.sources if s.fixed]
        return np.concatenate([s.proposal().copy() for s in srcs])

    def subscene(self, indices):
        """A Scene holding copies of the listed sources."""
        import copy
        sub = Scene([copy.deepcopy(self.sources[i]) for i in indices])
        sub.sky_zero = self.sky_zero
        return sub
