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
    """Amplitudes from per-radius RectBivariateSplines over an (n, rh) grid,
    as forcepho/sources.py:623-681 builds them."""

    def __init__(self, nsersic_grid, rh_grid, amplitudes, radii, smoothing=None):
        from scipy.interpolate import RectBivariateSpline
        A = np.asarray(amplitudes, dtype=np.float64)  # [n_n, n_rh, n_radii]
        self.radii = np.asarray(radii, dtype=np.float64)
        self.splines = [RectBivariateSpline(nsersic_grid, rh_grid, A[:, :, i], s=smoothing)
                        for i in range(A.shape[-1])]

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

    def __init__(self, sources=()):
        self.sources = list(sources)
        self.sky_zero = None  # set when coordinates are shifted to patch offsets
        for i, s in enumerate(self.sources):
            s.id = i

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
