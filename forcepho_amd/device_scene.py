"""Source records built on the device from parameter tensors.

``DeviceRecords`` maps parameter tensors theta [n, S, K] -- per proposal and
source the NB band fluxes then ra, dec, q, pa, sersic, rh (the order of
``Scene.get_all_source_params``, forcepho/sources.py:46-97) -- to the FpxSource
records (include/forcepho_amd.h) of n proposals in device memory with one
launch of ``fpx_build_sources``: ``Galaxy.proposal``
(forcepho/sources.py:607-621) for many proposals at once and without a host
round trip, so that a whole leapfrog trajectory or an autograd forward pass
stays on the device.  The kernel evaluates the mixture amplitudes and their
(sersic, rh) derivatives in double for two amplitude models:

* ``LogNormalMixtureAmplitudes`` (sources.py), the analytic model of the
  synthetic scenes;
* ``SplineMixtureAmplitudes``, forcepho's own model (sources.py:623-681):
  per-radius tensor-product B-splines (scipy's RectBivariateSpline or
  SmoothBivariateSpline of a fitted table), whose knots and coefficients
  (``spline.tck``) ``DeviceSplineAmplitudes`` lays out in device memory.

``stack_grad_device`` is forcepho's ``Posterior.stack_grad``
(forcepho/model.py:370-388) on Response tensors, one launch of
``fpx_stack_gradients``: band fluxes from their own band, the shape
parameters summed over bands, and the chi2 summed over bands.
"""

import ctypes

import numpy as np

from ._lib import lib, check, AmplitudeModel, AMP_LOGNORMAL, AMP_SPLINE, MAX_SPLINE_DEGREE
from .kernel_limits import MAXBANDS, MAXRADII, MAXSOURCES
from .sources import LogNormalMixtureAmplitudes, SplineMixtureAmplitudes

__all__ = ["DeviceRecords", "DeviceSplineAmplitudes", "stack_grad_device",
           "device_amplitude_models_ok"]

# float offsets of the FpxSource fields (include/forcepho_amd.h)
F_SHAPE, F_FLUX = 0, 6
F_AMP = 6 + MAXBANDS
F_DRH = F_AMP + MAXRADII
F_DN = F_DRH + MAXRADII
REC_FLOATS = F_DN + MAXRADII   # 66 floats = 264 bytes


def _torch():
    import torch
    return torch


def _stream(device):
    return _torch().cuda.current_stream(device).cuda_stream


def _model_kind(scene):
    """'lognormal' (one common set of radii), 'spline' (one shared spline
    model) or None for a scene's sources."""
    models = [s.amplitude_model for s in scene.sources]
    if not models:
        return None, None
    if all(isinstance(m, LogNormalMixtureAmplitudes) for m in models):
        r0 = models[0].radii
        if all(len(m.radii) == len(r0) and np.allclose(m.radii, r0) for m in models):
            return "lognormal", models[0]
        return None, None
    if all(m is models[0] for m in models) and isinstance(models[0], SplineMixtureAmplitudes):
        return "spline", models[0]
    return None, None


def device_amplitude_models_ok(scene):
    """True when DeviceRecords can evaluate the scene's amplitudes: every
    source uses LogNormalMixtureAmplitudes with one common set of radii, or
    every source shares one SplineMixtureAmplitudes."""
    return _model_kind(scene)[0] is not None


class DeviceSplineAmplitudes:
    """A SplineMixtureAmplitudes' tables in device memory, in the layout of
    FpxAmplitudeModel (kind FPX_AMP_SPLINE).

    The spline of radius i is s_i(x, y) = sum_ab c_i[a, b] Bx_a(x) By_b(y)
    with B-spline bases of degrees (kx, ky) on knots (tx_i, ty_i): tx [R, Tx]
    (each row padded with its last knot), tx_len [R], ty likewise, and
    c [R, Cx * Cy] with c_i[a, b] at a * Cy + b.
    """

    def __init__(self, model, device):
        torch = _torch()
        splines = model.splines
        self.R = len(splines)
        if not 1 <= self.R <= MAXRADII:
            raise ValueError(f"spline model has {self.R} radii (1..{MAXRADII} supported)")
        degs = {tuple(int(d) for d in s.degrees) for s in splines}
        if len(degs) != 1:
            raise ValueError("all radius splines must share their degrees")
        self.kx, self.ky = degs.pop()
        if max(self.kx, self.ky) > MAX_SPLINE_DEGREE or min(self.kx, self.ky) < 0:
            raise ValueError(f"spline degrees {self.kx, self.ky} outside 0..{MAX_SPLINE_DEGREE}")
        txs = [np.asarray(s.tck[0], np.float64) for s in splines]
        tys = [np.asarray(s.tck[1], np.float64) for s in splines]
        if any(len(t) < 2 * self.kx + 2 for t in txs) or any(len(t) < 2 * self.ky + 2 for t in tys):
            raise ValueError("a radius spline has fewer knots than its degree needs")
        nx = [len(t) - self.kx - 1 for t in txs]
        ny = [len(t) - self.ky - 1 for t in tys]
        Tx, Ty = max(len(t) for t in txs), max(len(t) for t in tys)
        self.Cx, self.Cy = max(nx), max(ny)
        tx = np.empty((self.R, Tx))
        ty = np.empty((self.R, Ty))
        c = np.zeros((self.R, self.Cx, self.Cy))
        for i, s in enumerate(splines):
            tx[i, :len(txs[i])], tx[i, len(txs[i]):] = txs[i], txs[i][-1]
            ty[i, :len(tys[i])], ty[i, len(tys[i]):] = tys[i], tys[i][-1]
            c[i, :nx[i], :ny[i]] = np.asarray(s.tck[2], np.float64).reshape(nx[i], ny[i])
        f64 = dict(dtype=torch.float64, device=device)
        i32 = dict(dtype=torch.int32, device=device)
        self.tx, self.ty = torch.tensor(tx, **f64), torch.tensor(ty, **f64)
        self.c = torch.tensor(c.reshape(self.R, self.Cx * self.Cy), **f64)
        self.tx_len = torch.tensor([len(t) for t in txs], **i32)
        self.ty_len = torch.tensor([len(t) for t in tys], **i32)

    def struct(self):
        """The FpxAmplitudeModel of these tables (addresses of this object's
        tensors: keep it alive while the struct is used)."""
        return AmplitudeModel(kind=AMP_SPLINE, n_radii=self.R, lnr=None, lnr_count=None,
                              kx=self.kx, ky=self.ky, tx_stride=self.tx.shape[1],
                              ty_stride=self.ty.shape[1], cx=self.Cx, cy=self.Cy,
                              tx=self.tx.data_ptr(), ty=self.ty.data_ptr(),
                              tx_len=self.tx_len.data_ptr(), ty_len=self.ty_len.data_ptr(),
                              coef=self.c.data_ptr())


class DeviceRecords:
    """Source records of n proposals, written from parameter tensors.

    Parameters
    ----------
    scenes : list of n Scenes (their amplitude models give each proposal's
             radii: LogNormalMixtureAmplitudes, or one SplineMixtureAmplitudes
             shared by every source of every scene)
    n_bands : list of n band counts
    records : uint8 device tensor of at least n * MAXSOURCES * 264 bytes, the
              proposals' FpxSource arrays back to back (written in place)
    S, NB : sources and bands of the parameter tensors (>= every scene's)
    """

    def __init__(self, scenes, n_bands, records, S, NB):
        torch = _torch()
        n = len(scenes)
        dev = records.device
        self.n, self.S, self.NB, self.K = n, int(S), int(NB), int(NB) + 6
        if not 1 <= self.S <= MAXSOURCES or not 1 <= self.NB <= MAXBANDS:
            raise ValueError("S or NB outside the kernel limits")
        if records.numel() < n * MAXSOURCES * REC_FLOATS * 4:
            raise ValueError("records buffer smaller than n * MAXSOURCES FpxSource records")
        kinds = [_model_kind(scene) for scene in scenes]
        if any(k is None for k, _ in kinds):
            raise TypeError("device records evaluate LogNormalMixtureAmplitudes or one "
                            "shared SplineMixtureAmplitudes per scene")
        self.spline = None
        if all(k == "spline" for k, _ in kinds):
            model = kinds[0][1]
            if any(m is not model for _, m in kinds):
                raise TypeError("every scene must share one SplineMixtureAmplitudes")
            self.spline = DeviceSplineAmplitudes(model, dev)
            self.model = self.spline.struct()
        elif any(k == "spline" for k, _ in kinds):
            raise TypeError("scenes mix spline and log-normal amplitude models")
        else:
            lnr = np.full((n, MAXRADII), np.log(1e-4))
            count = np.zeros(n, dtype=np.int32)
            for k, scene in enumerate(scenes):
                r = len(scene.sources[0].amplitude_model.radii)
                if not 1 <= r <= MAXRADII:
                    raise ValueError(f"{r} radii (1..{MAXRADII} supported)")
                lnr[k, :r] = scene.sources[0].amplitude_model._lnr
                count[k] = r
            self.lnr = torch.tensor(lnr, dtype=torch.float64, device=dev)
            self.lnr_count = torch.tensor(count, device=dev)
            self.model = AmplitudeModel(kind=AMP_LOGNORMAL, n_radii=0,
                                        lnr=self.lnr.data_ptr(),
                                        lnr_count=self.lnr_count.data_ptr())
        n_src = [len(scene.sources) for scene in scenes]
        if max(n_src) > self.S or max(int(b) for b in n_bands) > self.NB:
            raise ValueError("a scene has more sources or bands than S / NB")
        self.n_sources = torch.tensor(n_src, dtype=torch.int32, device=dev)
        self.n_bands = torch.tensor([int(b) for b in n_bands], dtype=torch.int32, device=dev)
        self.records = records
        self.props = records[:n * MAXSOURCES * REC_FLOATS * 4].view(torch.float32).view(
            n, MAXSOURCES, REC_FLOATS)
        self.props.zero_()

    def write(self, theta, lower=None, upper=None):
        """Write the records of parameters theta [n, S, K] (float64,
        contiguous, on the records' device): one fpx_build_sources launch on
        the current stream.  With ``lower`` and ``upper`` (same shape), theta
        is the unconstrained z of a bounded transform and the parameters are
        lower + sigmoid(z) (upper - lower)."""
        torch = _torch()
        shape = (self.n, self.S, self.K)
        for name, t in (("theta", theta), ("lower", lower), ("upper", upper)):
            if t is None and name != "theta":
                continue
            if (tuple(t.shape) != shape or t.dtype != torch.float64
                    or t.device != self.records.device or not t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous float64 {list(shape)} tensor on "
                                 f"{self.records.device}")
        if (lower is None) != (upper is None):
            raise ValueError("lower and upper come together")
        check(lib().fpx_build_sources(theta.data_ptr(),
                                      None if lower is None else lower.data_ptr(),
                                      None if upper is None else upper.data_ptr(),
                                      self.n, self.S, self.NB,
                                      self.n_sources.data_ptr(), self.n_bands.data_ptr(),
                                      ctypes.byref(self.model), self.records.data_ptr(),
                                      _stream(self.records.device)), "fpx_build_sources")


def stack_grad_device(resp, S, NB, chi2=None):
    """Response [n, MAXBANDS, MAXSOURCES, 7] (float32 device, contiguous) ->
    dchi2/dtheta [n, S, NB + 6] float64: each band flux from its own band,
    ra, dec, q, pa, sersic, rh summed over the bands b < NB.  With chi2
    [n, MAXBANDS] (float64) also returns the chi2 summed over those bands
    [n].  One fpx_stack_gradients launch on the current stream."""
    torch = _torch()
    n = resp.shape[0]
    if (tuple(resp.shape[1:]) != (MAXBANDS, MAXSOURCES, 7) or resp.dtype != torch.float32
            or not resp.is_contiguous()):
        raise ValueError("resp must be a contiguous float32 [n, MAXBANDS, MAXSOURCES, 7] tensor")
    grad = torch.empty((n, int(S), int(NB) + 6), dtype=torch.float64, device=resp.device)
    total = None
    if chi2 is not None:
        if (tuple(chi2.shape) != (n, MAXBANDS) or chi2.dtype != torch.float64
                or not chi2.is_contiguous()):
            raise ValueError("chi2 must be a contiguous float64 [n, MAXBANDS] tensor")
        total = torch.empty(n, dtype=torch.float64, device=resp.device)
    check(lib().fpx_stack_gradients(resp.data_ptr(), None if chi2 is None else chi2.data_ptr(),
                                    n, int(S), int(NB), grad.data_ptr(),
                                    None if total is None else total.data_ptr(),
                                    _stream(resp.device)), "fpx_stack_gradients")
    return grad if chi2 is None else (grad, total)
