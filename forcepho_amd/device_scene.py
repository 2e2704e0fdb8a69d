"""Source records built on the device from parameter tensors.

``DeviceRecords`` maps parameter tensors theta [n, S, K] -- per proposal and
source the NB band fluxes then ra, dec, q, pa, sersic, rh (the order of
``Scene.get_all_source_params``, forcepho/sources.py:46-97) -- to the FpxSource
records (include/forcepho_amd.h) of n proposals in device memory, with the
mixture amplitudes and their (sersic, rh) derivatives evaluated by torch
operations on the device.  It is ``Galaxy.proposal``
(forcepho/sources.py:607-621) for many proposals at once and without a host
round trip, so that a whole leapfrog trajectory or an autograd forward pass
stays on the device.

Two amplitude models are evaluated on the device:

* ``LogNormalMixtureAmplitudes`` (sources.py), the analytic model of the
  synthetic scenes;
* ``SplineMixtureAmplitudes``, forcepho's own model (sources.py:623-681):
  per-radius bicubic tensor-product B-splines (scipy's RectBivariateSpline or
  SmoothBivariateSpline of a fitted table).  ``DeviceSplineAmplitudes``
  evaluates the splines' knots and coefficients (``spline.tck``) with de
  Boor's recurrences for the B-spline basis and its first derivative
  (C. de Boor, A Practical Guide to Splines, 1978, algorithms BSPLVB and
  BSPLVD), arguments clamped to the knot range as FITPACK's evaluators
  (fpbisp, parder) do.

``stack_grad_device`` is ``model.stack_grad`` (forcepho/model.py:214-236) on
Response tensors: band fluxes from their own band, the shape parameters
summed over bands.
"""

import numpy as np

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
    """Device evaluation of a SplineMixtureAmplitudes: A_i(n, rh) and its
    partial derivatives for every radius i.

    The spline of radius i is s_i(x, y) = sum_ab c_i[a, b] Bx_a(x) By_b(y)
    with B-spline bases of degrees (kx, ky) on knots (tx_i, ty_i).  At a point
    only kx + 1 (ky + 1) basis functions are nonzero: those of the knot
    interval holding the point.
    """

    def __init__(self, model, device):
        torch = _torch()
        splines = model.splines
        self.R = len(splines)
        degs = {tuple(s.degrees) for s in splines}
        if len(degs) != 1:
            raise ValueError("all radius splines must share their degrees")
        self.kx, self.ky = degs.pop()
        txs = [np.asarray(s.tck[0], np.float64) for s in splines]
        tys = [np.asarray(s.tck[1], np.float64) for s in splines]
        nx = [len(t) - self.kx - 1 for t in txs]
        ny = [len(t) - self.ky - 1 for t in tys]
        Tx, Ty = max(len(t) for t in txs), max(len(t) for t in tys)
        Cx, Cy = max(nx), max(ny)
        tx = np.empty((self.R, Tx))
        ty = np.empty((self.R, Ty))
        c = np.zeros((self.R, Cx, Cy))
        for i, s in enumerate(splines):
            tx[i, :len(txs[i])], tx[i, len(txs[i]):] = txs[i], txs[i][-1]
            ty[i, :len(tys[i])], ty[i, len(tys[i]):] = tys[i], tys[i][-1]
            c[i, :nx[i], :ny[i]] = np.asarray(s.tck[2], np.float64).reshape(nx[i], ny[i])
        f64 = dict(dtype=torch.float64, device=device)
        self.tx, self.ty = torch.tensor(tx, **f64), torch.tensor(ty, **f64)
        self.c = torch.tensor(c.reshape(self.R, Cx * Cy), **f64)
        self.Cy = Cy
        lx = torch.tensor([len(t) for t in txs], device=device)
        ly = torch.tensor([len(t) for t in tys], device=device)
        # valid intervals: k <= l <= len(t) - k - 2
        self.lx_hi = (lx - self.kx - 2)[:, None]
        self.ly_hi = (ly - self.ky - 2)[:, None]
        ar = torch.arange(self.R, device=device)
        self.x_lo, self.x_hi = self.tx[:, self.kx][:, None], self.tx[ar, lx - self.kx - 1][:, None]
        self.y_lo, self.y_hi = self.ty[:, self.ky][:, None], self.ty[ar, ly - self.ky - 1][:, None]

    @staticmethod
    def _basis(t, l_hi, k, lo, hi, x):
        """Nonzero basis values and first derivatives at x [R, M]:
        (l [R, M], B [k+1][R, M], dB [k+1][R, M]) with B[a] = B_{l-k+a}."""
        torch = _torch()
        x = torch.minimum(torch.maximum(x, lo), hi)
        l = torch.searchsorted(t, x.contiguous(), right=True) - 1
        l = torch.minimum(torch.clamp(l, min=k), l_hi)

        def knot(offset):
            return torch.gather(t, 1, l + offset)

        h = [torch.ones_like(x)]
        lower = None
        for j in range(1, k + 1):
            if j == k:
                lower = h                      # degree k - 1 basis, for the derivative
            hh = h
            h = [torch.zeros_like(x) for _ in range(j + 1)]
            for n in range(1, j + 1):
                xb, xa = knot(n), knot(n - j)
                span = xb - xa
                w = torch.where(span != 0, hh[n - 1] / torch.where(span != 0, span, 1.0), 0.0)
                h[n - 1] = h[n - 1] + w * (xb - x)
                h[n] = w * (x - xa)
        if k == 0:
            return l, h, [torch.zeros_like(x)]
        d = [torch.zeros_like(x) for _ in range(k + 1)]
        for n in range(1, k + 1):
            xb, xa = knot(n), knot(n - k)
            span = xb - xa
            w = torch.where(span != 0, k * lower[n - 1] / torch.where(span != 0, span, 1.0), 0.0)
            d[n - 1] = d[n - 1] - w
            d[n] = w
        return l, h, d

    def __call__(self, nsersic, rh):
        """A, dA/dn, dA/drh, each [..., R], for parameter tensors of any
        shape [...] (float64)."""
        torch = _torch()
        shape = nsersic.shape
        R, kx, ky = self.R, self.kx, self.ky
        xs = nsersic.reshape(1, -1).expand(R, -1)
        ys = rh.reshape(1, -1).expand(R, -1)
        lx, bx, dbx = self._basis(self.tx, self.lx_hi, kx, self.x_lo, self.x_hi, xs)
        ly, by, dby = self._basis(self.ty, self.ly_hi, ky, self.y_lo, self.y_hi, ys)
        A = torch.zeros_like(xs)
        dA_dx = torch.zeros_like(xs)
        dA_dy = torch.zeros_like(xs)
        for a in range(kx + 1):
            row = (lx - kx + a) * self.Cy
            sx, dx = torch.zeros_like(xs), torch.zeros_like(xs)
            for b in range(ky + 1):
                cab = torch.gather(self.c, 1, row + (ly - ky + b))
                sx = sx + cab * by[b]
                dx = dx + cab * dby[b]
            A = A + bx[a] * sx
            dA_dx = dA_dx + dbx[a] * sx
            dA_dy = dA_dy + bx[a] * dx
        out = [v.transpose(0, 1).reshape(*shape, R) for v in (A, dA_dx, dA_dy)]
        return out[0], out[1], out[2]


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
        elif any(k == "spline" for k, _ in kinds):
            raise TypeError("scenes mix spline and log-normal amplitude models")
        lnr = np.full((n, MAXRADII), np.log(1e-4))
        rmask = np.zeros((n, MAXRADII), dtype=bool)
        fmask = np.zeros((n, self.S, self.NB))
        for k, (scene, nb) in enumerate(zip(scenes, n_bands)):
            model = scene.sources[0].amplitude_model
            if self.spline is None:
                r = len(model.radii)
                lnr[k, :r] = model._lnr
                rmask[k, :r] = True
            fmask[k, :len(scene.sources), :int(nb)] = 1.0
        f64 = dict(dtype=torch.float64, device=dev)
        self.lnr = torch.tensor(lnr, **f64)
        self.rmask = torch.tensor(rmask, device=dev)
        self.flux_mask = torch.tensor(fmask, **f64)
        self.records = records
        self.props = records[:n * MAXSOURCES * REC_FLOATS * 4].view(torch.float32).view(
            n, MAXSOURCES, REC_FLOATS)
        self.props.zero_()

    def amplitudes(self, nsersic, rh):
        """A, dA/dn, dA/drh [n, S, R] for parameters [n, S] (R = MAXRADII
        for the log-normal model, the spline model's radius count otherwise)."""
        if self.spline is not None:
            return self.spline(nsersic, rh)
        torch = _torch()
        sig = 0.6 + 0.25 * nsersic
        mu = torch.log(rh) - 0.15 * (nsersic - 1.0)
        z = (self.lnr[:, None, :] - mu[..., None]) / sig[..., None]
        z2 = torch.where(self.rmask[:, None, :], z * z, torch.full_like(z, float("inf")))
        w = torch.exp(-0.5 * (z2 - z2.min(dim=-1, keepdim=True).values))
        A = w / w.sum(dim=-1, keepdim=True)
        zz = torch.where(self.rmask[:, None, :], z, torch.zeros_like(z))
        dl_drh = zz / (sig * rh)[..., None]
        dl_dn = zz / sig[..., None] * (-0.15) + zz * zz * 0.25 / sig[..., None]
        dA_dn = A * (dl_dn - (A * dl_dn).sum(-1, keepdim=True))
        dA_drh = A * (dl_drh - (A * dl_drh).sum(-1, keepdim=True))
        return A, dA_dn, dA_drh

    def write(self, theta):
        """Write the records of parameters theta [n, S, K] (float64 device)."""
        S, NB = self.S, self.NB
        props = self.props
        props[:, :S, F_FLUX:F_FLUX + NB] = (theta[:, :, :NB] * self.flux_mask).float()
        props[:, :S, F_SHAPE:F_SHAPE + 6] = theta[:, :, NB:].float()
        A, dA_dn, dA_drh = self.amplitudes(theta[:, :, NB + 4], theta[:, :, NB + 5])
        R = A.shape[-1]
        props[:, :S, F_AMP:F_AMP + R] = A.float()
        props[:, :S, F_DRH:F_DRH + R] = dA_drh.float()
        props[:, :S, F_DN:F_DN + R] = dA_dn.float()


def stack_grad_device(resp, S, NB):
    """Response [n, MAXBANDS, MAXSOURCES, 7] (device) -> dchi2/dtheta
    [n, S, NB + 6] float64: each band flux from its own band, ra, dec, q, pa,
    sersic, rh summed over bands."""
    torch = _torch()
    d = resp[:, :, :S, :].double()
    return torch.cat([d[:, :NB, :, 0].permute(0, 2, 1), d[:, :, :, 1:].sum(dim=1)], dim=2)
