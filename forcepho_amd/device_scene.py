"""Source records built on the device from parameter tensors.

``DeviceRecords`` maps parameter tensors theta [n, S, K] -- per proposal and
source the NB band fluxes then ra, dec, q, pa, sersic, rh (the order of
``Scene.get_all_source_params``, forcepho/sources.py:46-97) -- to the FpxSource
records (include/forcepho_amd.h) of n proposals in device memory, with the
mixture amplitudes and their (sersic, rh) derivatives of
``LogNormalMixtureAmplitudes`` evaluated by torch operations.  It is
``Galaxy.proposal`` (forcepho/sources.py:607-621) for many proposals at once
and without a host round trip, so that a whole leapfrog trajectory or an
autograd forward pass stays on the device.

``stack_grad_device`` is ``model.stack_grad`` (forcepho/model.py:214-236) on
Response tensors: band fluxes from their own band, the shape parameters
summed over bands.
"""

import numpy as np

from .kernel_limits import MAXBANDS, MAXRADII, MAXSOURCES
from .sources import LogNormalMixtureAmplitudes

__all__ = ["DeviceRecords", "stack_grad_device", "device_amplitude_models_ok"]

# float offsets of the FpxSource fields (include/forcepho_amd.h)
F_SHAPE, F_FLUX = 0, 6
F_AMP = 6 + MAXBANDS
F_DRH = F_AMP + MAXRADII
F_DN = F_DRH + MAXRADII
REC_FLOATS = F_DN + MAXRADII   # 66 floats = 264 bytes


def _torch():
    import torch
    return torch


def device_amplitude_models_ok(scene):
    """True when every source of ``scene`` uses LogNormalMixtureAmplitudes
    with one common set of radii (what DeviceRecords evaluates)."""
    models = [s.amplitude_model for s in scene.sources]
    if not models or not all(isinstance(m, LogNormalMixtureAmplitudes) for m in models):
        return False
    r0 = models[0].radii
    return all(len(m.radii) == len(r0) and np.allclose(m.radii, r0) for m in models)


class DeviceRecords:
    """Source records of n proposals, written from parameter tensors.

    Parameters
    ----------
    scenes : list of n Scenes (their amplitude models give each proposal's
             radii; all LogNormalMixtureAmplitudes)
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
        lnr = np.full((n, MAXRADII), np.log(1e-4))
        rmask = np.zeros((n, MAXRADII), dtype=bool)
        fmask = np.zeros((n, self.S, self.NB))
        for k, (scene, nb) in enumerate(zip(scenes, n_bands)):
            if not device_amplitude_models_ok(scene):
                raise TypeError("device records implement LogNormalMixtureAmplitudes only")
            model = scene.sources[0].amplitude_model
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
        """A, dA/dn, dA/drh [n, S, MAXRADII] of the log-normal model
        (sources.LogNormalMixtureAmplitudes) for parameters [n, S]."""
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
        props[:, :S, F_AMP:F_AMP + MAXRADII] = A.float()
        props[:, :S, F_DRH:F_DRH + MAXRADII] = dA_drh.float()
        props[:, :S, F_DN:F_DN + MAXRADII] = dA_dn.float()


def stack_grad_device(resp, S, NB):
    """Response [n, MAXBANDS, MAXSOURCES, 7] (device) -> dchi2/dtheta
    [n, S, NB + 6] float64: each band flux from its own band, ra, dec, q, pa,
    sersic, rh summed over bands."""
    torch = _torch()
    d = resp[:, :, :S, :].double()
    return torch.cat([d[:, :NB, :, 0].permute(0, 2, 1), d[:, :, :, 1:].sum(dim=1)], dim=2)
