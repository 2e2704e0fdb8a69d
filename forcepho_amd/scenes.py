"""Seeded synthetic scenes for the benchmark and test workloads.

Workloads (all at 0.03"/pixel, 9 mixture radii, circular PSF Gaussians of
width 0.5-3 px with amplitudes summing to one, replicated for every radius
bin):
  single   1 band x 1 exposure, 64x64 px, 1 source
  multi    8 bands x 2 exposures, 128x128 px, 15 sources, 4 PSF Gaussians
  uneven   one patch of the batched mix: U{1..20} sources, U{1..8} bands,
           U{1..4} exposures per band, square stamps of side U{48..256},
           U{3..6} PSF Gaussians
Source draws: flux ~ LogNormal(ln 10, 0.5) per band, q ~ U[0.3, 1],
pa ~ U[-pi/2, pi/2], sersic ~ U[1, 5], rh ~ U[0.05, 0.3] arcsec, position
within +-2 px of the stamp centre.  Data = model + N(0, sigma) with sigma set
for S/N ~ 50 of the first source inside its half-light radius (the noise
recipe of forcepho's demo_utils.compute_noise_level), ierr = 1/sigma.

Pixel values are filled by a ``render(patch, proposal) -> model`` callable
so that the same draws serve the device path (``render_on_device``) and the
tests (which render with their own checker).
"""

import numpy as np

from .patches import Stamp, LinearWCS, PSF_DTYPE
from .sources import Galaxy, Scene, DEFAULT_RADII

__all__ = ["PIXEL_SCALE", "SKY_CENTER", "draw_patch_spec", "spec_single", "spec_multi",
           "spec_uneven", "build_patch", "render_on_device", "PatchSpec"]

PIXEL_SCALE = 0.03          # arcsec / pixel
SKY_CENTER = (53.0, -27.0)  # degrees
SNR = 50.0


class PatchSpec:
    """Everything drawn for one patch: stamps (data still zero), the true
    scene, the noise realisation per exposure and the sky reference."""

    def __init__(self, stamps, scene, noise, sky_zero):
        self.stamps, self.scene, self.noise, self.sky_zero = stamps, scene, noise, sky_zero


def _psf_block(rng, n_comp, n_radii):
    sig = rng.uniform(0.5, 3.0, size=n_comp)
    amp = rng.uniform(0.2, 1.0, size=n_comp)
    amp = amp / amp.sum()
    block = np.zeros(n_radii * n_comp, dtype=PSF_DTYPE)
    for r in range(n_radii):
        sl = slice(r * n_comp, (r + 1) * n_comp)
        block["amp"][sl] = amp
        block["Cxx"][sl] = sig ** 2
        block["Cyy"][sl] = sig ** 2
        block["sersic_bin"][sl] = r
    return block


def draw_patch_spec(rng, n_bands, n_exp_per_band, sides, n_sources, n_psf_comp,
                    radii=DEFAULT_RADII, pos_jitter=2.0, dither=0.5):
    """Draw one patch.  ``sides``: stamp side per band (square stamps);
    ``n_exp_per_band``, ``n_psf_comp``: per band."""
    n_radii = len(radii)
    ra0, dec0 = SKY_CENTER
    s_deg = PIXEL_SCALE / 3600.0
    cosd = np.cos(np.deg2rad(dec0))
    filters = [f"b{b}" for b in range(n_bands)]
    sources = []
    for i in range(n_sources):
        g = Galaxy(filters=filters, radii=radii)
        g.flux = rng.lognormal(np.log(10.0), 0.5, size=n_bands)
        g.q = rng.uniform(0.3, 1.0)
        g.pa = rng.uniform(-np.pi / 2, np.pi / 2)
        g.sersic = rng.uniform(1.0, 5.0)
        g.rhalf = rng.uniform(0.05, 0.3)
        dx, dy = rng.uniform(-pos_jitter, pos_jitter, size=2)
        g.ra = ra0 + dx * s_deg / cosd
        g.dec = dec0 + dy * s_deg
        sources.append(g)
    scene = Scene(sources)
    stamps, noise = [], []
    for b in range(n_bands):
        nside = int(sides[b])
        # noise per exposure (demo_utils.compute_noise_level with the first source)
        npix_h = np.pi * (sources[0].rhalf / PIXEL_SCALE) ** 2
        sigma = (sources[0].flux[b] / 2.0) / SNR / np.sqrt(npix_h)
        for k in range(int(n_exp_per_band[b])):
            crpix = ((nside - 1) / 2.0 + rng.uniform(-dither, dither),
                     (nside - 1) / 2.0 + rng.uniform(-dither, dither))
            wcs = LinearWCS((ra0, dec0), crpix, np.eye(2) * s_deg)
            psf = _psf_block(rng, int(n_psf_comp[b]), n_radii)
            st = Stamp(b, np.zeros((nside, nside), np.float32),
                       np.full((nside, nside), 1.0 / sigma, np.float32), wcs, psf)
            stamps.append(st)
            noise.append(rng.normal(0.0, sigma, size=(nside, nside)).astype(np.float32))
    return PatchSpec(stamps, scene, noise, np.array([ra0, dec0]))


def spec_single(seed):
    """1 band x 1 exposure x 64x64, 1 source, 3 PSF Gaussians."""
    rng = np.random.default_rng(seed)
    return draw_patch_spec(rng, 1, [1], [64], 1, [3])


def spec_multi(seed, side=128, n_bands=8, n_exp=2, n_sources=15, n_psf_comp=4):
    """8 bands x 2 exposures x 128x128, 15 sources, 4 PSF Gaussians."""
    rng = np.random.default_rng(seed)
    return draw_patch_spec(rng, n_bands, [n_exp] * n_bands, [side] * n_bands, n_sources,
                           [n_psf_comp] * n_bands)


def spec_uneven(rng):
    """One patch of the uneven batched mix."""
    n_src = int(rng.integers(1, 21))
    n_bands = int(rng.integers(1, 9))
    n_exp = rng.integers(1, 5, size=n_bands)
    side = int(rng.integers(48, 257))
    n_psf = rng.integers(3, 7, size=n_bands)
    return draw_patch_spec(rng, n_bands, n_exp, [side] * n_bands, n_src, n_psf)


def build_patch(spec, patch_cls, render, return_residual=False):
    """Pack a spec into ``patch_cls``, fill data = model(truth) + noise.

    ``render(patch, proposal)`` returns the model as a flat float array over
    the packed pixels.  Returns (patch, scene, theta_true)."""
    patch = patch_cls(return_residual=return_residual)
    patch.build_patch(spec.stamps, spec.scene, sky_zero=spec.sky_zero)
    proposal = spec.scene.get_proposal()
    model = np.asarray(render(patch, proposal), dtype=np.float64)
    data = np.zeros(patch.npix, dtype=np.float64)
    for e in range(patch.n_exp):
        start = patch.exposure_start[e]
        n = int(patch.original_sizes[e])
        flat_noise = spec.noise[e].reshape(-1)[patch.pixel_index[e]]
        data[start:start + n] = model[start:start + n] + flat_noise
    patch.data = data.astype(patch.pix_dtype)
    return patch, spec.scene, spec.scene.get_all_source_params().copy()


def render_on_device(patch, proposal):
    """Model image of ``proposal`` computed by the HIP kernel: evaluates with
    zero data and returns minus the residual."""
    from .proposal import HIPProposer
    saved = patch.data
    patch.data = np.zeros_like(saved)
    patch.return_residual = True
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    _, _, resid = prop.evaluate_proposal(proposal, patch=patch, unpack=False)
    patch.data = saved
    patch.return_residual = False
    patch.free()
    return -resid.astype(np.float64)
