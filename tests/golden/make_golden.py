"""Write the golden fixtures tests/golden/*.npz.

Each fixture holds a packed patch, a proposal and the outputs of the
reference's own CPU kernel (forcepho/src/compute_gaussians_kernel.cc,
compiled from its sources by oracle/Makefile into oracle/_ref) evaluated on
them: chi2 per band, dchi2/dparam [n_bands, n_sources, 7] (float32) and the
residual image.  Run in a tree where oracle/_ref is built:

    make -C oracle all && python tests/golden/make_golden.py
"""

import copy
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle  # noqa: E402
from forcepho_amd import scenes  # noqa: E402
from forcepho_amd.patches import StampPatch  # noqa: E402
from forcepho_amd.sources import DEFAULT_RADII  # noqa: E402

ARRAYS = ["xpix", "ypix", "data", "ierr", "exposure_start", "exposure_N", "band_start",
          "band_N", "rad2", "D", "crpix", "crval", "CW", "G", "n_psf_per_source",
          "psfgauss_start"]


def _spec(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "single":
        return scenes.spec_single(seed)
    if kind == "multi_small":
        return scenes.spec_multi(seed, side=32)
    if kind == "max_sources":
        radii = np.concatenate([[0.003], DEFAULT_RADII])  # MAXRADII = 10 radii
        return scenes.draw_patch_spec(rng, 1, [1], [48], 30, [3], radii=radii)
    if kind == "cutoff":
        # sources far from the stamp centre: most pairs exceed MAX_EXP_ARG
        return scenes.draw_patch_spec(rng, 1, [1], [64], 3, [2], pos_jitter=60.0)
    if kind == "padding":
        spec = scenes.draw_patch_spec(rng, 2, [1, 2], [37, 41], 2, [3, 2])
        return spec
    if kind == "uneven":
        return scenes.draw_patch_spec(rng, 3, [3, 1, 2], [40, 72, 56], 4, [2, 5, 3])
    if kind == "radius_zero":
        radii = np.concatenate([[0.0], DEFAULT_RADII[:8]])
        return scenes.draw_patch_spec(rng, 1, [2], [48], 2, [3], radii=radii)
    raise ValueError(kind)


CASES = [("single", 1), ("single", 7), ("multi_small", 2), ("max_sources", 3),
         ("cutoff", 4), ("padding", 5), ("uneven", 6), ("radius_zero", 8),
         ("psf_rule", 9), ("zero_flux", 10)]


def make_case(kind, seed):
    base = {"psf_rule": "uneven", "zero_flux": "uneven"}.get(kind, kind)
    spec = _spec(base, seed)
    patch, scene, theta = scenes.build_patch(spec, StampPatch, oracle.render_model)
    proposal = scene.get_proposal()
    # move away from the truth so gradients are not ~0
    proposal["q"] *= np.float32(0.93)
    proposal["pa"] += np.float32(0.07)
    proposal["fluxes"] *= np.float32(1.05)
    if kind == "psf_rule":
        # give every source but the first a different PSF block: the kernel
        # must still use the first source's block (compute_gaussians_kernel.cc:100)
        pg = patch.psfgauss
        for e in range(patch.n_exp):
            n = int(patch.n_psf_per_source[patch.bands[e]])
            s0 = int(patch.psfgauss_start[e])
            for s in range(1, int(patch.n_sources)):
                blk = pg[s0 + s * n:s0 + (s + 1) * n]
                blk["Cxx"] *= 3.0
                blk["amp"] *= 0.5
                blk["xcen"] += 0.4
    if kind == "zero_flux":
        proposal["fluxes"][1, 0] = 0.0   # source 1 has no flux in band 0
    chi2, derivs, resid = oracle.evaluate_reference(patch, proposal, return_residual=True)
    out = {name: np.asarray(getattr(patch, name)) for name in ARRAYS}
    out["psfgauss"] = np.asarray(patch.psfgauss).view(np.uint8).reshape(-1)
    out["n_sources"] = np.int32(patch.n_sources)
    out["n_radii"] = np.int32(patch.n_radii)
    out["n_bands"] = np.int32(patch.n_bands)
    out["original_sizes"] = np.asarray(patch.original_sizes)
    out["proposal"] = np.ascontiguousarray(proposal).view(np.uint8).reshape(-1)
    out["chi2"] = chi2
    out["dchi2"] = derivs
    out["residual"] = resid
    return out


def main():
    mod = oracle.reference_module()
    if mod is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle all` where the "
                         "reference sources are present")
    for kind, seed in CASES:
        out = make_case(kind, seed)
        fn = os.path.join(HERE, f"{kind}_s{seed}.npz")
        np.savez_compressed(fn, **out)
        print(fn, os.path.getsize(fn), "bytes", "chi2", out["chi2"][:2])


if __name__ == "__main__":
    main()
