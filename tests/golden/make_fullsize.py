"""Write tests/golden/fullsize/multi_band_fp64.npz: the float64 oracle's chi2,
gradient and tolerance magnitudes for the full-size multi-band benchmark patch
(8 bands x 2 exposures x 128x128 px, 15 sources, 9 radii x 4 PSF Gaussians;
scenes.spec_multi(0), data rendered by the C restatement, proposal with every
q scaled by 0.97).  About 90 s of CPU:

    make -C oracle all && python tests/golden/make_fullsize.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import fpx_testutil as U  # noqa: E402
import oracle  # noqa: E402
from forcepho_amd import scenes  # noqa: E402
from forcepho_amd.patches import SimplePatch  # noqa: E402

OUT = os.path.join(HERE, "fullsize", "multi_band_fp64.npz")


def build():
    """The patch and proposal the fixture describes."""
    spec = scenes.spec_multi(0)
    patch, scene, _ = scenes.build_patch(spec, SimplePatch, oracle.render_model)
    proposal = scene.get_proposal()
    proposal["q"] *= np.float32(0.97)
    return patch, proposal


def data_digest(patch):
    """Order-independent float64 digest of the packed data and weights."""
    d = patch.data.astype(np.float64)
    w = patch.ierr.astype(np.float64)
    return np.array([d.sum(), (d * d).sum(), w.sum(), (d * w).sum()])


def main():
    patch, proposal = build()
    c64, g64, mass, smass = U.fp64_reference(patch, proposal)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    np.savez_compressed(OUT, chi2=c64, dchi2=g64, mass=mass, smass=smass,
                        digest=data_digest(patch))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
