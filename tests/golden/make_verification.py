"""Write the fixtures of the reference's end-to-end verification case.

forcepho's verification (tests/verification/verify_reference.py) compares the
fast model of one Sersic galaxy with an FP64 image rendered by the reference's
slow Python model (tests/verification/make_reference.py).  Its inputs are data
files of the reference tree; this script copies or extracts them into
tests/golden/verification/ so the tests run where the reference tree is absent:

* ``reference_f200w_sersic2.2_rhalf0.10.fits`` -- copy of
  tests/verification/data/reference-2021Nov30_f200w_sersic=2.2_rhalf=0.10.fits
  (image, its seven derivative images, the one-row source catalog);
* ``psf_f200w_ng4m0.npz`` -- the 4-Gaussian PSF table (BINTABLE extension) of
  tests/verification/data/psf_jwst_oct21_ng4m0_f200w.fits, the PSF file the
  image header names (PSFFILE);
* ``sersic_mog_smooth0.0150.npz`` -- datasets nsersic, rh, amplitudes, radii of
  tests/data/sersic_mog_model.smooth=0.0150.h5, the mixture table;
* ``test_image_data.fits`` and ``test_image_psf.h5`` -- copies of
  tests/data/test_image/sersic1.0_rhalf0.150_snr050_noise1_{data.fits,psf.h5},
  the inputs of tests/cswap_test.py.

Run from the repository root where /root/reference exists:
    python tests/golden/make_verification.py
"""

import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

REF = "/root/reference/tests"
OUT = os.path.join(HERE, "verification")


def main():
    from forcepho_amd.formats import fits
    from forcepho_amd.formats.hdf5 import File
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(REF, "verification/data/"
                                 "reference-2021Nov30_f200w_sersic=2.2_rhalf=0.10.fits"),
                    os.path.join(OUT, "reference_f200w_sersic2.2_rhalf0.10.fits"))
    psf = fits.getdata(os.path.join(REF, "verification/data/psf_jwst_oct21_ng4m0_f200w.fits"), -1)
    np.savez(os.path.join(OUT, "psf_f200w_ng4m0.npz"),
             **{name: np.asarray(psf[name]) for name in psf.dtype.names})
    with File(os.path.join(REF, "data/sersic_mog_model.smooth=0.0150.h5")) as f:
        np.savez(os.path.join(OUT, "sersic_mog_smooth0.0150.npz"),
                 nsersic=f["nsersic"][:], rh=f["rh"][:], amplitudes=f["amplitudes"][:],
                 radii=f["radii"][:])
    shutil.copyfile(os.path.join(REF, "data/test_image/sersic1.0_rhalf0.150_snr050_noise1_data.fits"),
                    os.path.join(OUT, "test_image_data.fits"))
    shutil.copyfile(os.path.join(REF, "data/test_image/sersic1.0_rhalf0.150_snr050_noise1_psf.h5"),
                    os.path.join(OUT, "test_image_psf.h5"))
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
