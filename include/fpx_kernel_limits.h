/* fpx_kernel_limits.h -- compile-time limits shared by the HIP kernels, the
 * C-ABI structs and the Python host code (forcepho_amd/kernel_limits.py parses
 * the #define lines below, the same way forcepho/kernel_limits.py:21-38 parses
 * forcepho/src/kernel_limits.h).
 *
 * The first block must equal the reference's values (forcepho/src/kernel_limits.h:5-26):
 * they size the Source and Response structs, whose byte layouts are part of
 * the drop-in boundary.
 */
#ifndef FPX_KERNEL_LIMITS_H
#define FPX_KERNEL_LIMITS_H

/* Maximum number of bands in one patch. */
#define MAXBANDS 30

/* Maximum number of active sources in one patch. */
#define MAXSOURCES 30

/* On-sky parameters per source that get a chi2 derivative:
 * flux, ra, dec, q, pa, sersic, rh. */
#define NPARAMS 7

/* A Gaussian is skipped at a pixel when its exponent argument
 * dx^T F dx exceeds this value (the model uses exp(-0.5 * arg)). */
#define MAX_EXP_ARG 36.0

/* Maximum number of Sersic mixture radii per source. */
#define MAXRADII 10

/* ---- limits of this implementation (no counterpart in the reference) ---- */

/* Wavefront width on CDNA4; exposures are padded to a multiple of it. */
#define FPX_WAVE 64

/* Maximum PSF Gaussians per source per exposure (radii x PSF components). */
#define FPX_MAX_PSF_PER_SOURCE 120

#endif
