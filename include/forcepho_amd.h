/* forcepho_amd.h -- C ABI of the MI355X chi2 + gradient library (libfpx.so).
 *
 * Everything here is plain C: pointers, fixed-width integers and sizes.  The
 * Python host code binds it with ctypes (forcepho_amd/_lib.py); any other
 * language can bind it the same way.
 *
 * Device pointers are HIP device addresses (hipMalloc / torch tensors on a
 * ROCm device).  Host pointers are marked "host".  Every entry point returns
 * 0 on success and a negative FPX_E* code on failure; fpx_last_error() then
 * describes the failure (thread-local, valid until the next call on that
 * thread).
 */
#ifndef FORCEPHO_AMD_H
#define FORCEPHO_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "fpx_kernel_limits.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FPX_ABI_VERSION 3

/* ------------------------------------------------------------------------ */
/* Wire structs.  Byte layouts equal the reference's (checked by           */
/* static_assert in the library and against numpy dtypes in tests).        */
/* ------------------------------------------------------------------------ */

/* One PSF Gaussian, already tagged with the Sersic radius bin it convolves.
 * Replaces PSFSourceGaussian (forcepho/src/patch.cc:103-119); field order is
 * storage.PSF_COLS + sersic_bin (forcepho/patches/storage.py:27-28). 28 bytes. */
typedef struct FpxPSFSourceGaussian {
    float amp;
    float xcen, ycen;
    float Cxx, Cyy, Cxy;
    int32_t sersic_radius_bin;
} FpxPSFSourceGaussian;

/* Pixel data and per-exposure metadata of one patch, all arrays device
 * resident.  Replaces Patch (forcepho/src/patch.cc:23-100; numpy dtype
 * forcepho/patches/patch.py:44-73). 152 bytes.
 *   data, ierr, xpix, ypix, residual : float[npix], exposures concatenated in
 *       band order, each exposure padded to a multiple of FPX_WAVE pixels;
 *       residual may be NULL (then no residual image is written)
 *   exposure_start, exposure_N : int32[n_exp]
 *   band_start, band_N          : int16[n_bands]  (indices into exposures)
 *   rad2                        : float[n_radii]   (arcsec^2)
 *   D, CW                       : float[n_exp][n_sources][2][2]
 *   crpix, crval                : float[n_exp][n_sources][2]
 *   G                           : float[n_exp]
 *   n_psf_per_source            : int32[n_bands]
 *   psfgauss                    : FpxPSFSourceGaussian[...]
 *   psfgauss_start              : int32[n_exp]
 * As in the reference, every source of exposure e uses the PSF block that
 * starts at psfgauss[psfgauss_start[e]]. */
typedef struct FpxPatch {
    float *data;
    float *ierr;
    float *xpix;
    float *ypix;
    float *residual;
    int32_t *exposure_start;
    int32_t *exposure_N;
    int16_t *band_start;
    int16_t *band_N;
    int32_t n_sources;
    int32_t n_radii;
    float *rad2;
    float *D;
    float *crpix;
    float *crval;
    float *CW;
    float *G;
    int32_t *n_psf_per_source;
    FpxPSFSourceGaussian *psfgauss;
    int32_t *psfgauss_start;
} FpxPatch;

/* On-sky parameters of one Sersic source.  Replaces Source
 * (forcepho/src/proposal.cu:13-27; numpy dtype forcepho/proposal.py:229-240).
 * 264 bytes.  ra/dec are degrees relative to the patch reference
 * coordinates, pa radians, rh arcsec. */
typedef struct FpxSource {
    float ra, dec;
    float q, pa;
    float nsersic, rh;
    float fluxes[MAXBANDS];
    float mixture_amplitudes[MAXRADII];
    float damplitude_drh[MAXRADII];
    float damplitude_dnsersic[MAXRADII];
} FpxSource;

/* dchi2/dparam of one band: [MAXSOURCES][NPARAMS] with parameter order
 * (flux, ra, dec, q, pa, sersic, rh).  Replaces Response
 * (forcepho/src/proposal.cu:33-39). 840 bytes. */
typedef struct FpxResponse {
    float dchi2_dparam[MAXSOURCES * NPARAMS];
} FpxResponse;

/* One unit of kernel work: a run of at most pixels-per-tile consecutive
 * pixels of one exposure.  Built on the host by fpx_plan_tiles.  16 bytes.
 *   parts, part : a tile shared by `parts` workgroups (2 <= parts <=
 *     FPX_MAX_TILE_PARTS) appears `parts` times in a row in the table, with
 *     part = 0, 1, ..., parts - 1 and the same exposure, band and pixels.
 *     Piece `part` takes the part-th of `parts` equal runs of the tile's
 *     surviving ImageGaussians (in their (source, component) order) in both
 *     passes; the pieces exchange their partial models through the slot's
 *     split buffer, piece 0 adds the chi2 and writes the residual.  For the
 *     few tiles that carry several times the mean work (Gaussians crowded
 *     near a few pixels), so that no single workgroup holds the launch back.
 *     parts 0 or 1: the whole tile (what fpx_plan_tiles writes).  The pieces
 *     wait for each other, so they must be able to run at the same time: a
 *     persistent launch, whose queue hands them out together. */
typedef struct FpxTile {
    int32_t exposure;
    int16_t band;
    uint8_t part;
    uint8_t parts;
    int32_t pix_start;   /* absolute index into the patch pixel arrays */
    int32_t pix_count;
} FpxTile;

#define FPX_MAX_TILE_PARTS 16

/* One evaluation in a launch: a patch, a proposal for it and where the
 * results go.  An array of these lives in device memory; a launch may hold
 * many (many patches, many HMC chains on one patch, ...).
 *   accum  : double[fpx_accum_doubles(n_bands)], zero before the first launch;
 *            the kernel leaves it zero again after each launch
 *   status : NULL, or a device uint32 word, zero before the launch.  A tile
 *            that breaks a limit (FPX_STATUS_* bits below) is left out and
 *            ORs its bit into the word; the output kernel then writes NaN
 *            into pchi2 of every band of the slot.  The word keeps its bits
 *            until the caller clears it (fpx_evaluate_sync clears it and
 *            returns FPX_E_ARG).  The reference has no such checks
 *            (exceeding its limits is undefined, kernel_limits.h).
 *   pixel_weight : NULL for chi2 mode.  Non-NULL selects vector-Jacobian mode:
 *            no model pass, pchi2 is written as 0 and pdchi2_dp receives
 *            sum_pix pixel_weight[pix] * d model[pix] / d param.
 *   flags  : FPX_SLOT_MODEL_ONLY (chi2 mode): the model pass alone; the
 *            residual and pchi2 are written, pdchi2_dp is zero.  Other bits
 *            reserved, 0.  No counterpart in the reference, whose kernel
 *            always runs both passes; it serves forward passes that need
 *            the residual image and not the gradient.
 *   split  : NULL when no tile of the table is split (parts <= 1), else
 *            device memory of fpx_split_bytes(n_tiles, cfg) bytes, zero
 *            before the first launch that uses it and left zero by every
 *            launch (after one that sets FPX_STATUS_SPLIT_WAIT, zero it
 *            again).  Slots that may run at the same time need their own.
 * 88 bytes. */
typedef struct FpxSlot {
    const FpxPatch *patch;
    const FpxSource *proposal;
    const FpxTile *tiles;
    const float *pixel_weight;
    double *pchi2;        /* device double[n_bands] */
    float *pdchi2_dp;     /* device FpxResponse[n_bands] */
    double *accum;
    uint32_t *status;
    int32_t n_bands;
    int32_t n_tiles;
    int32_t flags;
    int32_t reserved;
    float *split;
} FpxSlot;

#define FPX_SLOT_MODEL_ONLY 1
/* FpxSlot.flags bit (chi2 mode, one slot per launch, no output kernel):
 * every tile stores its own sums, without atomics, at
 * accum[tile * (1 + MAXSOURCES * NPARAMS) + j] -- j = 0 its chi2, j = 1 +
 * NPARAMS * s + k its gradient of source s, parameter k, for j < 1 + NPARAMS
 * * n_sources -- and the caller adds the tiles of each band (accum then
 * holds n_tiles rows; it may be mapped host memory).  A tile left out, or
 * one that finds the status word set, stores NaN as its chi2.  For small
 * patches, where the output kernel's launch would cost more than the
 * evaluation; no counterpart in the reference. */
#define FPX_SLOT_TILE_PARTIALS 2

/* Bits of FpxSlot.status. */
#define FPX_STATUS_LIMIT 1u       /* a tile broke a size limit: n_sources, n_psf_per_source of
                                     its band, n_radii, its band index, its pixel count (tile
                                     table planned for a larger tile), or max_gauss below
                                     n_sources x n_psf_per_source */
#define FPX_STATUS_RADIUS_BIN 2u  /* a PSF entry's sersic_radius_bin is outside [0, n_radii) */
#define FPX_STATUS_SPLIT_WAIT 4u  /* a piece of a split tile waited FPX_SPLIT_WAIT_MS for the
                                     others and went on without their partial models (the
                                     pieces were not running together) */
#define FPX_SPLIT_WAIT_MS 50

/* Bits of the `persistent` argument of the launches that take one (1 and 0
 * keep their meaning).
 *   FPX_LAUNCH_PERSISTENT : a resident grid that pulls tiles from the queue
 *   FPX_LAUNCH_SPLIT      : the tile tables may hold split tiles (FpxTile
 *                           parts >= 2); needs FPX_LAUNCH_PERSISTENT.  Without
 *                           it a split tile is left out (FPX_STATUS_LIMIT). */
#define FPX_LAUNCH_PERSISTENT 1
#define FPX_LAUNCH_SPLIT 2

/* A mixture-amplitude model that fpx_build_sources evaluates on the device:
 * the mixture amplitudes of a source and their derivatives with respect to
 * sersic and rh.  Array pointers are device pointers.
 *   kind FPX_AMP_LOGNORMAL: a log-normal weight over ln radius,
 *     sig = 0.6 + 0.25 n, mu = ln rh - 0.15 (n - 1),
 *     A_i = exp(-z_i^2 / 2) / sum_j exp(-z_j^2 / 2), z_i = (lnr_i - mu) / sig;
 *     lnr [n_prop][MAXRADII] the ln radii of each proposal, lnr_count [n_prop]
 *     how many are used (the rest of A is zero).
 *   kind FPX_AMP_SPLINE: forcepho's per-radius bicubic (any degree <= 5)
 *     tensor-product B-splines of (sersic, rh) (forcepho/sources.py:623-681,
 *     scipy's tck): radius i has knots tx[i][0..tx_len[i]) (row stride
 *     tx_stride), ty likewise, coefficients coef[i][a * cy + b] (row stride
 *     cx * cy); arguments are clamped to the knot range as FITPACK does.
 * 88 bytes. */
typedef struct FpxAmplitudeModel {
    int32_t kind;
    int32_t n_radii;          /* spline: radii of the model (<= MAXRADII) */
    const double *lnr;
    const int32_t *lnr_count;
    int32_t kx, ky;
    int32_t tx_stride, ty_stride;
    int32_t cx, cy;
    const double *tx, *ty;
    const int32_t *tx_len, *ty_len;
    const double *coef;
} FpxAmplitudeModel;

#define FPX_AMP_LOGNORMAL 0
#define FPX_AMP_SPLINE 1
#define FPX_MAX_SPLINE_DEGREE 5

/* Error codes. */
#define FPX_OK 0
#define FPX_E_ARG -1      /* invalid argument or limit exceeded */
#define FPX_E_HIP -2      /* HIP runtime error */
#define FPX_E_NODEV -3    /* no usable gfx950 device */

/* Struct ids for fpx_struct_size. */
#define FPX_STRUCT_PATCH 0
#define FPX_STRUCT_SOURCE 1
#define FPX_STRUCT_RESPONSE 2
#define FPX_STRUCT_PSFGAUSS 3
#define FPX_STRUCT_TILE 4
#define FPX_STRUCT_SLOT 5
#define FPX_STRUCT_AMPMODEL 6

/* ------------------------------------------------------------------------ */
/* Host-only helpers (no GPU needed).                                       */
/* ------------------------------------------------------------------------ */

int fpx_abi_version(void);
const char *fpx_last_error(void);

/* sizeof() of the wire structs, for layout checks against the host language.
 * Returns 0 for an unknown id. */
size_t fpx_struct_size(int which);

/* Doubles of accumulator space one FpxSlot needs for n_bands bands. */
size_t fpx_accum_doubles(int32_t n_bands);

/* Bytes of the split buffer (FpxSlot.split) of a slot whose tile table holds
 * n_tiles entries for kernel configuration cfg: a counter per entry and the
 * partial model of every piece.  0 for an invalid cfg or n_tiles <= 0. */
size_t fpx_split_bytes(int32_t n_tiles, int32_t cfg);

/* Split every exposure of a packed patch into tiles of at most pix_per_tile
 * pixels (pix_per_tile a multiple of FPX_WAVE).  Inputs are the host copies of
 * the patch index arrays.  Writes at most max_tiles tiles to tiles_out (host)
 * and returns the number of tiles (>= 0), or FPX_E_ARG.  Passing
 * tiles_out == NULL returns the count only. */
int fpx_plan_tiles(const int32_t *exposure_start, const int32_t *exposure_N,
                   const int16_t *band_start, const int16_t *band_N,
                   int32_t n_bands, int32_t pix_per_tile,
                   FpxTile *tiles_out, int32_t max_tiles);

/* Kernel configurations (tile shape).  cfg in [0, fpx_n_configs()). */
int fpx_n_configs(void);
int fpx_config_info(int32_t cfg, int32_t *waves_per_group,
                    int32_t *pixels_per_lane, int32_t *pixels_per_tile);

/* Bytes of LDS one workgroup of configuration cfg uses on gfx950 when the
 * largest exposure of a launch holds max_gauss ImageGaussians (n_sources x
 * n_psf_per_source); the Gaussians that reach a tile are held in chunks of
 * what fits a budget of 10 KiB per wavefront.  Returns FPX_E_ARG if cfg is
 * invalid. */
int fpx_lds_bytes(int32_t cfg, int32_t max_gauss, size_t *bytes);

/* ------------------------------------------------------------------------ */
/* Device entry points.  stream is a hipStream_t (NULL = default stream).   */
/* All launches are asynchronous with respect to the host.                  */
/* ------------------------------------------------------------------------ */

/* Select and check the device (must be gfx950).  Optional: every device
 * entry point checks the current device on first use. */
int fpx_init(int32_t device);

/* Compute units, max LDS bytes per workgroup and peak engine clock (kHz) of
 * the current device. */
int fpx_device_info(int32_t *n_cu, int32_t *lds_per_group, int32_t *clock_khz);

/* chi2 and dchi2/dparam for one proposal on one patch (two kernels on
 * `stream`: the tiles, then the per-slot output pass).
 * Replaces the PyCUDA launch of EvaluateProposal<<<n_bands, 1024, 48000>>>
 * (forcepho/proposal.py:133-136, forcepho/src/compute_gaussians_kernel.cu:213-312)
 * and, per band, the C++ EvaluateProposal(band, patch, proposal)
 * (forcepho/src/compute_gaussians_kernel.cc:227-315).
 *   slot      : device pointer to one FpxSlot (its tiles built with the
 *               pixels_per_tile of cfg)
 *   n_tiles   : the slot's n_tiles (host copy, sizes the grid)
 *   max_gauss : n_sources x max over bands of n_psf_per_source (host copy)
 * Results: pchi2[b] (double, sum over the band's pixels of
 * (r*ierr - d*ierr)*(r*ierr + d*ierr)), pdchi2_dp[b] (float Response) and,
 * when patch->residual != NULL, residual[pix] = data - model.  With no
 * queue counter to find the last workgroup, this launch is two kernels (the
 * tile kernel, then the output kernel); the entry points that take a queue
 * write a one-slot launch's outputs from its last workgroup. */
int fpx_evaluate_proposal(const FpxSlot *slot, int32_t n_tiles,
                          int32_t max_gauss, int32_t cfg, void *stream);

/* One evaluation with host buffers, synchronous: proposal_bytes of the
 * proposal are copied from proposal_host to proposal_dev (the slot's
 * proposal buffer), the evaluation runs (with a queue: one kernel whose last
 * workgroup writes the outputs, persistent or not; without: the tile kernel
 * and the output kernel), and out_bytes from out_dev (the slot's
 * pchi2 / pdchi2_dp region) are copied to out_host before the call returns.
 * status_dev is the slot's status word or NULL; when it lies inside the
 * copied region its value is read from the copy, else it is copied back on
 * its own.  A nonzero status is cleared and the call returns FPX_E_ARG.
 * The PyCUDA call EvaluateProposal(patch, cuda.In(proposal),
 * cuda.Out(chi2), cuda.Out(derivs)) of forcepho/proposal.py:133-136 as one
 * native call.  Host buffers may be pinned (hipHostMalloc) or pageable. */
int fpx_evaluate_sync(const FpxSlot *slot, int32_t n_tiles, int32_t max_gauss,
                      int32_t cfg, int32_t persistent, uint32_t *queue,
                      void *proposal_dev, const void *proposal_host,
                      size_t proposal_bytes, const void *out_dev, void *out_host,
                      size_t out_bytes, uint32_t *status_dev, void *stream);

/* ------------------------------------------------------------------------ */
/* Bound patches: the reference's one-call launch, with the tile table,     */
/* accumulator, queue and status word owned by the library.                 */
/* ------------------------------------------------------------------------ */

/* A patch bound for evaluation (opaque): its tile table, FP64 accumulator,
 * work queue, status word and staging buffers on one device. */
typedef struct FpxEvaluator FpxEvaluator;

/* Bind a device-resident patch for evaluation on the current device.
 * patch_dev is the device address of an FpxPatch whose arrays are device
 * resident (what forcepho's send_patchstruct_to_device uploads,
 * forcepho/patches/device_patch.py:310-341) and n_bands the number of bands
 * (the reference's grid size, forcepho/proposal.py:133-136).  The call reads
 * the struct and its index arrays back (synchronously), checks every limit
 * the kernel relies on (n_sources in [1, MAXSOURCES], n_bands in
 * [1, MAXBANDS], n_radii in [1, MAXRADII], n_psf_per_source in
 * [1, FPX_MAX_PSF_PER_SOURCE], every radius bin the kernel reads in
 * [0, n_radii), exposure starts and sizes non-negative), plans the tile
 * table for kernel configuration cfg (-1: the library default) and
 * allocates the evaluator's device buffers.  Evaluations read the struct at
 * patch_dev each time, so swapping pixel buffers (swap_on_device,
 * forcepho/patches/device_patch.py:343-364) needs no new bind; bind again
 * when the packed metadata changes (another scene, other exposures).
 *   persistent : 1 = a resident grid pulling tiles from a queue,
 *                0 = one workgroup per tile
 * No counterpart in the reference, whose launch needs no set-up beyond the
 * uploaded struct; the tile table is what this kernel adds. */
int fpx_bind(const FpxPatch *patch_dev, int32_t n_bands, int32_t cfg, int32_t persistent,
             FpxEvaluator **out);

/* fpx_bind with the caller's tile table (host FpxTile[n_tiles], built for
 * kernel configuration cfg) instead of a planned one: a balanced order, or
 * split tiles (FpxTile parts >= 2, which need persistent = 1; the evaluator
 * then owns the split buffer).  The table is checked against the patch's
 * exposures (each entry inside one exposure of its band, at most the
 * configuration's pixels per tile, every split tile's pieces in a row) and
 * copied; that it covers each pixel once is the caller's to ensure.  No
 * counterpart in the reference. */
int fpx_bind_tiles(const FpxPatch *patch_dev, int32_t n_bands, int32_t cfg, int32_t persistent,
                   const FpxTile *tiles_host, int32_t n_tiles, FpxEvaluator **out);

/* chi2 and dchi2/dparam of one proposal on a bound patch, asynchronous on
 * stream.  Replaces the kernel EvaluateProposal(patch, proposal, pchi2,
 * pdchi2_dp) (forcepho/src/compute_gaussians_kernel.cu:212-214, launched
 * <<<n_bands, 1024, 48000>>> by forcepho/proposal.py:133-136).
 *   proposal_dev  : device FpxSource[n_sources]
 *   pchi2_dev     : device double[n_bands]
 *   pdchi2_dp_dev : device FpxResponse[n_bands]
 * The residual image is written when the patch struct's residual is not
 * NULL.  No allocation and no synchronisation with the host; calls on one
 * evaluator share its buffers and must be ordered on one stream.  A limit
 * broken after the bind (the struct at patch_dev changed) leaves NaN in
 * pchi2 and sets the evaluator's status (fpx_evaluator_status). */
int fpx_evaluate(FpxEvaluator *ev, const FpxSource *proposal_dev, double *pchi2_dev,
                 FpxResponse *pdchi2_dp_dev, void *stream);

/* The vector-Jacobian product of the model image on a bound patch, the
 * autograd backward: pdchi2_dp_dev[b] receives sum_pix pixel_weight[pix] *
 * d model[pix] / d param over band b's pixels (pixel_weight_dev: device
 * float[npix] over the packed pixels), pchi2_dev[b] is written 0; the
 * gradient pass alone, the Gaussians recomputed from the proposal.
 * Asynchronous, as fpx_evaluate.  No counterpart in the reference, whose
 * kernel always weights by the residual. */
int fpx_evaluate_vjp(FpxEvaluator *ev, const FpxSource *proposal_dev, const float *pixel_weight_dev,
                     double *pchi2_dev, FpxResponse *pdchi2_dp_dev, void *stream);

/* The model pass alone on a bound patch: the residual image (written when
 * the patch struct's residual is not NULL) and pchi2_dev; pdchi2_dp_dev is
 * written 0.  Asynchronous, as fpx_evaluate.  For forward passes that need
 * the residual and not the gradient; no counterpart in the reference. */
int fpx_evaluate_model(FpxEvaluator *ev, const FpxSource *proposal_dev, double *pchi2_dev,
                       FpxResponse *pdchi2_dp_dev, void *stream);

/* The same evaluation with host buffers, synchronous: n_sources records of
 * proposal_host in, chi2_host[n_bands] and resp_host[n_bands] out (resp_host
 * may be NULL), the status checked and cleared (FPX_E_ARG if a tile was left
 * out).  forcepho's GPUProposer.evaluate_proposal kernel call with its
 * cuda.In / cuda.Out arguments (forcepho/proposal.py:133-136) as one native
 * call; n_sources must be the bound patch's.  Returns once the results are
 * in the host buffers, which the launch signals through mapped memory; the
 * kernel may then still be exiting, so the stream is not idle (later work
 * on it is ordered after the kernel as usual). */
int fpx_evaluate_host(FpxEvaluator *ev, const FpxSource *proposal_host, int32_t n_sources,
                      double *chi2_host, FpxResponse *resp_host, void *stream);

/* Synchronise stream, then read and clear the evaluator's status word:
 * FPX_OK if every tile of the evaluations since the last check was
 * evaluated, else FPX_E_ARG (fpx_last_error names the broken limit). */
int fpx_evaluator_status(FpxEvaluator *ev, void *stream);

/* Tile count, kernel configuration and max_gauss (n_sources x the largest
 * n_psf_per_source) of a bound patch; any output may be NULL. */
int fpx_evaluator_info(const FpxEvaluator *ev, int32_t *n_tiles, int32_t *cfg,
                       int32_t *max_gauss);

/* Free an evaluator (after its last evaluation has completed). */
int fpx_release(FpxEvaluator *ev);

/* Bytes of the static device array fpx_evaluate_static works in: a status
 * word, the queue counter and the FP64 accumulators of MAXBANDS bands. */
#define FPX_STATIC_SCRATCH_BYTES (256 + MAXBANDS * (1 + MAXSOURCES * NPARAMS) * 8)

/* fpx_evaluate with no allocation at all.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The caller
 * passes the tile table (fpx_plan_tiles for configuration cfg, copied to
 * device memory it owns) and the outputs; the accumulator, queue and status
 * word live in a static device array of FPX_STATIC_SCRATCH_BYTES that every
 * GPU the library runs on holds until the process ends, cleared by a fill
 * kernel at the start of every call.
 *   n_bands   : at most MAXBANDS (larger is refused with FPX_E_ARG)
 *   max_gauss : n_sources x the largest n_psf_per_source
 *   cfg       : -1 for the library default
 * Asynchronous on stream, like fpx_evaluate; a tile that breaks a limit is
 * left out and pchi2 is written as NaN.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads)
 * race and must not be issued. */
int fpx_evaluate_static(const FpxPatch *patch_dev, const FpxTile *tiles_dev, int32_t n_tiles,
                        int32_t n_bands, int32_t max_gauss, const FpxSource *proposal_dev,
                        double *pchi2_dev, FpxResponse *pdchi2_dp_dev, int32_t cfg,
                        int32_t persistent, void *stream);

/* Many evaluations in one launch (uneven patches, many chains).  With one
 * slot and a queue the tile kernel's last workgroup writes the outputs (one
 * kernel); otherwise the output kernel follows the tile kernel.
 *   slots        : device FpxSlot[n_slots]
 *   item_offsets : device int32[n_slots + 1], prefix sums of the slots'
 *                  n_tiles (item_offsets[0] = 0); may be NULL if n_slots == 1
 *   n_items      : item_offsets[n_slots] (host copy)
 *   max_gauss    : max over slots of n_sources x max n_psf_per_source
 *   persistent   : 0 = one workgroup per tile; 1 (FPX_LAUNCH_PERSISTENT) =
 *                  a resident grid that pulls tiles from the queue counter;
 *                  | FPX_LAUNCH_SPLIT when a tile table holds split tiles
 *   queue        : device uint32[2], zero before the first launch that uses
 *                  it and left zero after each: the work counter of a
 *                  persistent launch and the done counter that finds the
 *                  last workgroup (persistent, or one slot with outputs);
 *                  may be NULL for a launch that is not persistent (one
 *                  slot then takes the output kernel too).  Launches that
 *                  may run at the same time must not share one. */
int fpx_evaluate_batch(const FpxSlot *slots, int32_t n_slots,
                       const int32_t *item_offsets, int32_t n_items,
                       int32_t max_gauss, int32_t cfg, int32_t persistent,
                       uint32_t *queue, void *stream);

/* The two halves of fpx_evaluate_batch, for callers that time or overlap
 * them (no counterpart in the reference, whose single kernel does both):
 * fpx_evaluate_tiles launches only the tile kernel, which adds every tile's
 * chi2 and gradient sums into the slots' FP64 accumulators; fpx_finalize
 * writes pchi2 / Response of every slot from its accumulator and zeroes it.
 * Launches of fpx_evaluate_tiles before one fpx_finalize accumulate. */
int fpx_evaluate_tiles(const FpxSlot *slots, int32_t n_slots,
                       const int32_t *item_offsets, int32_t n_items,
                       int32_t max_gauss, int32_t cfg, int32_t persistent,
                       uint32_t *queue, void *stream);
int fpx_finalize(const FpxSlot *slots, int32_t n_slots, void *stream);

/* Measurement helper with no counterpart in the reference: counts, per
 * slot, the (pixel, ImageGaussian) pairs inside the cutoff (dx^T F dx <=
 * MAX_EXP_ARG, the test of common_kernel.cc:73,108) for the slots' current
 * proposals, testing every Gaussian at every pixel.  Slots, item_offsets and
 * n_items as for fpx_evaluate_batch; adds each slot's count to
 * pairs_in[slot] (device, n_slots values) and, if item_pairs is not NULL,
 * each tile's count to item_pairs[item] (device, n_items values); the caller
 * zeroes both. */
int fpx_count_pairs(const FpxSlot *slots, int32_t n_slots,
                    const int32_t *item_offsets, int32_t n_items,
                    int32_t max_gauss, unsigned long long *pairs_in,
                    uint32_t *item_pairs, void *stream);

/* Host-side planning helper with no counterpart in the reference: the
 * number of workgroups of configuration cfg that are resident on the whole
 * current device at once for a launch with max_gauss (the persistent grid
 * size).  The dispatcher hands out the workgroups of a launch round-robin
 * over the compute units, so a non-persistent launch of at most that many
 * tiles normally runs tile i on compute unit i mod n_cu (fpx_device_info). */
int fpx_resident_groups(int32_t cfg, int32_t max_gauss, int32_t *groups);

/* Source records of n_prop proposals from their parameter vectors, on the
 * device: Galaxy.proposal (forcepho/sources.py:607-621, the amplitudes from
 * SplineMixtureAmplitudes, :657-681) for many proposals in one launch, so
 * that an HMC step or an autograd pass builds its records without the host.
 *   theta   : device double [n_prop][n_src_max][n_bands_max + 6], per source
 *             the band fluxes then ra, dec, q, pa, sersic, rh
 *   lower, upper : NULL, or device double arrays of theta's shape: theta
 *             then holds the unconstrained z of forcepho's BoundedTransform
 *             (forcepho/model.py:596-637) and the parameters are
 *             lower + (upper - lower) / (1 + exp(-z))
 *   n_sources, n_bands : device int32 [n_prop], the sources and bands of
 *             each proposal (fluxes of later bands are written as 0, records
 *             of later sources as all zero); NULL = n_src_max / n_bands_max
 *   model   : host pointer to the amplitude model (its arrays on the device)
 *   records : device FpxSource [n_prop][MAXSOURCES]; the first n_src_max
 *             records of each proposal are written
 * The values are computed in double and rounded to float. */
int fpx_build_sources(const double *theta, const double *lower, const double *upper,
                      int32_t n_prop, int32_t n_src_max, int32_t n_bands_max,
                      const int32_t *n_sources, const int32_t *n_bands,
                      const FpxAmplitudeModel *model, FpxSource *records,
                      void *stream);

/* dchi2/dtheta in the parameter order of fpx_build_sources, stacked from
 * per-band responses as forcepho's Posterior.stack_grad does
 * (forcepho/model.py:370-388): the flux of band b from band b's response,
 * ra, dec, q, pa, sersic, rh summed over the bands b < n_bands_max; and the
 * chi2 summed over those bands.
 *   resp     : device FpxResponse [n_prop][MAXBANDS]
 *   chi2     : device double [n_prop][MAXBANDS]
 *   grad     : device double [n_prop][n_src_max][n_bands_max + 6] or NULL
 *   chi2_sum : device double [n_prop] or NULL */
int fpx_stack_gradients(const FpxResponse *resp, const double *chi2, int32_t n_prop,
                        int32_t n_src_max, int32_t n_bands_max, double *grad,
                        double *chi2_sum, void *stream);

/* Potential, gradient and leapfrog update of n_prop HMC chains in the
 * unconstrained z of forcepho's BoundedTransform (forcepho/model.py:596-637),
 * from the per-band responses and chi2 of the chains' last evaluation (no
 * counterpart in the reference, whose sampler does this on the host for one
 * chain).  Per chain k, with theta = lower + sigmoid(z) (upper - lower) and
 * J = dtheta/dz = (upper - lower) sigmoid(z) (1 - sigmoid(z)):
 *   U[k] = chi2_k / 2 - sum mask ln J   (chi2 summed over bands b < n_bands_max)
 *   g    = mask (dchi2/dtheta J / 2 - (1 - 2 sigmoid(z)))
 * (the negative log posterior of FastPosterior.evaluate, model.py:302-340,
 * with a flat prior; dchi2/dtheta stacked as fpx_stack_gradients does); then,
 * if p is not NULL, p -= kick eps[k] g, and with drift z += eps[k] p mask.
 *   lower, upper, mask, z, p, g : device double [n_prop][n_src_max][n_bands_max + 6]
 *   eps, U   : device double [n_prop]
 *   resp     : device FpxResponse [n_prop][MAXBANDS]; chi2: device double [n_prop][MAXBANDS] */
int fpx_leapfrog_update(const double *lower, const double *upper, const double *mask,
                        const double *eps, const FpxResponse *resp, const double *chi2,
                        int32_t n_prop, int32_t n_src_max, int32_t n_bands_max,
                        double kick, int32_t drift, double *z, double *p, double *U,
                        double *g, void *stream);

/* Metropolis test of one HMC transition of n_prop chains of n_per
 * parameters each (no counterpart in the reference, whose NUTS sampler,
 * littlemcmc, picks among the points of one chain's trajectory on the
 * host): H = U + |p|^2 / 2 at the start (U0, p0) and the end
 * (U1, p1) of the trajectory; accept[k] = ln u[k] < H0 - H1 (a NaN energy
 * rejects); a rejected chain's z is set back to z0.
 *   U0, U1, u : device double [n_prop]; p0, p1, z0, z : device double
 *   [n_prop][n_per]; accept : device int32 [n_prop] */
int fpx_hmc_metropolis(const double *U0, const double *U1, const double *p0,
                       const double *p1, const double *u, const double *z0,
                       int32_t n_prop, int32_t n_per, double *z, int32_t *accept,
                       void *stream);

/* Throughput probe of the VALU pipes used by the hot loop: kind 0 = FP32
 * FMA and kind 1 = v_exp_f32, 8 independent chains of `iters` steps per lane
 * (dependency latency shows); kind 2 = FMA with 32 chains and kind 3 =
 * v_exp_f32 with 16 chains (the issue rate, given 8 wavefronts per SIMD:
 * n_threads = 256 x 8 x 4 x compute units or more).  Writes one float per
 * thread to out (device, n_threads floats): lane 0 of each wavefront the
 * shader-clock cycles of its loop, lane 1 the 100 MHz reference ticks of the
 * same interval.  Returns the operations the launch performs in *ops; the
 * caller may also time the launch. */
int fpx_probe_valu(int32_t kind, int32_t iters, float *out, int32_t n_threads,
                   double *ops, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FORCEPHO_AMD_H */
