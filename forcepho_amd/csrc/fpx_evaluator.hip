// fpx_evaluator.hip -- bound patches (FpxEvaluator): the reference's one-call
// EvaluateProposal(patch, proposal, pchi2, pdchi2_dp) launch
// (forcepho/src/compute_gaussians_kernel.cu:212-214) on a patch whose tile
// table, FP64 accumulator, work queue and status word the library owns.
//
// fpx_bind reads the packed patch's struct and index arrays back once,
// checks every limit the kernel relies on and plans the tile table; an
// evaluation then passes its slot to the kernels by value, so it touches no
// host memory, allocates nothing and can be recorded into a graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "forcepho_amd.h"
#include "fpx_internal.h"

// One device allocation per evaluator, each part 256-byte aligned:
//   io     status word (uint32, 8 bytes with padding), then two unused
//          output areas of the io layout below
//   queue  uint32[2] work-queue counter of persistent launches
//   prop   FpxSource[MAXSOURCES] (fpx_evaluate_host's proposal)
//   accum  double[fpx_accum_doubles(n_bands)]
//   tiles  FpxTile[n_tiles]
//   split  fpx_split_bytes(n_tiles, cfg), when the tile table holds split
//          tiles (fpx_bind_tiles)
// and one mapped pinned host block: fpx_evaluate_host's outputs in the io
// layout (pchi2 double[MAXBANDS] at kIoChi2, FpxResponse[MAXBANDS] at
// kIoResp), written there by the output kernel itself (no copy back), then
// the proposal's host staging, then -- for patches of at most
// kPartialTiles tiles -- one row of sums per tile (FPX_SLOT_TILE_PARTIALS).
//
// fpx_evaluate_host on such a small patch is one kernel: the tiles read the
// proposal in place from the staging and store their sums in the host rows,
// which the host adds up; the output kernel's launch and the copies would
// cost more than the evaluation itself.  fpx_evaluate_host does not wait for
// the stream: the launch's last workgroup stores the call's sequence number
// into the first word of the host block once every result is visible to the
// host, and the host spins on that word (a stream synchronisation takes
// some 4 us longer to return).
struct FpxEvaluator {
  int device = 0;
  const FpxPatch *patch = nullptr;
  int32_t n_bands = 0, n_sources = 0, max_gauss = 0, cfg = 0, n_tiles = 0;
  int32_t launch_flags = 0;  // FPX_LAUNCH_PERSISTENT, FPX_LAUNCH_SPLIT
  unsigned char *block = nullptr;
  unsigned char *host = nullptr;      // mapped pinned block (host address)
  unsigned char *host_dev = nullptr;  // the same block's device address
  uint32_t *status = nullptr;
  double *chi2 = nullptr;
  FpxResponse *resp = nullptr;
  uint32_t *queue = nullptr;
  FpxSource *prop = nullptr;
  double *accum = nullptr;
  FpxTile *tiles = nullptr;
  float *split = nullptr;          // split buffer, or NULL
  size_t o_stage = 0, o_part = 0;  // offsets in the host block
  std::vector<int32_t> tile_band;  // per tile, when the partials path is used
  uint32_t seq = 0;                // sequence number of the last host-path call
};

namespace {

constexpr size_t kIoFlag = 0;  // host block: the completion flag
constexpr size_t kIoChi2 = 8;
constexpr int kPartialTiles = 64;
constexpr size_t kAccRow = 1 + MAXSOURCES * NPARAMS;  // doubles per tile row
constexpr size_t kIoResp = kIoChi2 + sizeof(double) * MAXBANDS;
constexpr size_t kIoBytes = kIoResp + sizeof(FpxResponse) * MAXBANDS;

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int arg_err(const std::string &msg) { return fpx::report_error(FPX_E_ARG, msg.c_str()); }

#define EV_HIP(call)                                              \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) return fpx::report_hip_error(_e, #call); \
  } while (0)

template <typename T>
int read_back(std::vector<T> &dst, const T *src, size_t n) {
  dst.assign(n, T());
  if (n == 0) return FPX_OK;
  if (!src) return arg_err("a patch array the kernel reads is NULL");
  EV_HIP(hipMemcpy(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  return FPX_OK;
}

int check_device(const FpxEvaluator *ev) {
  int dev = -1;
  EV_HIP(hipGetDevice(&dev));
  if (dev != ev->device) return arg_err("the evaluator was bound on another device");
  return FPX_OK;
}

void release(FpxEvaluator *ev) {
  if (!ev) return;
  // a host-path call returns once its results are visible, which can be
  // before its kernel has exited
  (void)hipDeviceSynchronize();
  if (ev->block) (void)hipFree(ev->block);
  if (ev->host) (void)hipHostFree(ev->host);
  delete ev;
}

FpxSlot slot_of(const FpxEvaluator *ev, const FpxSource *prop, double *pchi2, FpxResponse *resp) {
  FpxSlot s;
  std::memset(&s, 0, sizeof(s));
  s.patch = ev->patch;
  s.proposal = prop;
  s.tiles = ev->tiles;
  s.pixel_weight = nullptr;
  s.pchi2 = pchi2;
  s.pdchi2_dp = reinterpret_cast<float *>(resp);
  s.accum = ev->accum;
  s.status = ev->status;
  s.n_bands = ev->n_bands;
  s.n_tiles = ev->n_tiles;
  s.split = ev->split;
  return s;
}

// Check a caller's tile table against the patch's exposures (index arrays
// read back by bind): every entry inside one exposure of its band, at most
// ppt pixels, and every split tile's pieces 0..parts-1 in a row with the
// same pixels.  Sets *split when some tile is split.
int check_tiles(const FpxTile *t, int32_t n, int ppt, int n_bands, int n_exp,
                const std::vector<int16_t> &band_start, const std::vector<int16_t> &band_N,
                const std::vector<int32_t> &exp_start, const std::vector<int32_t> &exp_N,
                bool *split) {
  *split = false;
  for (int32_t k = 0; k < n; ++k) {
    const FpxTile &x = t[k];
    if (x.band < 0 || x.band >= n_bands) return arg_err("tile table: band out of range");
    if (x.exposure < band_start[x.band] || x.exposure >= band_start[x.band] + band_N[x.band] ||
        x.exposure >= n_exp)
      return arg_err("tile table: exposure not of the tile's band");
    if (x.pix_count < 0 || x.pix_count > ppt)
      return arg_err("tile table: pixel count outside [0, pixels per tile of cfg]");
    if (x.pix_start < exp_start[x.exposure] ||
        (long)x.pix_start + x.pix_count > (long)exp_start[x.exposure] + exp_N[x.exposure])
      return arg_err("tile table: pixels outside the tile's exposure");
  }
  // runs: a whole tile has part 0; a split tile's pieces 0..parts-1 follow
  // each other with the same exposure, band, pixels and piece count
  for (int32_t k = 0; k < n;) {
    const FpxTile &x = t[k];
    const int parts = x.parts > 1 ? x.parts : 1;
    if (parts > FPX_MAX_TILE_PARTS) return arg_err("tile table: more than FPX_MAX_TILE_PARTS pieces");
    if (k + parts > n) return arg_err("tile table: a split tile's pieces are cut off");
    for (int q = 0; q < parts; ++q) {
      const FpxTile &y = t[k + q];
      if (y.part != q || y.parts != x.parts || y.exposure != x.exposure || y.band != x.band ||
          y.pix_start != x.pix_start || y.pix_count != x.pix_count)
        return arg_err("tile table: pieces of a split tile must follow each other, part 0..parts-1");
    }
    *split = *split || parts > 1;
    k += parts;
  }
  return FPX_OK;
}

// tiles_in: the caller's tile table (host), or NULL to plan one.
int bind(const FpxPatch *patch_dev, int32_t n_bands, int32_t cfg, int32_t persistent,
         FpxEvaluator *ev, const FpxTile *tiles_in = nullptr, int32_t n_in = 0) {
  if (!patch_dev) return arg_err("NULL patch");
  if (n_bands < 1 || n_bands > MAXBANDS) return arg_err("n_bands out of range [1, MAXBANDS]");
  if (cfg < 0) cfg = fpx::kDefaultConfig;
  const int ppt = fpx::config_pixels(cfg);
  if (ppt <= 0) return arg_err("invalid kernel configuration");
  EV_HIP(hipGetDevice(&ev->device));
  FpxPatch P;
  EV_HIP(hipMemcpy(&P, patch_dev, sizeof(P), hipMemcpyDeviceToHost));
  if (P.n_sources < 1 || P.n_sources > MAXSOURCES)
    return arg_err("n_sources out of range [1, MAXSOURCES]");
  if (P.n_radii < 1 || P.n_radii > MAXRADII) return arg_err("n_radii out of range [1, MAXRADII]");
  if (!P.data || !P.ierr || !P.xpix || !P.ypix || !P.rad2 || !P.D || !P.crpix || !P.crval ||
      !P.CW || !P.G)
    return arg_err("a patch array the kernel reads is NULL");
  std::vector<int16_t> band_start, band_N;
  std::vector<int32_t> n_psf;
  int rc = read_back(band_start, P.band_start, n_bands);
  if (!rc) rc = read_back(band_N, P.band_N, n_bands);
  if (!rc) rc = read_back(n_psf, P.n_psf_per_source, n_bands);
  if (rc) return rc;
  int n_exp = 0, max_psf = 0;
  for (int b = 0; b < n_bands; ++b) {
    if (band_start[b] < 0 || band_N[b] < 0) return arg_err("negative band_start or band_N");
    if (n_psf[b] < 1 || n_psf[b] > FPX_MAX_PSF_PER_SOURCE)
      return arg_err("n_psf_per_source out of range [1, FPX_MAX_PSF_PER_SOURCE]");
    n_exp = std::max(n_exp, band_start[b] + band_N[b]);
    max_psf = std::max(max_psf, n_psf[b]);
  }
  std::vector<int32_t> exp_start, exp_N, psf_start;
  rc = read_back(exp_start, P.exposure_start, n_exp);
  if (!rc) rc = read_back(exp_N, P.exposure_N, n_exp);
  if (!rc) rc = read_back(psf_start, P.psfgauss_start, n_exp);
  if (rc) return rc;
  // the PSF entries the kernel reads: psfgauss[psfgauss_start[e] + p],
  // p < n_psf_per_source of e's band (every source uses that block)
  long psf_extent = 0;
  for (int b = 0; b < n_bands; ++b)
    for (int k = 0; k < band_N[b]; ++k) {
      const int e = band_start[b] + k;
      if (exp_start[e] < 0 || exp_N[e] < 0) return arg_err("negative exposure start or size");
      if (psf_start[e] < 0) return arg_err("negative psfgauss_start");
      psf_extent = std::max<long>(psf_extent, (long)psf_start[e] + n_psf[b]);
    }
  std::vector<FpxPSFSourceGaussian> psf;
  rc = read_back(psf, P.psfgauss, (size_t)psf_extent);
  if (rc) return rc;
  for (int b = 0; b < n_bands; ++b)
    for (int k = 0; k < band_N[b]; ++k) {
      const int e = band_start[b] + k;
      for (int p = 0; p < n_psf[b]; ++p) {
        const int bin = psf[psf_start[e] + p].sersic_radius_bin;
        if (bin < 0 || bin >= P.n_radii)
          return arg_err("a PSF entry's sersic_radius_bin is outside [0, n_radii)");
      }
    }
  // tile table: the caller's, checked, or planned here
  bool has_split = false;
  int n_tiles = 0;
  std::vector<FpxTile> tiles;
  if (tiles_in) {
    if (n_in < 0) return arg_err("negative tile count");
    rc = check_tiles(tiles_in, n_in, ppt, n_bands, n_exp, band_start, band_N, exp_start, exp_N,
                     &has_split);
    if (rc) return rc;
    if (has_split && !persistent)
      return arg_err("split tiles need a persistent evaluator (persistent = 1)");
    n_tiles = n_in;
    tiles.assign(tiles_in, tiles_in + n_in);
    if (tiles.empty()) tiles.resize(1);
  } else {
    n_tiles = fpx_plan_tiles(exp_start.data(), exp_N.data(), band_start.data(), band_N.data(),
                             n_bands, ppt, nullptr, 0);
    if (n_tiles < 0) return n_tiles;
    tiles.resize((size_t)std::max(n_tiles, 1));
    if (n_tiles > 0) {
      rc = fpx_plan_tiles(exp_start.data(), exp_N.data(), band_start.data(), band_N.data(),
                          n_bands, ppt, tiles.data(), n_tiles);
      if (rc < 0) return rc;
    }
  }
  ev->patch = patch_dev;
  ev->n_bands = n_bands;
  ev->n_sources = P.n_sources;
  ev->max_gauss = P.n_sources * max_psf;
  ev->cfg = cfg;
  ev->launch_flags = (persistent ? FPX_LAUNCH_PERSISTENT : 0) | (has_split ? FPX_LAUNCH_SPLIT : 0);
  ev->n_tiles = n_tiles;
  // device block
  const size_t o_queue = align256(kIoBytes);
  const size_t o_prop = o_queue + 256;
  const size_t o_accum = o_prop + align256(sizeof(FpxSource) * MAXSOURCES);
  const size_t o_tiles = o_accum + align256(sizeof(double) * fpx_accum_doubles(n_bands));
  const size_t o_split = o_tiles + align256(sizeof(FpxTile) * tiles.size());
  const size_t bytes = o_split + (has_split ? align256(fpx_split_bytes(n_tiles, cfg)) : 0);
  EV_HIP(hipMalloc((void **)&ev->block, bytes));
  EV_HIP(hipMemset(ev->block, 0, o_tiles));
  if (has_split) EV_HIP(hipMemset(ev->block + o_split, 0, bytes - o_split));
  EV_HIP(hipMemcpy(ev->block + o_tiles, tiles.data(), sizeof(FpxTile) * tiles.size(),
                   hipMemcpyHostToDevice));
  ev->status = reinterpret_cast<uint32_t *>(ev->block);
  ev->chi2 = reinterpret_cast<double *>(ev->block + kIoChi2);
  ev->resp = reinterpret_cast<FpxResponse *>(ev->block + kIoResp);
  ev->queue = reinterpret_cast<uint32_t *>(ev->block + o_queue);
  ev->prop = reinterpret_cast<FpxSource *>(ev->block + o_prop);
  ev->accum = reinterpret_cast<double *>(ev->block + o_accum);
  ev->tiles = reinterpret_cast<FpxTile *>(ev->block + o_tiles);
  ev->split = has_split ? reinterpret_cast<float *>(ev->block + o_split) : nullptr;
  // mapped pinned host block: the outputs of fpx_evaluate_host, the
  // proposal staging, and the tile rows of a small patch
  ev->o_stage = align256(kIoBytes);
  ev->o_part = ev->o_stage + align256(sizeof(FpxSource) * MAXSOURCES);
  size_t host_bytes = ev->o_part;
  if (n_tiles > 0 && n_tiles <= kPartialTiles) {
    host_bytes += sizeof(double) * kAccRow * n_tiles;
    ev->tile_band.resize(n_tiles);
    for (int t = 0; t < n_tiles; ++t) ev->tile_band[t] = tiles[t].band;
  }
  EV_HIP(hipHostMalloc((void **)&ev->host, host_bytes, hipHostMallocMapped));
  EV_HIP(hipHostGetDevicePointer((void **)&ev->host_dev, ev->host, 0));
  // pinned blocks are recycled: the flag must not start out holding a
  // sequence number
  std::memset(ev->host, 0, host_bytes);
  return FPX_OK;
}

// The next sequence number of a host-path call (never 0); the flag is
// cleared to 0 before the launch that is to store it.
uint32_t next_seq(FpxEvaluator *ev) {
  if (++ev->seq == 0) ev->seq = 1;
  __atomic_store_n(reinterpret_cast<uint32_t *>(ev->host + kIoFlag), 0u, __ATOMIC_RELEASE);
  return ev->seq;
}

// Wait until the launch that carries `seq` has stored it in the host block's
// flag: spin on the mapped word for up to 50 ms, then fall back to a stream
// synchronisation (which also reports a failed launch).
int wait_flag(FpxEvaluator *ev, uint32_t seq, hipStream_t s) {
  const uint32_t *flag = reinterpret_cast<const uint32_t *>(ev->host + kIoFlag);
  const auto t0 = std::chrono::steady_clock::now();
  while (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq) {
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      EV_HIP(hipStreamSynchronize(s));
      if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq)
        return fpx::report_error(FPX_E_HIP, "the evaluation did not signal completion");
      break;
    }
  }
  return FPX_OK;
}

int read_status(FpxEvaluator *ev, hipStream_t s, uint32_t status) {
  if (status == 0) return FPX_OK;
  EV_HIP(hipMemsetAsync(ev->status, 0, sizeof(uint32_t), s));
  // pieces that timed out may have left their split counters behind
  if ((status & FPX_STATUS_SPLIT_WAIT) && ev->split)
    EV_HIP(hipMemsetAsync(ev->split, 0, fpx_split_bytes(ev->n_tiles, ev->cfg), s));
  EV_HIP(hipStreamSynchronize(s));
  return arg_err(fpx::status_text(status));
}

// The partials path of fpx_evaluate_host (n_tiles <= kPartialTiles).
int evaluate_small(FpxEvaluator *ev, double *chi2_host, FpxResponse *resp_host, hipStream_t s) {
  FpxSlot slot = slot_of(ev, reinterpret_cast<const FpxSource *>(ev->host_dev + ev->o_stage),
                         nullptr, nullptr);
  slot.accum = reinterpret_cast<double *>(ev->host_dev + ev->o_part);
  slot.flags = FPX_SLOT_TILE_PARTIALS;
  const uint32_t seq = next_seq(ev);
  int rc = fpx::launch_one(slot, ev->n_tiles, ev->max_gauss, ev->cfg, ev->launch_flags, ev->queue,
                           s,
                           false, reinterpret_cast<uint32_t *>(ev->host_dev + kIoFlag), seq);
  if (rc) return rc;
  rc = wait_flag(ev, seq, s);
  if (rc) return rc;
  const double *rows = reinterpret_cast<const double *>(ev->host + ev->o_part);
  const int nv = 1 + NPARAMS * ev->n_sources;
  std::vector<double> sum((size_t)ev->n_bands * nv, 0.0);
  for (int t = 0; t < ev->n_tiles; ++t) {
    double *dst = sum.data() + (size_t)ev->tile_band[t] * nv;
    const double *src = rows + (size_t)t * kAccRow;
    for (int j = 0; j < nv; ++j) dst[j] += src[j];
  }
  bool nan = false;
  for (int b = 0; b < ev->n_bands; ++b) nan |= sum[(size_t)b * nv] != sum[(size_t)b * nv];
  if (nan) {
    uint32_t status = 0;
    EV_HIP(hipMemcpy(&status, ev->status, sizeof(status), hipMemcpyDeviceToHost));
    rc = read_status(ev, s, status);
    if (rc) return rc;
  }
  for (int b = 0; b < ev->n_bands; ++b) {
    chi2_host[b] = sum[(size_t)b * nv];
    if (resp_host) {
      float *g = resp_host[b].dchi2_dparam;
      std::memset(g, 0, sizeof(FpxResponse));
      for (int j = 1; j < nv; ++j) g[j - 1] = (float)sum[(size_t)b * nv + j];
    }
  }
  return FPX_OK;
}

}  // namespace

extern "C" {

int fpx_bind(const FpxPatch *patch_dev, int32_t n_bands, int32_t cfg, int32_t persistent,
             FpxEvaluator **out) {
  if (!out) return arg_err("NULL output handle");
  *out = nullptr;
  FpxEvaluator *ev = new FpxEvaluator();
  const int rc = bind(patch_dev, n_bands, cfg, persistent, ev);
  if (rc) {
    release(ev);
    return rc;
  }
  *out = ev;
  return FPX_OK;
}

int fpx_bind_tiles(const FpxPatch *patch_dev, int32_t n_bands, int32_t cfg, int32_t persistent,
                   const FpxTile *tiles_host, int32_t n_tiles, FpxEvaluator **out) {
  if (!out) return arg_err("NULL output handle");
  *out = nullptr;
  if (!tiles_host && n_tiles != 0) return arg_err("NULL tile table");
  FpxEvaluator *ev = new FpxEvaluator();
  static const FpxTile kNone = {};
  const int rc = bind(patch_dev, n_bands, cfg, persistent, ev, tiles_host ? tiles_host : &kNone,
                      n_tiles);
  if (rc) {
    release(ev);
    return rc;
  }
  *out = ev;
  return FPX_OK;
}

int fpx_evaluate(FpxEvaluator *ev, const FpxSource *proposal_dev, double *pchi2_dev,
                 FpxResponse *pdchi2_dp_dev, void *stream) {
  if (!ev || !proposal_dev || !pchi2_dev || !pdchi2_dp_dev)
    return arg_err("NULL evaluator, proposal or output");
  const int rc = check_device(ev);
  if (rc) return rc;
  return fpx::launch_one(slot_of(ev, proposal_dev, pchi2_dev, pdchi2_dp_dev), ev->n_tiles,
                         ev->max_gauss, ev->cfg, ev->launch_flags, ev->queue, stream);
}

int fpx_evaluate_vjp(FpxEvaluator *ev, const FpxSource *proposal_dev, const float *pixel_weight_dev,
                     double *pchi2_dev, FpxResponse *pdchi2_dp_dev, void *stream) {
  if (!ev || !proposal_dev || !pixel_weight_dev || !pchi2_dev || !pdchi2_dp_dev)
    return arg_err("NULL evaluator, proposal, pixel weights or output");
  const int rc = check_device(ev);
  if (rc) return rc;
  FpxSlot slot = slot_of(ev, proposal_dev, pchi2_dev, pdchi2_dp_dev);
  slot.pixel_weight = pixel_weight_dev;
  return fpx::launch_one(slot, ev->n_tiles, ev->max_gauss, ev->cfg, ev->launch_flags, ev->queue,
                         stream);
}

int fpx_evaluate_model(FpxEvaluator *ev, const FpxSource *proposal_dev, double *pchi2_dev,
                       FpxResponse *pdchi2_dp_dev, void *stream) {
  if (!ev || !proposal_dev || !pchi2_dev || !pdchi2_dp_dev)
    return arg_err("NULL evaluator, proposal or output");
  const int rc = check_device(ev);
  if (rc) return rc;
  FpxSlot slot = slot_of(ev, proposal_dev, pchi2_dev, pdchi2_dp_dev);
  slot.flags = FPX_SLOT_MODEL_ONLY;
  return fpx::launch_one(slot, ev->n_tiles, ev->max_gauss, ev->cfg, ev->launch_flags, ev->queue,
                         stream);
}

int fpx_evaluate_host(FpxEvaluator *ev, const FpxSource *proposal_host, int32_t n_sources,
                      double *chi2_host, FpxResponse *resp_host, void *stream) {
  if (!ev || !proposal_host || !chi2_host) return arg_err("NULL evaluator, proposal or output");
  if (n_sources != ev->n_sources)
    return arg_err("proposal length differs from the bound patch's n_sources");
  int rc = check_device(ev);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  unsigned char *stage = ev->host + ev->o_stage;
  const size_t pbytes = sizeof(FpxSource) * (size_t)n_sources;
  std::memcpy(stage, proposal_host, pbytes);
  if (!ev->tile_band.empty()) return evaluate_small(ev, chi2_host, resp_host, s);
  EV_HIP(hipMemcpyAsync(ev->prop, stage, pbytes, hipMemcpyHostToDevice, s));
  // the launch's output pass (its last workgroup) writes pchi2 and the
  // responses into the mapped host block directly
  double *chi2_out = reinterpret_cast<double *>(ev->host_dev + kIoChi2);
  FpxResponse *resp_out = reinterpret_cast<FpxResponse *>(ev->host_dev + kIoResp);
  if (ev->n_tiles > 0) {
    const uint32_t seq = next_seq(ev);
    rc = fpx::launch_one(slot_of(ev, ev->prop, chi2_out, resp_out), ev->n_tiles, ev->max_gauss,
                         ev->cfg, ev->launch_flags, ev->queue, stream, true,
                         reinterpret_cast<uint32_t *>(ev->host_dev + kIoFlag), seq);
    if (rc) return rc;
    rc = wait_flag(ev, seq, s);
    if (rc) return rc;
  } else {  // no tiles: the output kernel alone writes zeros
    rc = fpx::launch_one(slot_of(ev, ev->prop, chi2_out, resp_out), 0, ev->max_gauss, ev->cfg,
                         ev->launch_flags, ev->queue, stream);
    if (rc) return rc;
    EV_HIP(hipStreamSynchronize(s));
  }
  const double *chi2 = reinterpret_cast<const double *>(ev->host + kIoChi2);
  // a set status word makes the output pass write NaN chi2: only then is
  // the word read (and cleared) from the device
  bool nan = false;
  for (int b = 0; b < ev->n_bands; ++b) nan |= chi2[b] != chi2[b];
  if (nan) {
    uint32_t status = 0;
    EV_HIP(hipMemcpy(&status, ev->status, sizeof(status), hipMemcpyDeviceToHost));
    rc = read_status(ev, s, status);
    if (rc) return rc;
  }
  std::memcpy(chi2_host, chi2, sizeof(double) * ev->n_bands);
  if (resp_host) std::memcpy(resp_host, ev->host + kIoResp, sizeof(FpxResponse) * ev->n_bands);
  return FPX_OK;
}

int fpx_evaluator_status(FpxEvaluator *ev, void *stream) {
  if (!ev) return arg_err("NULL evaluator");
  int rc = check_device(ev);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  EV_HIP(hipStreamSynchronize(s));
  uint32_t status = 0;
  EV_HIP(hipMemcpy(&status, ev->status, sizeof(status), hipMemcpyDeviceToHost));
  return read_status(ev, s, status);
}

int fpx_evaluator_info(const FpxEvaluator *ev, int32_t *n_tiles, int32_t *cfg,
                       int32_t *max_gauss) {
  if (!ev) return arg_err("NULL evaluator");
  if (n_tiles) *n_tiles = ev->n_tiles;
  if (cfg) *cfg = ev->cfg;
  if (max_gauss) *max_gauss = ev->max_gauss;
  return FPX_OK;
}

int fpx_release(FpxEvaluator *ev) {
  release(ev);
  return FPX_OK;
}

}  // extern "C"
