// fpx_static.hip -- fpx_evaluate_static: the evaluation of fpx_evaluate with
// no allocation, for callers that cannot allocate device memory (code under
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The buffers the kernels need beyond the caller's inputs and outputs -- the
// per-band FP64 accumulator, the work-queue counter of persistent launches
// and the status word -- live in `scratch`, a static device array of
// kScratchBytes that every GPU the library runs on holds until the process
// ends.  Its device address is looked up once per device with
// hipGetSymbolAddress; host code uses only that address, never the array's
// name (which on the host is a stand-in, not a device address).  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, two threads) race and must not be issued.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "forcepho_amd.h"
#include "fpx_internal.h"

namespace {

// [0, 4) status word, [8, 16) queue counter (2 x uint32), [256, ...) the
// FP64 accumulator of up to MAXBANDS bands
constexpr size_t kStatusOff = 0;
constexpr size_t kQueueOff = 8;
constexpr size_t kAccumOff = 256;
constexpr size_t kAccumDoubles = (size_t)MAXBANDS * (1 + MAXSOURCES * NPARAMS);
constexpr size_t kScratchBytes = kAccumOff + kAccumDoubles * sizeof(double);
static_assert(kScratchBytes == FPX_STATIC_SCRATCH_BYTES, "scratch size in the header");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Zero the status word, the queue and the first n_doubles of the accumulator
// (a fill kernel rather than a memset, so that a caller may record the call
// into a graph).
__global__ __launch_bounds__(256) void clear_scratch(unsigned char *base, int n_doubles) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_doubles) reinterpret_cast<double *>(base + kAccumOff)[i] = 0.0;
  if (i < 4) reinterpret_cast<uint32_t *>(base)[i] = 0u;  // status, pad, queue[0..1]
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (looked up on the first call on
// that device).
int scratch_address(unsigned char **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fpx::report_hip_error(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fpx::report_error(FPX_E_NODEV, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fpx::report_hip_error(e, "hipGetSymbolAddress(scratch)");
    g_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_ptr[dev];
  return FPX_OK;
}

}  // namespace

extern "C" {

int fpx_evaluate_static(const FpxPatch *patch_dev, const FpxTile *tiles_dev, int32_t n_tiles,
                        int32_t n_bands, int32_t max_gauss, const FpxSource *proposal_dev,
                        double *pchi2_dev, FpxResponse *pdchi2_dp_dev, int32_t cfg,
                        int32_t persistent, void *stream) {
  if (!patch_dev || !proposal_dev || !pchi2_dev || !pdchi2_dp_dev)
    return fpx::report_error(FPX_E_ARG, "NULL patch, proposal or output");
  if (n_tiles < 0 || (n_tiles > 0 && !tiles_dev))
    return fpx::report_error(FPX_E_ARG, "n_tiles negative, or tiles NULL");
  if (n_bands < 1 || n_bands > MAXBANDS)
    return fpx::report_error(FPX_E_ARG,
                             "n_bands out of range [1, MAXBANDS]: the static scratch holds the "
                             "accumulators of at most MAXBANDS bands");
  if (cfg < 0) cfg = fpx::kDefaultConfig;
  if (fpx::config_pixels(cfg) <= 0) return fpx::report_error(FPX_E_ARG, "invalid kernel configuration");
  unsigned char *base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  const int n_doubles = (int)fpx_accum_doubles(n_bands);
  hipLaunchKernelGGL(clear_scratch, dim3((n_doubles + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, base, n_doubles);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fpx::report_hip_error(e, "clear_scratch");
  FpxSlot slot;
  std::memset(&slot, 0, sizeof(slot));
  slot.patch = patch_dev;
  slot.proposal = proposal_dev;
  slot.tiles = tiles_dev;
  slot.pchi2 = pchi2_dev;
  slot.pdchi2_dp = reinterpret_cast<float *>(pdchi2_dp_dev);
  slot.accum = reinterpret_cast<double *>(base + kAccumOff);
  slot.status = reinterpret_cast<uint32_t *>(base + kStatusOff);
  slot.n_bands = n_bands;
  slot.n_tiles = n_tiles;
  return fpx::launch_one(slot, n_tiles, max_gauss, cfg, persistent ? FPX_LAUNCH_PERSISTENT : 0,
                         reinterpret_cast<uint32_t *>(base + kQueueOff), stream);
}

}  // extern "C"
