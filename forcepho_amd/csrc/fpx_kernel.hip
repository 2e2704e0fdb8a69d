// fpx_kernel.hip -- the chi2 + gradient kernel for gfx950 and its C ABI.
//
// Work decomposition.  A launch evaluates one or more FpxSlots (a patch plus
// a proposal).  Each slot's pixels are cut into tiles of 64*PPL pixels of one
// exposure (fpx_plan_tiles).  One workgroup of W wavefronts owns one tile at a
// time: every wavefront holds the same 64*PPL pixels (PPL per lane) and the
// W wavefronts split the exposure's ImageGaussian list between them, so the
// parallelism is pixels x Gaussians rather than pixels alone (the
// reference runs one block per band: compute_gaussians_kernel.cu:209,262).
//
// Per tile:
//   prologue  one lane per (source, PSF component) tests whether the
//             Gaussian can reach the tile's pixel box and the survivors are
//             compacted in order; their ImageGaussian records (fpx_gauss.h)
//             are built into LDS, in chunks when they exceed the LDS budget.
//   pass 1    model = sum_g amp_g G_g H_g over the wave's Gaussians (each
//             wavefront an equal contiguous range of the records)
//             (common_kernel.cc:57-83), partial models combined through LDS,
//             residual r = d - model, chi2 += (r ierr - d ierr)(r ierr + d ierr)
//             in FP64 (compute_gaussians_kernel.cc:286-293), weight
//             w2 = -2 r ierr^2.
//   pass 2    per Gaussian, 12 sums over the lane's pixels of w2 G and
//             w2 G H times (1, dx, dy, dx^2, dy^2, dx dy), mapped by the
//             record's coefficients onto 7 per-source accumulators of
//             w2 * d model / d theta (common_kernel.cc:95-143; the ra/dec
//             Jacobian CW is applied once per source after the Gaussian
//             loop), wave-reduced into per-wave LDS rows.
//   combine   per-wave rows summed in fixed order, one FP64 global atomic per
//             value into the slot's per-band accumulator.  pchi2 / Response
//             are written from the accumulators, which are re-zeroed, by a
//             second kernel (finalize_kernel, same stream: the kernel boundary
//             orders it after every tile) or, in a launch of one slot, by the
//             workgroup that finishes last (found with a device-scope done
//             count, kModeOutputs).
// A Gaussian is skipped at a pixel when dx^T F dx > MAX_EXP_ARG
// (common_kernel.cc:73,108).  Two shortcuts skip work that this test would
// discard anyway: the prologue builds records only for Gaussians whose
// MAX_EXP_ARG box reaches the tile's pixel bounding box, and a Gaussian whose
// exponent exceeds the limit at every pixel of a wavefront stops after the
// exponent (wave-uniform branch).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "forcepho_amd.h"
#include "fpx_gauss.h"
#include "fpx_internal.h"

static_assert(sizeof(FpxPatch) == 152, "Patch layout (patch.cc:23-100)");
static_assert(offsetof(FpxPatch, n_sources) == 72, "Patch.n_sources offset");
static_assert(offsetof(FpxPatch, rad2) == 80, "Patch.rad2 offset");
static_assert(offsetof(FpxPatch, psfgauss_start) == 144, "Patch.psfgauss_start offset");
static_assert(sizeof(FpxSource) == 264, "Source layout (proposal.cu:13-27)");
static_assert(offsetof(FpxSource, fluxes) == 24, "Source.fluxes offset");
static_assert(offsetof(FpxSource, mixture_amplitudes) == 144, "Source.mixture_amplitudes");
static_assert(offsetof(FpxSource, damplitude_drh) == 184, "Source.damplitude_drh");
static_assert(offsetof(FpxSource, damplitude_dnsersic) == 224, "Source.damplitude_dnsersic");
static_assert(sizeof(FpxResponse) == 840, "Response layout (proposal.cu:33-39)");
static_assert(sizeof(FpxPSFSourceGaussian) == 28, "PSFSourceGaussian layout (patch.cc:103-119)");
static_assert(sizeof(FpxTile) == 16, "tile layout");
static_assert(offsetof(FpxTile, part) == 6 && offsetof(FpxTile, parts) == 7, "tile parts");
static_assert(sizeof(FpxSlot) == 88, "slot layout");
static_assert(offsetof(FpxSlot, split) == 80, "FpxSlot.split offset");

namespace fpx {

constexpr int kAcc = 1 + MAXSOURCES * NPARAMS;        // doubles per band accumulator
// dx^T F dx > MAX_EXP_ARG  <=>  log2 G < kLog2Min (log2 G = kExp2Scale dx^T F dx)
constexpr float kLog2Min = kExp2Scale * (float)MAX_EXP_ARG;

// Split buffer of a slot (FpxSlot.split): a uint32 counter per entry of the
// tile table, padded to 256 bytes, then a row of 64*PPL floats per entry,
// the partial model of the piece at that entry.  A split tile's counter is
// the one at its piece 0.
__host__ __device__ constexpr size_t split_counter_bytes(int n_tiles) {
  return ((size_t)n_tiles * sizeof(uint32_t) + 255) / 256 * 256;
}
// time limit of a piece's wait for the others, in s_memrealtime ticks (100 MHz)
constexpr unsigned long long kSplitWaitTicks = 100000ull * FPX_SPLIT_WAIT_MS;

// ------------------------------------------------------------------------
// wavefront helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ bool wave_any(bool p) {
  return __builtin_amdgcn_ballot_w64(p) != 0;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ float uniform_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// Lanes below this one whose bit is set in mask.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// ------------------------------------------------------------------------
// Address spaces.  Reads of data that the kernel never writes and that every
// lane reads at the same address (slot table, tile records, the patch
// header and its per-exposure scalars) go through the constant address
// space, i.e. scalar loads: they need no wait on vector memory and land in
// SGPRs.  Per-lane reads go through the global address space, whose loads
// take an SGPR base and need no flat-address handling.
// ------------------------------------------------------------------------
#define FPX_CONST_AS __attribute__((address_space(4)))
#define FPX_GLOBAL_AS __attribute__((address_space(1)))
template <typename T>
using cptr = const FPX_CONST_AS T *;
template <typename T>
using gptr = const FPX_GLOBAL_AS T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
  return (cptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<T> as_global(const T *p) {
  return (gptr<T>)p;
}

// A struct read word by word through a pointer of another address space
// (the implicit copy of a struct binds only generic references).
template <typename T, typename WordPtr>
__device__ __forceinline__ T load_words(WordPtr src) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  T v;
  uint32_t *dst = reinterpret_cast<uint32_t *>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
  return v;
}
template <typename T>
__device__ __forceinline__ T load_const(const T *p, int i) {
  return load_words<T>((const FPX_CONST_AS uint32_t *)(p + i));
}
template <typename T>
__device__ __forceinline__ T load_global(const T *p, int i) {
  return load_words<T>((const FPX_GLOBAL_AS uint32_t *)(p + i));
}

// The fields of a patch that a tile reads, from one scalar read of the
// header.
struct PatchView {
  gptr<float> data, ierr, xpix, ypix, rad2, D, crpix, crval, CW;
  cptr<float> G;
  cptr<int32_t> n_psf_per_source, psfgauss_start;
  const FpxPSFSourceGaussian *psfgauss;  // read with load_global
  FPX_GLOBAL_AS float *residual;
  int n_sources, n_radii;
};

__device__ __forceinline__ PatchView patch_view(const FpxPatch *P) {
  const cptr<FpxPatch> h = as_const(P);
  PatchView v;
  v.data = as_global<float>(h->data);
  v.ierr = as_global<float>(h->ierr);
  v.xpix = as_global<float>(h->xpix);
  v.ypix = as_global<float>(h->ypix);
  v.rad2 = as_global<float>(h->rad2);
  v.D = as_global<float>(h->D);
  v.crpix = as_global<float>(h->crpix);
  v.crval = as_global<float>(h->crval);
  v.CW = as_global<float>(h->CW);
  v.G = as_const<float>(h->G);
  v.n_psf_per_source = as_const<int32_t>(h->n_psf_per_source);
  v.psfgauss_start = as_const<int32_t>(h->psfgauss_start);
  v.psfgauss = h->psfgauss;
  v.residual = (FPX_GLOBAL_AS float *)h->residual;
  v.n_sources = h->n_sources;
  v.n_radii = h->n_radii;
  return v;
}

// ------------------------------------------------------------------------
// LDS carve-up of one workgroup
//   records  gauss_cap x kRec floats   ImageGaussians of the current chunk
//                                       of the tile's surviving Gaussians
//   part     W x 64*PPL floats         per-wave partial models
//   geo      MAXSOURCES x kGeoFloats   per-source geometry of the exposure
//                                       (T, dT/dq, dT/dpa, centre)
//   cwl      MAXSOURCES x 4 floats     CW of every source for the exposure
//   rad2     kRad2 floats              the patch's radius-bin variances
//   rows     W x kRowBytes             per-wave accumulators: chi2 (double)
//                                       then MAXSOURCES*NPARAMS gradient sums
//   index    index_cap + 1 uint16      cull flags, then the survivors'
//                                       (source, component) indices in order
//                                       and their count (all below 2^16)
// ------------------------------------------------------------------------
constexpr int kRowBytes = 8 + MAXSOURCES * NPARAMS * 4;
static_assert(MAXSOURCES * FPX_MAX_PSF_PER_SOURCE < 65536, "LDS index entries are 16-bit");

constexpr int kRad2 = 16;  // >= MAXRADII
constexpr int kExpoFloats = MAXSOURCES * (kGeoFloats + 4) + kRad2;

__host__ __device__ constexpr size_t lds_bytes_for(int waves, int ppl, int gauss_cap,
                                                   int index_cap) {
  return (size_t)gauss_cap * kRec * sizeof(float) + (size_t)waves * ppl * 64 * sizeof(float) +
         (size_t)kExpoFloats * sizeof(float) + (size_t)waves * kRowBytes +
         (size_t)(index_cap + 1) * sizeof(uint16_t);
}

struct Lds {
  float *recs;
  float *part;
  float *geo;
  float *cwl;
  float *rad2;
  unsigned char *rows;
  uint16_t *index;
};

__device__ __forceinline__ Lds carve(unsigned char *smem, int waves, int ppl, int gauss_cap) {
  Lds l;
  l.recs = reinterpret_cast<float *>(smem);
  l.part = l.recs + (size_t)gauss_cap * kRec;
  l.geo = l.part + waves * 64 * ppl;
  l.cwl = l.geo + MAXSOURCES * kGeoFloats;
  l.rad2 = l.cwl + MAXSOURCES * 4;
  l.rows = reinterpret_cast<unsigned char *>(l.rad2 + kRad2);
  l.index = reinterpret_cast<uint16_t *>(l.rows + waves * kRowBytes);
  return l;
}

__device__ __forceinline__ double &row_chi2(unsigned char *row) {
  return *reinterpret_cast<double *>(row);
}
__device__ __forceinline__ float *row_grad(unsigned char *row) {
  return reinterpret_cast<float *>(row + 8);
}

// Pixel bounding box of the tile (all wavefronts hold the same pixels);
// empty (x0 > x1) when no pixel is valid.
struct Box {
  float x0, x1, y0, y1;
};

// Radius-bin variance of a PSF entry from the LDS copy.  Only entries whose
// bin the cull has checked against [0, n_radii) get here.
__device__ __forceinline__ float bin_covar(const Lds &l, const FpxPSFSourceGaussian &psf) {
  return l.rad2[psf.sersic_radius_bin];
}

// Cull the n_src x n_psf Gaussians of the exposure against the tile's box:
// every thread flags one Gaussian per round of W * 64, then the flags are
// compacted into the list of survivors in (source, component) order.
// Returns the number of survivors.  A PSF entry whose radius bin lies
// outside [0, n_radii) never survives and sets FPX_STATUS_RADIUS_BIN in the
// slot's status word.  psf0 is the PSF entry of Gaussian threadIdx.x (when
// below n), loaded by the caller ahead of time.
//
// Compaction: workgroups of 8 or more wavefronts keep their flags in
// registers (one bit per round; n <= MAXSOURCES * FPX_MAX_PSF_PER_SOURCE =
// 3600 gives at most 64 rounds x wavefronts) and write the list in parallel
// from per-wavefront ballot counts and a prefix over them; narrower
// workgroups, with fewer wavefronts to share the work and more rounds, flag
// into LDS and compact with wavefront 0 alone, which measured faster there.
template <int W>
__device__ int cull_tile(const FpxPSFSourceGaussian *psfg, int pstart, int n_src, int n_psf,
                         int n_radii, const Box &box, const Lds &l, int wave, int lane,
                         uint32_t *status, const FpxPSFSourceGaussian &psf0) {
  constexpr int kThreads = W * 64;
  constexpr bool kParallel = W >= 8;
  const int n = n_src * n_psf;
  const int t = threadIdx.x;
  uint16_t *index = l.index;
  auto reaches = [&](int lg, const FpxPSFSourceGaussian &psf) {
    const int s = lg / n_psf;
    const bool bin_ok = psf.sersic_radius_bin >= 0 && psf.sersic_radius_bin < n_radii;
    if (!bin_ok && status != nullptr) atomicOr(status, (uint32_t)FPX_STATUS_RADIUS_BIN);
    return bin_ok && gauss_reaches(load_geometry(l.geo + s * kGeoFloats), psf, bin_covar(l, psf),
                                   box.x0, box.x1, box.y0, box.y1);
  };
  __syncthreads();  // previous users of the index and l.part are done; geometry is in LDS
  if constexpr (kParallel) {
    static_assert(((MAXSOURCES * FPX_MAX_PSF_PER_SOURCE + kThreads - 1) / kThreads) * W <= 64,
                  "one flag bit per round and one count per (round, wavefront) lane");
    int *counts = reinterpret_cast<int *>(l.part);  // [round][wave]
    const int rounds = (n + kThreads - 1) / kThreads;
    uint64_t keep = 0;
    for (int k = 0; k < rounds; ++k) {
      const int lg = k * kThreads + t;
      bool in = false;
      if (lg < n) in = reaches(lg, k == 0 ? psf0 : load_global(psfg, pstart + lg - (lg / n_psf) * n_psf));
      keep |= (in ? 1ull : 0ull) << k;
      const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
      if (lane == 0) counts[k * W + wave] = __builtin_popcountll(m);
    }
    __syncthreads();
    const int n_blocks = rounds * W;
    const int c = lane < n_blocks ? counts[lane] : 0;
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    const int excl = incl - c;
    for (int k = 0; k < rounds; ++k) {
      const bool in = ((keep >> k) & 1ull) != 0;
      const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
      const int base = __shfl(excl, k * W + wave, 64);
      if (in) index[base + lanes_below(m)] = (uint16_t)(k * kThreads + t);
    }
    return __builtin_amdgcn_readfirstlane(total);
  } else {
    if (t < n) index[t] = reaches(t, psf0) ? 1 : 0;
    for (int lg = t + kThreads; lg < n; lg += kThreads)
      index[lg] = reaches(lg, load_global(psfg, pstart + lg - (lg / n_psf) * n_psf)) ? 1 : 0;
    __syncthreads();
    if (wave == 0) {
      int count = 0;
      for (int b0 = 0; b0 < n; b0 += 64) {
        const int lg = b0 + lane;
        const bool keep = lg < n && index[lg] != 0;
        const unsigned long long m = __builtin_amdgcn_ballot_w64(keep);
        // every lane has read its flag above; writes land at or below b0 + lane
        if (keep) index[count + lanes_below(m)] = lg;
        count += __builtin_popcountll(m);
      }
      if (lane == 0) index[n] = count;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(l.index[n]);
  }
}

// Build the records of survivors [first, first + count) into LDS.
template <int W>
__device__ void build_records(const FpxPSFSourceGaussian *psfg, gptr<FpxSource> props,
                              int pstart, float gcal,
                              int band, int n_psf, int first, int count, const Lds &l) {
  __syncthreads();  // previous users of the records are done
  for (int i = threadIdx.x; i < count; i += W * 64) {
    const int lg = l.index[first + i];
    const int s = lg / n_psf;
    const FpxPSFSourceGaussian psf = load_global(psfg, pstart + lg - s * n_psf);
    build_record(load_geometry(l.geo + s * kGeoFloats), props + s, psf, bin_covar(l, psf), gcal,
                 band, s, l.recs + i * kRec);
  }
  __syncthreads();
}

// The pixel offsets from a Gaussian centre at a lane's PPL pixels and their
// products.  The passes compute these only when the centre changes from one
// record to the next: all radius bins of a PSF component share its centre,
// and the patch packer orders each exposure's PSF entries by centre
// (patches/patch.py), so a source's records come in runs of one centre.
template <int PPL>
__device__ __forceinline__ void offsets_at_pixels(const float4 r0, const float (&xp)[PPL],
                                                  const float (&yp)[PPL], float (&dx)[PPL],
                                                  float (&dy)[PPL], float (&dxx)[PPL],
                                                  float (&dyy)[PPL], float (&dxy)[PPL]) {
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    dx[i] = xp[i] - r0.x;
    dy[i] = yp[i] - r0.y;
    dxx[i] = dx[i] * dx[i];
    dyy[i] = dy[i] * dy[i];
    dxy[i] = dx[i] * dy[i];
  }
}

// log2 of one Gaussian at a lane's PPL pixels and the cutoff flag.  Returns a
// bit per pixel slot i (a 64-pixel block of the tile) that is set when some
// lane of the wavefront has pixel i inside the cutoff.
template <int PPL>
__device__ __forceinline__ unsigned gauss_log2(const float4 r0, const float cyy,
                                               const float (&dxx)[PPL], const float (&dyy)[PPL],
                                               const float (&dxy)[PPL], float (&lg)[PPL],
                                               bool (&on)[PPL]) {
  unsigned blocks = 0;
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    lg[i] = fmaf(r0.z, dxx[i], fmaf(r0.w, dxy[i], cyy * dyy[i]));
    on[i] = !(lg[i] < kLog2Min);
    blocks |= wave_any(on[i]) ? (1u << i) : 0u;
  }
  return blocks;
}

// Bits of a record's centre as a wave-uniform pair, to detect a change of
// centre with a scalar compare (all lanes read the same record).
__device__ __forceinline__ long long centre_key(const float4 r0) {
  const unsigned x = __builtin_amdgcn_readfirstlane(__builtin_bit_cast(unsigned, r0.x));
  const unsigned y = __builtin_amdgcn_readfirstlane(__builtin_bit_cast(unsigned, r0.y));
  return (long long)(((unsigned long long)x << 32) | y);
}
// no record has this centre (two NaN halves)
constexpr long long kNoCentre = (long long)0x7fc000017fc00001ull;

// A wavefront's share of a chunk's n records: the contiguous range
// [n w / W, n (w + 1) / W) of wavefront w.  Equal counts (to one record)
// balance the wavefronts better than dealing runs of records round-robin,
// whose last round leaves some wavefronts a run more than others, and a
// contiguous range holds few sources and keeps runs of one centre together
// (fewer per-source flushes and offset computations).
__device__ __forceinline__ int share_begin(int n, int W, int wave) { return (n * wave) / W; }

// Pass 1: add this wave's share of records [0, n) to the lane's model.
template <int PPL>
__device__ __forceinline__ void model_pass(const float *__restrict__ recs, int wave, int W,
                                           int n, const float (&xp)[PPL],
                                           const float (&yp)[PPL], float (&m)[PPL]) {
  const float4 *R4 = reinterpret_cast<const float4 *>(recs);
  float dx[PPL], dy[PPL], dxx[PPL], dyy[PPL], dxy[PPL];
  long long centre = kNoCentre;
  const int g_end = share_begin(n, W, wave + 1);
  for (int g = share_begin(n, W, wave); g < g_end; ++g) {
    const float4 r0 = R4[g * (kRec / 4) + 0];  // xc yc cxx cxy
    const float4 r1 = R4[g * (kRec / 4) + 1];  // cyy h0 sxx sxy
    const float4 r2 = R4[g * (kRec / 4) + 2];  // syy amp src dA/dflux
    const long long key = centre_key(r0);
    if (key != centre) {
      offsets_at_pixels<PPL>(r0, xp, yp, dx, dy, dxx, dyy, dxy);
      centre = key;
    }
    float lg[PPL];
    bool on[PPL];
    const unsigned blocks = gauss_log2<PPL>(r0, r1.x, dxx, dyy, dxy, lg, on);
    if (blocks == 0) continue;
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
      if (!(blocks & (1u << i))) continue;  // no lane has pixel i inside
      // An empty statement the compiler may not speculate: without it the
      // short body below is if-converted into selects and computed for every
      // block (2% of the kernel at the multi-band workload).  It emits no
      // instruction.
      __asm__ volatile("");
      const float G = on[i] ? __builtin_amdgcn_exp2f(lg[i]) : 0.0f;
      const float H = fmaf(r1.z, dxx[i], fmaf(r1.w, dxy[i], fmaf(r2.x, dyy[i], r1.y)));
      m[i] = fmaf(G * r2.y, H, m[i]);
    }
  }
}

// Sum 8 per-lane values over the wavefront with halving exchanges (10
// shuffles instead of 8 x 6): the xor-32, -16 and -8 steps each keep half of
// the remaining values and add the partner's copy of that half, so that lane
// l then holds value k(l) = 4 bit5(l) + 2 bit4(l) + bit3(l) summed over the 8
// lanes that differ from l in bits 3-5; the xor-4, -2 and -1 steps finish the
// sum.  Returns the wavefront total of value k(lane).
__device__ __forceinline__ float wave_sum8(const float (&v)[8], int lane) {
  const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0, b3 = (lane & 8) != 0;
  float w4[4], w2[2];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float send = b5 ? v[j] : v[4 + j];
    const float keep = b5 ? v[4 + j] : v[j];
    w4[j] = keep + __shfl_xor(send, 32, 64);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const float send = b4 ? w4[j] : w4[2 + j];
    const float keep = b4 ? w4[2 + j] : w4[j];
    w2[j] = keep + __shfl_xor(send, 16, 64);
  }
  const float send = b3 ? w2[0] : w2[1];
  const float keep = b3 ? w2[1] : w2[0];
  float t = keep + __shfl_xor(send, 8, 64);
  t += __shfl_xor(t, 4, 64);
  t += __shfl_xor(t, 2, 64);
  t += __shfl_xor(t, 1, 64);
  return t;
}

// Reduce one source's accumulators over the wavefront and add them to the
// wave's row: flux, ra, dec (the pixel-space x/y sums mapped through the
// ra/dec Jacobian CW of the source), q, pa, sersic, rh.
__device__ __forceinline__ void flush_source(int s, float aF, float aX, float aY, float aQ,
                                             float aP, float aN, float aR,
                                             const float *__restrict__ cwl,
                                             unsigned char *__restrict__ wrow, int lane) {
  const float *cw = cwl + 4 * s;  // [v11 v12 v21 v22] = dpix/dsky
  const float v[8] = {aF, cw[0] * aX + cw[2] * aY, cw[1] * aX + cw[3] * aY, aQ,
                      aP, aN, aR, 0.0f};
  const float t = wave_sum8(v, lane);
  const int k = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
  if ((lane & 7) == 0 && k < NPARAMS) row_grad(wrow)[NPARAMS * s + k] += t;
}

// Pass 2: gradient accumulation over this wave's share of records [0, n).  Per Gaussian the lane sums gw m_k and gw H m_k over its pixels
// (m = 1, dx, dy, dx^2, dy^2, dx dy) and maps the 12 sums to the source's 7
// accumulators with the record's coefficients; a source's sums are flushed
// when the next record belongs to another source.
template <int PPL>
__device__ __forceinline__ void grad_pass(const float *__restrict__ recs, int wave, int W,
                                          int n, const float (&xp)[PPL], const float (&yp)[PPL],
                                          const float (&w2)[PPL], const float *__restrict__ cwl,
                                          unsigned char *__restrict__ wrow, int lane) {
  const float4 *R4 = reinterpret_cast<const float4 *>(recs);
  float aF = 0.f, aX = 0.f, aY = 0.f, aQ = 0.f, aP = 0.f, aN = 0.f, aR = 0.f;
  bool touched = false;
  int cur = -1;
  float dx[PPL], dy[PPL], dxx[PPL], dyy[PPL], dxy[PPL];
  long long centre = kNoCentre;
  const int g_end = share_begin(n, W, wave + 1);
  for (int g = share_begin(n, W, wave); g < g_end; ++g) {
    const float4 *rg = R4 + g * (kRec / 4);
    const float4 r0 = rg[0];  // xc yc cxx cxy
    const float4 r1 = rg[1];  // cyy h0 sxx sxy
    const float4 r2 = rg[2];  // syy amp src dA/dflux
    const int src = __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, r2.z));
    if (src != cur) {
      if (touched) flush_source(cur, aF, aX, aY, aQ, aP, aN, aR, cwl, wrow, lane);
      aF = aX = aY = aQ = aP = aN = aR = 0.f;
      touched = false;
      cur = src;
    }
    const long long key = centre_key(r0);
    if (key != centre) {
      offsets_at_pixels<PPL>(r0, xp, yp, dx, dy, dxx, dyy, dxy);
      centre = key;
    }
    float lg[PPL];
    bool on[PPL];
    const unsigned blocks = gauss_log2<PPL>(r0, r1.x, dxx, dyy, dxy, lg, on);
    if (blocks == 0) continue;
    touched = true;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f, a5 = 0.f;
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f, b4 = 0.f, b5 = 0.f;
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
      if (!(blocks & (1u << i))) continue;  // no lane has pixel i inside
      const float G = on[i] ? __builtin_amdgcn_exp2f(lg[i]) : 0.0f;
      const float gw = G * w2[i];
      const float H = fmaf(r1.z, dxx[i], fmaf(r1.w, dxy[i], fmaf(r2.x, dyy[i], r1.y)));
      const float gh = gw * H;
      a0 += gw;
      a1 = fmaf(gw, dx[i], a1);
      a2 = fmaf(gw, dy[i], a2);
      a3 = fmaf(gw, dxx[i], a3);
      a4 = fmaf(gw, dyy[i], a4);
      a5 = fmaf(gw, dxy[i], a5);
      b0 += gh;
      b1 = fmaf(gh, dx[i], b1);
      b2 = fmaf(gh, dy[i], b2);
      b3 = fmaf(gh, dxx[i], b3);
      b4 = fmaf(gh, dyy[i], b4);
      b5 = fmaf(gh, dxy[i], b5);
    }
    const float4 r3 = rg[3];  // dA/dsersic dA/drh ex1 ex2
    const float4 r4 = rg[4];  // ex3 ex4 ey2 ey4
    const float4 r5 = rg[5];  // dA/dq kq0 kq3 kq4
    const float4 r6 = rg[6];  // kq5 hq3 hq4 hq5
    const float4 r7 = rg[7];  // dA/dpa kp0 kp3 kp4
    const float4 r8 = rg[8];  // kp5 hp3 hp4 hp5
    aF = fmaf(r2.w, b0, aF);
    aN = fmaf(r3.x, b0, aN);
    aR = fmaf(r3.y, b0, aR);
    aX = fmaf(r3.z, b1, fmaf(r3.w, b2, fmaf(r4.x, a1, fmaf(r4.y, a2, aX))));
    aY = fmaf(r3.w, b1, fmaf(r4.z, b2, fmaf(r4.y, a1, fmaf(r4.w, a2, aY))));
    aQ = fmaf(r5.x, b0, fmaf(r5.y, a0, fmaf(r5.z, a3, fmaf(r5.w, a4, aQ))));
    aQ = fmaf(r6.x, a5, fmaf(r6.y, b3, fmaf(r6.z, b4, fmaf(r6.w, b5, aQ))));
    aP = fmaf(r7.x, b0, fmaf(r7.y, a0, fmaf(r7.z, a3, fmaf(r7.w, a4, aP))));
    aP = fmaf(r8.x, a5, fmaf(r8.y, b3, fmaf(r8.z, b4, fmaf(r8.w, b5, aP))));
  }
  if (touched) flush_source(cur, aF, aX, aY, aQ, aP, aN, aR, cwl, wrow, lane);
}

// One tile: returns false when the exposure's n_psf breaks a size limit
// (FPX_STATUS_LIMIT is then set and the tile contributes nothing), true
// otherwise, the tile's sums being in the wave rows.
// Arrive at a split tile's counter and wait until all `parts` pieces have
// arrived (thread 0 polls; the other wavefronts wait at the barrier).  The
// rows and the counter are only ever accessed with device-scope atomic loads
// and stores, which are coherent across the XCDs by themselves; what the
// exchange needs is order: every wavefront waits until its row stores are
// performed before the barrier, so they precede the arrival, and the rows
// are read only after the count is seen (a control dependency).  A release /
// acquire pair would add a writeback and an invalidation of the whole L2 of
// the XCD for every piece, which the other workgroups there pay for.  The
// wait is bounded by kSplitWaitTicks: returns false when the time ran out.
__device__ bool split_arrive_and_wait(uint32_t *counter, int parts) {
  __shared__ int s_split_ok;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    int ok = 1;
    while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           (uint32_t)parts) {
      if (__builtin_amdgcn_s_memrealtime() - t0 > kSplitWaitTicks) {
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    s_split_ok = ok;
  }
  __syncthreads();
  return s_split_ok != 0;
}

// Results of tile_work: the tile was left out (a size limit), evaluated, or
// evaluated by a piece that arrived at its split counter (and must leave it).
enum TileResult { kTileLeftOut = 0, kTileDone = 1, kTileArrived = 2 };

template <int W, int PPL>
__device__ TileResult tile_work(const FpxSlot &slot, const FpxTile &tile, const PatchView &P,
                                gptr<FpxSource> props, int e, int band, int n_src,
                                int n_radii, int gauss_cap, int index_cap, int local, int part,
                                int parts, unsigned char *smem) {
  // The tile's Gaussians that reach it are processed in chunks of gauss_cap
  // records (one chunk when they all fit).
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // the wave index is uniform across the wavefront; saying so puts the
  // Gaussian loop bounds below in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const Lds l = carve(smem, W, PPL, gauss_cap);
  const bool vjp = slot.pixel_weight != nullptr;
  const int pix_count = tile.pix_count;

  // Every load that needs only the tile record and the patch header is
  // requested here, before any of them is waited for, so that the tile pays
  // one memory latency for all of them: the exposure's n_psf, PSF offset and
  // calibration, the pixels (with data and weights in chi2 mode, whose
  // latency then also overlaps the cull, the records and the model pass;
  // the pixel weights in vjp mode), the radius-bin variances and the
  // geometry inputs of one source per thread (MAXSOURCES and MAXRADII are
  // below the 64 threads of a wavefront).  Indices are clamped into range
  // instead of branched on.
  const int n_psf = P.n_psf_per_source[band];
  const int pstart = P.psfgauss_start[e];
  const float gcal = P.G[e];
  const float rad2_v = P.rad2[min(tid, n_radii - 1)];
  const int js = min(tid, n_src - 1);
  const GeometryInputs gin =
      load_geometry_inputs(P.D, P.CW, P.crpix, P.crval, props + js, n_src, e, js);
  gptr<float> pixel_weight = as_global<float>(slot.pixel_weight);
  float xp[PPL], yp[PPL], w2[PPL], dat[PPL], ier[PPL];
  int pix[PPL];
  bool valid[PPL];
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    const int off = i * 64 + lane;
    valid[i] = off < pix_count;
    pix[i] = tile.pix_start + off;
    xp[i] = yp[i] = dat[i] = ier[i] = w2[i] = 0.0f;
    if (i * 64 < pix_count) {  // some lane of the block holds a pixel
      const int pc = valid[i] ? pix[i] : tile.pix_start;
      const float x = P.xpix[pc], y = P.ypix[pc];
      xp[i] = valid[i] ? x : 0.0f;
      yp[i] = valid[i] ? y : 0.0f;
      if (!vjp) {
        const float d = P.data[pc], ie = P.ierr[pc];
        dat[i] = valid[i] ? d : 0.0f;
        ier[i] = valid[i] ? ie : 0.0f;
      } else {
        const float pw = pixel_weight[pc];
        w2[i] = valid[i] ? pw : 0.0f;
      }
    }
  }

  if (!(n_psf >= 1 && n_psf <= FPX_MAX_PSF_PER_SOURCE && n_src * n_psf <= index_cap)) {
    if (tid == 0 && slot.status != nullptr) atomicOr(slot.status, (uint32_t)FPX_STATUS_LIMIT);
    return kTileLeftOut;
  }
  // the cull's first PSF entry, requested before the geometry is computed
  const int lg0 = min(tid, n_src * n_psf - 1);
  const FpxPSFSourceGaussian psf0 = load_global(P.psfgauss, pstart + lg0 - (lg0 / n_psf) * n_psf);

  for (int j = tid; j < W * kRowBytes / 4; j += W * 64) reinterpret_cast<float *>(l.rows)[j] = 0.0f;
  // per-source geometry of the exposure, its CW and the radius-bin variances,
  // once per tile (the cull and the records read them from LDS)
  {
    SourceGeometry g;
    geometry_from<true>(gin, g);
    if (tid < n_src) {
      store_geometry(l.geo + tid * kGeoFloats, g);
#pragma unroll
      for (int c = 0; c < 4; ++c) l.cwl[4 * tid + c] = gin.CW[c];
    }
    if (tid < n_radii) l.rad2[tid] = rad2_v;
  }

  Box box;
  {
    constexpr float kInf = __builtin_huge_valf();
    float x0 = kInf, x1 = -kInf, y0 = kInf, y1 = -kInf;
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
      if (valid[i]) {
        x0 = fminf(x0, xp[i]);
        x1 = fmaxf(x1, xp[i]);
        y0 = fminf(y0, yp[i]);
        y1 = fmaxf(y1, yp[i]);
      }
    }
    box.x0 = uniform_f(wave_min(x0));
    box.x1 = uniform_f(wave_max(x1));
    box.y0 = uniform_f(wave_min(y0));
    box.y1 = uniform_f(wave_max(y1));
  }

  const int n_surv = cull_tile<W>(P.psfgauss, pstart, n_src, n_psf, n_radii, box, l, wave, lane,
                                  slot.status, psf0);
  // this piece's run of the survivors (all of them when the tile is whole)
  const int r_first = (n_surv * part) / parts;
  const int r_count = (n_surv * (part + 1)) / parts - r_first;
  const int n_chunks = (r_count + gauss_cap - 1) / gauss_cap;
  auto chunk_size = [&](int c) { return min(gauss_cap, r_count - c * gauss_cap); };
  bool arrived = false;
  if (!vjp) {
    float m[PPL];
#pragma unroll
    for (int i = 0; i < PPL; ++i) m[i] = 0.0f;
    // chunks in reverse, so that the first (full) chunk is the one left in
    // LDS for the gradient pass, which then rebuilds only the others
    for (int c = n_chunks - 1; c >= 0; --c) {
      const int n_act = chunk_size(c);
      build_records<W>(P.psfgauss, props, pstart, gcal, band, n_psf,
                       r_first + c * gauss_cap, n_act, l);
      model_pass<PPL>(l.recs, wave, W, n_act, xp, yp, m);
    }
#pragma unroll
    for (int i = 0; i < PPL; ++i) l.part[(wave * PPL + i) * 64 + lane] = m[i];
    __syncthreads();
    float model_at[PPL];
    if (parts > 1) {
      // A piece of a split tile: its partial model (the sum over its
      // wavefronts) to its row of the split buffer, then, once every piece
      // has arrived, the model is the sum of the pieces' rows in piece order
      // (the same sum whichever piece finishes first).
      constexpr int kPix = 64 * PPL;
      uint32_t *counter = reinterpret_cast<uint32_t *>(slot.split) + (local - part);
      const float *rows = reinterpret_cast<const float *>(
          reinterpret_cast<const unsigned char *>(slot.split) + split_counter_bytes(slot.n_tiles));
      float *mine = const_cast<float *>(rows) + (size_t)local * kPix;
      for (int j = tid; j < kPix; j += W * 64) {
        float sum = 0.0f;
#pragma unroll
        for (int w = 0; w < W; ++w) sum += l.part[(w * PPL + (j >> 6)) * 64 + (j & 63)];
        __hip_atomic_store(mine + j, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (!split_arrive_and_wait(counter, parts) && tid == 0 && slot.status != nullptr)
        atomicOr(slot.status, (uint32_t)FPX_STATUS_SPLIT_WAIT);
      arrived = true;
      const float *first_row = rows + (size_t)(local - part) * kPix;
#pragma unroll
      for (int i = 0; i < PPL; ++i) {
        float sum = 0.0f;
        for (int k = 0; k < parts; ++k)
          sum += __hip_atomic_load(first_row + (size_t)k * kPix + i * 64 + lane, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        model_at[i] = sum;
      }
    } else {
#pragma unroll
      for (int i = 0; i < PPL; ++i) {
        float sum = 0.0f;
#pragma unroll
        for (int w = 0; w < W; ++w) sum += l.part[(w * PPL + i) * 64 + lane];
        model_at[i] = sum;
      }
    }
    double chi2 = 0.0;
    // piece 0 of a split tile writes the residual and adds the chi2
    FPX_GLOBAL_AS float *resid = part == 0 ? P.residual : nullptr;
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
      const float model = model_at[i];
      const float d = dat[i];
      const float ie = ier[i];
      const float r = d - model;
      if (wave == 0 && valid[i] && resid != nullptr) resid[pix[i]] = r;
      const float ri = r * ie;
      const float di = d * ie;
      chi2 += ((double)ri - (double)di) * ((double)ri + (double)di);
      w2[i] = -2.0f * (ri * ie);
    }
    if (wave == 0 && part == 0) {
      chi2 = wave_sum(chi2);
      if (lane == 0) row_chi2(l.rows) += chi2;
    }
  }
  const TileResult done = arrived ? kTileArrived : kTileDone;

  // a model-only slot stops at its residual and chi2 (the flag is the same
  // for the whole workgroup, and process_tile's barrier follows)
  if (!vjp && (__builtin_amdgcn_readfirstlane(slot.flags) & FPX_SLOT_MODEL_ONLY)) return done;

  bool wave_has_weight = false;
#pragma unroll
  for (int i = 0; i < PPL; ++i) wave_has_weight |= (w2[i] != 0.0f);
  wave_has_weight = wave_any(wave_has_weight);

  unsigned char *wrow = l.rows + wave * kRowBytes;
  // first chunk first: after the model pass its records are still in LDS
  for (int c = 0; c < n_chunks; ++c) {
    const int n_act = chunk_size(c);
    if (c != 0 || vjp)
      build_records<W>(P.psfgauss, props, pstart, gcal, band, n_psf, r_first + c * gauss_cap,
                       n_act, l);
    if (wave_has_weight)
      grad_pass<PPL>(l.recs, wave, W, n_act, xp, yp, w2, l.cwl, wrow, lane);
  }
  return done;
}

// Add the tile's per-wave rows to the slot accumulator (band < 0: nothing
// to add).  No fence or completion count: the outputs are written by
// finalize_kernel, launched after this kernel on the same stream.  With
// FPX_SLOT_TILE_PARTIALS the sums are stored instead at
// accum[local * kAcc + j] (tile `local`'s own row, no atomics), for the
// caller to add up; a tile left out (band < 0), or one that finds the slot's
// status word set, stores NaN as its chi2.
template <int W>
__device__ void finish_tile(const FpxSlot &slot, int local, int band, int n_src,
                            unsigned char *smem, int gauss_cap, int ppl) {
  const int tid = threadIdx.x;
  const Lds l = carve(smem, W, ppl, gauss_cap);
  if (__builtin_amdgcn_readfirstlane(slot.flags) & FPX_SLOT_TILE_PARTIALS) {
    double *acc = slot.accum + (size_t)local * kAcc;
    const uint32_t st = slot.status == nullptr
                            ? 0u
                            : __hip_atomic_load(slot.status, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    const bool left_out = band < 0 || st != 0u;
    const int nv = 1 + NPARAMS * n_src;
    for (int j = tid; j < nv; j += W * 64) {
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        unsigned char *row = l.rows + w * kRowBytes;
        v += (j == 0) ? row_chi2(row) : (double)row_grad(row)[j - 1];
      }
      acc[j] = (j == 0 && left_out) ? __builtin_nan("") : v;
    }
  } else if (band >= 0) {
    double *acc = slot.accum + (size_t)band * kAcc;
    const int nv = 1 + NPARAMS * n_src;
    for (int j = tid; j < nv; j += W * 64) {
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        unsigned char *row = l.rows + w * kRowBytes;
        v += (j == 0) ? row_chi2(row) : (double)row_grad(row)[j - 1];
      }
      if (v != 0.0) unsafeAtomicAdd(acc + j, v);
    }
  }
  __syncthreads();  // the rows are reused by the next tile
}

// Outputs of every slot from its accumulator, which is left zero: pchi2[b]
// and the float Response of each band; pchi2 is NaN in every band of a slot
// whose status word is set (a tile broke a limit and was left out).  Grid
// (slot, chunk of kFinalizeThreads accumulator values): every thread moves
// at most one value, so the latency of the agent-scope load and store is
// paid once, not once per value.  slots == NULL: the one slot `one`, passed
// by value.
constexpr int kFinalizeThreads = 256;
constexpr int kFinalizeChunks = (MAXBANDS * kAcc + kFinalizeThreads - 1) / kFinalizeThreads;

// Output value j of a slot (j < n_bands * kAcc) from its accumulator value v
// (read by the caller), which is zeroed; st is the slot's status word.
__device__ __forceinline__ void store_output(const FpxSlot &slot, int j, double v, uint32_t st) {
  const int b = j / kAcc;
  const int k = j - b * kAcc;
  if (k == 0)
    slot.pchi2[b] = st == 0u ? v : __builtin_nan("");
  else
    slot.pdchi2_dp[b * (MAXSOURCES * NPARAMS) + (k - 1)] = (float)v;
  __hip_atomic_store(slot.accum + j, 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t load_status(const FpxSlot &slot) {
  return slot.status == nullptr
             ? 0u
             : __hip_atomic_load(slot.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void slot_output(const FpxSlot &slot, int j) {
  const double v = __hip_atomic_load(slot.accum + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  store_output(slot, j, v, j % kAcc == 0 ? load_status(slot) : 0u);
}

// All outputs of a slot from one workgroup of `threads` threads: each thread
// reads kBatch accumulator values before it writes any, so that the loads'
// latency is paid once per kBatch values.
template <int kBatch>
__device__ void slot_outputs(const FpxSlot &slot, int threads) {
  const int nv = min(slot.n_bands, MAXBANDS) * kAcc;
  const uint32_t st = load_status(slot);
  for (int j0 = threadIdx.x; j0 < nv; j0 += kBatch * threads) {
    double v[kBatch];
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const int j = j0 + i * threads;
      v[i] = j < nv ? __hip_atomic_load(slot.accum + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
      const int j = j0 + i * threads;
      if (j < nv) store_output(slot, j, v[i], st);
    }
  }
}

__global__ __launch_bounds__(kFinalizeThreads) void finalize_kernel(
    const FpxSlot *__restrict__ slots, FpxSlot one, int n_slots) {
  const FpxSlot slot = slots != nullptr ? slots[blockIdx.x] : one;
  const int nb = min(slot.n_bands, MAXBANDS);
  const int j = blockIdx.y * kFinalizeThreads + threadIdx.x;
  if (j < nb * kAcc) slot_output(slot, j);
}

// kSplit: the kernel variant that evaluates split tiles (FpxTile.parts >=
// 2).  The other variant leaves such a tile out (FPX_STATUS_LIMIT), and its
// code has no trace of the pieces' bookkeeping, which costs registers.
template <int W, int PPL, bool kSplit>
__device__ void process_tile(const FpxSlot &slot, int local, int gauss_cap, int index_cap,
                             unsigned char *smem) {
  constexpr int kPix = 64 * PPL;
  // The tile record and the patch header are read with scalar loads: the
  // sizes and loop bounds below are in SGPRs.
  const FpxTile tile = load_const(slot.tiles, local);
  const PatchView P = patch_view(slot.patch);
  const gptr<FpxSource> props = as_global(slot.proposal);
  const int e = tile.exposure;
  const int band = tile.band;
  const int n_src = P.n_sources;
  const int n_radii = P.n_radii;
  const int pix_count = tile.pix_count;
  const int parts = kSplit && tile.parts > 1 ? (int)tile.parts : 1;
  const int part = parts > 1 ? (int)tile.part : 0;
  // A tile whose sizes break the kernel's assumptions contributes nothing and
  // sets FPX_STATUS_LIMIT in the slot's status word, which makes the output
  // kernel write NaN chi2 (every lane sees the same values, so the whole
  // workgroup agrees).  The sizes read here come from the tile record and the
  // patch header; tile_work checks the exposure's n_psf.  A piece of a split
  // tile needs the split buffer and its whole run of entries in the table.
  bool ok = band >= 0 && band < slot.n_bands && band < MAXBANDS && n_src >= 1 &&
            n_src <= MAXSOURCES && n_radii >= 1 && n_radii <= MAXRADII && pix_count >= 0 &&
            pix_count <= kPix &&
            (kSplit ? (parts == 1 ||
                       (parts <= FPX_MAX_TILE_PARTS && part < parts && part <= local &&
                        slot.split != nullptr && local - part + parts <= slot.n_tiles))
                    : tile.parts <= 1);
  if (!ok && threadIdx.x == 0 && slot.status != nullptr)
    atomicOr(slot.status, (uint32_t)FPX_STATUS_LIMIT);
  TileResult res = kTileLeftOut;
  if (ok)
    res = tile_work<W, PPL>(slot, tile, P, props, e, band, n_src, n_radii, gauss_cap, index_cap,
                            local, part, parts, smem);
  __syncthreads();
  // A piece leaves its split counter after every wavefront has read the
  // pieces' rows (the barrier above); the piece that makes the count 2 *
  // parts is the last to leave and sets it back to zero for the next launch.
  if (res == kTileArrived && threadIdx.x == 0) {
    uint32_t *counter = reinterpret_cast<uint32_t *>(slot.split) + (local - part);
    const uint32_t before =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (before == 2u * (uint32_t)parts - 1u)
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const bool used = res != kTileLeftOut;
  finish_tile<W>(slot, local, used ? band : -1, used ? n_src : 0, smem, gauss_cap, PPL);
}

// Launch modes (bits of the kernel's `mode`).  kModePersistent: a resident
// grid pulls tiles from queue[0].  kModeOutputs: the last workgroup to finish
// writes the outputs of the launch's one slot (what finalize_kernel does
// otherwise); it is found with the done counter queue[1].  kModeFlag: that
// last workgroup, after everything else, stores `seq` into *flag (memory the
// host reads, mapped) with a system-scope release, so that the host can see
// the launch's results without waiting for the stream.  In every mode the
// last workgroup leaves queue[0] and queue[1] zero.
constexpr int kModePersistent = 1;
constexpr int kModeOutputs = 2;
constexpr int kModeFlag = 4;

// kOne: the single slot `one`, passed by value (no slot table in device
// memory; slots is not read).  A compile-time variant rather than a runtime
// choice, which would keep both slots' fields live (8 more VGPRs at 8x4, one
// wavefront per SIMD fewer).
//
// Shapes of up to 4 pixels per lane and 8 wavefronts are held to 128 VGPRs
// (4 wavefronts per SIMD), where they fit without spilling; the others get
// what they need.
template <int W, int PPL>
constexpr int kMinWavesPerEU = (PPL <= 4 && W <= 8) ? 4 : 1;

template <int W, int PPL, bool kOne, bool kSplit>
__global__ __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(kMinWavesPerEU<W, PPL>)))
void chi2_grad_kernel(
    const FpxSlot *__restrict__ slots, FpxSlot one, int n_slots,
    const int32_t *__restrict__ item_offsets, int n_items, int gauss_cap, int index_cap,
    int mode, uint32_t *__restrict__ queue, uint32_t *__restrict__ flag, uint32_t seq) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item;
  const bool persistent = (mode & kModePersistent) != 0;
  for (int it = 0;; ++it) {
    int item;
    if (persistent) {
      if (threadIdx.x == 0) s_item = (int)atomicAdd(queue, 1u);
      __syncthreads();
      item = __builtin_amdgcn_readfirstlane(s_item);
      __syncthreads();
    } else {
      if (it > 0) break;
      item = blockIdx.x;
    }
    if (item >= n_items) break;
    int si = 0;
    int local = item;
    if (!kOne && n_slots > 1) {
      int lo = 0, hi = n_slots - 1;  // largest si with item_offsets[si] <= item
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (as_const(item_offsets)[mid] <= item) lo = mid; else hi = mid - 1;
      }
      si = lo;
      local = item - as_const(item_offsets)[si];
    }
    const FpxSlot slot = kOne ? one : load_const(slots, si);
    process_tile<W, PPL, kSplit>(slot, local, gauss_cap, index_cap, smem);
  }
  if (mode == 0) return;
  // Count the workgroup done.  With outputs or a flag, every wave first waits
  // for its stores and atomics to be performed (the vector memory counter
  // drains only then), the barrier orders all of them before the count, and
  // the count is an acquire-release: at device scope when the last workgroup
  // writes the outputs (it sees every other workgroup's sums; its own output
  // stores reach the host through the flag's system-scope release below), at
  // system scope when every workgroup left its tiles' sums in host memory
  // for the host to read once it sees the flag (FPX_SLOT_TILE_PARTIALS).
  const bool outputs = (mode & kModeOutputs) != 0;
  const bool signal = (mode & kModeFlag) != 0;
  const bool partials =
      signal && (__builtin_amdgcn_readfirstlane((kOne ? one : load_const(slots, 0)).flags) &
                 FPX_SLOT_TILE_PARTIALS);
  if (outputs || signal) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned d;
    if (partials)
      d = __hip_atomic_fetch_add(queue + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_SYSTEM);
    else if (outputs || signal)
      d = __hip_atomic_fetch_add(queue + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    else
      d = atomicAdd(queue + 1, 1u);
    s_item = (int)(d == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_item) return;
  if (outputs) slot_outputs<4>(kOne ? one : load_const(slots, 0), W * 64);
  if (signal) {  // the outputs are stored before the flag
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    queue[0] = 0u;
    queue[1] = 0u;
    if (signal) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ------------------------------------------------------------------------
// Census of the (pixel, ImageGaussian) pairs inside the cutoff, i.e. those
// for which the reference's per-pixel loops go past their
// `exparg > MAX_EXP_ARG` test (common_kernel.cc:67-73, 103-108), with the
// reference's own exponent expression, over the packed pixels (padding
// included, as the reference's loops run over it).  One workgroup per tile;
// every Gaussian of the exposure is tested at every pixel (no culling).  Adds the
// count of each slot to pairs_in[slot] and, if item_pairs is given, each
// tile's count to item_pairs[item].  A measurement of the workload, used to
// report the roofline and to balance tiles over compute units; not part of
// the evaluation.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void census_kernel(
    const FpxSlot *__restrict__ slots, int n_slots, const int32_t *__restrict__ item_offsets,
    int n_items, int index_cap, unsigned long long *__restrict__ pairs_in,
    unsigned *__restrict__ item_pairs) {
  extern __shared__ __align__(16) unsigned char smem[];
  float *shape = reinterpret_cast<float *>(smem);  // index_cap x (xc yc fxx fxy fyy)
  const int item = blockIdx.x;
  int si = 0, local = item;
  if (n_slots > 1) {
    int lo = 0, hi = n_slots - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (item_offsets[mid] <= item) lo = mid; else hi = mid - 1;
    }
    si = lo;
    local = item - item_offsets[si];
  }
  const FpxSlot slot = slots[si];
  const FpxTile tile = slot.tiles[local];
  const FpxPatch *P = slot.patch;
  const int e = __builtin_amdgcn_readfirstlane(tile.exposure);
  const int band = __builtin_amdgcn_readfirstlane(tile.band);
  const int n_src = __builtin_amdgcn_readfirstlane(P->n_sources);
  if (band < 0 || band >= slot.n_bands || band >= MAXBANDS) return;
  if (tile.parts > 1 && tile.part != 0) return;  // a split tile is counted at its piece 0
  const int n_psf = __builtin_amdgcn_readfirstlane(P->n_psf_per_source[band]);
  if (n_src < 1 || n_src > MAXSOURCES || n_psf < 1 || n_src * n_psf > index_cap) return;
  const int n = n_src * n_psf;
  for (int g = threadIdx.x; g < n; g += blockDim.x) {
    const int s = g / n_psf;
    float *o = shape + 5 * g;
    image_gaussian_shape(P, slot.proposal, e, s, g - s * n_psf, o[0], o[1], o[2], o[3], o[4]);
  }
  __syncthreads();
  unsigned count = 0;
  for (int k = threadIdx.x; k < tile.pix_count; k += blockDim.x) {
    const float xp = P->xpix[tile.pix_start + k];
    const float yp = P->ypix[tile.pix_start + k];
    for (int g = 0; g < n; ++g) {
      const float *o = shape + 5 * g;
      const float dx = xp - o[0];
      const float dy = yp - o[1];
      const float vx = o[2] * dx + o[3] * dy;
      const float vy = o[4] * dy + o[3] * dx;
      const float exparg = dx * vx + dy * vy;
      count += (exparg > (float)MAX_EXP_ARG) ? 0u : 1u;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) count += __shfl_xor(count, o, 64);
  if ((threadIdx.x & 63) == 0 && count != 0) {
    atomicAdd(pairs_in + si, (unsigned long long)count);
    if (item_pairs) atomicAdd(item_pairs + item, count);
  }
}

// ------------------------------------------------------------------------
// VALU throughput probes: KIND 0 runs C independent FP32 FMA chains per
// lane, KIND 1 C independent v_exp_f32 chains.  With few chains a wavefront
// waits on each result (a latency measurement); with 16-32 chains and 8
// wavefronts per SIMD the issue rate itself is reached.  Lane 0 of every
// wavefront stores the elapsed shader-clock cycles (s_memtime) and lane 1 the
// elapsed 100 MHz reference ticks (s_memrealtime) of its loop; other lanes
// store the chain sum so the loop cannot be removed.
template <int KIND, int C>
__global__ __launch_bounds__(256) void probe_kernel(float *out, int iters) {
  float a[C];
#pragma unroll
  for (int k = 0; k < C; ++k) a[k] = (float)((threadIdx.x + k) & 63) * (1.0f / 64.0f);
  const float b = 0.999f, c = 1e-4f;
  const unsigned long long t0 = __builtin_amdgcn_s_memtime();
  const unsigned long long r0 = __builtin_amdgcn_s_memrealtime();
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < C; ++k) a[k] = KIND == 0 ? fmaf(a[k], b, c) : __builtin_amdgcn_exp2f(-a[k]);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < C; ++k) s += a[k];
  const unsigned long long t1 = __builtin_amdgcn_s_memtime();
  const unsigned long long r1 = __builtin_amdgcn_s_memrealtime();
  const int lane = threadIdx.x & 63;
  float v = s;
  if (lane == 0) v = (float)(t1 - t0);
  if (lane == 1) v = (float)(r1 - r0);
  out[blockIdx.x * blockDim.x + threadIdx.x] = v;
}

}  // namespace fpx

// ==========================================================================
// host side
// ==========================================================================
namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_err(hipError_t e, const char *where) {
  return set_err(FPX_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

// Text of a nonzero slot status word (FPX_STATUS_* bits).
std::string status_message(uint32_t status) {
  std::string m = "the kernel left tiles out:";
  if (status & FPX_STATUS_LIMIT)
    m += " a tile broke a size limit (n_sources, n_psf_per_source, n_radii, band, pixel count, "
         "or max_gauss below n_sources x n_psf_per_source);";
  if (status & FPX_STATUS_RADIUS_BIN) m += " a PSF entry's radius bin is outside [0, n_radii);";
  if (status & FPX_STATUS_SPLIT_WAIT)
    m += " a piece of a split tile timed out waiting for the others (not a persistent launch, "
         "or the split buffer was not zero);";
  if (status & ~(uint32_t)(FPX_STATUS_LIMIT | FPX_STATUS_RADIUS_BIN | FPX_STATUS_SPLIT_WAIT))
    m += " unknown status bits;";
  return m;
}

#define FPX_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hip_err(_e, #call);   \
  } while (0)

// LDS budget of a workgroup: 10 KiB per wavefront, so that four wavefronts
// per SIMD (16 per CU) fit in the 160 KiB of a CU; a tile whose surviving
// Gaussians exceed the records that fit is processed in chunks.
constexpr int kLdsPerWave = 10 * 1024;
// static LDS of the chi2 kernel (the queue item and the split-tile wait
// flag), for the host-only size query; launches use the kernel's reported
// size
constexpr int kStaticLds = 16;

// A tile shape and its four kernels, fn[2 * split + one]: one = 0 reads a
// slot table, 1 takes one slot by value; split = 1 evaluates split tiles.
constexpr int kVariants = 4;
struct KernelCfg {
  int waves;
  int ppl;
  const void *fn[kVariants];
};

#define FPX_CFG(W, P)                                                               \
  {W, P,                                                                            \
   {reinterpret_cast<const void *>(&fpx::chi2_grad_kernel<W, P, false, false>),     \
    reinterpret_cast<const void *>(&fpx::chi2_grad_kernel<W, P, true, false>),      \
    reinterpret_cast<const void *>(&fpx::chi2_grad_kernel<W, P, false, true>),      \
    reinterpret_cast<const void *>(&fpx::chi2_grad_kernel<W, P, true, true>)}}
const KernelCfg kConfigs[] = {
    FPX_CFG(4, 1), FPX_CFG(4, 2), FPX_CFG(4, 4), FPX_CFG(8, 8),
    FPX_CFG(16, 4), FPX_CFG(8, 4), FPX_CFG(2, 4), FPX_CFG(4, 8),
    FPX_CFG(8, 1),
};
#undef FPX_CFG
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

struct DeviceState {
  bool checked = false;
  bool ok = false;
  int n_cu = 0;
  int lds_max = 0;
  int clock_khz = 0;
  // per kernel, index kVariants * configuration + variant
  size_t lds_set[kVariants * kNumConfigs] = {};
  // resident workgroups per CU, cached per kernel and LDS size, so that a
  // launch recorded into a graph issues no occupancy query
  size_t occ_lds[kVariants * kNumConfigs] = {};
  int occ_per_cu[kVariants * kNumConfigs] = {};
  size_t census_lds = 0;
  int static_lds[kVariants * kNumConfigs];  // static LDS of each kernel, -1 until queried
  DeviceState() {
    for (int &v : static_lds) v = -1;
  }
};

std::mutex g_mu;
DeviceState g_dev[64];

int current_device_state(DeviceState **out) {
  int dev = 0;
  FPX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(FPX_E_NODEV, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState &st = g_dev[dev];
  if (!st.checked) {
    st.checked = true;
    hipDeviceProp_t prop;
    FPX_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      return set_err(FPX_E_NODEV, std::string("device is ") + prop.gcnArchName +
                                      ", this library is built for gfx950 only");
    }
    st.n_cu = prop.multiProcessorCount;
    st.lds_max = (int)prop.sharedMemPerBlock;
    st.clock_khz = prop.clockRate;
    st.ok = true;
  }
  if (!st.ok) return set_err(FPX_E_NODEV, "current device is not a usable gfx950");
  *out = &st;
  return FPX_OK;
}

int check_cfg(int32_t cfg) {
  if (cfg < 0 || cfg >= kNumConfigs) return set_err(FPX_E_ARG, "invalid kernel configuration");
  return FPX_OK;
}

// Records per LDS chunk: all max_gauss if they fit the workgroup's budget
// (kLdsPerWave per wavefront, less the kernel's static LDS), else what fits,
// but at least kMinChunk when the device's per-workgroup limit allows it.  0
// if not even one fits.
constexpr int kMinChunk = 16;
int pick_gauss_cap(const KernelCfg &kc, int max_gauss, int lds_max, int static_lds) {
  const long fixed = (long)fpx::lds_bytes_for(kc.waves, kc.ppl, 0, max_gauss);
  const long per = (long)(fpx::kRec * sizeof(float));
  lds_max -= static_lds;
  const long budget = std::min<long>(lds_max, (long)kc.waves * kLdsPerWave - static_lds);
  long fit = (budget - fixed) / per;
  if (fit < std::min<long>(max_gauss, kMinChunk)) fit = ((long)lds_max - fixed) / per;
  return (int)std::max<long>(0, std::min<long>(max_gauss, fit));
}

// Workgroups of configuration cfg, kernel variant v, with `lds` bytes of
// dynamic LDS that fit one CU at once (cached, so that a launch recorded into
// a graph makes no occupancy query).
int groups_per_cu(DeviceState *st, int cfg, int v, size_t lds, int *per_cu) {
  const int k = kVariants * cfg + v;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (st->occ_lds[k] == lds && st->occ_per_cu[k] > 0) {
      *per_cu = st->occ_per_cu[k];
      return FPX_OK;
    }
  }
  int n = 0;
  FPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kConfigs[cfg].fn[v],
                                                       kConfigs[cfg].waves * 64, lds));
  if (n < 1) n = 1;
  std::lock_guard<std::mutex> lock(g_mu);
  st->occ_lds[k] = lds;
  st->occ_per_cu[k] = n;
  *per_cu = n;
  return FPX_OK;
}

// Dynamic LDS bytes and records per chunk of a launch of configuration cfg,
// kernel variant v.
int launch_lds(DeviceState *st, int cfg, int v, int max_gauss, size_t *lds, int *cap,
               int *index_cap) {
  const KernelCfg &kc = kConfigs[cfg];
  const int k = kVariants * cfg + v;
  int static_lds = 0;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (st->static_lds[k] < 0) {
      hipFuncAttributes attr;
      FPX_HIP(hipFuncGetAttributes(&attr, kc.fn[v]));
      st->static_lds[k] = (int)attr.sharedSizeBytes;
    }
    static_lds = st->static_lds[k];
  }
  *cap = pick_gauss_cap(kc, max_gauss, st->lds_max, static_lds);
  if (*cap < 1) return set_err(FPX_E_ARG, "max_gauss too large for the LDS of one workgroup");
  *index_cap = max_gauss;
  *lds = fpx::lds_bytes_for(kc.waves, kc.ppl, *cap, *index_cap);
  return FPX_OK;
}

// Launch the tile kernel (and, with finalize, the output pass) over the slot
// table `slots` (device) or, when slots is NULL, the single slot *one (host
// struct, passed by value as a kernel argument).  The output pass is the
// tile kernel's last workgroup when the launch holds one slot and has a
// queue counter (kModeOutputs: one kernel), else finalize_kernel launched
// after it.
//
// launch_flags: FPX_LAUNCH_PERSISTENT, FPX_LAUNCH_SPLIT (the kernel variant
// that evaluates split tiles; its pieces wait for each other, which only a
// persistent launch guarantees to end well).
int launch(const FpxSlot *slots, const FpxSlot *one, int32_t n_slots,
           const int32_t *item_offsets, int32_t n_items, int32_t max_gauss, int32_t cfg,
           int32_t launch_flags, uint32_t *queue, void *stream, bool finalize_outputs = true,
           uint32_t *flag = nullptr, uint32_t seq = 0) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (launch_flags & ~(int32_t)(FPX_LAUNCH_PERSISTENT | FPX_LAUNCH_SPLIT))
    return set_err(FPX_E_ARG, "unknown launch flags");
  const bool persistent = (launch_flags & FPX_LAUNCH_PERSISTENT) != 0;
  const bool split = (launch_flags & FPX_LAUNCH_SPLIT) != 0;
  if (split && !persistent)
    return set_err(FPX_E_ARG, "split tiles need a persistent launch (FPX_LAUNCH_PERSISTENT)");
  if ((!slots && !one) || n_slots <= 0 || n_items < 0)
    return set_err(FPX_E_ARG, "empty or invalid slot list");
  if (!slots && n_slots != 1) return set_err(FPX_E_ARG, "a slot passed by value is one slot");
  if (n_slots > 1 && !item_offsets) return set_err(FPX_E_ARG, "item_offsets required for n_slots > 1");
  if (max_gauss <= 0 || max_gauss > MAXSOURCES * FPX_MAX_PSF_PER_SOURCE)
    return set_err(FPX_E_ARG, "max_gauss out of range [1, MAXSOURCES*FPX_MAX_PSF_PER_SOURCE]");
  if (persistent && !queue) return set_err(FPX_E_ARG, "persistent launch needs a queue counter");
  if (flag && !queue) return set_err(FPX_E_ARG, "a completion flag needs a queue counter");
  if (flag && n_items == 0) return set_err(FPX_E_ARG, "a completion flag needs a tile to launch");
  DeviceState *st = nullptr;
  rc = current_device_state(&st);
  if (rc) return rc;
  FpxSlot byval;
  std::memset(&byval, 0, sizeof(byval));
  if (!slots) byval = *one;
  void *fargs[] = {(void *)&slots, (void *)&byval, (void *)&n_slots};
  const void *finalize = reinterpret_cast<const void *>(&fpx::finalize_kernel);
  if (n_items == 0) {  // no pixels: chi2 and gradients are zero
    if (finalize_outputs)
      FPX_HIP(hipLaunchKernel(finalize, dim3(n_slots, fpx::kFinalizeChunks),
                              dim3(fpx::kFinalizeThreads), fargs, 0, (hipStream_t)stream));
    return FPX_OK;
  }
  const KernelCfg &kc = kConfigs[cfg];
  const int v = (split ? 2 : 0) + (slots ? 0 : 1);
  int cap = 0, index_cap = 0;
  size_t lds = 0;
  rc = launch_lds(st, cfg, v, max_gauss, &lds, &cap, &index_cap);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (st->lds_set[kVariants * cfg + v] < lds) {
      FPX_HIP(hipFuncSetAttribute(kc.fn[v], hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
      st->lds_set[kVariants * cfg + v] = lds;
    }
  }
  const int threads = kc.waves * 64;
  int grid = n_items;
  if (persistent) {
    int per_cu = 0;
    rc = groups_per_cu(st, cfg, v, lds, &per_cu);
    if (rc) return rc;
    grid = std::min<long>((long)per_cu * st->n_cu, (long)n_items);
  }
  const bool fused = finalize_outputs && n_slots == 1 && queue != nullptr;
  int mode = (persistent ? fpx::kModePersistent : 0) | (fused ? fpx::kModeOutputs : 0) |
             (flag ? fpx::kModeFlag : 0);
  void *args[] = {(void *)&slots,     (void *)&byval,      (void *)&n_slots,
                  (void *)&item_offsets, (void *)&n_items, (void *)&cap,
                  (void *)&index_cap, (void *)&mode,       (void *)&queue,
                  (void *)&flag,      (void *)&seq};
  FPX_HIP(hipLaunchKernel(kc.fn[v], dim3(grid), dim3(threads), args, lds, (hipStream_t)stream));
  if (finalize_outputs && !fused)
    FPX_HIP(hipLaunchKernel(finalize, dim3(n_slots, fpx::kFinalizeChunks),
                            dim3(fpx::kFinalizeThreads), fargs, 0, (hipStream_t)stream));
  return FPX_OK;
}

}  // namespace

namespace fpx {
int report_error(int code, const char *msg) { return set_err(code, msg); }
int report_hip_error(hipError_t e, const char *where) { return hip_err(e, where); }
int launch_one(const FpxSlot &slot, int32_t n_tiles, int32_t max_gauss, int32_t cfg,
               int32_t launch_flags, uint32_t *queue, void *stream, bool finalize_outputs,
               uint32_t *flag, uint32_t seq) {
  return launch(nullptr, &slot, 1, nullptr, n_tiles, max_gauss, cfg, launch_flags, queue, stream,
                finalize_outputs, flag, seq);
}
std::string status_text(uint32_t status) { return status_message(status); }
int config_pixels(int32_t cfg) {
  return (cfg >= 0 && cfg < kNumConfigs) ? 64 * kConfigs[cfg].ppl : 0;
}
}  // namespace fpx

extern "C" {

int fpx_abi_version(void) { return FPX_ABI_VERSION; }

const char *fpx_last_error(void) { return g_err.c_str(); }

size_t fpx_struct_size(int which) {
  switch (which) {
    case FPX_STRUCT_PATCH: return sizeof(FpxPatch);
    case FPX_STRUCT_SOURCE: return sizeof(FpxSource);
    case FPX_STRUCT_RESPONSE: return sizeof(FpxResponse);
    case FPX_STRUCT_PSFGAUSS: return sizeof(FpxPSFSourceGaussian);
    case FPX_STRUCT_TILE: return sizeof(FpxTile);
    case FPX_STRUCT_SLOT: return sizeof(FpxSlot);
    case FPX_STRUCT_AMPMODEL: return sizeof(FpxAmplitudeModel);
    default: return 0;
  }
}

size_t fpx_accum_doubles(int32_t n_bands) {
  return n_bands > 0 ? (size_t)n_bands * fpx::kAcc : 0;
}

int fpx_plan_tiles(const int32_t *exposure_start, const int32_t *exposure_N,
                   const int16_t *band_start, const int16_t *band_N, int32_t n_bands,
                   int32_t pix_per_tile, FpxTile *tiles_out, int32_t max_tiles) {
  if (!exposure_start || !exposure_N || !band_start || !band_N)
    return set_err(FPX_E_ARG, "null index array");
  if (n_bands <= 0 || n_bands > MAXBANDS) return set_err(FPX_E_ARG, "n_bands out of range [1, MAXBANDS]");
  if (pix_per_tile <= 0 || pix_per_tile % FPX_WAVE != 0)
    return set_err(FPX_E_ARG, "pix_per_tile must be a positive multiple of FPX_WAVE");
  long count = 0;
  for (int b = 0; b < n_bands; ++b) {
    if (band_N[b] < 0 || band_start[b] < 0) return set_err(FPX_E_ARG, "negative band index");
    for (int k = 0; k < band_N[b]; ++k) {
      const int e = band_start[b] + k;
      const int n = exposure_N[e];
      if (n < 0 || exposure_start[e] < 0) return set_err(FPX_E_ARG, "negative exposure size or start");
      for (int off = 0; off < n; off += pix_per_tile) {
        if (tiles_out) {
          if (count >= max_tiles) return set_err(FPX_E_ARG, "tiles_out too small");
          FpxTile &t = tiles_out[count];
          t.exposure = e;
          t.band = (int16_t)b;
          t.part = 0;
          t.parts = 0;
          t.pix_start = exposure_start[e] + off;
          t.pix_count = std::min(pix_per_tile, n - off);
        }
        ++count;
        if (count > INT32_MAX) return set_err(FPX_E_ARG, "too many tiles");
      }
    }
  }
  return (int)count;
}

int fpx_n_configs(void) { return kNumConfigs; }

size_t fpx_split_bytes(int32_t n_tiles, int32_t cfg) {
  if (n_tiles <= 0 || cfg < 0 || cfg >= kNumConfigs) return 0;
  return fpx::split_counter_bytes(n_tiles) +
         (size_t)n_tiles * 64 * kConfigs[cfg].ppl * sizeof(float);
}

int fpx_config_info(int32_t cfg, int32_t *waves_per_group, int32_t *pixels_per_lane,
                    int32_t *pixels_per_tile) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (waves_per_group) *waves_per_group = kConfigs[cfg].waves;
  if (pixels_per_lane) *pixels_per_lane = kConfigs[cfg].ppl;
  if (pixels_per_tile) *pixels_per_tile = 64 * kConfigs[cfg].ppl;
  return FPX_OK;
}

int fpx_lds_bytes(int32_t cfg, int32_t max_gauss, size_t *bytes) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (!bytes || max_gauss < 0) return set_err(FPX_E_ARG, "bad argument");
  const KernelCfg &kc = kConfigs[cfg];
  const int cap = pick_gauss_cap(kc, max_gauss, 160 * 1024, kStaticLds);
  *bytes = fpx::lds_bytes_for(kc.waves, kc.ppl, cap, max_gauss);
  return FPX_OK;
}

int fpx_resident_groups(int32_t cfg, int32_t max_gauss, int32_t *groups) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (!groups) return set_err(FPX_E_ARG, "null output");
  if (max_gauss <= 0 || max_gauss > MAXSOURCES * FPX_MAX_PSF_PER_SOURCE)
    return set_err(FPX_E_ARG, "max_gauss out of range [1, MAXSOURCES*FPX_MAX_PSF_PER_SOURCE]");
  DeviceState *st = nullptr;
  rc = current_device_state(&st);
  if (rc) return rc;
  int cap = 0, index_cap = 0, per_cu = 0;
  size_t lds = 0;
  rc = launch_lds(st, cfg, 0, max_gauss, &lds, &cap, &index_cap);
  if (rc) return rc;
  rc = groups_per_cu(st, cfg, 0, lds, &per_cu);
  if (rc) return rc;
  *groups = per_cu * st->n_cu;
  return FPX_OK;
}

int fpx_init(int32_t device) {
  int n = 0;
  FPX_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_err(FPX_E_NODEV, "no such device");
  FPX_HIP(hipSetDevice(device));
  DeviceState *st = nullptr;
  return current_device_state(&st);
}

int fpx_device_info(int32_t *n_cu, int32_t *lds_per_group, int32_t *clock_khz) {
  DeviceState *st = nullptr;
  int rc = current_device_state(&st);
  if (rc) return rc;
  if (n_cu) *n_cu = st->n_cu;
  if (lds_per_group) *lds_per_group = st->lds_max;
  if (clock_khz) *clock_khz = st->clock_khz;
  return FPX_OK;
}

int fpx_evaluate_proposal(const FpxSlot *slot, int32_t n_tiles, int32_t max_gauss,
                          int32_t cfg, void *stream) {
  return launch(slot, nullptr, 1, nullptr, n_tiles, max_gauss, cfg, 0, nullptr, stream);
}

int fpx_evaluate_sync(const FpxSlot *slot, int32_t n_tiles, int32_t max_gauss, int32_t cfg,
                      int32_t persistent, uint32_t *queue, void *proposal_dev,
                      const void *proposal_host, size_t proposal_bytes, const void *out_dev,
                      void *out_host, size_t out_bytes, uint32_t *status_dev, void *stream) {
  if (!proposal_dev || !proposal_host || !out_dev || !out_host)
    return set_err(FPX_E_ARG, "NULL proposal or output buffer");
  if (proposal_bytes > sizeof(FpxSource) * MAXSOURCES)
    return set_err(FPX_E_ARG, "proposal larger than MAXSOURCES records");
  if (out_bytes == 0) return set_err(FPX_E_ARG, "no output bytes to copy back");
  const hipStream_t s = (hipStream_t)stream;
  if (proposal_bytes > 0)
    FPX_HIP(hipMemcpyAsync(proposal_dev, proposal_host, proposal_bytes, hipMemcpyHostToDevice, s));
  const int rc = launch(slot, nullptr, 1, nullptr, n_tiles, max_gauss, cfg, persistent,
                        queue, stream);
  if (rc) return rc;
  FPX_HIP(hipMemcpyAsync(out_host, out_dev, out_bytes, hipMemcpyDeviceToHost, s));
  // the status word: from the copied outputs when it lies among them, else
  // copied on its own
  const uintptr_t o0 = (uintptr_t)out_dev, sp = (uintptr_t)status_dev;
  const bool inside = status_dev && sp >= o0 && sp + sizeof(uint32_t) <= o0 + out_bytes;
  uint32_t status = 0;
  if (status_dev && !inside)
    FPX_HIP(hipMemcpyAsync(&status, status_dev, sizeof(status), hipMemcpyDeviceToHost, s));
  FPX_HIP(hipStreamSynchronize(s));
  if (inside) std::memcpy(&status, (const unsigned char *)out_host + (sp - o0), sizeof(status));
  if (status != 0) {
    FPX_HIP(hipMemsetAsync(status_dev, 0, sizeof(uint32_t), s));
    FPX_HIP(hipStreamSynchronize(s));
    return set_err(FPX_E_ARG, status_message(status));
  }
  return FPX_OK;
}

int fpx_evaluate_batch(const FpxSlot *slots, int32_t n_slots, const int32_t *item_offsets,
                       int32_t n_items, int32_t max_gauss, int32_t cfg, int32_t persistent,
                       uint32_t *queue, void *stream) {
  return launch(slots, nullptr, n_slots, item_offsets, n_items, max_gauss, cfg,
                persistent, queue, stream);
}

int fpx_evaluate_tiles(const FpxSlot *slots, int32_t n_slots, const int32_t *item_offsets,
                       int32_t n_items, int32_t max_gauss, int32_t cfg, int32_t persistent,
                       uint32_t *queue, void *stream) {
  return launch(slots, nullptr, n_slots, item_offsets, n_items, max_gauss, cfg,
                persistent, queue, stream, false);
}

int fpx_finalize(const FpxSlot *slots, int32_t n_slots, void *stream) {
  if (!slots || n_slots <= 0) return set_err(FPX_E_ARG, "empty or invalid slot list");
  DeviceState *st = nullptr;
  int rc = current_device_state(&st);
  if (rc) return rc;
  FpxSlot none;
  std::memset(&none, 0, sizeof(none));
  void *fargs[] = {(void *)&slots, (void *)&none, (void *)&n_slots};
  FPX_HIP(hipLaunchKernel(reinterpret_cast<const void *>(&fpx::finalize_kernel),
                          dim3(n_slots, fpx::kFinalizeChunks), dim3(fpx::kFinalizeThreads), fargs,
                          0, (hipStream_t)stream));
  return FPX_OK;
}

int fpx_count_pairs(const FpxSlot *slots, int32_t n_slots, const int32_t *item_offsets,
                    int32_t n_items, int32_t max_gauss, unsigned long long *pairs_in,
                    uint32_t *item_pairs, void *stream) {
  if (!slots || n_slots <= 0 || n_items < 0 || !pairs_in)
    return set_err(FPX_E_ARG, "empty or invalid slot list");
  if (n_slots > 1 && !item_offsets) return set_err(FPX_E_ARG, "item_offsets required for n_slots > 1");
  if (max_gauss <= 0 || max_gauss > MAXSOURCES * FPX_MAX_PSF_PER_SOURCE)
    return set_err(FPX_E_ARG, "max_gauss out of range [1, MAXSOURCES*FPX_MAX_PSF_PER_SOURCE]");
  if (n_items == 0) return FPX_OK;
  DeviceState *st = nullptr;
  int rc = current_device_state(&st);
  if (rc) return rc;
  const size_t lds = (size_t)max_gauss * 5 * sizeof(float);
  if ((long)lds > st->lds_max) return set_err(FPX_E_ARG, "max_gauss too large for the LDS");
  const void *fn = reinterpret_cast<const void *>(&fpx::census_kernel);
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (st->census_lds < lds) {
      FPX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      st->census_lds = lds;
    }
  }
  int index_cap = max_gauss;
  void *args[] = {(void *)&slots,     (void *)&n_slots,  (void *)&item_offsets,
                  (void *)&n_items,   (void *)&index_cap, (void *)&pairs_in,
                  (void *)&item_pairs};
  FPX_HIP(hipLaunchKernel(fn, dim3(n_items), dim3(256), args, lds, (hipStream_t)stream));
  return FPX_OK;
}

int fpx_probe_valu(int32_t kind, int32_t iters, float *out, int32_t n_threads, double *ops,
                   void *stream) {
  if (!out || n_threads <= 0 || n_threads % 256 != 0 || iters <= 0)
    return set_err(FPX_E_ARG, "probe needs n_threads a positive multiple of 256 and iters > 0");
  DeviceState *st = nullptr;
  int rc = current_device_state(&st);
  if (rc) return rc;
  const dim3 grid(n_threads / 256), block(256);
  const hipStream_t s = (hipStream_t)stream;
  int chains = 8;
  switch (kind) {
    case 0: hipLaunchKernelGGL((fpx::probe_kernel<0, 8>), grid, block, 0, s, out, iters); break;
    case 1: hipLaunchKernelGGL((fpx::probe_kernel<1, 8>), grid, block, 0, s, out, iters); break;
    case 2:
      chains = 32;
      hipLaunchKernelGGL((fpx::probe_kernel<0, 32>), grid, block, 0, s, out, iters);
      break;
    case 3:
      chains = 16;
      hipLaunchKernelGGL((fpx::probe_kernel<1, 16>), grid, block, 0, s, out, iters);
      break;
    default:
      return set_err(FPX_E_ARG, "probe kind must be 0 (fma x8), 1 (exp x8), 2 (fma x32) or 3 (exp x16)");
  }
  FPX_HIP(hipGetLastError());
  if (ops) *ops = (double)n_threads * (double)iters * chains;
  return FPX_OK;
}

}  // extern "C"
