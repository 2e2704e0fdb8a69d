// Source records from parameter vectors, and gradient stacking, on the
// device.
//
// build_sources_kernel is Galaxy.proposal (forcepho/sources.py:607-621) for
// many proposals at once: a group of 16 lanes per (proposal, source) writes
// the FpxSource record, one lane per radius, with the mixture amplitudes and their sersic / rh
// derivatives evaluated in double:
//   * the log-normal weight model of the synthetic scenes, in closed form;
//   * forcepho's SplineMixtureAmplitudes (sources.py:657-681), per-radius
//     tensor-product B-splines given by scipy's (tx, ty, c, kx, ky): the
//     nonzero basis functions and their first derivatives come from de Boor's
//     recurrences (C. de Boor, A Practical Guide to Splines, 1978, BSPLVB and
//     BSPLVD), with the argument clamped to the knot range and the knot
//     interval clamped to [k, len - k - 2], as FITPACK's fpbisp and parder
//     evaluate.
// stack_kernel is Posterior.stack_grad (forcepho/model.py:370-388) over many
// proposals, plus the chi2 summed over bands.
//
// Both are small: a few hundred flops per source.  They exist so that a
// leapfrog step or an autograd pass is a handful of launches.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "forcepho_amd.h"
#include "fpx_internal.h"

namespace fpx {
namespace {

constexpr int kMaxDeg = FPX_MAX_SPLINE_DEGREE;

// Nonzero B-spline basis functions of degree k (<= kMaxDeg) at x on the
// knots t[0, len), and their first derivatives: b[a] and d[a] belong to
// basis function l - k + a.  Loops run to kMaxDeg with the degree as a
// predicate so that the small arrays stay in registers.
struct Basis {
  int l;
  double b[kMaxDeg + 1];
  double d[kMaxDeg + 1];
};

__device__ void bspline_basis(const double *__restrict__ t, int len, int k, double x, Basis &out) {
  // clamp into the knot range as fitpack does; comparisons rather than
  // fmin/fmax, which would turn a NaN into the first knot (NaN stays NaN and
  // gives NaN values, as scipy's splines do)
  x = x < t[k] ? t[k] : (x > t[len - k - 1] ? t[len - k - 1] : x);
  // the knot interval: the largest l <= len - k - 2 with t[l] <= x, at least k
  int l = k;
  for (int j = k + 1; j <= len - k - 2; ++j)
    if (t[j] <= x) l = j;
  double h[kMaxDeg + 1], lower[kMaxDeg + 1];
#pragma unroll
  for (int a = 0; a <= kMaxDeg; ++a) {
    h[a] = a == 0 ? 1.0 : 0.0;
    lower[a] = 0.0;
  }
#pragma unroll
  for (int j = 1; j <= kMaxDeg; ++j) {
    if (j > k) break;
    double prev[kMaxDeg + 1];
#pragma unroll
    for (int a = 0; a <= kMaxDeg; ++a) {
      prev[a] = h[a];
      if (j == k) lower[a] = h[a];  // degree k - 1, for the derivative
      h[a] = 0.0;
    }
#pragma unroll
    for (int n = 1; n <= kMaxDeg; ++n) {
      if (n > j) break;
      const double xb = t[l + n], xa = t[l + n - j];
      const double span = xb - xa;
      const double w = span != 0.0 ? prev[n - 1] / span : 0.0;
      h[n - 1] += w * (xb - x);
      h[n] = w * (x - xa);
    }
  }
  double d[kMaxDeg + 1];
#pragma unroll
  for (int a = 0; a <= kMaxDeg; ++a) d[a] = 0.0;
#pragma unroll
  for (int n = 1; n <= kMaxDeg; ++n) {
    if (n > k) break;
    const double xb = t[l + n], xa = t[l + n - k];
    const double span = xb - xa;
    const double w = span != 0.0 ? k * lower[n - 1] / span : 0.0;
    d[n - 1] -= w;
    d[n] = w;
  }
  out.l = l;
#pragma unroll
  for (int a = 0; a <= kMaxDeg; ++a) {
    out.b[a] = h[a];
    out.d[a] = d[a];
  }
}

// Amplitude, d/dsersic and d/drh of radius i of a spline model at (n, rh).
__device__ void spline_amplitude(const FpxAmplitudeModel &m, int i, double n, double rh,
                                 double &A, double &dA_dn, double &dA_drh) {
  Basis bx, by;
  bspline_basis(m.tx + (size_t)i * m.tx_stride, m.tx_len[i], m.kx, n, bx);
  bspline_basis(m.ty + (size_t)i * m.ty_stride, m.ty_len[i], m.ky, rh, by);
  const double *c = m.coef + (size_t)i * m.cx * m.cy;
  A = dA_dn = dA_drh = 0.0;
#pragma unroll
  for (int a = 0; a <= kMaxDeg; ++a) {
    if (a > m.kx) break;
    const double *row = c + (size_t)(bx.l - m.kx + a) * m.cy + (by.l - m.ky);
    double s = 0.0, ds = 0.0;
#pragma unroll
    for (int b = 0; b <= kMaxDeg; ++b) {
      if (b > m.ky) break;
      s += row[b] * by.b[b];
      ds += row[b] * by.d[b];
    }
    A += bx.b[a] * s;
    dA_dn += bx.d[a] * s;
    dA_drh += bx.b[a] * ds;
  }
}

// A record is written by a group of kGroup lanes: lane r evaluates radius r
// (r < MAXRADII) and the group's sums and minimum over radii are butterfly
// shuffles, so the latency of the double exp / log / divides is paid once per
// record instead of once per radius.
constexpr int kGroup = 16;
static_assert(kGroup >= MAXRADII && 64 % kGroup == 0, "a record's radii fit one lane group");

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int m = kGroup / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, kGroup);
  return v;
}

__device__ __forceinline__ double group_min(double v) {
#pragma unroll
  for (int m = kGroup / 2; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m, kGroup));
  return v;
}

// 1 / (1 + exp(-x)): the logistic function of forcepho's BoundedTransform
// (forcepho/model.py:596-637).
__device__ __forceinline__ double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

__global__ __launch_bounds__(64) void build_sources_kernel(
    const double *__restrict__ theta, const double *__restrict__ lower,
    const double *__restrict__ upper, int n_prop, int S, int NB,
    const int32_t *__restrict__ n_sources, const int32_t *__restrict__ n_bands,
    const FpxAmplitudeModel model, FpxSource *__restrict__ records) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int idx = gid / kGroup;  // (proposal, source); the same for the whole group
  const int r = gid - idx * kGroup;
  if (idx >= n_prop * S) return;
  const int k = idx / S;
  const int s = idx - k * S;
  const int K = NB + 6;
  const double *th = theta + (size_t)idx * K;
  const double *lo = lower ? lower + (size_t)idx * K : nullptr;
  const double *hi = upper ? upper + (size_t)idx * K : nullptr;
  // parameter j of the source: theta itself, or lower + sigmoid(z) (upper -
  // lower) when theta holds the unconstrained z
  auto param = [&](int j) -> double {
    return lo ? lo[j] + sigmoid(th[j]) * (hi[j] - lo[j]) : th[j];
  };
  FpxSource *rec = records + (size_t)k * MAXSOURCES + s;
  float *dst = reinterpret_cast<float *>(rec);
  constexpr int kFloats = (int)(sizeof(FpxSource) / sizeof(float));
  const int ns = n_sources ? n_sources[k] : S;
  const int nb = n_bands ? n_bands[k] : NB;
  if (s >= ns) {  // past the proposal's sources: an all-zero record
    for (int j = r; j < kFloats; j += kGroup) dst[j] = 0.0f;
    return;
  }
  // ra dec q pa nsersic rh, then the band fluxes
  for (int j = r; j < 6 + MAXBANDS; j += kGroup) {
    const int b = j - 6;
    dst[j] = j < 6 ? (float)param(NB + j) : ((b < nb && b < NB) ? (float)param(b) : 0.0f);
  }
  const double n = param(NB + 4), rh = param(NB + 5);
  double A = 0.0, dA_dn = 0.0, dA_drh = 0.0;
  if (model.kind == FPX_AMP_SPLINE) {
    if (r < model.n_radii) spline_amplitude(model, r, n, rh, A, dA_dn, dA_drh);
  } else {
    const int R = min(model.lnr_count[k], MAXRADII);
    const bool valid = r < R;
    const double sig = 0.6 + 0.25 * n;
    const double mu = log(rh) - 0.15 * (n - 1.0);
    // lanes past the radii read radius MAXRADII - 1 (in bounds) and discard it
    const double lnr = model.lnr[(size_t)k * MAXRADII + min(r, MAXRADII - 1)];
    const double z = valid ? (lnr - mu) / sig : 0.0;
    const double z2min = group_min(valid ? z * z : __builtin_inf());
    const double w = valid ? exp(-0.5 * (z * z - z2min)) : 0.0;
    A = w / group_sum(w);
    const double l_rh = z / (sig * rh);
    const double l_n = z / sig * (-0.15) + z * z * 0.25 / sig;
    dA_drh = A * (l_rh - group_sum(A * l_rh));
    dA_dn = A * (l_n - group_sum(A * l_n));
  }
  if (r < MAXRADII) {
    rec->mixture_amplitudes[r] = (float)A;
    rec->damplitude_drh[r] = (float)dA_drh;
    rec->damplitude_dnsersic[r] = (float)dA_dn;
  }
}

__global__ __launch_bounds__(64) void stack_kernel(const float *__restrict__ resp,
                                                   const double *__restrict__ chi2, int n_prop,
                                                   int S, int NB, double *__restrict__ grad,
                                                   double *__restrict__ chi2_sum) {
  const int K = NB + 6;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (grad && idx < n_prop * S * K) {
    const int k = idx / (S * K);
    const int rem = idx - k * S * K;
    const int s = rem / K;
    const int j = rem - s * K;
    const float *r = resp + (size_t)k * MAXBANDS * MAXSOURCES * NPARAMS + s * NPARAMS;
    double v;
    if (j < NB) {
      v = (double)r[j * MAXSOURCES * NPARAMS];
    } else {
      const int p = j - NB + 1;
      v = 0.0;
      for (int b = 0; b < NB; ++b) v += (double)r[b * MAXSOURCES * NPARAMS + p];
    }
    grad[idx] = v;
  }
  if (chi2_sum && idx < n_prop) {
    double v = 0.0;
    for (int b = 0; b < NB; ++b) v += chi2[(size_t)idx * MAXBANDS + b];
    chi2_sum[idx] = v;
  }
}

// U(z), dU/dz and a leapfrog kick / drift of every chain (one workgroup per
// chain).  With theta = lower + sigmoid(z) (upper - lower) (forcepho's
// BoundedTransform, model.py:596-637):
//   U    = chi2 / 2 - sum mask ln(dtheta/dz)
//   dU/dz = mask (dchi2/dtheta dtheta/dz / 2 - (1 - 2 sigmoid(z)))
// the negative log posterior of FastPosterior.evaluate (model.py:302-340,
// -lnp with the Jacobian term) in z, dchi2/dtheta stacked from the per-band
// responses as stack_kernel does.  Then, if p is given, p -= kick eps dU/dz
// and, with drift, z += eps p mask.
constexpr int kStepThreads = 256;

__global__ __launch_bounds__(kStepThreads) void leapfrog_kernel(
    const double *__restrict__ lower, const double *__restrict__ upper,
    const double *__restrict__ mask, const double *__restrict__ eps,
    const float *__restrict__ resp, const double *__restrict__ chi2, int S, int NB, double kick,
    int drift, double *__restrict__ z, double *__restrict__ p, double *__restrict__ U,
    double *__restrict__ g) {
  __shared__ double part[kStepThreads / 64];
  const int k = blockIdx.x;
  const int K = NB + 6;
  const int n = S * K;
  const size_t base = (size_t)k * n;
  const float *r = resp + (size_t)k * MAXBANDS * MAXSOURCES * NPARAMS;
  const double e = eps[k];
  double lndet = 0.0;
  for (int i = threadIdx.x; i < n; i += kStepThreads) {
    const int s = i / K;
    const int j = i - s * K;
    const double zi = z[base + i];
    const double m = mask[base + i];
    const double sz = sigmoid(zi);
    const double jac = (upper[base + i] - lower[base + i]) * sz * (1.0 - sz);
    double gt;
    if (j < NB) {
      gt = (double)r[(j * MAXSOURCES + s) * NPARAMS];
    } else {
      gt = 0.0;
      for (int b = 0; b < NB; ++b) gt += (double)r[(b * MAXSOURCES + s) * NPARAMS + (j - NB + 1)];
    }
    const double gi = (0.5 * gt * jac - (1.0 - 2.0 * sz)) * m;
    g[base + i] = gi;
    if (m != 0.0) lndet += log(jac) * m;
    if (p) {
      const double pi = p[base + i] - kick * e * gi;
      p[base + i] = pi;
      if (drift) z[base + i] = zi + e * pi * m;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lndet += __shfl_xor(lndet, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = lndet;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0, c = 0.0;
    for (int w = 0; w < kStepThreads / 64; ++w) tot += part[w];
    for (int b = 0; b < NB; ++b) c += chi2[(size_t)k * MAXBANDS + b];
    U[k] = 0.5 * c - tot;
  }
}

// Metropolis test of one HMC transition per chain (one workgroup per chain):
// H = U + |p|^2 / 2 at the start (U0, p0) and at the end (U1, p1) of the
// trajectory; accept when ln u < H0 - H1 (a NaN energy rejects); a rejected
// chain's position goes back to z0.
__global__ __launch_bounds__(kStepThreads) void metropolis_kernel(
    const double *__restrict__ U0, const double *__restrict__ U1,
    const double *__restrict__ p0, const double *__restrict__ p1,
    const double *__restrict__ u, const double *__restrict__ z0, int n, double *__restrict__ z,
    int32_t *__restrict__ accept) {
  __shared__ double part[2][kStepThreads / 64];
  __shared__ int ok;
  const int k = blockIdx.x;
  const size_t base = (size_t)k * n;
  double k0 = 0.0, k1 = 0.0;
  for (int i = threadIdx.x; i < n; i += kStepThreads) {
    k0 += p0[base + i] * p0[base + i];
    k1 += p1[base + i] * p1[base + i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    k0 += __shfl_xor(k0, off, 64);
    k1 += __shfl_xor(k1, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = k0;
    part[1][threadIdx.x >> 6] = k1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < kStepThreads / 64; ++w) {
      s0 += part[0][w];
      s1 += part[1][w];
    }
    const double H0 = U0[k] + 0.5 * s0;
    const double H1 = U1[k] + 0.5 * s1;
    ok = log(u[k]) < H0 - H1 ? 1 : 0;  // false for NaN
    accept[k] = ok;
  }
  __syncthreads();
  if (!ok)
    for (int i = threadIdx.x; i < n; i += kStepThreads) z[base + i] = z0[base + i];
}

int check_launch(const char *where) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? FPX_OK : report_hip_error(e, where);
}

}  // namespace
}  // namespace fpx

extern "C" {

int fpx_build_sources(const double *theta, const double *lower, const double *upper,
                      int32_t n_prop, int32_t n_src_max, int32_t n_bands_max,
                      const int32_t *n_sources, const int32_t *n_bands,
                      const FpxAmplitudeModel *model, FpxSource *records, void *stream) {
  using fpx::report_error;
  if (n_prop < 0) return report_error(FPX_E_ARG, "n_prop < 0");
  if (n_src_max < 1 || n_src_max > MAXSOURCES)
    return report_error(FPX_E_ARG, "n_src_max out of range [1, MAXSOURCES]");
  if (n_bands_max < 1 || n_bands_max > MAXBANDS)
    return report_error(FPX_E_ARG, "n_bands_max out of range [1, MAXBANDS]");
  if (!model) return report_error(FPX_E_ARG, "no amplitude model");
  if ((lower == nullptr) != (upper == nullptr))
    return report_error(FPX_E_ARG, "lower and upper bounds come together");
  const FpxAmplitudeModel m = *model;
  if (m.kind == FPX_AMP_LOGNORMAL) {
    if (!m.lnr || !m.lnr_count) return report_error(FPX_E_ARG, "log-normal model without radii");
  } else if (m.kind == FPX_AMP_SPLINE) {
    if (m.n_radii < 1 || m.n_radii > MAXRADII)
      return report_error(FPX_E_ARG, "spline n_radii out of range [1, MAXRADII]");
    if (m.kx < 0 || m.ky < 0 || m.kx > FPX_MAX_SPLINE_DEGREE || m.ky > FPX_MAX_SPLINE_DEGREE)
      return report_error(FPX_E_ARG, "spline degree out of range [0, FPX_MAX_SPLINE_DEGREE]");
    if (!m.tx || !m.ty || !m.tx_len || !m.ty_len || !m.coef)
      return report_error(FPX_E_ARG, "spline model with a NULL table");
    if (m.tx_stride < 2 * m.kx + 2 || m.ty_stride < 2 * m.ky + 2 || m.cx < m.kx + 1 ||
        m.cy < m.ky + 1)
      return report_error(FPX_E_ARG, "spline tables smaller than their degrees allow");
  } else {
    return report_error(FPX_E_ARG, "unknown amplitude model kind");
  }
  if (n_prop == 0) return FPX_OK;
  if (!theta || !records) return report_error(FPX_E_ARG, "NULL theta or records");
  const long threads = (long)n_prop * n_src_max * fpx::kGroup;
  if (threads > (1L << 30)) return report_error(FPX_E_ARG, "too many records for one launch");
  const int grid = (int)((threads + 63) / 64);
  hipLaunchKernelGGL(fpx::build_sources_kernel, dim3(grid), dim3(64), 0, (hipStream_t)stream,
                     theta, lower, upper, n_prop, n_src_max, n_bands_max, n_sources, n_bands, m,
                     records);
  return fpx::check_launch("build_sources_kernel");
}

int fpx_stack_gradients(const FpxResponse *resp, const double *chi2, int32_t n_prop,
                        int32_t n_src_max, int32_t n_bands_max, double *grad, double *chi2_sum,
                        void *stream) {
  using fpx::report_error;
  if (n_prop < 0) return report_error(FPX_E_ARG, "n_prop < 0");
  if (n_src_max < 1 || n_src_max > MAXSOURCES)
    return report_error(FPX_E_ARG, "n_src_max out of range [1, MAXSOURCES]");
  if (n_bands_max < 1 || n_bands_max > MAXBANDS)
    return report_error(FPX_E_ARG, "n_bands_max out of range [1, MAXBANDS]");
  if ((grad && !resp) || (chi2_sum && !chi2))
    return report_error(FPX_E_ARG, "an output without its input");
  if (n_prop == 0 || (!grad && !chi2_sum)) return FPX_OK;
  const long threads = std::max<long>(grad ? (long)n_prop * n_src_max * (n_bands_max + 6) : 0,
                                      chi2_sum ? (long)n_prop : 0);
  if (threads > (1L << 30)) return report_error(FPX_E_ARG, "too many values for one launch");
  const int grid = (int)((threads + 63) / 64);
  hipLaunchKernelGGL(fpx::stack_kernel, dim3(grid), dim3(64), 0, (hipStream_t)stream,
                     reinterpret_cast<const float *>(resp), chi2, n_prop, n_src_max, n_bands_max,
                     grad, chi2_sum);
  return fpx::check_launch("stack_kernel");
}

int fpx_leapfrog_update(const double *lower, const double *upper, const double *mask,
                        const double *eps, const FpxResponse *resp, const double *chi2,
                        int32_t n_prop, int32_t n_src_max, int32_t n_bands_max, double kick,
                        int32_t drift, double *z, double *p, double *U, double *g,
                        void *stream) {
  using fpx::report_error;
  if (n_prop < 0) return report_error(FPX_E_ARG, "n_prop < 0");
  if (n_src_max < 1 || n_src_max > MAXSOURCES)
    return report_error(FPX_E_ARG, "n_src_max out of range [1, MAXSOURCES]");
  if (n_bands_max < 1 || n_bands_max > MAXBANDS)
    return report_error(FPX_E_ARG, "n_bands_max out of range [1, MAXBANDS]");
  if (!lower || !upper || !mask || !eps || !resp || !chi2 || !z || !U || !g)
    return report_error(FPX_E_ARG, "NULL array");
  if (drift && !p) return report_error(FPX_E_ARG, "drift needs the momenta");
  if (n_prop == 0) return FPX_OK;
  hipLaunchKernelGGL(fpx::leapfrog_kernel, dim3(n_prop), dim3(fpx::kStepThreads), 0,
                     (hipStream_t)stream, lower, upper, mask, eps,
                     reinterpret_cast<const float *>(resp), chi2, n_src_max, n_bands_max, kick,
                     drift ? 1 : 0, z, p, U, g);
  return fpx::check_launch("leapfrog_kernel");
}

int fpx_hmc_metropolis(const double *U0, const double *U1, const double *p0,
                       const double *p1, const double *u, const double *z0, int32_t n_prop,
                       int32_t n_per, double *z, int32_t *accept, void *stream) {
  using fpx::report_error;
  if (n_prop < 0 || n_per < 0) return report_error(FPX_E_ARG, "negative size");
  if (!U0 || !U1 || !p0 || !p1 || !u || !z0 || !z || !accept)
    return report_error(FPX_E_ARG, "NULL array");
  if (n_prop == 0) return FPX_OK;
  hipLaunchKernelGGL(fpx::metropolis_kernel, dim3(n_prop), dim3(fpx::kStepThreads), 0,
                     (hipStream_t)stream, U0, U1, p0, p1, u, z0, n_per, z, accept);
  return fpx::check_launch("metropolis_kernel");
}

}  // extern "C"
