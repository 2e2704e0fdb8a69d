// Source records from parameter vectors, and gradient stacking, on the
// device.
//
// build_sources_kernel is Galaxy.proposal (forcepho/sources.py:607-621) for
// many proposals at once: one thread per (proposal, source) writes the
// FpxSource record, with the mixture amplitudes and their sersic / rh
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
  x = fmin(fmax(x, t[k]), t[len - k - 1]);
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

__global__ __launch_bounds__(64) void build_sources_kernel(
    const double *__restrict__ theta, int n_prop, int S, int NB,
    const int32_t *__restrict__ n_sources, const int32_t *__restrict__ n_bands,
    const FpxAmplitudeModel model, FpxSource *__restrict__ records) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_prop * S) return;
  const int k = idx / S;
  const int s = idx - k * S;
  const int K = NB + 6;
  const double *th = theta + (size_t)idx * K;
  FpxSource *rec = records + (size_t)k * MAXSOURCES + s;
  const int ns = n_sources ? n_sources[k] : S;
  const int nb = n_bands ? n_bands[k] : NB;
  if (s >= ns) {  // past the proposal's sources: an all-zero record
    float *dst = reinterpret_cast<float *>(rec);
    for (int j = 0; j < (int)(sizeof(FpxSource) / sizeof(float)); ++j) dst[j] = 0.0f;
    return;
  }
  rec->ra = (float)th[NB];
  rec->dec = (float)th[NB + 1];
  rec->q = (float)th[NB + 2];
  rec->pa = (float)th[NB + 3];
  rec->nsersic = (float)th[NB + 4];
  rec->rh = (float)th[NB + 5];
  for (int b = 0; b < MAXBANDS; ++b) rec->fluxes[b] = (b < nb && b < NB) ? (float)th[b] : 0.0f;
  const double n = th[NB + 4], rh = th[NB + 5];
  if (model.kind == FPX_AMP_SPLINE) {
    for (int i = 0; i < MAXRADII; ++i) {
      double A = 0.0, dA_dn = 0.0, dA_drh = 0.0;
      if (i < model.n_radii) spline_amplitude(model, i, n, rh, A, dA_dn, dA_drh);
      rec->mixture_amplitudes[i] = (float)A;
      rec->damplitude_drh[i] = (float)dA_drh;
      rec->damplitude_dnsersic[i] = (float)dA_dn;
    }
    return;
  }
  const int R = min(model.lnr_count[k], MAXRADII);
  const double *lnr = model.lnr + (size_t)k * MAXRADII;
  const double sig = 0.6 + 0.25 * n;
  const double mu = log(rh) - 0.15 * (n - 1.0);
  double z[MAXRADII];
  double z2min = __builtin_inf();
#pragma unroll
  for (int i = 0; i < MAXRADII; ++i) {
    z[i] = i < R ? (lnr[i] - mu) / sig : 0.0;
    if (i < R) z2min = fmin(z2min, z[i] * z[i]);
  }
  double w[MAXRADII], wsum = 0.0;
#pragma unroll
  for (int i = 0; i < MAXRADII; ++i) {
    w[i] = i < R ? exp(-0.5 * (z[i] * z[i] - z2min)) : 0.0;
    wsum += w[i];
  }
  double A[MAXRADII], l_rh[MAXRADII], l_n[MAXRADII], s_rh = 0.0, s_n = 0.0;
#pragma unroll
  for (int i = 0; i < MAXRADII; ++i) {
    A[i] = w[i] / wsum;
    l_rh[i] = z[i] / (sig * rh);
    l_n[i] = z[i] / sig * (-0.15) + z[i] * z[i] * 0.25 / sig;
    s_rh += A[i] * l_rh[i];
    s_n += A[i] * l_n[i];
  }
#pragma unroll
  for (int i = 0; i < MAXRADII; ++i) {
    rec->mixture_amplitudes[i] = (float)A[i];
    rec->damplitude_drh[i] = (float)(A[i] * (l_rh[i] - s_rh));
    rec->damplitude_dnsersic[i] = (float)(A[i] * (l_n[i] - s_n));
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

int check_launch(const char *where) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? FPX_OK : report_hip_error(e, where);
}

}  // namespace
}  // namespace fpx

extern "C" {

int fpx_build_sources(const double *theta, int32_t n_prop, int32_t n_src_max,
                      int32_t n_bands_max, const int32_t *n_sources, const int32_t *n_bands,
                      const FpxAmplitudeModel *model, FpxSource *records, void *stream) {
  using fpx::report_error;
  if (n_prop < 0) return report_error(FPX_E_ARG, "n_prop < 0");
  if (n_src_max < 1 || n_src_max > MAXSOURCES)
    return report_error(FPX_E_ARG, "n_src_max out of range [1, MAXSOURCES]");
  if (n_bands_max < 1 || n_bands_max > MAXBANDS)
    return report_error(FPX_E_ARG, "n_bands_max out of range [1, MAXBANDS]");
  if (!model) return report_error(FPX_E_ARG, "no amplitude model");
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
  const long threads = (long)n_prop * n_src_max;
  if (threads > (1L << 30)) return report_error(FPX_E_ARG, "too many records for one launch");
  const int grid = (int)((threads + 63) / 64);
  hipLaunchKernelGGL(fpx::build_sources_kernel, dim3(grid), dim3(64), 0, (hipStream_t)stream,
                     theta, n_prop, n_src_max, n_bands_max, n_sources, n_bands, m, records);
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

}  // extern "C"
