// fpx_gauss.h -- device-side construction of one ImageGaussian record: the
// convolution of one source mixture Gaussian with one PSF Gaussian of one
// exposure, plus the image->sky Jacobian folded into the coefficients the
// pixel passes use.
//
// Mathematics follows forcepho/src/compute_gaussians_kernel.cc:45-114
// (ConvolveOneSource) and forcepho/src/common_kernel.cc:175-240
// (GetGaussianAndJacobian):
//   T = D R(pa) S(q),  dT/dq = D R dS/dq,  dT/dpa = D dR/dpa S
//   Sigma = rad2[bin] T T^T + C_psf,  F = Sigma^-1
//   amp = G a_psf sqrt(det F) / (2 pi) * flux_band * a_mix[bin]
//   dF/dq = -F dSigma/dq F with dSigma/dq = rad2 (T dT^T + dT T^T)
//   dA/dq = amp/2 tr(Sigma dF/dq)   (same for pa)
// The record is laid out for the two pixel passes (fpx_kernel.hip), 24
// floats (6 float4):
//   [0..3]   xc yc hx hy      centre and the half-widths of the box that holds
//                             the whole dx^T F dx <= MAX_EXP_ARG ellipse
//   [4..8]   fxx fxy fyy h0 amp                       model pass
//   [9..22]  amp/12 dA/dflux dA/dsersic dA/drh
//            dA/dq kq dFxx/dq dFyy/dq dFxy/dq
//            dA/dpa kp dFxx/dpa dFyy/dpa dFxy/dpa    gradient pass
//   [23]     source index (int bits)
// with h0 = 1 - (fxx + fyy)/24 and kq = -(dFxx/dq + dFyy/dq)/2 (kp
// likewise), so that the pixel integral factor
// H = 1 + (vx^2 + vy^2 - fxx - fyy)/24 of common_kernel.cc:77 is
// H = h0 + (vx^2 + vy^2)/24.  The ellipse dx^T Sigma^-1 dx <= 36 spans
// +-6 sqrt(Sigma_xx) in x (sqrt(Sigma_yy) in y); hx, hy add a relative
// 1e-3 and an absolute 0.01 px margin so that no pixel the float32 test
// would keep lies outside the box.
#pragma once

#include <hip/hip_runtime.h>
#include "forcepho_amd.h"

namespace fpx {

constexpr int kRec = 24;  // floats per ImageGaussian record

enum RecField {
  kXc = 0, kYc, kHx, kHy,
  kFxx, kFxy, kFyy, kH0, kAmp,
  kAmp12, kDAF, kDAN, kDAR,
  kDAQ, kKQ, kQxx, kQyy, kQxy,
  kDAP, kKP, kPxx, kPyy, kPxy,
  kSrc
};

struct M22 {
  float a, b, c, d;  // [[a, b], [c, d]]
};

__device__ __forceinline__ M22 mat_mul(const M22 &x, const M22 &y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d,
          x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}

// -F dS F for symmetric F = [[fxx, fxy], [fxy, fyy]] and symmetric
// dS = [[sxx, sxy], [sxy, syy]]; returns (xx, yy, xy).
__device__ __forceinline__ void neg_sandwich(float fxx, float fxy, float fyy,
                                             float sxx, float sxy, float syy,
                                             float &oxx, float &oyy, float &oxy) {
  const float m11 = fxx * sxx + fxy * sxy;  // (F dS)
  const float m12 = fxx * sxy + fxy * syy;
  const float m21 = fxy * sxx + fyy * sxy;
  const float m22 = fxy * sxy + fyy * syy;
  oxx = -(m11 * fxx + m12 * fxy);
  oxy = -(m11 * fxy + m12 * fyy);
  oyy = -(m21 * fxy + m22 * fyy);
}

// Per (exposure, source) geometry: T = D R(pa) S(q) and its q / pa
// derivatives, the source centre in pixels CW (sky - crval) + crpix.
struct SourceGeometry {
  M22 T, dTq, dTp;
  float xs, ys;
};

template <bool kDerivs>
__device__ __forceinline__ void source_geometry(const FpxPatch *__restrict__ P,
                                                const FpxSource *__restrict__ src, int e,
                                                int s, SourceGeometry &g) {
  const int base = P->n_sources * e + s;
  const float *Dp = P->D + 4 * base;
  const float *CWp = P->CW + 4 * base;
  const float *crpix = P->crpix + 2 * base;
  const float *crval = P->crval + 2 * base;
  const float q = src->q;
  float sn, cs;
  sincosf(src->pa, &sn, &cs);
  const M22 D = {Dp[0], Dp[1], Dp[2], Dp[3]};
  const M22 R = {cs, -sn, sn, cs};
  const float iq = 1.0f / q;
  const M22 S = {iq, 0.f, 0.f, q};
  const M22 DR = mat_mul(D, R);
  g.T = mat_mul(DR, S);
  if (kDerivs) {
    const M22 dR = {-sn, -cs, cs, -sn};
    const M22 dS = {-1.0f / (q * q), 0.f, 0.f, 1.0f};
    g.dTq = mat_mul(DR, dS);
    g.dTp = mat_mul(mat_mul(D, dR), S);
  }
  const float sx = src->ra - crval[0];
  const float sy = src->dec - crval[1];
  g.xs = CWp[0] * sx + CWp[1] * sy + crpix[0];
  g.ys = CWp[2] * sx + CWp[3] * sy + crpix[1];
}

// Sigma = covar T T^T + C_psf
__device__ __forceinline__ void convolved_sigma(const M22 &T, float covar,
                                                const FpxPSFSourceGaussian &psf, float &Sxx,
                                                float &Sxy, float &Syy) {
  Sxx = covar * (T.a * T.a + T.b * T.b) + psf.Cxx;
  Sxy = covar * (T.a * T.c + T.b * T.d) + psf.Cxy;
  Syy = covar * (T.c * T.c + T.d * T.d) + psf.Cyy;
}

// Half-width of the box holding dx^T Sigma^-1 dx <= MAX_EXP_ARG along an
// axis of variance S, with a relative 1e-3 and an absolute 0.01 px margin.
__device__ __forceinline__ float cutoff_halfwidth(float S) {
  return sqrtf((float)MAX_EXP_ARG) * sqrtf(fmaxf(S, 0.0f)) * 1.001f + 0.01f;
}

// True when the MAX_EXP_ARG box of Gaussian (s, p) of exposure e can reach
// a pixel of [x0, x1] x [y0, y1].
__device__ __forceinline__ bool gauss_reaches(const FpxPatch *__restrict__ P,
                                              const FpxSource *__restrict__ props, int e,
                                              int s, int p, float x0, float x1, float y0,
                                              float y1) {
  SourceGeometry g;
  source_geometry<false>(P, props + s, e, s, g);
  const FpxPSFSourceGaussian psf = P->psfgauss[P->psfgauss_start[e] + p];
  float Sxx, Sxy, Syy;
  convolved_sigma(g.T, P->rad2[psf.sersic_radius_bin], psf, Sxx, Sxy, Syy);
  const float xc = g.xs + psf.xcen;
  const float yc = g.ys + psf.ycen;
  const float gx = fmaxf(fmaxf(x0 - xc, xc - x1), 0.0f);
  const float gy = fmaxf(fmaxf(y0 - yc, yc - y1), 0.0f);
  return gx <= cutoff_halfwidth(Sxx) && gy <= cutoff_halfwidth(Syy);
}

// Build the record of source s, PSF component p (0 <= p < n_psf_per_source[band])
// for exposure e of band `band`.  PSF lookup ignores the source index, as
// the reference does (compute_gaussians_kernel.cc:100, .cu:77).  The pad
// slot holds the source index (as int bits).
__device__ __forceinline__ void build_record(const FpxPatch *__restrict__ P,
                                             const FpxSource *__restrict__ props,
                                             int e, int band, int s, int p,
                                             float *__restrict__ out) {
  const FpxSource *src = props + s;
  SourceGeometry g;
  source_geometry<true>(P, src, e, s, g);
  const M22 &T = g.T, &dTq = g.dTq, &dTp = g.dTp;
  const float gcal = P->G[e];
  const FpxPSFSourceGaussian psf = P->psfgauss[P->psfgauss_start[e] + p];
  const int rb = psf.sersic_radius_bin;
  const float covar = P->rad2[rb];
  const float amp_s = src->mixture_amplitudes[rb];
  const float da_dn = src->damplitude_dnsersic[rb];
  const float da_dr = src->damplitude_drh[rb];
  const float flux = src->fluxes[band];

  float Sxx, Sxy, Syy;
  convolved_sigma(T, covar, psf, Sxx, Sxy, Syy);
  const float idet = 1.0f / (Sxx * Syy - Sxy * Sxy);
  const float fxx = Syy * idet;
  const float fxy = -Sxy * idet;
  const float fyy = Sxx * idet;
  const float detF = fxx * fyy - fxy * fxy;
  const float tmp = gcal * psf.amp * sqrtf(detF) * 0.15915494309189535f;  // 1/(2 pi)
  const float amp = tmp * flux * amp_s;

  // dSigma/dq and dSigma/dpa = covar (T dT^T + dT T^T)
  const float qsxx = covar * 2.0f * (T.a * dTq.a + T.b * dTq.b);
  const float qsxy = covar * (T.a * dTq.c + T.b * dTq.d + T.c * dTq.a + T.d * dTq.b);
  const float qsyy = covar * 2.0f * (T.c * dTq.c + T.d * dTq.d);
  const float psxx = covar * 2.0f * (T.a * dTp.a + T.b * dTp.b);
  const float psxy = covar * (T.a * dTp.c + T.b * dTp.d + T.c * dTp.a + T.d * dTp.b);
  const float psyy = covar * 2.0f * (T.c * dTp.c + T.d * dTp.d);
  float qxx, qyy, qxy, pxx, pyy, pxy;
  neg_sandwich(fxx, fxy, fyy, qsxx, qsxy, qsyy, qxx, qyy, qxy);
  neg_sandwich(fxx, fxy, fyy, psxx, psxy, psyy, pxx, pyy, pxy);
  const float dAQ = 0.5f * amp * (Sxx * qxx + 2.0f * Sxy * qxy + Syy * qyy);
  const float dAP = 0.5f * amp * (Sxx * pxx + 2.0f * Sxy * pxy + Syy * pyy);

  const float h0 = 1.0f - (fxx + fyy) * (1.0f / 24.0f);
  out[kXc] = g.xs + psf.xcen;
  out[kYc] = g.ys + psf.ycen;
  out[kHx] = cutoff_halfwidth(Sxx);
  out[kHy] = cutoff_halfwidth(Syy);
  out[kFxx] = fxx;
  out[kFxy] = fxy;
  out[kFyy] = fyy;
  out[kH0] = h0;
  out[kAmp] = amp;
  out[kAmp12] = amp * (1.0f / 12.0f);
  out[kDAF] = tmp * amp_s;
  out[kDAN] = tmp * flux * da_dn;
  out[kDAR] = tmp * flux * da_dr;
  out[kDAQ] = dAQ;
  out[kKQ] = -0.5f * (qxx + qyy);
  out[kQxx] = qxx;
  out[kQyy] = qyy;
  out[kQxy] = qxy;
  out[kDAP] = dAP;
  out[kKP] = -0.5f * (pxx + pyy);
  out[kPxx] = pxx;
  out[kPyy] = pyy;
  out[kPxy] = pxy;
  out[kSrc] = __builtin_bit_cast(float, s);
}

}  // namespace fpx
