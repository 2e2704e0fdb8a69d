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
// The record is laid out for the two pixel passes (fpx_kernel.hip), 36
// floats (9 float4).  With dx = x - xc, dy = y - yc and the pixel sums
//   a_k = sum_pix gw m_k,  b_k = sum_pix gw H m_k,  m = (1, dx, dy, dx^2, dy^2, dx dy)
// (gw = G * pixel weight, G the Gaussian, H the pixel-integral factor) the
// gradient of one Gaussian is a fixed linear map of (a, b); the record holds
// that map's coefficients so the per-pixel work is the 12 sums only:
//   [0..8]   xc yc  cxx cxy cyy  h0  sxx sxy syy
//            exponent log2 G = cxx dx^2 + cxy dx dy + cyy dy^2
//            (c = -log2(e)/2 (fxx, 2 fxy, fyy)); H = h0 + sxx dx^2 + sxy dx dy
//            + syy dy^2 (h0 = 1 - (fxx + fyy)/24, s = ((fxx^2 + fxy^2),
//            2 fxy (fxx + fyy), (fxy^2 + fyy^2))/24), i.e. the
//            H = 1 + (vx^2 + vy^2 - fxx - fyy)/24, v = F dx of common_kernel.cc:77
//   [9]      amp                                     model pass: amp G H
//   [10]     source index (int bits)
//   [11..13] dA/dflux dA/dsersic dA/drh              x b0
//   [14..17] ex1 ex2 ex3 ex4, [18..19] ey2 ey4        d/dx: ex1 b1 + ex2 b2 + ex3 a1
//            + ex4 a2;  d/dy: ex2 b1 + ey2 b2 + ex4 a1 + ey4 a2
//   [20..27] dA/dq kq0 kq3 kq4 kq5 hq3 hq4 hq5       d/dq: dA/dq b0 + kq0 a0 + kq3 a3
//            + kq4 a4 + kq5 a5 + hq3 b3 + hq4 b4 + hq5 b5
//   [28..35] the same eight for pa
// Expanding the per-pixel terms of common_kernel.cc:95-143 gives, with
// amp12 = amp/12, dF/dq = [[qxx, qxy], [qxy, qyy]] and kq = -(qxx + qyy)/2:
//   ex1 = amp fxx, ex2 = amp fxy, ey2 = amp fyy, ex3 = -2 amp sxx,
//   ex4 = -amp sxy, ey4 = -2 amp syy,
//   kq0 = amp12 kq, kq3 = amp12 (qxx fxx + qxy fxy), kq4 = amp12 (qyy fyy + qxy fxy),
//   kq5 = amp12 ((qxx + qyy) fxy + qxy (fxx + fyy)),
//   hq3 = -amp qxx / 2, hq4 = -amp qyy / 2, hq5 = -amp qxy.
#pragma once

#include <hip/hip_runtime.h>
#include "forcepho_amd.h"

namespace fpx {

constexpr int kRec = 36;  // floats per ImageGaussian record
constexpr float kExp2Scale = -0.72134752044448170368f;  // -0.5 * log2(e)

enum RecField {
  kXc = 0, kYc, kCxx, kCxy, kCyy, kH0, kSxx, kSxy, kSyy,
  kAmp, kSrc,
  kDAF, kDAN, kDAR,
  kEx1, kEx2, kEx3, kEx4, kEy2, kEy4,
  kDAQ, kQ0, kQ3, kQ4, kQ5, kHQ3, kHQ4, kHQ5,
  kDAP, kP0, kP3, kP4, kP5, kHP3, kHP4, kHP5
};
static_assert(kHP5 + 1 == kRec, "record layout");

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

// The inputs of a SourceGeometry: the source's q, pa and sky position, and
// the exposure's D, CW, crpix and crval for that source.  Loaded apart from
// the arithmetic so that a caller can issue these loads together with its
// others and wait once.
struct GeometryInputs {
  float q, pa, ra, dec;
  float D[4], CW[4], crpix[2], crval[2];
};

// (FloatPtr and SrcPtr: pointers to float and FpxSource in any address
// space.)
template <typename FloatPtr, typename SrcPtr>
__device__ __forceinline__ GeometryInputs load_geometry_inputs(FloatPtr D, FloatPtr CW,
                                                               FloatPtr crpix, FloatPtr crval,
                                                               SrcPtr src, int n_sources, int e,
                                                               int s) {
  GeometryInputs in;
  const int base = n_sources * e + s;
  in.q = src->q;
  in.pa = src->pa;
  in.ra = src->ra;
  in.dec = src->dec;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    in.D[k] = D[4 * base + k];
    in.CW[k] = CW[4 * base + k];
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    in.crpix[k] = crpix[2 * base + k];
    in.crval[k] = crval[2 * base + k];
  }
  return in;
}

template <bool kDerivs>
__device__ __forceinline__ void geometry_from(const GeometryInputs &in, SourceGeometry &g) {
  const float q = in.q;
  float sn, cs;
  sincosf(in.pa, &sn, &cs);
  const M22 D = {in.D[0], in.D[1], in.D[2], in.D[3]};
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
  const float sx = in.ra - in.crval[0];
  const float sy = in.dec - in.crval[1];
  g.xs = in.CW[0] * sx + in.CW[1] * sy + in.crpix[0];
  g.ys = in.CW[2] * sx + in.CW[3] * sy + in.crpix[1];
}

template <bool kDerivs>
__device__ __forceinline__ void source_geometry(const FpxPatch *__restrict__ P,
                                                const FpxSource *__restrict__ src, int e,
                                                int s, SourceGeometry &g) {
  geometry_from<kDerivs>(
      load_geometry_inputs(P->D, P->CW, P->crpix, P->crval, src, P->n_sources, e, s), g);
}

// A SourceGeometry kept in kGeoFloats floats (LDS): T, dT/dq, dT/dpa, centre.
constexpr int kGeoFloats = 16;

__device__ __forceinline__ void store_geometry(float *__restrict__ dst, const SourceGeometry &g) {
  const M22 *m[3] = {&g.T, &g.dTq, &g.dTp};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dst[4 * k + 0] = m[k]->a;
    dst[4 * k + 1] = m[k]->b;
    dst[4 * k + 2] = m[k]->c;
    dst[4 * k + 3] = m[k]->d;
  }
  dst[12] = g.xs;
  dst[13] = g.ys;
}

__device__ __forceinline__ SourceGeometry load_geometry(const float *__restrict__ src) {
  SourceGeometry g;
  g.T = {src[0], src[1], src[2], src[3]};
  g.dTq = {src[4], src[5], src[6], src[7]};
  g.dTp = {src[8], src[9], src[10], src[11]};
  g.xs = src[12];
  g.ys = src[13];
  return g;
}

// Sigma = covar T T^T + C_psf
__device__ __forceinline__ void convolved_sigma(const M22 &T, float covar,
                                                const FpxPSFSourceGaussian &psf, float &Sxx,
                                                float &Sxy, float &Syy) {
  Sxx = covar * (T.a * T.a + T.b * T.b) + psf.Cxx;
  Sxy = covar * (T.a * T.c + T.b * T.d) + psf.Cxy;
  Syy = covar * (T.c * T.c + T.d * T.d) + psf.Cyy;
}

// F = Sigma^-1 of a symmetric 2x2 Sigma.
__device__ __forceinline__ void inverse_sigma(float Sxx, float Sxy, float Syy, float &fxx,
                                              float &fxy, float &fyy) {
  const float idet = 1.0f / (Sxx * Syy - Sxy * Sxy);
  fxx = Syy * idet;
  fxy = -Sxy * idet;
  fyy = Sxx * idet;
}

// Centre and F of ImageGaussian (s, p) of exposure e.  The radius bin is
// clamped into the patch's table (the evaluation kernel instead refuses a
// bin outside it, through the slot's status word).
__device__ __forceinline__ void image_gaussian_shape(const FpxPatch *__restrict__ P,
                                                     const FpxSource *__restrict__ props,
                                                     int e, int s, int p, float &xc, float &yc,
                                                     float &fxx, float &fxy, float &fyy) {
  SourceGeometry g;
  source_geometry<false>(P, props + s, e, s, g);
  const FpxPSFSourceGaussian psf = P->psfgauss[P->psfgauss_start[e] + p];
  const int n_radii = min(max(P->n_radii, 1), MAXRADII);
  const int bin = min(max(psf.sersic_radius_bin, 0), n_radii - 1);
  float Sxx, Sxy, Syy;
  convolved_sigma(g.T, P->rad2[bin], psf, Sxx, Sxy, Syy);
  inverse_sigma(Sxx, Sxy, Syy, fxx, fxy, fyy);
  xc = g.xs + psf.xcen;
  yc = g.ys + psf.ycen;
}

// Half-width of the box holding dx^T Sigma^-1 dx <= MAX_EXP_ARG along an
// axis of variance S, with a relative 1e-3 and an absolute 0.01 px margin.
// A NaN variance gives a NaN half-width (a comparison, not fmaxf, clamps
// negative S: fmaxf would turn NaN into 0).
__device__ __forceinline__ float cutoff_halfwidth(float S) {
  return sqrtf((float)MAX_EXP_ARG) * sqrtf(S < 0.0f ? 0.0f : S) * 1.001f + 0.01f;
}

// True when the MAX_EXP_ARG box of the convolution of the source (geometry
// g) with PSF Gaussian psf (radius bin variance covar) can reach a pixel of
// [x0, x1] x [y0, y1].  A box that cannot be decided (NaN from NaN source
// parameters) counts as reaching: the reference's cutoff test `exparg >
// MAX_EXP_ARG` is false for NaN (common_kernel.cc:73,108), so NaN parameters
// make every pixel's model, and chi2, NaN there too.
__device__ __forceinline__ bool gauss_reaches(const SourceGeometry &g,
                                              const FpxPSFSourceGaussian &psf, float covar,
                                              float x0, float x1, float y0, float y1) {
  float Sxx, Sxy, Syy;
  convolved_sigma(g.T, covar, psf, Sxx, Sxy, Syy);
  const float xc = g.xs + psf.xcen;
  const float yc = g.ys + psf.ycen;
  const float gx = fmaxf(fmaxf(x0 - xc, xc - x1), 0.0f);
  const float gy = fmaxf(fmaxf(y0 - yc, yc - y1), 0.0f);
  return !(gx > cutoff_halfwidth(Sxx)) && !(gy > cutoff_halfwidth(Syy));
}

// Build the record of source s (geometry g for the exposure, with dT/dq and
// dT/dpa) convolved with PSF Gaussian psf of the exposure (radius bin
// variance covar, exposure calibration gcal) in band `band`.  The caller
// takes psf from the exposure's PSF block whatever the source, as the
// reference does (compute_gaussians_kernel.cc:100, .cu:77).  kSrc holds the
// source index (as int bits).  SrcPtr: a pointer to FpxSource in any
// address space.
template <typename SrcPtr>
__device__ __forceinline__ void build_record(const SourceGeometry &g, SrcPtr src,
                                             const FpxPSFSourceGaussian &psf, float covar,
                                             float gcal, int band, int s,
                                             float *__restrict__ out) {
  const M22 &T = g.T, &dTq = g.dTq, &dTp = g.dTp;
  const int rb = psf.sersic_radius_bin;
  const float amp_s = src->mixture_amplitudes[rb];
  const float da_dn = src->damplitude_dnsersic[rb];
  const float da_dr = src->damplitude_drh[rb];
  const float flux = src->fluxes[band];

  float Sxx, Sxy, Syy;
  convolved_sigma(T, covar, psf, Sxx, Sxy, Syy);
  float fxx, fxy, fyy;
  inverse_sigma(Sxx, Sxy, Syy, fxx, fxy, fyy);
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
  const float tr = fxx + fyy;
  const float sxx = (fxx * fxx + fxy * fxy) * (1.0f / 24.0f);
  const float sxy = 2.0f * fxy * tr * (1.0f / 24.0f);
  const float syy = (fxy * fxy + fyy * fyy) * (1.0f / 24.0f);
  const float amp12 = amp * (1.0f / 12.0f);
  out[kXc] = g.xs + psf.xcen;
  out[kYc] = g.ys + psf.ycen;
  out[kCxx] = kExp2Scale * fxx;
  out[kCxy] = kExp2Scale * (2.0f * fxy);
  out[kCyy] = kExp2Scale * fyy;
  out[kH0] = h0;
  out[kSxx] = sxx;
  out[kSxy] = sxy;
  out[kSyy] = syy;
  out[kAmp] = amp;
  out[kSrc] = __builtin_bit_cast(float, s);
  out[kDAF] = tmp * amp_s;
  out[kDAN] = tmp * flux * da_dn;
  out[kDAR] = tmp * flux * da_dr;
  out[kEx1] = amp * fxx;
  out[kEx2] = amp * fxy;
  out[kEx3] = -2.0f * amp * sxx;
  out[kEx4] = -amp * sxy;
  out[kEy2] = amp * fyy;
  out[kEy4] = -2.0f * amp * syy;
  out[kDAQ] = dAQ;
  out[kQ0] = amp12 * (-0.5f * (qxx + qyy));
  out[kQ3] = amp12 * (qxx * fxx + qxy * fxy);
  out[kQ4] = amp12 * (qyy * fyy + qxy * fxy);
  out[kQ5] = amp12 * ((qxx + qyy) * fxy + qxy * tr);
  out[kHQ3] = -0.5f * amp * qxx;
  out[kHQ4] = -0.5f * amp * qyy;
  out[kHQ5] = -amp * qxy;
  out[kDAP] = dAP;
  out[kP0] = amp12 * (-0.5f * (pxx + pyy));
  out[kP3] = amp12 * (pxx * fxx + pxy * fxy);
  out[kP4] = amp12 * (pyy * fyy + pxy * fxy);
  out[kP5] = amp12 * ((pxx + pyy) * fxy + pxy * tr);
  out[kHP3] = -0.5f * amp * pxx;
  out[kHP4] = -0.5f * amp * pyy;
  out[kHP5] = -amp * pxy;
}

}  // namespace fpx
