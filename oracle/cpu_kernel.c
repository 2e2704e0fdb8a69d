/* cpu_kernel.c -- TEST INFRASTRUCTURE, not part of the product.
 *
 * A plain-C restatement of the reference CPU chi2 + gradient evaluation,
 * used as the checker of the HIP kernel and as the CPU baseline.  It is the
 * algorithm of
 *   forcepho/src/compute_gaussians_kernel.cc:45-114   (ConvolveOneSource)
 *   forcepho/src/compute_gaussians_kernel.cc:227-315  (EvaluateProposal)
 *   forcepho/src/common_kernel.cc:57-83               (ComputeResidualImage)
 *   forcepho/src/common_kernel.cc:95-143              (ComputeGaussianDerivative)
 *   forcepho/src/common_kernel.cc:175-240             (GetGaussianAndJacobian)
 *   forcepho/src/matrix22.cc:70-193                   (2x2 algebra)
 * with the same loop order (exposure -> pixel -> all Gaussians -> chi2 ->
 * per source Gaussians), the same float32 expression order and the same
 * accumulation types (chi2 in double, gradients sequential in float), so
 * that built without FMA contraction it reproduces the reference's float32
 * results bit for bit (checked against the compiled reference in tests).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "forcepho_amd.h"

typedef struct { float a, b, c, d; } m22;  /* [[a b] [c d]] = v11 v12 v21 v22 */

static m22 m_mul(m22 x, m22 y) {
    m22 r = {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d,
             x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
    return r;
}
static m22 m_scale(float s, m22 x) { m22 r = {s * x.a, s * x.b, s * x.c, s * x.d}; return r; }
static m22 m_add(m22 x, m22 y) { m22 r = {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; return r; }
static float m_det(m22 x) { return x.a * x.d - x.b * x.c; }
static float m_trace(m22 x) { return x.a + x.d; }
static m22 m_inv(m22 x) {
    float idet = 1.0f / m_det(x);
    m22 r = {x.d * idet, -x.b * idet, -x.c * idet, x.a * idet};
    return r;
}
/* A A^T */
static m22 m_aat(m22 A) {
    float t = A.a * A.c + A.b * A.d;
    m22 r = {A.a * A.a + A.b * A.b, t, t, A.c * A.c + A.d * A.d};
    return r;
}
/* A B^T + B A^T */
static m22 m_symabt(m22 A, m22 B) {
    float t = A.a * B.c + A.b * B.d + A.c * B.a + A.d * B.b;
    m22 r = {2.0f * (A.a * B.a + A.b * B.b), t, t, 2.0f * (A.c * B.c + A.d * B.d)};
    return r;
}
/* A B A, expanded term by term as matrix22.cc:144-150 */
static m22 m_aba(m22 A, m22 B) {
    m22 r;
    r.a = A.a * A.a * B.a + A.a * A.c * B.b + A.a * A.b * B.c + A.b * A.c * B.d;
    r.b = A.a * A.b * B.a + A.a * A.d * B.b + A.b * A.b * B.c + A.b * A.d * B.d;
    r.c = A.a * A.c * B.a + A.c * A.c * B.b + A.a * A.d * B.c + A.c * A.d * B.d;
    r.d = A.b * A.c * B.a + A.c * A.d * B.b + A.b * A.d * B.c + A.d * A.d * B.d;
    return r;
}

/* One on-image Gaussian and its 15 Jacobian terms (common_kernel.cc:20-46). */
typedef struct {
    float amp, xcen, ycen, fxx, fyy, fxy;
    float dA_dFlux, dx_dAlpha, dy_dAlpha, dx_dDelta, dy_dDelta;
    float dA_dQ, dFxx_dQ, dFyy_dQ, dFxy_dQ;
    float dA_dPA, dFxx_dPA, dFyy_dPA, dFxy_dPA;
    float dA_dSersic, dA_drh;
} oracle_gauss;

static void convolve_source(int band, int exposure, int g, int n_psf,
                            const FpxPatch *patch, const FpxSource *sources,
                            oracle_gauss *out) {
    const FpxSource *src = sources + g;
    int psf0 = patch->psfgauss_start[exposure];
    float G = patch->G[exposure];
    int dcw = 4 * (patch->n_sources * exposure + g);
    int cr = 2 * (patch->n_sources * exposure + g);
    float smean0 = src->ra - patch->crval[cr];
    float smean1 = src->dec - patch->crval[cr + 1];

    m22 D = {patch->D[dcw], patch->D[dcw + 1], patch->D[dcw + 2], patch->D[dcw + 3]};
    float c = cosf(src->pa), s = sinf(src->pa);
    m22 R = {c, -s, s, c};
    m22 S = {1.0f / src->q, 0.0f, 0.0f, src->q};
    m22 dS = {-1.0f / (src->q * src->q), 0.0f, 0.0f, 1.0f};
    m22 dR = {-s, -c, c, -s};
    m22 T = m_mul(m_mul(D, R), S);
    m22 dTq = m_mul(m_mul(D, R), dS);
    m22 dTp = m_mul(m_mul(D, dR), S);
    m22 CW = {patch->CW[dcw], patch->CW[dcw + 1], patch->CW[dcw + 2], patch->CW[dcw + 3]};
    float cx = CW.a * smean0 + CW.b * smean1;
    float cy = CW.c * smean0 + CW.d * smean1;
    float xs = cx + patch->crpix[cr];
    float ys = cy + patch->crpix[cr + 1];

    for (int p = 0; p < n_psf; p++) {
        const FpxPSFSourceGaussian *pg = patch->psfgauss + psf0 + p;
        oracle_gauss *o = out + g * n_psf + p;
        int rb = pg->sersic_radius_bin;
        float covar = patch->rad2[rb];
        float amp_s = src->mixture_amplitudes[rb];
        float da_dn = src->damplitude_dnsersic[rb];
        float da_dr = src->damplitude_drh[rb];
        float flux = src->fluxes[band];

        m22 scov = m_scale(covar, m_aat(T));
        m22 psfc = {pg->Cxx, pg->Cxy, pg->Cxy, pg->Cyy};
        m22 Sig = m_add(scov, psfc);
        m22 f = m_inv(Sig);
        float detF = m_det(f);
        o->fxx = f.a;
        o->fxy = f.c;
        o->fyy = f.d;
        o->xcen = xs + pg->xcen;
        o->ycen = ys + pg->ycen;
        float tmp = G * pg->amp * sqrtf(detF) * (1.0f / (2.0f * (float)M_PI));
        o->amp = tmp * flux * amp_s;
        m22 dSq = m_scale(covar, m_symabt(T, dTq));
        m22 dSp = m_scale(covar, m_symabt(T, dTp));
        m22 dFq = m_scale(-1.0f, m_aba(f, dSq));
        m22 dFp = m_scale(-1.0f, m_aba(f, dSp));
        o->dA_dQ = 0.5f * o->amp * m_trace(m_mul(Sig, dFq));
        o->dA_dPA = 0.5f * o->amp * m_trace(m_mul(Sig, dFp));
        o->dA_dFlux = tmp * amp_s;
        o->dA_dSersic = tmp * flux * da_dn;
        o->dA_drh = tmp * flux * da_dr;
        o->dx_dAlpha = CW.a;
        o->dy_dAlpha = CW.c;
        o->dx_dDelta = CW.b;
        o->dy_dDelta = CW.d;
        o->dFxx_dQ = dFq.a;
        o->dFyy_dQ = dFq.d;
        o->dFxy_dQ = dFq.c;
        o->dFxx_dPA = dFp.a;
        o->dFyy_dPA = dFp.d;
        o->dFxy_dPA = dFp.c;
    }
}

static float pixel_residual(float xp, float yp, float data, const oracle_gauss *g, int n) {
    float r = data;
    for (int i = 0; i < n; i++, g++) {
        float dx = xp - g->xcen;
        float dy = yp - g->ycen;
        float vx = g->fxx * dx + g->fxy * dy;
        float vy = g->fyy * dy + g->fxy * dx;
        float ea = dx * vx + dy * vy;
        if (ea > (float)MAX_EXP_ARG) continue;
        float Gp = expf(-0.5f * ea);
        float H = 1.0f + (vx * vx + vy * vy - g->fxx - g->fyy) / 24.f;
        float C = g->amp * Gp * H;
        r -= C;
    }
    return r;
}

static void pixel_gradient(float xp, float yp, float rw, const oracle_gauss *g, int n,
                           float *d) {
    for (int i = 0; i < n; i++, g++) {
        float dx = xp - g->xcen;
        float dy = yp - g->ycen;
        float vx = g->fxx * dx + g->fxy * dy;
        float vy = g->fyy * dy + g->fxy * dx;
        float ea = dx * vx + dy * vy;
        if (ea > (float)MAX_EXP_ARG) continue;
        float Gp = expf(-0.5f * ea);
        float H = 1.0f + (vx * vx + vy * vy - g->fxx - g->fyy) * (1.0f / 24.0f);
        float dC_dA = -2.f * rw * Gp;
        float c_h = dC_dA * g->amp;
        dC_dA *= H;
        float C = c_h * H;
        float dC_dx = C * vx;
        float dC_dy = C * vy;
        float dC_dfx = -0.5f * C * dx * dx;
        float dC_dfy = -0.5f * C * dy * dy;
        float dC_dfxy = -1.0f * C * dx * dy;
        dC_dx -= c_h * (g->fxx * vx + g->fxy * vy) * (1.0f / 12.0f);
        dC_dy -= c_h * (g->fyy * vy + g->fxy * vx) * (1.0f / 12.0f);
        dC_dfx -= c_h * (1.0f - 2.0f * dx * vx) * (1.0f / 24.0f);
        dC_dfy -= c_h * (1.0f - 2.0f * dy * vy) * (1.0f / 24.0f);
        dC_dfxy += c_h * (dy * vx + dx * vy) * (1.0f / 12.0f);
        d[0] += g->dA_dFlux * dC_dA;
        d[1] += g->dx_dAlpha * dC_dx + g->dy_dAlpha * dC_dy;
        d[2] += g->dx_dDelta * dC_dx + g->dy_dDelta * dC_dy;
        d[3] += g->dA_dQ * dC_dA + dC_dfx * g->dFxx_dQ + dC_dfxy * g->dFxy_dQ + dC_dfy * g->dFyy_dQ;
        d[4] += g->dA_dPA * dC_dA + dC_dfx * g->dFxx_dPA + dC_dfxy * g->dFxy_dPA + dC_dfy * g->dFyy_dPA;
        d[5] += g->dA_dSersic * dC_dA;
        d[6] += g->dA_drh * dC_dA;
    }
}

/* chi2 and dchi2/dparam of one band.  patch holds HOST pointers.
 * dchi2_out: float[MAXSOURCES * NPARAMS] (zeroed here).  Returns 0, or -1 on
 * allocation failure / limit violation. */
int oracle_evaluate_band(int band, const FpxPatch *patch, const FpxSource *sources,
                         double *chi2_out, float *dchi2_out) {
    int n_src = patch->n_sources;
    int n_psf = patch->n_psf_per_source[band];
    int n_total = n_src * n_psf;
    if (n_src < 1 || n_src > MAXSOURCES || n_psf < 1) return -1;
    oracle_gauss *gs = (oracle_gauss *)malloc(sizeof(oracle_gauss) * (size_t)n_total);
    if (!gs) return -1;
    for (int j = 0; j < MAXSOURCES * NPARAMS; j++) dchi2_out[j] = 0.0f;
    double chi2 = 0.0;
    float d[NPARAMS];
    for (int k = 0; k < patch->band_N[band]; k++) {
        int e = patch->band_start[band] + k;
        for (int g = 0; g < n_src; g++) convolve_source(band, e, g, n_psf, patch, sources, gs);
        for (int p = 0; p < patch->exposure_N[e]; p++) {
            int pix = patch->exposure_start[e] + p;
            float xp = patch->xpix[pix];
            float yp = patch->ypix[pix];
            float data = patch->data[pix];
            float ierr = patch->ierr[pix];
            float r = pixel_residual(xp, yp, data, gs, n_total);
            if (patch->residual) patch->residual[pix] = r;
            r *= ierr;
            data *= ierr;
            double c2 = ((double)r - (double)data);
            c2 *= ((double)r + (double)data);
            chi2 += c2;
            r *= ierr;
            for (int g = 0; g < n_src; g++) {
                for (int j = 0; j < NPARAMS; j++) d[j] = 0.0f;
                pixel_gradient(xp, yp, r, gs + g * n_psf, n_psf, d);
                for (int j = 0; j < NPARAMS; j++) dchi2_out[g * NPARAMS + j] += d[j];
            }
        }
    }
    free(gs);
    *chi2_out = chi2;
    return 0;
}
