"""TEST INFRASTRUCTURE -- an independent float64 formulation of the chi2 and
gradient, used to adjudicate float32 tolerances.  Not the product.

It follows the reference's pure-Python model rather than its C++ kernel:
  * convolution and Jacobian as forcepho/slow/gaussmodel.py:271-331
    (_get_gaussian_gradients: Sigma = T rad2 T^T + C_psf, F = Sigma^-1,
    K = flux G a_mix a_psf sqrt(det F)/(2 pi), dF = -F dSigma F,
    dA/dq = K/(2 det F) * det F tr(Sigma dF/dq)), here without the divisions
    by flux and a_mix so that zero fluxes stay finite;
  * the pixel model and its derivatives as forcepho/slow/gaussmodel.py:349-475
    (_compute_gaussian with the second-order pixel-integration term);
  * the kernel's conventions: skip where dx^T F dx > MAX_EXP_ARG
    (common_kernel.cc:73,108), every source of exposure e uses the PSF block
    at psfgauss_start[e] (compute_gaussians_kernel.cc:100), chi2 summed as
    (r ierr - d ierr)(r ierr + d ierr) (compute_gaussians_kernel.cc:286-293).
"""

import numpy as np

MAX_EXP_ARG = 36.0
NPARAMS = 7


def _rot(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, -s], [s, c]]), np.array([[-s, -c], [c, -s]])


def image_gaussians(patch, proposal, e, band):
    """float64 ImageGaussians of exposure e: dict of (n_src*n_psf,) arrays plus
    ``jac`` (n, 7, 6): rows d/d(flux, alpha, delta, q, pa, sersic, rh), columns
    d(A, x, y, Fxx, Fyy, Fxy)."""
    n_src = int(patch.n_sources)
    n_psf = int(patch.n_psf_per_source[band])
    G = float(patch.G[e])
    p0 = int(patch.psfgauss_start[e])
    psf = patch.psfgauss[p0:p0 + n_psf]
    out = {k: np.zeros(n_src * n_psf) for k in ("amp", "xcen", "ycen", "fxx", "fxy", "fyy")}
    jac = np.zeros((n_src * n_psf, NPARAMS, 6))
    for s in range(n_src):
        src = proposal[s]
        D = patch.D[e, s].astype(np.float64)
        CW = patch.CW[e, s].astype(np.float64)
        crpix = patch.crpix[e, s].astype(np.float64)
        crval = patch.crval[e, s].astype(np.float64)
        q, pa = float(src["q"]), float(src["pa"])
        R, dR = _rot(pa)
        S = np.diag([1.0 / q, q])
        dS = np.diag([-1.0 / q ** 2, 1.0])
        T = D @ R @ S
        dTq = D @ R @ dS
        dTp = D @ dR @ S
        sky = np.array([float(src["ra"]), float(src["dec"])]) - crval
        cen = CW @ sky + crpix
        flux = float(src["fluxes"][band])
        for p in range(n_psf):
            g = s * n_psf + p
            rb = int(psf["sersic_bin"][p])
            scov = float(patch.rad2[rb]) * np.eye(2)
            pcov = np.array([[psf["Cxx"][p], psf["Cxy"][p]], [psf["Cxy"][p], psf["Cyy"][p]]],
                            dtype=np.float64)
            Sig = T @ scov @ T.T + pcov
            F = np.linalg.inv(Sig)
            detF = np.linalg.det(F)
            am = float(src["mixture_amplitudes"][rb])
            al = float(psf["amp"][p])
            unit = G * al * np.sqrt(detF) / (2 * np.pi)   # K / (flux am)
            K = flux * am * unit
            dSq = T @ scov @ dTq.T + dTq @ scov @ T.T
            dSp = T @ scov @ dTp.T + dTp @ scov @ T.T
            dFq = -F @ dSq @ F
            dFp = -F @ dSp @ F
            ddetF_q = detF * np.trace(Sig @ dFq)
            ddetF_p = detF * np.trace(Sig @ dFp)
            out["amp"][g] = K
            out["xcen"][g] = cen[0] + float(psf["xcen"][p])
            out["ycen"][g] = cen[1] + float(psf["ycen"][p])
            out["fxx"][g], out["fxy"][g], out["fyy"][g] = F[0, 0], F[1, 0], F[1, 1]
            J = jac[g]
            J[0, 0] = am * unit
            J[1, 1], J[1, 2] = CW[0, 0], CW[1, 0]
            J[2, 1], J[2, 2] = CW[0, 1], CW[1, 1]
            J[3, 0] = K / (2 * detF) * ddetF_q
            J[3, 3:] = dFq[0, 0], dFq[1, 1], dFq[1, 0]
            J[4, 0] = K / (2 * detF) * ddetF_p
            J[4, 3:] = dFp[0, 0], dFp[1, 1], dFp[1, 0]
            J[5, 0] = flux * unit * float(src["damplitude_dnsersic"][rb])
            J[6, 0] = flux * unit * float(src["damplitude_drh"][rb])
    out["jac"] = jac
    return out


def _pixel_terms(ig, xp, yp):
    """C and dC/d(A, x, y, Fxx, Fyy, Fxy) of every (pixel, Gaussian) pair."""
    dx = xp[:, None] - ig["xcen"][None, :]
    dy = yp[:, None] - ig["ycen"][None, :]
    fxx, fxy, fyy, amp = ig["fxx"], ig["fxy"], ig["fyy"], ig["amp"]
    vx = fxx * dx + fxy * dy
    vy = fyy * dy + fxy * dx
    ea = dx * vx + dy * vy
    on = ~(ea > MAX_EXP_ARG)
    Gp = np.where(on, np.exp(-0.5 * np.where(on, ea, 0.0)), 0.0)
    H = 1.0 + (vx * vx + vy * vy - fxx - fyy) / 24.0
    C = amp * Gp * H
    c_h = amp * Gp
    dC = np.empty(C.shape + (6,))
    dC[..., 0] = Gp * H
    dC[..., 1] = C * vx - c_h * (fxx * vx + fxy * vy) / 12.0
    dC[..., 2] = C * vy - c_h * (fyy * vy + fxy * vx) / 12.0
    dC[..., 3] = -0.5 * C * dx * dx - c_h * (1.0 - 2.0 * dx * vx) / 24.0
    dC[..., 4] = -0.5 * C * dy * dy - c_h * (1.0 - 2.0 * dy * vy) / 24.0
    dC[..., 5] = -1.0 * C * dx * dy + c_h * (dy * vx + dx * vy) / 12.0
    return C, dC


def evaluate(patch, proposal, pixel_weight=None, chunk=2048, mass=False):
    """float64 (chi2[n_bands], dchi2[n_bands, n_src, 7], residual[npix]).

    With ``pixel_weight`` (float[npix]) returns instead, in place of dchi2,
    sum_pix pixel_weight * d model / d theta (chi2 still computed).
    With ``mass=True`` also returns sum_pix sum_g |w_pix * d C_g / d theta|
    [n_bands, n_src, 7]: the magnitude of what the gradient sums, which sets
    the rounding floor of any float32 evaluation of it."""
    nb, ns = int(patch.n_bands), int(patch.n_sources)
    chi2 = np.zeros(nb)
    grad = np.zeros((nb, ns, NPARAMS))
    resid = np.zeros(patch.npix)
    xpix = patch.xpix.astype(np.float64)
    ypix = patch.ypix.astype(np.float64)
    data = patch.data.astype(np.float64)
    ierr = patch.ierr.astype(np.float64)
    gmass = np.zeros((nb, ns, NPARAMS))
    for b in range(nb):
        n_psf = int(patch.n_psf_per_source[b])
        for k in range(int(patch.band_N[b])):
            e = int(patch.band_start[b]) + k
            ig = image_gaussians(patch, proposal, e, b)
            start, n = int(patch.exposure_start[e]), int(patch.exposure_N[e])
            acc = np.zeros((ns * n_psf, 6))
            macc = np.zeros((ns * n_psf, NPARAMS))
            for c0 in range(start, start + n, chunk):
                c1 = min(c0 + chunk, start + n)
                C, dC = _pixel_terms(ig, xpix[c0:c1], ypix[c0:c1])
                model = C.sum(axis=1)
                r = data[c0:c1] - model
                resid[c0:c1] = r
                ri, di = r * ierr[c0:c1], data[c0:c1] * ierr[c0:c1]
                chi2[b] += np.sum((ri - di) * (ri + di))
                if pixel_weight is None:
                    w = -2.0 * r * ierr[c0:c1] ** 2
                else:
                    w = np.asarray(pixel_weight[c0:c1], dtype=np.float64)
                acc += np.einsum("p,pgf->gf", w, dC)
                if mass:
                    per = np.abs(np.einsum("pgf,gkf->pgk", dC, ig["jac"]))
                    macc += np.einsum("p,pgk->gk", np.abs(w), per)
            gk = np.einsum("gkf,gf->gk", ig["jac"], acc)
            grad[b] += gk.reshape(ns, n_psf, NPARAMS).sum(axis=1)
            gmass[b] += macc.reshape(ns, n_psf, NPARAMS).sum(axis=1)
    if mass:
        return chi2, grad, resid, gmass
    return chi2, grad, resid
