"""Linear flux solve conditional on source shapes (mirrors
forcepho/fitting.py:572-620).

With every shape parameter fixed the model is linear in the fluxes:
``model_b = sum_i flux_ib X_ib`` where X_ib is source i's unit-flux image in
band b (``design_matrix``, forcepho/patches/device_patch.py:185-238, rendered
by the HIP kernel one source at a time).  The maximum-likelihood fluxes solve
the weighted normal equations (X W)(X W)^T f = (X W)(y W), W = ierr, and
(X W)(X W)^T is the flux precision matrix.
"""

import numpy as np

__all__ = ["optimize_one_band", "optimize_fluxes"]


def optimize_one_band(X, w, y, fixedX=0):
    """ML fluxes and precision matrix of one band.

    Parameters
    ----------
    X : (n_source, n_pix) unit-flux models
    w : (n_pix,) inverse errors
    y : (n_pix,) data
    fixedX : (n_pix,) or 0, model of fixed sources to subtract from y

    Returns
    -------
    flux : (n_source, 1)
    precision : (n_source, n_source)
    """
    Xp = np.asarray(X, dtype=np.float64) * w
    yp = (np.asarray(y, dtype=np.float64) - fixedX) * w
    ata = Xp @ Xp.T
    flux = np.linalg.solve(ata, Xp @ yp[:, None])
    return flux, ata


def optimize_fluxes(patcher, active, return_all=False):
    """ML fluxes and flux precision matrices of the active sources in every
    band, from the patch's device-side data (which holds data minus any
    subtracted fixed sources).

    Returns
    -------
    fluxes : list over ``patcher.bandlist`` of (n_source,) arrays
    precisions : list of (n_source, n_source) arrays
    (plus the design matrices, data and weights with ``return_all``)
    """
    if not hasattr(active, "sources"):
        active = patcher.set_scene(active)
    Xes, ys = patcher.design_matrix(active)
    ws = patcher.split_band("ierr")
    fluxes, precisions = [], []
    for X, y, w in zip(Xes, ys, ws):
        flux, precision = optimize_one_band(X, w.astype(np.float64), y)
        fluxes.append(np.atleast_1d(np.squeeze(flux)))
        precisions.append(precision)
    if return_all:
        return fluxes, precisions, Xes, ys, ws
    return fluxes, precisions
