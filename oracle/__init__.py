"""TEST INFRASTRUCTURE -- the checker of the HIP kernel, never the product.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this
package.  It wraps
  * ``_build/liboracle_cpu.so``: the C restatement of the reference CPU
    kernel (cpu_kernel.c), bit-identical to the reference's float32 results;
  * ``_ref/compute_gaussians_kernel*.so``: the reference's own CPU kernel
    compiled from its sources by oracle/Makefile (present only where the
    reference tree was available when building);
  * ``fp64``: an independent float64 numpy formulation (oracle/fp64.py).
Both compiled checkers take a host-side Patch struct of raw array pointers,
exactly as forcepho's CPUPatchMixin builds it
(forcepho/patches/device_patch.py:446-516), and are called once per band as
forcepho's CPUProposer does (forcepho/proposal.py:204-215).
"""

import ctypes
import glob
import importlib.util
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_LIB = os.path.join(HERE, "_build", "liboracle_cpu.so")

MAXSOURCES, NPARAMS = 30, 7

_PDT = np.dtype([(n, np.uintp) for n in ("data", "ierr", "xpix", "ypix", "residual",
                                           "exposure_start", "exposure_N", "band_start",
                                           "band_N")]
                + [("n_sources", np.int32), ("n_radii", np.int32)]
                + [(n, np.uintp) for n in ("rad2", "D", "crpix", "crval", "CW", "G",
                                           "n_psf_per_source", "psfgauss",
                                           "psfgauss_start")], align=True)
_ARRAYS = {"data": np.float32, "ierr": np.float32, "xpix": np.float32, "ypix": np.float32,
           "exposure_start": np.int32, "exposure_N": np.int32, "band_start": np.int16,
           "band_N": np.int16, "rad2": np.float32, "D": np.float32, "crpix": np.float32,
           "crval": np.float32, "CW": np.float32, "G": np.float32,
           "n_psf_per_source": np.int32, "psfgauss_start": np.int32}

_cpu = None


def build():
    """Compile the C checker (and the reference build where possible)."""
    subprocess.run(["make", "-C", HERE, "all"], check=True, capture_output=True)


def _cpu_lib():
    global _cpu
    if _cpu is None:
        if not os.path.exists(CPU_LIB):
            build()
        _cpu = ctypes.CDLL(CPU_LIB)
        _cpu.oracle_evaluate_band.restype = ctypes.c_int
        _cpu.oracle_evaluate_band.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p]
    return _cpu


def host_patch_struct(patch, residual=None):
    """A host Patch struct (numpy record) pointing at the patch's arrays.
    Returns (struct, keepalive) -- keep ``keepalive`` referenced while the
    struct is in use."""
    keep = {}
    ps = np.zeros(1, dtype=_PDT)
    for name, dt in _ARRAYS.items():
        arr = np.ascontiguousarray(getattr(patch, name), dtype=dt)
        keep[name] = arr
        ps[name] = arr.ctypes.data
    pg = np.ascontiguousarray(patch.psfgauss)
    assert pg.dtype.itemsize == 28
    keep["psfgauss"] = pg
    ps["psfgauss"] = pg.ctypes.data
    if residual is not None:
        assert residual.dtype == np.float32 and residual.flags.c_contiguous
        keep["residual"] = residual
        ps["residual"] = residual.ctypes.data
    ps["n_sources"] = int(patch.n_sources)
    ps["n_radii"] = int(patch.n_radii)
    keep["struct"] = ps
    return ps, keep


def _proposal_bytes(proposal):
    raw = np.ascontiguousarray(proposal)
    assert raw.dtype.itemsize == 264
    return raw


def evaluate_cpu(patch, proposal, return_residual=False):
    """C restatement: (chi2 f64[n_bands], derivs f32[n_bands, n_src, 7],
    residual f32[npix] or None)."""
    L = _cpu_lib()
    resid = np.zeros(patch.npix, dtype=np.float32) if return_residual else None
    ps, keep = host_patch_struct(patch, resid)
    prop = _proposal_bytes(proposal)
    nb, ns = int(patch.n_bands), int(patch.n_sources)
    chi2 = np.zeros(nb, dtype=np.float64)
    derivs = np.zeros((nb, ns, NPARAMS), dtype=np.float32)
    out = np.zeros(MAXSOURCES * NPARAMS, dtype=np.float32)
    c = ctypes.c_double()
    for b in range(nb):
        rc = L.oracle_evaluate_band(b, ps.ctypes.data, prop.ctypes.data, ctypes.byref(c),
                                    out.ctypes.data)
        if rc != 0:
            raise RuntimeError("oracle_evaluate_band failed")
        chi2[b] = c.value
        derivs[b] = out.reshape(MAXSOURCES, NPARAMS)[:ns]
    del keep
    return chi2, derivs, resid


def reference_module():
    """The reference's own compiled CPU kernel, or None if it was not built."""
    hits = glob.glob(os.path.join(HERE, "_ref", "compute_gaussians_kernel*.so"))
    if not hits:
        return None
    spec = importlib.util.spec_from_file_location("compute_gaussians_kernel", hits[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def evaluate_reference(patch, proposal, return_residual=False, module=None):
    """Same outputs as evaluate_cpu, computed by the reference's compiled
    EvaluateProposal (one call per band)."""
    mod = module or reference_module()
    if mod is None:
        raise RuntimeError("reference CPU kernel not built (oracle/_ref)")
    resid = np.zeros(patch.npix, dtype=np.float32) if return_residual else None
    ps, keep = host_patch_struct(patch, resid)
    prop = _proposal_bytes(proposal)
    nb, ns = int(patch.n_bands), int(patch.n_sources)
    chi2 = np.zeros(nb, dtype=np.float64)
    derivs = np.zeros((nb, ns, NPARAMS), dtype=np.float32)
    for b in range(nb):
        r = mod.EvaluateProposal(b, int(ps.ctypes.data), int(prop.ctypes.data))
        chi2[b] = r.chi2
        derivs[b] = np.asarray(r.dchi2_dp).reshape(MAXSOURCES, NPARAMS)[:ns]
    del keep
    return chi2, derivs, resid


def render_model(patch, proposal):
    """Model image (float64 over packed pixels) from the C restatement:
    evaluation with zero data, negated residual."""
    saved = patch.data
    patch.data = np.zeros_like(saved)
    try:
        _, _, resid = evaluate_cpu(patch, proposal, return_residual=True)
    finally:
        patch.data = saved
    return -resid.astype(np.float64)
