"""The C ABI bound the way a forcepho maintainer would bind it: ctypes over
libfpx.so and the HIP runtime, device memory from hipMalloc, no torch.  The
patch arrays and the Patch struct are uploaded as forcepho's GPUPatchMixin
uploads them (forcepho/patches/device_patch.py:257-341).  The bound-patch
entry points (fpx_bind once per packed patch, then one fpx_evaluate_host
per evaluation: the reference's one PyCUDA call, forcepho/proposal.py:
133-136) are the binding INTEGRATION.md shows; the lower-level slot entry
(tile table from fpx_plan_tiles, one fpx_evaluate_sync per evaluation) is
checked too.  Results are checked against the golden fixture made by the
reference's CPU kernel and against this package's own HIPProposer on the
same patch."""

import ctypes

import numpy as np
import pytest

import fpx_testutil as U
from forcepho_amd import _lib
from forcepho_amd.kernel_limits import MAXBANDS, MAXSOURCES, NPARAMS
from forcepho_amd.proposal import response_struct_dtype, source_struct_dtype

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class RawDevice:
    """hipMalloc / hipMemcpy / hipFree through ctypes (the runtime libfpx.so
    itself is linked against)."""

    def __init__(self):
        _lib.lib()                                    # loads libfpx.so and its HIP runtime
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 8))) == 0
        self.ptrs.append(p.value)
        return p.value

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(ctypes.c_void_p(dst), arr.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(arr.nbytes), H2D) == 0

    def d2h(self, arr, src):
        assert self.hip.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src),
                                  ctypes.c_size_t(arr.nbytes), D2H) == 0

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(ctypes.c_void_p(p))
        self.ptrs = []


def _upload_patch(dev, patch):
    """Device copies of the patch's pointer fields and the Patch struct
    pointing at them; returns (struct address, {name: address})."""
    pdt = patch.patch_struct_dtype
    ptrs = {}
    for name in pdt.names:
        if pdt[name] != patch.ptr_dtype or name == "residual":
            continue
        arr = np.ascontiguousarray(getattr(patch, name))
        ptrs[name] = dev.malloc(arr.nbytes)
        dev.h2d(ptrs[name], arr)
    ptrs["residual"] = dev.malloc(4 * patch.npix)
    ps = np.zeros(1, dtype=pdt)
    for name in pdt.names:
        ps[name] = ptrs[name] if pdt[name] == patch.ptr_dtype else getattr(patch, name)
    addr = dev.malloc(ps.nbytes)
    dev.h2d(addr, ps)
    return addr, ptrs


def _evaluate_raw(dev, patch, proposal, cfg=5):
    L = _lib.lib()
    ppt = ctypes.c_int32()
    assert L.fpx_config_info(cfg, None, None, ctypes.byref(ppt)) == 0
    d_patch, ptrs = _upload_patch(dev, patch)
    idx = [np.ascontiguousarray(patch.exposure_start, np.int32),
           np.ascontiguousarray(patch.exposure_N, np.int32),
           np.ascontiguousarray(patch.band_start, np.int16),
           np.ascontiguousarray(patch.band_N, np.int16)]
    args = [a.ctypes.data for a in idx] + [int(patch.n_bands), ppt.value]
    n_tiles = L.fpx_plan_tiles(*args, None, 0)
    assert n_tiles > 0
    tiles = np.zeros(n_tiles, dtype=_lib.TILE_DTYPE)
    assert L.fpx_plan_tiles(*args, tiles.ctypes.data, n_tiles) == n_tiles
    d_tiles = dev.malloc(tiles.nbytes)
    dev.h2d(d_tiles, tiles)
    n_acc = L.fpx_accum_doubles(MAXBANDS)
    d_accum = dev.malloc(8 * n_acc)
    dev.h2d(d_accum, np.zeros(n_acc))
    out_bytes = 8 * MAXBANDS + response_struct_dtype.itemsize * MAXBANDS
    d_out = dev.malloc(out_bytes)
    d_prop = dev.malloc(source_struct_dtype.itemsize * MAXSOURCES)
    nb, ns = int(patch.n_bands), int(patch.n_sources)
    slot = np.zeros(1, dtype=_lib.SLOT_DTYPE)
    slot["patch"], slot["proposal"], slot["tiles"] = d_patch, d_prop, d_tiles
    slot["pchi2"], slot["pdchi2_dp"] = d_out, d_out + 8 * MAXBANDS
    slot["accum"], slot["n_bands"], slot["n_tiles"] = d_accum, nb, n_tiles
    d_slot = dev.malloc(slot.nbytes)
    dev.h2d(d_slot, slot)
    prop = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
    out = np.zeros(out_bytes, dtype=np.uint8)
    n_out = 8 * MAXBANDS + response_struct_dtype.itemsize * nb
    max_gauss = ns * int(np.max(patch.n_psf_per_source))
    rc = L.fpx_evaluate_sync(d_slot, n_tiles, max_gauss, cfg, 0, None, d_prop, prop.ctypes.data,
                             prop.nbytes, d_out, out.ctypes.data, n_out, None, None)
    assert rc == 0, L.fpx_last_error()
    chi2 = out[:8 * nb].view(np.float64).copy()
    derivs = out[8 * MAXBANDS:n_out].view(np.float32).reshape(nb, MAXSOURCES, NPARAMS)[:, :ns]
    resid = np.zeros(patch.npix, dtype=np.float32)
    dev.d2h(resid, ptrs["residual"])
    return chi2, derivs.copy(), resid


@pytest.mark.parametrize("name", ["multi_small_s2", "padding_s5"])
def test_raw_ctypes_binding_matches_fixture_and_host_layer(name):
    from forcepho_amd.proposal import HIPProposer
    print(f"[gpu] raw ctypes binding on {name}", flush=True)
    patch, proposal, expected = U.load_golden(name, return_residual=True)
    dev = RawDevice()
    try:
        chi2, derivs, resid = _evaluate_raw(dev, patch, proposal)
        chi2b, derivsb, residb = _evaluate_raw(dev, patch, proposal)   # a second call
    finally:
        dev.free()
    rel = np.abs(chi2 - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL
    dscale = np.max(np.abs(patch.data)) + 1e-30
    assert np.max(np.abs(resid - expected["residual"])) <= 2e-6 * dscale
    np.testing.assert_allclose(chi2b, chi2, rtol=1e-12)
    np.testing.assert_array_equal(residb, resid)
    # the torch-backed host layer on the same patch: same kernel, same results
    patch.send_to_device()
    c, d, r = HIPProposer(patch=patch, config=5, persistent=False).evaluate_proposal(
        proposal, patch=patch, unpack=False)
    np.testing.assert_allclose(chi2, c, rtol=1e-12)
    np.testing.assert_allclose(derivs, d, rtol=1e-6, atol=1e-6 * np.abs(d).max())
    np.testing.assert_array_equal(resid, r)


def _evaluate_bound(dev, patch, proposal, cfg=-1):
    """The maintainer's binding: fpx_bind on the uploaded struct, then
    fpx_evaluate_host per call."""
    L = _lib.lib()
    d_patch, ptrs = _upload_patch(dev, patch)
    ev = ctypes.c_void_p()
    rc = L.fpx_bind(d_patch, int(patch.n_bands), cfg, 1, ctypes.byref(ev))
    assert rc == 0, L.fpx_last_error()
    try:
        nb, ns = int(patch.n_bands), int(patch.n_sources)
        prop = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
        chi2 = np.zeros(nb)
        resp = np.zeros(nb, dtype=response_struct_dtype)
        out = []
        for _ in range(2):
            rc = L.fpx_evaluate_host(ev, prop.ctypes.data, ns, chi2.ctypes.data, resp.ctypes.data,
                                     None)
            assert rc == 0, L.fpx_last_error()
            resid = np.zeros(patch.npix, dtype=np.float32)
            dev.d2h(resid, ptrs["residual"])
            derivs = resp["dchi2_dparam"].reshape(nb, MAXSOURCES, NPARAMS)[:, :ns]
            out.append((chi2.copy(), derivs.astype(np.float64), resid))
        # the asynchronous form with device buffers, then the status check
        d_prop = dev.malloc(prop.nbytes)
        dev.h2d(d_prop, prop)
        d_chi2 = dev.malloc(8 * nb)
        d_resp = dev.malloc(response_struct_dtype.itemsize * nb)
        assert L.fpx_evaluate(ev, d_prop, d_chi2, d_resp, None) == 0, L.fpx_last_error()
        assert L.fpx_evaluator_status(ev, None) == 0, L.fpx_last_error()
        c_async = np.zeros(nb)
        dev.d2h(c_async, d_chi2)
        r_async = np.zeros(nb, dtype=response_struct_dtype)
        dev.d2h(r_async, d_resp)
        out.append((c_async, r_async["dchi2_dparam"].reshape(nb, MAXSOURCES, NPARAMS)[:, :ns]
                    .astype(np.float64), None))
        n_tiles, used_cfg, max_gauss = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        assert L.fpx_evaluator_info(ev, ctypes.byref(n_tiles), ctypes.byref(used_cfg),
                                    ctypes.byref(max_gauss)) == 0
        assert max_gauss.value == ns * int(np.max(patch.n_psf_per_source))
        assert n_tiles.value > 0
    finally:
        L.fpx_release(ev)
    return out


@pytest.mark.parametrize("name", ["multi_small_s2", "padding_s5", "max_sources_s3"])
def test_bound_patch_binding_matches_fixture(name):
    """fpx_bind + fpx_evaluate_host through raw ctypes against the reference
    CPU kernel's fixture: chi2 to 1e-6, residual to 2e-6 of the data scale,
    gradients within the float64 band; the asynchronous fpx_evaluate gives
    the same values."""
    print(f"[gpu] bound-patch binding on {name}", flush=True)
    patch, proposal, expected = U.load_golden(name, return_residual=True)
    dev = RawDevice()
    try:
        (c1, d1, r1), (c2, d2, r2), (c3, d3, _) = _evaluate_bound(dev, patch, proposal)
    finally:
        dev.free()
    rel = np.abs(c1 - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL
    dscale = np.max(np.abs(patch.data)) + 1e-30
    assert np.max(np.abs(r1 - expected["residual"])) <= 2e-6 * dscale
    _, g64, mass, smass = U.fp64_reference(patch, proposal)
    assert U.grad_errors(d1, g64, mass, smass).max() <= 1.0
    np.testing.assert_allclose(c2, c1, rtol=1e-12)
    np.testing.assert_array_equal(r2, r1)
    np.testing.assert_allclose(c3, c1, rtol=1e-12)
    np.testing.assert_allclose(d3, d1, rtol=1e-5, atol=1e-6 * np.abs(d1).max())


def test_bind_refuses_out_of_range_radius_bin():
    """fpx_bind reads the packed metadata back and refuses a PSF entry whose
    radius bin lies outside the patch's table, before any kernel runs."""
    patch, proposal, expected = U.load_golden("single_s1", return_residual=True)
    patch.psfgauss["sersic_bin"][int(patch.psfgauss_start[0])] = int(patch.n_radii)
    L = _lib.lib()
    dev = RawDevice()
    try:
        d_patch, _ = _upload_patch(dev, patch)
        ev = ctypes.c_void_p()
        assert L.fpx_bind(d_patch, int(patch.n_bands), -1, 1, ctypes.byref(ev)) == -1
        assert b"radius_bin" in L.fpx_last_error() and ev.value is None
    finally:
        dev.free()


def test_bound_patch_host_path_on_many_tiles():
    """A patch of more than 64 tiles (8 bands x 2 exposures x 64x64 px, 256
    tiles of 256 px) takes fpx_evaluate_host's copy-and-output-kernel path
    instead of the per-tile partials of small patches; both agree with the
    torch-backed proposer."""
    import torch
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.proposal import HIPProposer
    import oracle
    print("[gpu] bound patch, 256 tiles", flush=True)
    spec = scenes.spec_multi(3, side=64)
    patch, scene, _ = scenes.build_patch(spec, SimplePatch, oracle.render_model)
    patch.return_residual = False
    patch.send_to_device()
    proposal = scene.get_proposal()
    proposal["q"] *= np.float32(0.97)
    L = _lib.lib()
    ev = ctypes.c_void_p()
    assert L.fpx_bind(patch.device_patch.data_ptr(), int(patch.n_bands), 5, 1, ctypes.byref(ev)) == 0
    try:
        n_tiles = ctypes.c_int32()
        assert L.fpx_evaluator_info(ev, ctypes.byref(n_tiles), None, None) == 0
        assert n_tiles.value > 64
        nb, ns = int(patch.n_bands), int(patch.n_sources)
        prop = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
        chi2 = np.zeros(nb)
        resp = np.zeros(nb, dtype=response_struct_dtype)
        assert L.fpx_evaluate_host(ev, prop.ctypes.data, ns, chi2.ctypes.data, resp.ctypes.data,
                                   None) == 0, L.fpx_last_error()
    finally:
        L.fpx_release(ev)
    c, d = HIPProposer(patch=patch, config=5).evaluate_proposal(proposal, patch=patch)
    np.testing.assert_allclose(chi2, c, rtol=1e-12)
    g = resp["dchi2_dparam"].reshape(nb, MAXSOURCES, NPARAMS)[:, :ns].astype(np.float64)
    assert U.bandmax_rel(g, d).max() <= 1e-6


def test_bound_patch_small_path_reports_device_radius_bin():
    """After the bind, a radius bin outside the table written into the
    device copy of the PSF block: the small-patch path's tiles leave the
    Gaussian out, store NaN, and fpx_evaluate_host returns FPX_E_ARG (the
    device status word read back); with the bin restored it is right again."""
    import torch
    print("[gpu] bound patch, device radius bin", flush=True)
    patch, proposal, expected = U.load_golden("single_s1", return_residual=False)
    patch.send_to_device()
    L = _lib.lib()
    ev = ctypes.c_void_p()
    assert L.fpx_bind(patch.device_patch.data_ptr(), 1, -1, 1, ctypes.byref(ev)) == 0
    try:
        prop = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
        chi2 = np.zeros(1)
        resp = np.zeros(1, dtype=response_struct_dtype)
        call = lambda: L.fpx_evaluate_host(ev, prop.ctypes.data, 1, chi2.ctypes.data,
                                           resp.ctypes.data, None)
        assert call() == 0
        good = chi2.copy()
        raw = patch.device_ptrs["psfgauss"]
        off = int(patch.psfgauss_start[0]) * 28 + 24
        saved = raw[off:off + 4].clone()
        raw[off:off + 4].copy_(torch.tensor([int(patch.n_radii)], dtype=torch.int32,
                                            device=raw.device).view(torch.uint8))
        torch.cuda.synchronize()
        assert call() == -1 and b"radius bin" in L.fpx_last_error()
        raw[off:off + 4].copy_(saved)
        torch.cuda.synchronize()
        assert call() == 0, L.fpx_last_error()
        np.testing.assert_allclose(chi2, good, rtol=1e-12)
        assert np.max(np.abs(good - expected["chi2"]) / np.abs(expected["chi2"])) <= U.CHI2_RTOL
    finally:
        L.fpx_release(ev)


@pytest.mark.parametrize("name", ["multi_small_s2", "padding_s5"])
def test_bound_patch_vjp_and_model_only(name):
    """The bound patch's vector-Jacobian entry (fpx_evaluate_vjp) against the
    float64 oracle's sum_pix w dmodel/dtheta for the same pixel weights, with
    chi2 written 0; its model-only entry (fpx_evaluate_model) gives
    fpx_evaluate's chi2 and residual image, with a zero gradient."""
    import oracle.fp64 as fp64
    print(f"[gpu] bound-patch vjp and model-only entries on {name}", flush=True)
    patch, proposal, expected = U.load_golden(name, return_residual=True)
    L = _lib.lib()
    dev = RawDevice()
    ev = ctypes.c_void_p()
    try:
        d_patch, ptrs = _upload_patch(dev, patch)
        assert L.fpx_bind(d_patch, int(patch.n_bands), -1, 1, ctypes.byref(ev)) == 0
        nb, ns = int(patch.n_bands), int(patch.n_sources)
        prop = np.ascontiguousarray(proposal, dtype=source_struct_dtype)
        d_prop = dev.malloc(prop.nbytes)
        dev.h2d(d_prop, prop)
        d_chi2 = dev.malloc(8 * nb)
        d_resp = dev.malloc(response_struct_dtype.itemsize * nb)

        def outputs():
            c = np.zeros(nb)
            dev.d2h(c, d_chi2)
            r = np.zeros(nb, dtype=response_struct_dtype)
            dev.d2h(r, d_resp)
            return c, r["dchi2_dparam"].reshape(nb, MAXSOURCES, NPARAMS)[:, :ns].astype(np.float64)

        def residual():
            r = np.zeros(patch.npix, dtype=np.float32)
            dev.d2h(r, ptrs["residual"])
            return r

        assert L.fpx_evaluate(ev, d_prop, d_chi2, d_resp, None) == 0, L.fpx_last_error()
        assert L.fpx_evaluator_status(ev, None) == 0
        c_full, _ = outputs()
        r_full = residual()
        # model-only: the same chi2 and residual, no gradient
        dev.h2d(ptrs["residual"], np.zeros(patch.npix, dtype=np.float32))
        assert L.fpx_evaluate_model(ev, d_prop, d_chi2, d_resp, None) == 0, L.fpx_last_error()
        assert L.fpx_evaluator_status(ev, None) == 0
        c_model, g_model = outputs()
        np.testing.assert_allclose(c_model, c_full, rtol=1e-12)
        np.testing.assert_array_equal(residual(), r_full)
        assert np.all(g_model == 0)
        # vector-Jacobian product against the float64 oracle
        w = np.linspace(-1.0, 1.0, patch.npix).astype(np.float32)
        d_w = dev.malloc(w.nbytes)
        dev.h2d(d_w, w)
        assert L.fpx_evaluate_vjp(ev, d_prop, d_w, d_chi2, d_resp, None) == 0, L.fpx_last_error()
        assert L.fpx_evaluator_status(ev, None) == 0
        c_vjp, g_vjp = outputs()
        assert np.all(c_vjp == 0)
        _, g64, _, mass = fp64.evaluate(patch, proposal, pixel_weight=w.astype(np.float64),
                                        mass=True)
        tol = U.GRAD_RTOL * np.abs(g64) + U.GRAD_FLOOR * mass
        assert np.all(np.abs(g_vjp - g64) <= tol), np.max(np.abs(g_vjp - g64) / np.maximum(tol, 1e-300))
        assert (L.fpx_evaluate_vjp(ev, d_prop, None, d_chi2, d_resp, None) == -1
                and b"pixel weights" in L.fpx_last_error())
    finally:
        if ev.value:
            L.fpx_release(ev)
        dev.free()
