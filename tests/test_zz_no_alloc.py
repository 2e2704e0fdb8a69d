"""fpx_evaluate_static, the allocation-free evaluation, through the C ABI:
two calls in a row on one GPU (the 8-band multi_small patch, then the
1-band single patch: the second call finds the first one's scratch
contents) against fpx_evaluate on a bound patch (the entry point it is the
allocation-free form of) and against the reference CPU kernel's fixtures.
Device memory for inputs and outputs comes from torch; the entry point
itself allocates nothing."""

import ctypes

import numpy as np
import pytest

import fpx_testutil as U
from forcepho_amd import _lib
from forcepho_amd.kernel_limits import MAXBANDS, MAXSOURCES, NPARAMS
from forcepho_amd.proposal import source_struct_dtype


def test_static_entry_refuses_too_many_bands():
    """More bands than the static scratch holds is refused before any
    device work (no GPU needed)."""
    L = _lib.lib()
    p = ctypes.c_void_p(8)
    rc = L.fpx_evaluate_static(p, p, 1, MAXBANDS + 1, 27, p, p, p, -1, 1, None)
    assert rc == -1 and b"MAXBANDS" in L.fpx_last_error()
    assert L.fpx_evaluate_static(None, p, 1, 1, 27, p, p, p, -1, 1, None) == -1
    assert L.fpx_evaluate_static(p, None, 1, 1, 27, p, p, p, -1, 1, None) == -1
    assert L.fpx_evaluate_static(p, p, 1, 1, 27, p, p, p, 99, 1, None) == -1


def _both(name, cfg, persistent):
    """chi2 and gradients of fixture `name` from fpx_evaluate (bound patch)
    and from fpx_evaluate_static (caller's tile table, static scratch)."""
    import torch
    L = _lib.lib()
    patch, proposal, expected = U.load_golden(name, return_residual=False)
    patch.send_to_device()
    nb, ns = int(patch.n_bands), int(patch.n_sources)
    stream = torch.cuda.current_stream().cuda_stream
    raw = np.ascontiguousarray(proposal, dtype=source_struct_dtype).view(np.uint8).reshape(-1)
    prop = torch.from_numpy(raw.copy()).cuda()
    f64 = dict(dtype=torch.float64, device="cuda")
    f32 = dict(dtype=torch.float32, device="cuda")
    # the existing entry point: a bound patch
    ev = ctypes.c_void_p()
    assert L.fpx_bind(patch.device_patch.data_ptr(), nb, cfg, persistent, ctypes.byref(ev)) == 0
    try:
        chi2_a = torch.zeros(nb, **f64)
        resp_a = torch.zeros(nb * MAXSOURCES * NPARAMS, **f32)
        assert L.fpx_evaluate(ev, prop.data_ptr(), chi2_a.data_ptr(), resp_a.data_ptr(),
                              stream) == 0, L.fpx_last_error()
        assert L.fpx_evaluator_status(ev, stream) == 0, L.fpx_last_error()
    finally:
        L.fpx_release(ev)
    # the allocation-free form: the caller plans and holds the tile table
    ppt = ctypes.c_int32()
    assert L.fpx_config_info(cfg, None, None, ctypes.byref(ppt)) == 0
    idx = [np.ascontiguousarray(patch.exposure_start, np.int32),
           np.ascontiguousarray(patch.exposure_N, np.int32),
           np.ascontiguousarray(patch.band_start, np.int16),
           np.ascontiguousarray(patch.band_N, np.int16)]
    args = [a.ctypes.data for a in idx] + [nb, ppt.value]
    n_tiles = L.fpx_plan_tiles(*args, None, 0)
    tiles = np.zeros(n_tiles, dtype=_lib.TILE_DTYPE)
    assert L.fpx_plan_tiles(*args, tiles.ctypes.data, n_tiles) == n_tiles
    tiles_dev = torch.from_numpy(tiles.view(np.uint8).reshape(-1).copy()).cuda()
    chi2_b = torch.full((nb,), -1.0, **f64)
    resp_b = torch.full((nb * MAXSOURCES * NPARAMS,), -1.0, **f32)
    rc = L.fpx_evaluate_static(patch.device_patch.data_ptr(), tiles_dev.data_ptr(), n_tiles, nb,
                               patch.max_gauss(), prop.data_ptr(), chi2_b.data_ptr(),
                               resp_b.data_ptr(), cfg, persistent, stream)
    assert rc == 0, L.fpx_last_error()
    torch.cuda.synchronize()
    shape = (nb, MAXSOURCES, NPARAMS)
    out = [(c.cpu().numpy(), r.cpu().numpy().reshape(shape)[:, :ns].astype(np.float64))
           for c, r in ((chi2_a, resp_a), (chi2_b, resp_b))]
    return patch, proposal, expected, out


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,persistent", [(5, 1), (2, 0)])
def test_static_entry_matches_bound_evaluation(cfg, persistent):
    print(f"[gpu] fpx_evaluate_static, config {cfg} persistent {persistent}", flush=True)
    for name in ("multi_small_s2", "single_s1"):          # the second call is smaller
        patch, proposal, expected, ((ca, da), (cb, db)) = _both(name, cfg, persistent)
        np.testing.assert_allclose(cb, ca, rtol=1e-12)
        assert U.bandmax_rel(db, da).max() <= 1e-6, name
        # to the parity bar against the reference CPU kernel and float64
        assert np.max(np.abs(cb - expected["chi2"]) / np.abs(expected["chi2"])) <= U.CHI2_RTOL
        assert U.bandmax_rel(db, expected["dchi2"]).max() <= U.GRAD_BANDMAX_RTOL
        _, g64, mass, smass = U.fp64_reference(patch, proposal)
        assert U.grad_errors(db, g64, mass, smass).max() <= 1.0, name
