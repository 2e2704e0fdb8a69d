"""The C ABI library loads without a GPU, exports every declared symbol, and
its wire structs match the host dtypes and the reference layouts."""

import ctypes
import os
import re

import numpy as np

from forcepho_amd import _lib, kernel_limits
from forcepho_amd.patches import PatchBase, PSF_DTYPE
from forcepho_amd.proposal import source_struct_dtype, response_struct_dtype

import fpx_testutil as U

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "forcepho_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fpx_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported_and_bound():
    L = _lib.lib()
    decl = declared_symbols()
    assert len(decl) >= 12
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTED_SYMBOLS) == decl


def test_struct_sizes_match_dtypes():
    L = _lib.lib()
    assert L.fpx_abi_version() == 3
    assert L.fpx_struct_size(_lib.STRUCT_PATCH) == PatchBase.pdt.itemsize == 152
    assert L.fpx_struct_size(_lib.STRUCT_SOURCE) == source_struct_dtype.itemsize == 264
    assert L.fpx_struct_size(_lib.STRUCT_RESPONSE) == response_struct_dtype.itemsize == 840
    assert L.fpx_struct_size(_lib.STRUCT_PSFGAUSS) == PSF_DTYPE.itemsize == 28
    assert L.fpx_struct_size(_lib.STRUCT_TILE) == _lib.TILE_DTYPE.itemsize == 16
    assert L.fpx_struct_size(_lib.STRUCT_SLOT) == _lib.SLOT_DTYPE.itemsize == 88
    assert _lib.SLOT_DTYPE.fields["split"][1] == 80
    assert _lib.TILE_DTYPE.fields["part"][1] == 6 and _lib.TILE_DTYPE.fields["parts"][1] == 7
    assert L.fpx_struct_size(_lib.STRUCT_AMPMODEL) == ctypes.sizeof(_lib.AmplitudeModel) == 88
    assert _lib.AmplitudeModel.tx.offset == 48 and _lib.AmplitudeModel.coef.offset == 80
    assert L.fpx_struct_size(99) == 0


def test_reference_field_offsets():
    """Offsets of the reference's Patch and Source structs
    (forcepho/src/patch.cc:23-100, forcepho/src/proposal.cu:13-27)."""
    pdt = PatchBase.pdt
    expect = {"data": 0, "ierr": 8, "xpix": 16, "ypix": 24, "residual": 32,
              "exposure_start": 40, "exposure_N": 48, "band_start": 56, "band_N": 64,
              "n_sources": 72, "n_radii": 76, "rad2": 80, "D": 88, "crpix": 96,
              "crval": 104, "CW": 112, "G": 120, "n_psf_per_source": 128, "psfgauss": 136,
              "psfgauss_start": 144}
    assert {k: pdt.fields[k][1] for k in pdt.names} == expect
    sdt = source_struct_dtype
    assert {k: sdt.fields[k][1] for k in sdt.names} == {
        "ra": 0, "dec": 4, "q": 8, "pa": 12, "nsersic": 16, "rh": 20, "fluxes": 24,
        "mixture_amplitudes": 144, "damplitude_drh": 184, "damplitude_dnsersic": 224}


def test_kernel_limits_parsed():
    assert kernel_limits.MAXBANDS == 30
    assert kernel_limits.MAXSOURCES == 30
    assert kernel_limits.NPARAMS == 7
    assert kernel_limits.MAX_EXP_ARG == 36.0
    assert kernel_limits.MAXRADII == 10
    assert kernel_limits.FPX_WAVE == 64


def _plan(exp_start, exp_n, band_start, band_n, ppt):
    L = _lib.lib()
    arrs = [np.asarray(exp_start, np.int32), np.asarray(exp_n, np.int32),
            np.asarray(band_start, np.int16), np.asarray(band_n, np.int16)]
    ptrs = [a.ctypes.data for a in arrs]
    n = L.fpx_plan_tiles(*ptrs, len(band_n), ppt, None, 0)
    if n < 0:
        return n, None
    tiles = np.zeros(max(n, 1), dtype=_lib.TILE_DTYPE)
    assert L.fpx_plan_tiles(*ptrs, len(band_n), ppt, tiles.ctypes.data, n) == n
    return n, tiles[:n]


def test_plan_tiles_covers_every_pixel_once():
    exp_n = [64, 192, 320, 128]
    exp_start = np.concatenate([[0], np.cumsum(exp_n)[:-1]])
    n, tiles = _plan(exp_start, exp_n, [0, 1], [1, 3], 128)
    assert n == 1 + 2 + 3 + 1
    covered = np.zeros(sum(exp_n), int)
    for t in tiles:
        assert 0 < t["pix_count"] <= 128
        s = t["pix_start"]
        covered[s:s + t["pix_count"]] += 1
        e = t["exposure"]
        assert exp_start[e] <= s and s + t["pix_count"] <= exp_start[e] + exp_n[e]
        assert t["band"] == (0 if e == 0 else 1)
        assert t["part"] == 0 and t["parts"] == 0      # whole tiles
    assert np.all(covered == 1)


def test_split_bytes():
    """Split buffer: a counter per table entry (padded to 256 bytes) and a
    partial-model row of pixels-per-tile floats per entry."""
    L = _lib.lib()
    for cfg in range(L.fpx_n_configs()):
        ppt = ctypes.c_int32()
        assert L.fpx_config_info(cfg, None, None, ctypes.byref(ppt)) == 0
        for n in (1, 63, 64, 65, 1000):
            assert L.fpx_split_bytes(n, cfg) == -(-4 * n // 256) * 256 + 4 * n * ppt.value
    assert L.fpx_split_bytes(0, 0) == 0
    assert L.fpx_split_bytes(10, -1) == 0
    assert L.fpx_split_bytes(10, L.fpx_n_configs()) == 0


def test_plan_tiles_rejects_bad_sizes():
    assert _plan([0], [64], [0], [1], 100)[0] < 0        # not a multiple of 64
    assert _plan([0], [64], [0], [1], 0)[0] < 0
    L = _lib.lib()
    assert b"multiple" in L.fpx_last_error()
    one = np.zeros(1, np.int32)
    ptr = one.ctypes.data
    assert L.fpx_plan_tiles(ptr, ptr, None, ptr, 1, 64, None, 0) < 0


def test_configs_and_lds():
    L = _lib.lib()
    n = L.fpx_n_configs()
    assert n == U.N_CONFIGS
    w, p, t = (ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32())
    for cfg in range(n):
        assert L.fpx_config_info(cfg, ctypes.byref(w), ctypes.byref(p), ctypes.byref(t)) == 0
        assert t.value == 64 * p.value and w.value >= 1
        size, small = ctypes.c_size_t(), ctypes.c_size_t()
        assert L.fpx_lds_bytes(cfg, 540, ctypes.byref(size)) == 0
        assert L.fpx_lds_bytes(cfg, 27, ctypes.byref(small)) == 0
        # a workgroup stays within a CU's 160 KiB; few Gaussians need less
        assert 27 * 36 * 4 < small.value <= size.value <= 160 * 1024
        # the budget is 10 KiB per wavefront unless the records must spill over it
        assert size.value <= max(10 * 1024 * w.value, small.value + 16 * 36 * 4)
    assert L.fpx_config_info(n, None, None, None) < 0
    assert L.fpx_accum_doubles(8) == 8 * (1 + 30 * 7)


def test_launch_argument_checks_need_no_device():
    """Invalid arguments are rejected before any HIP call."""
    L = _lib.lib()
    assert L.fpx_evaluate_proposal(None, 1, 10, 0, None) < 0
    assert L.fpx_evaluate_batch(ctypes.c_void_p(8), 2, None, 4, 10, 0, 0, None, None) < 0
    assert b"item_offsets" in L.fpx_last_error()
    assert L.fpx_evaluate_batch(ctypes.c_void_p(8), 1, None, 4, 10, 99, 0, None, None) < 0
    assert L.fpx_evaluate_batch(ctypes.c_void_p(8), 1, None, 4, 0, 0, 0, None, None) < 0
    assert L.fpx_evaluate_batch(ctypes.c_void_p(8), 1, None, 4, 10, 0, 1, None, None) < 0


def test_evaluate_sync_argument_checks():
    """fpx_evaluate_sync refuses missing buffers and oversized copies before
    any device work (no GPU needed)."""
    L = _lib.lib()
    p = ctypes.c_void_p(8)

    def call(prop_dev=p, prop_host=p, nbytes=264, out_dev=p, out_host=p, out_bytes=8):
        rc = L.fpx_evaluate_sync(p, 1, 1, 0, 0, None, prop_dev, prop_host, nbytes, out_dev,
                                 out_host, out_bytes, None, None)
        return rc, L.fpx_last_error().decode()

    assert call(prop_dev=None) == (-1, "NULL proposal or output buffer")
    assert call(out_host=None)[0] == -1
    assert call(nbytes=264 * 31) == (-1, "proposal larger than MAXSOURCES records")
    assert call(out_bytes=0) == (-1, "no output bytes to copy back")


def test_evaluator_argument_checks_need_no_device():
    """The bound-patch entry points refuse NULL handles and out-of-range
    arguments before any HIP call."""
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.fpx_bind(None, 1, -1, 1, ctypes.byref(out)) == -1
    assert L.fpx_last_error() == b"NULL patch" and out.value is None
    assert L.fpx_bind(ctypes.c_void_p(8), 0, -1, 1, ctypes.byref(out)) == -1
    assert b"n_bands" in L.fpx_last_error()
    assert L.fpx_bind(ctypes.c_void_p(8), 31, -1, 1, ctypes.byref(out)) == -1
    assert L.fpx_bind(ctypes.c_void_p(8), 2, 99, 1, ctypes.byref(out)) == -1
    assert b"configuration" in L.fpx_last_error()
    assert L.fpx_bind(ctypes.c_void_p(8), 2, -1, 1, None) == -1
    p = ctypes.c_void_p(8)
    assert L.fpx_evaluate(None, p, p, p, None) == -1
    assert L.fpx_evaluate_host(None, p, 1, p, p, None) == -1
    assert L.fpx_evaluator_status(None, None) == -1
    assert L.fpx_evaluator_info(None, None, None, None) == -1
    assert L.fpx_release(None) == 0


def test_patch_limit_checks():
    """check_limits refuses what the kernel would leave out: too many PSF
    Gaussians per source, a radius bin outside the table, too many sources
    or radii; a packed synthetic patch passes."""
    import copy
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    import oracle
    spec = scenes.spec_single(3)
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, oracle.render_model)
    patch.check_limits()
    bad = copy.deepcopy(patch)
    bad.psfgauss["sersic_bin"][int(bad.psfgauss_start[0])] = bad.n_radii
    with np.testing.assert_raises_regex(ValueError, "radius bins"):
        bad.check_limits()
    bad = copy.deepcopy(patch)
    bad.psfgauss["sersic_bin"][int(bad.psfgauss_start[0])] = -1
    with np.testing.assert_raises_regex(ValueError, "radius bins"):
        bad.check_limits()
    bad = copy.deepcopy(patch)
    bad.n_psf_per_source = np.array([kernel_limits.FPX_MAX_PSF_PER_SOURCE + 1], np.int32)
    with np.testing.assert_raises_regex(ValueError, "n_psf_per_source"):
        bad.check_limits()
    bad = copy.deepcopy(patch)
    bad.n_sources = kernel_limits.MAXSOURCES + 1
    with np.testing.assert_raises_regex(ValueError, "n_sources"):
        bad.check_limits()
    bad = copy.deepcopy(patch)
    bad.n_radii = kernel_limits.MAXRADII + 1
    with np.testing.assert_raises_regex(ValueError, "n_radii"):
        bad.check_limits()
