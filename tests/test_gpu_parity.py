"""Parity of the HIP kernel with the reference CPU kernel (golden fixtures),
the C restatement and the float64 oracle.  Calls go through the C ABI."""

import functools

import numpy as np
import pytest

import fpx_testutil as U
import oracle

pytestmark = pytest.mark.gpu

NAMES = U.golden_names()


@pytest.fixture(scope="module")
def fp64_cache():
    return {}


def _fp64(name, patch, proposal, cache):
    if name not in cache:
        cache[name] = U.fp64_reference(patch, proposal)
    return cache[name]


def _evaluate(name, config=None, return_residual=True, persistent=True):
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden(name, return_residual=return_residual)
    patch.send_to_device()
    prop = HIPProposer(patch=patch, config=config, persistent=persistent)
    out = prop.evaluate_proposal(proposal, patch=patch, unpack=False)
    return patch, proposal, expected, out


@pytest.mark.parametrize("name", NAMES)
def test_fixture_parity(name, fp64_cache):
    print(f"[gpu] fixture {name}", flush=True)
    patch, proposal, expected, (chi2, dchi2, resid) = _evaluate(name)
    # chi2 against the reference CPU kernel
    rel = np.abs(chi2 - expected["chi2"]) / np.maximum(np.abs(expected["chi2"]), 1e-300)
    assert np.all(rel <= U.CHI2_RTOL), (name, rel)
    # gradients against the float64 oracle
    c64, g64, mass, smass = _fp64(name, patch, proposal, fp64_cache)
    err = U.grad_errors(dchi2, g64, mass, smass)
    assert err.max() <= 1.0, (name, err.max(), np.unravel_index(err.argmax(), err.shape))
    # and against the reference CPU kernel directly: within 1e-4 of each
    # band's largest gradient, and every entry within its float64 tolerance
    # band (the bound each kernel meets against float64, not their sum)
    rel_ref = U.bandmax_rel(dchi2, expected["dchi2"])
    direct = U.grad_errors(dchi2, expected["dchi2"], mass, smass)
    print(f"{name}: vs reference max {rel_ref.max():.2e} of the band's largest, "
          f"{direct.max():.3f} of the tolerance; vs float64 {err.max():.3f}", flush=True)
    assert rel_ref.max() <= U.GRAD_BANDMAX_RTOL, (name, rel_ref.max())
    assert direct.max() <= 1.0, (name, direct.max())
    # residual image against the reference, relative to the data scale
    dscale = np.max(np.abs(patch.data)) + 1e-30
    assert np.max(np.abs(resid - expected["residual"])) <= 2e-6 * dscale, name


@pytest.mark.parametrize("persistent", [True, False])
@pytest.mark.parametrize("config", range(U.N_CONFIGS))
def test_all_configs_agree(config, persistent):
    name = "multi_small_s2"
    print(f"[gpu] config {config} persistent {persistent} on {name}", flush=True)
    patch, proposal, expected, (chi2, dchi2, resid) = _evaluate(name, config=config,
                                                                persistent=persistent)
    c64, g64, mass, smass = U.fp64_reference(patch, proposal)
    rel = np.abs(chi2 - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL
    assert U.grad_errors(dchi2, g64, mass, smass).max() <= 1.0
    dscale = np.max(np.abs(patch.data)) + 1e-30
    assert np.max(np.abs(resid - expected["residual"])) <= 2e-6 * dscale


def test_autotune_keeps_results():
    """The plain proposer's autotuner leaves a (config, persistent) choice
    whose results still match the reference."""
    from forcepho_amd.proposal import HIPProposer
    name = "multi_small_s2"
    patch, proposal, expected = U.load_golden(name, return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    timings = prop.autotune(proposal, patch, repeats=2, configs=[0, 2, 5])
    assert len(timings) == 6 and all(t > 0 for t in timings.values())
    chi2, dchi2 = prop.evaluate_proposal(proposal, patch=patch)
    c64, g64, mass, smass = U.fp64_reference(patch, proposal)
    assert np.max(np.abs(chi2 - expected["chi2"]) / np.abs(expected["chi2"])) <= U.CHI2_RTOL
    assert U.grad_errors(dchi2, g64, mass, smass).max() <= 1.0


@pytest.mark.parametrize("persistent", [True, False])
def test_balanced_tiles_keep_results(persistent):
    """Reordering the tile table for load balance changes no result beyond
    summation order.  The proposal is a device tensor, so that the
    evaluations run over the proposer's own tile table (a numpy proposal on
    a patch this small takes the bound evaluator's host path); that path
    gives the same results too."""
    import torch
    from forcepho_amd.proposal import HIPProposer
    name = "multi_small_s2"
    patch, proposal, expected = U.load_golden(name, return_residual=True)
    patch.send_to_device()
    prop = HIPProposer(patch=patch, config=2, persistent=persistent)
    pdev = torch.from_numpy(np.ascontiguousarray(proposal).view(np.uint8).reshape(-1).copy()).cuda()
    before = prop.evaluate_proposal(pdev, patch=patch, unpack=False)
    # reordering alone (split tiles sum the model in another order; their
    # parity is tests/test_gpu_split.py's)
    perm = prop.balance(proposal, patch, split=False)
    n_tiles = patch.device_tiles(prop.pix_per_tile)[1]
    assert sorted(perm.tolist()) == list(range(n_tiles))
    after = prop.evaluate_proposal(pdev, patch=patch, unpack=False)
    host = prop.evaluate_proposal(proposal, patch=patch, unpack=False)
    np.testing.assert_allclose(host[0], before[0], rtol=1e-12)
    np.testing.assert_allclose(host[1], before[1], rtol=1e-5, atol=1e-6 * np.abs(before[1]).max())
    np.testing.assert_array_equal(host[2], before[2])
    np.testing.assert_allclose(after[0], before[0], rtol=1e-12)
    np.testing.assert_allclose(after[1], before[1], rtol=1e-5,
                               atol=1e-6 * np.abs(before[1]).max())
    np.testing.assert_array_equal(after[2], before[2])
    rel = np.abs(after[0] - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL


def test_repeatable_and_no_residual():
    """Two evaluations agree; without a residual buffer the same chi2 and
    gradients come back (verify_reference.py:133-134 repeatability)."""
    name = "uneven_s6"
    print(f"[gpu] repeat {name}", flush=True)
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden(name, return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    a = prop.evaluate_proposal(proposal, patch=patch)
    b = prop.evaluate_proposal(proposal, patch=patch)
    assert len(a) == 2
    np.testing.assert_allclose(a[0], b[0], rtol=1e-14)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-6, atol=1e-6 * np.abs(a[1]).max())
    rel = np.abs(a[0] - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL


def test_chi2_convention():
    """sum (r^2 - d^2) ierr^2 == chi2 (verify_reference.py:141-143) using the
    kernel's own residual image."""
    patch, proposal, expected, (chi2, dchi2, resid) = _evaluate("padding_s5")
    r = resid.astype(np.float64) * patch.ierr
    d = patch.data.astype(np.float64) * patch.ierr
    per_pix = (r - d) * (r + d)
    bands = np.split(per_pix, np.cumsum(patch.band_N_pix[:-1]))
    np.testing.assert_allclose(chi2, [x.sum() for x in bands], rtol=1e-6)


@pytest.mark.parametrize("config", range(U.N_CONFIGS))
def test_chunked_records_every_config(config):
    """30 sources: the surviving records exceed one LDS chunk for the small
    configurations, so they are built and consumed in several chunks."""
    name = "max_sources_s3"
    print(f"[gpu] chunks config {config} on {name}", flush=True)
    patch, proposal, expected, (chi2, dchi2, resid) = _evaluate(name, config=config)
    c64, g64, mass, smass = _fp64(name, patch, proposal, {})
    rel = np.abs(chi2 - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL
    assert U.grad_errors(dchi2, g64, mass, smass).max() <= 1.0
    dscale = np.max(np.abs(patch.data)) + 1e-30
    assert np.max(np.abs(resid - expected["residual"])) <= 2e-6 * dscale


def test_thirty_bands():
    """MAXBANDS bands of one small exposure each: band indexing and the
    per-band accumulators at their limit, against the C restatement."""
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    from forcepho_amd.proposal import HIPProposer
    rng = np.random.default_rng(30)
    spec = scenes.draw_patch_spec(rng, 30, [1] * 30, [16] * 30, 2, [3] * 30)
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, oracle.render_model)
    patch.send_to_device()
    proposal = scene.get_proposal()
    proposal["q"] *= np.float32(0.95)
    chi2, dchi2 = HIPProposer(patch=patch).evaluate_proposal(proposal, patch=patch)
    c32, _, _ = oracle.evaluate_cpu(patch, proposal)
    assert chi2.shape == (30,) and dchi2.shape == (30, 2, 7)
    assert np.max(np.abs(chi2 - c32) / np.abs(c32)) <= U.CHI2_RTOL
    c64, g64, mass, smass = U.fp64_reference(patch, proposal)
    assert U.grad_errors(dchi2, g64, mass, smass).max() <= 1.0


@functools.lru_cache(maxsize=1)
def _max_gaussians_case():
    """30 sources x 120 PSF entries (10 radii x 12 PSF components) in one
    24x24 exposure: 3600 Gaussians, MAXSOURCES x FPX_MAX_PSF_PER_SOURCE."""
    from forcepho_amd import scenes
    from forcepho_amd.patches import SimplePatch
    rng = np.random.default_rng(120)
    radii = 0.005 * 2.0 ** np.arange(10)
    spec = scenes.draw_patch_spec(rng, 1, [1], [24], 30, [12], radii=radii)
    patch, scene, theta = scenes.build_patch(spec, SimplePatch, oracle.render_model)
    proposal = scene.get_proposal()
    proposal["q"] *= np.float32(0.95)
    c32, _, _ = oracle.evaluate_cpu(patch, proposal)
    return patch, proposal, c32, U.fp64_reference(patch, proposal)


@pytest.mark.parametrize("config", range(U.N_CONFIGS))
def test_most_gaussians_per_exposure(config):
    """An exposure at the limit of Gaussians the cull takes (30 sources x 120
    PSF entries): several cull rounds per tile and the longest survivor list,
    in every configuration, against the C restatement and the FP64 model.
    Both the numpy path (a bound evaluator, slot by value) and a device
    proposal (the proposer's slot table)."""
    import torch
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, c32, (c64, g64, mass, smass) = _max_gaussians_case()
    assert patch.max_gauss() == 30 * 120
    print(f"[gpu] 3600 Gaussians, config {config}", flush=True)
    patch.send_to_device()
    prop = HIPProposer(patch=patch, config=config)
    pdev = torch.from_numpy(np.ascontiguousarray(proposal).view(np.uint8).reshape(-1).copy()).cuda()
    for p in (proposal, pdev):
        chi2, dchi2 = prop.evaluate_proposal(p, patch=patch)
        assert np.max(np.abs(chi2 - c32) / np.abs(c32)) <= U.CHI2_RTOL
        assert U.grad_errors(dchi2, g64, mass, smass).max() <= 1.0


@pytest.mark.parametrize("field", ["q", "pa", "ra", "dec"])
def test_nan_geometry_gives_nan_chi2(field):
    """A NaN in one source's position, axis ratio or angle gives NaN chi2 in
    every band and non-finite gradients, as the reference CPU kernel does
    (tests/test_oracle.py pins it): the cull must keep a Gaussian whose box
    it cannot decide rather than drop the source.  A clean evaluation right
    after still matches the reference (the accumulators were cleared)."""
    import torch
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden("multi_small_s2", return_residual=False)
    patch.send_to_device()
    bad = proposal.copy()
    bad[field][3] = np.nan
    as_dev = lambda p: torch.from_numpy(
        np.ascontiguousarray(p).view(np.uint8).reshape(-1).copy()).cuda()
    for config in (5, 2, 8):
        print(f"[gpu] NaN {field}, config {config}", flush=True)
        prop = HIPProposer(patch=patch, config=config)
        for arg in (bad, as_dev(bad)):
            chi2, dchi2 = prop.evaluate_proposal(arg, patch=patch)
            assert np.all(np.isnan(chi2)), (config, chi2)
            assert not np.all(np.isfinite(dchi2))
        for arg in (proposal, as_dev(proposal)):
            chi2, dchi2 = prop.evaluate_proposal(arg, patch=patch)
            rel = np.abs(chi2 - expected["chi2"]) / np.abs(expected["chi2"])
            assert rel.max() <= U.CHI2_RTOL


def test_launch_without_tiles_writes_zeros():
    """A launch with no tiles (no pixels) still writes the outputs: chi2 and
    gradients of an empty pixel set are zero."""
    import torch
    from forcepho_amd._lib import lib, check
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden("single_s1", return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch, persistent=False)
    prop.launch(proposal, patch)           # builds the proposer's device slot
    torch.cuda.synchronize()
    assert np.all(prop._chi2[:int(patch.n_bands)].cpu().numpy() != 0)
    check(lib().fpx_evaluate_proposal(prop._slot.data_ptr(), 0, patch.max_gauss(), prop.config,
                                      torch.cuda.current_stream().cuda_stream), "empty launch")
    torch.cuda.synchronize()
    assert np.all(prop._chi2[:int(patch.n_bands)].cpu().numpy() == 0)
    assert np.all(prop._resp.cpu().numpy() == 0)


def test_finalize_needs_a_launch():
    """The output pass reads the slot a launch built: before any launch (a
    numpy evaluation of a small patch builds none) finalize refuses on the
    host instead of launching over an empty slot."""
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden("single_s1", return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    prop.evaluate_proposal(proposal, patch=patch)
    with pytest.raises(RuntimeError, match="launch first"):
        prop.finalize()


@pytest.mark.parametrize("persistent", [True, False])
def test_tiles_then_finalize(persistent):
    """fpx_evaluate_tiles accumulates, fpx_finalize writes and clears: two tile
    launches then one output pass give twice one evaluation, and the next
    full evaluation is back to one."""
    import torch
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden("multi_small_s2", return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch, persistent=persistent)
    chi2, d = prop.evaluate_proposal(proposal, patch=patch)
    prop.launch(proposal, patch, outputs=False)
    prop.launch(proposal, patch, outputs=False)
    prop.finalize()
    torch.cuda.synchronize()
    nb, ns = int(patch.n_bands), int(patch.n_sources)
    chi2_2 = prop._chi2[:nb].cpu().numpy()
    d_2 = prop._resp.view(-1, 30, 7)[:nb, :ns].cpu().numpy()
    np.testing.assert_allclose(chi2_2, 2 * chi2, rtol=1e-12)
    np.testing.assert_allclose(d_2, 2 * d, rtol=1e-5, atol=1e-6 * np.abs(d).max())
    again, _ = prop.evaluate_proposal(proposal, patch=patch)
    np.testing.assert_allclose(again, chi2, rtol=1e-12)


def test_full_size_multi_band_against_fp64():
    """The 8-band x 2-exposure x 128x128, 15-source benchmark patch at full
    size (1.4e8 pairs): chi2 and gradients against the float64 oracle's
    values committed in tests/golden/fullsize (tests/golden/make_fullsize.py),
    with the same tolerance band as the fixtures."""
    import sys
    from forcepho_amd.proposal import HIPProposer
    sys.path.insert(0, U.GOLDEN_DIR)
    import make_fullsize as F
    print("[gpu] full-size multi-band", flush=True)
    ref = np.load(F.OUT)
    patch, proposal = F.build()
    np.testing.assert_allclose(F.data_digest(patch), ref["digest"], rtol=1e-12)
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    prop.autotune(proposal, patch, repeats=1, configs=[2, 5])
    chi2, d = prop.evaluate_proposal(proposal, patch=patch)
    assert np.max(np.abs(chi2 - ref["chi2"]) / np.abs(ref["chi2"])) <= U.CHI2_RTOL
    err = U.grad_errors(d, ref["dchi2"], ref["mass"], ref["smass"])
    print("full-size gradient error / tolerance: max", err.max(), "median", np.median(err))
    assert err.max() <= 1.0
    # the reference's own CPU kernel on the same patch and proposal: the GPU
    # within 1e-4 of each band's largest gradient of it, and the GPU's error
    # tail against float64 no worse than twice the reference kernel's own
    mod = oracle.reference_module()
    if mod is None:
        pytest.skip("the reference CPU kernel is not built (oracle/_ref)")
    c_ref, d_ref, _ = oracle.evaluate_reference(patch, proposal, module=mod)
    d_ref = d_ref.astype(np.float64)
    g64 = ref["dchi2"]
    assert U.bandmax_rel(d, d_ref).max() <= U.GRAD_BANDMAX_RTOL
    rel = lambda g: np.abs(g - g64) / np.where(g64 != 0, np.abs(g64), 1e-300)
    q = [50, 99]
    gpu_tail, ref_tail = np.percentile(rel(d), q), np.percentile(rel(d_ref), q)
    gpu_band, ref_band = U.bandmax_rel(d, g64).max(), U.bandmax_rel(d_ref, g64).max()
    gpu_tol = err.max()
    ref_tol = U.grad_errors(d_ref, g64, ref["mass"], ref["smass"]).max()
    print(f"per-entry relative error p50/p99: GPU {gpu_tail}, reference {ref_tail}; "
          f"of the band's largest: GPU {gpu_band:.2e}, reference {ref_band:.2e}; "
          f"of the tolerance: GPU {gpu_tol:.3f}, reference {ref_tol:.3f}", flush=True)
    assert gpu_tail[1] <= 2.0 * ref_tail[1]
    assert gpu_band <= 2.0 * ref_band
    assert gpu_tol <= max(2.0 * ref_tol, 0.5)


@pytest.mark.parametrize("name", ["multi_small_s2", "padding_s5"])
def test_pixel_order_independence(name):
    """Pixels shuffled at random within each exposure (tiles then hold
    scattered pixels): chi2, gradients and the residual image (put back in
    the original order) agree with the packed order, up to float32
    summation order."""
    from forcepho_amd.proposal import HIPProposer
    patch, proposal, expected = U.load_golden(name, return_residual=True)
    patch.send_to_device()
    c0, d0, r0 = HIPProposer(patch=patch).evaluate_proposal(proposal, patch=patch, unpack=False)
    shuffled, _, _ = U.load_golden(name, return_residual=True)
    rng = np.random.default_rng(4)
    perm = np.arange(shuffled.npix)
    for e in range(int(shuffled.n_exp)):
        s, n = int(shuffled.exposure_start[e]), int(shuffled.exposure_N[e])
        perm[s:s + n] = s + rng.permutation(n)
    for f in ("xpix", "ypix", "data", "ierr"):
        setattr(shuffled, f, np.ascontiguousarray(getattr(shuffled, f)[perm]))
    shuffled.send_to_device()
    c1, d1, r1 = HIPProposer(patch=shuffled).evaluate_proposal(proposal, patch=shuffled,
                                                                unpack=False)
    np.testing.assert_allclose(c1, c0, rtol=1e-6)
    _, g64, mass, smass = U.fp64_reference(patch, proposal)
    assert U.grad_errors(d1, g64, mass, smass).max() <= 1.0
    assert np.abs(d1 - d0).max() <= 1e-4 * np.abs(d0).max()
    back = np.empty_like(r1)
    back[perm] = r1
    assert np.abs(back - r0).max() <= 2e-6 * np.abs(patch.data).max()


@pytest.mark.parametrize("persistent", [True, False])
def test_too_small_max_gauss_is_an_error(persistent):
    """A max_gauss below n_sources x n_psf_per_source makes the kernel leave
    every tile out: fpx_evaluate_sync returns FPX_E_ARG and the chi2 it
    copied back is NaN, not a chi2 missing tiles.  The status word is
    cleared, so the next evaluation is right again."""
    import torch
    from forcepho_amd._lib import lib
    from forcepho_amd.proposal import HIPProposer, _CHI2_END
    print(f"[gpu] too-small max_gauss, persistent {persistent}", flush=True)
    patch, proposal, expected = U.load_golden("multi_small_s2", return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch, persistent=persistent)
    chi2, d = prop.evaluate_proposal(proposal, patch=patch)
    # the device slot fpx_evaluate_sync is given below is built by a launch
    # (a numpy evaluation of a patch this small takes the bound evaluator's
    # host path and builds none)
    prop.launch(proposal, patch)
    torch.cuda.synchronize()
    n_tiles = patch.device_tiles(prop.pix_per_tile)[1]
    nb = int(patch.n_bands)
    raw = np.ascontiguousarray(proposal).view(np.uint8)
    L = lib()
    rc = L.fpx_evaluate_sync(prop._slot.data_ptr(), n_tiles, patch.max_gauss() - 1, prop.config,
                             int(persistent), prop._queue.data_ptr(), prop._prop.data_ptr(),
                             raw.ctypes.data, raw.size, prop._out.data_ptr(),
                             prop._out_h.data_ptr(), _CHI2_END, prop._out.data_ptr(),
                             torch.cuda.current_stream().cuda_stream)
    assert rc == -1, rc
    assert b"size limit" in L.fpx_last_error()
    assert np.all(np.isnan(prop._chi2_np[:nb]))
    assert int(prop._status.cpu().item()) == 0          # cleared by the call
    again, d2 = prop.evaluate_proposal(proposal, patch=patch)
    np.testing.assert_allclose(again, chi2, rtol=1e-12)
    rel = np.abs(again - expected["chi2"]) / np.abs(expected["chi2"])
    assert rel.max() <= U.CHI2_RTOL


def test_device_radius_bin_out_of_range_is_an_error():
    """A radius bin outside [0, n_radii) in the device copy of the PSF block
    (the host copy, which check_limits reads, is valid) is caught by the
    kernel's own check: the Gaussian is left out, the status word set, and
    evaluate_proposal raises; batched slots report it the same way."""
    import torch
    from forcepho_amd._lib import FpxError
    from forcepho_amd.batch import BatchEvaluator
    from forcepho_amd.proposal import HIPProposer
    print("[gpu] device radius bin out of range", flush=True)
    patch, proposal, expected = U.load_golden("single_s1", return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    good, _ = prop.evaluate_proposal(proposal, patch=patch)
    raw = patch.device_ptrs["psfgauss"]
    off = int(patch.psfgauss_start[0]) * 28 + 24        # sersic_radius_bin of the first entry
    saved = raw[off:off + 4].clone()
    raw[off:off + 4].copy_(torch.tensor([int(patch.n_radii)], dtype=torch.int32,
                                        device=raw.device).view(torch.uint8))
    with pytest.raises(FpxError, match="left tiles out"):
        prop.evaluate_proposal(proposal, patch=patch)
    ev = BatchEvaluator([patch], chains=2)
    ev.set_proposals([proposal, proposal])
    ev.launch()
    with pytest.raises(FpxError, match="slots"):
        ev.results()
    raw[off:off + 4].copy_(saved)
    again, _ = prop.evaluate_proposal(proposal, patch=patch)
    np.testing.assert_allclose(again, good, rtol=1e-12)
    ev.launch()
    (c0, _), (c1, _) = ev.results()
    np.testing.assert_allclose(c0, good, rtol=1e-12)
