"""The one-kernel paths of single-patch evaluations, against the launch over
the proposer's own slot:

* evaluate_proposal with a numpy proposal on a patch of at most
  HOST_PATH_TILES tiles goes through the patch's bound evaluator
  (fpx_evaluate_host: the proposal read from mapped host memory, per-tile
  sums left there for the host, completion seen through a mapped flag).
  Many calls in a row with a different proposal each must each match the
  device-proposal launch of the same proposal: a result read before the
  kernel had written it, or a stale proposal, shows up as a mismatch.
* Bound evaluators created and released one after the other (their mapped
  blocks are recycled by the runtime) each give the fixture's chi2.
* A launch of one slot writes its outputs from its last workgroup; a
  BatchEvaluator of one patch and one chain takes that path too and must
  agree with the proposer."""

import numpy as np
import pytest

import fpx_testutil as U

pytestmark = pytest.mark.gpu


def _device(proposal):
    import torch
    raw = np.ascontiguousarray(proposal).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).cuda()


@pytest.mark.parametrize("name", ["single_s1", "multi_small_s2", "max_sources_s3"])
def test_host_path_calls_in_a_row(name):
    from forcepho_amd.proposal import HIPProposer, HOST_PATH_TILES
    print(f"[gpu] host path, 60 proposals in a row on {name}", flush=True)
    patch, proposal, expected = U.load_golden(name, return_residual=False)
    patch.send_to_device()
    prop = HIPProposer(patch=patch)
    assert patch.device_tiles(prop.pix_per_tile)[1] <= HOST_PATH_TILES
    rng = np.random.default_rng(7)
    for k in range(60):
        p = proposal.copy()
        p["q"] *= np.float32(1.0 + 0.05 * rng.uniform(-1, 1))
        p["fluxes"] *= np.float32(1.0 + 0.2 * rng.uniform(-1, 1))
        host = prop.evaluate_proposal(p, patch=patch)
        dev = prop.evaluate_proposal(_device(p), patch=patch)
        np.testing.assert_allclose(host[0], dev[0], rtol=1e-12, err_msg=f"call {k}")
        assert U.bandmax_rel(host[1], dev[1]).max() <= 1e-6, k


def test_evaluators_bound_one_after_another():
    from forcepho_amd.proposal import HIPProposer
    print("[gpu] fresh bound evaluators in a row", flush=True)
    for k in range(12):
        name = ("single_s1", "multi_small_s2", "padding_s5")[k % 3]
        patch, proposal, expected = U.load_golden(name, return_residual=False)
        patch.send_to_device()
        chi2, _ = HIPProposer(patch=patch).evaluate_proposal(proposal, patch=patch)
        rel = np.abs(chi2 - expected["chi2"]) / np.abs(expected["chi2"])
        assert rel.max() <= U.CHI2_RTOL, (k, name, rel.max())
        patch.free()


@pytest.mark.parametrize("persistent", [True, False])
def test_one_slot_batch_matches_proposer(persistent):
    from forcepho_amd.batch import BatchEvaluator
    from forcepho_amd.proposal import HIPProposer
    print(f"[gpu] one-slot batch, persistent {persistent}", flush=True)
    patch, proposal, expected = U.load_golden("multi_small_s2", return_residual=False)
    patch.send_to_device()
    ref = HIPProposer(patch=patch).evaluate_proposal(_device(proposal), patch=patch)
    ev = BatchEvaluator([patch], chains=1, persistent=persistent)
    ev.set_proposals([proposal])
    for _ in range(3):   # the last workgroup leaves the queue and accumulators zero
        ev.launch()
        (c, d), = ev.results()
        np.testing.assert_allclose(c, ref[0], rtol=1e-12)
        assert U.bandmax_rel(d, ref[1]).max() <= 1e-6
