"""The committed rocprof summaries that bench.py reads hold a record for every
kernel configuration the autotuner may select on the multi-band workload
(HBM traffic and issued VALU work), and for the configurations the batched
workload selects; the records agree with each other where they overlap."""

import os
import sys

import pytest

from forcepho_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_every_configuration_has_multi_band_records():
    for cfg in range(_lib.lib().fpx_n_configs()):
        assert bench.traffic(cfg) is not None, cfg
        e = bench.executed(cfg, 1e-4)
        assert e is not None, cfg
        assert e["flop_per_launch"] > 0 and e["exp_per_launch"] > 0


@pytest.mark.parametrize("path", [bench.VALU_FILE, bench.BATCH_VALU_FILE])
def test_executed_work_fields(path):
    """Issued flop = 64 lanes x (2 FMA + MUL + ADD) and exponentials = 64 x
    TRANS, from the raw counters of the same record; FP32 arithmetic is a
    part of all VALU instructions."""
    import json
    rows = [json.loads(l) for l in open(path) if l.startswith("{")]
    assert rows
    for r in rows:
        assert r["executed_flop_per_launch"] == pytest.approx(
            64 * (2 * r["fma"] + r["mul"] + r["add"]), rel=1e-6)
        assert r["exp_per_launch"] == pytest.approx(64 * r["trans"], rel=1e-6)
        assert r["fma"] + r["mul"] + r["add"] + r["trans"] < r["valu"]


def test_batched_records_cover_the_selected_shapes():
    # the batched launch selects 2x4 (config 6) or 4x4 / 8x4 on this workload
    for cfg in (2, 5, 6):
        assert bench.traffic(cfg, bench.BATCH_TRAFFIC_FILE) is not None, cfg
        assert bench.executed(cfg, 1e-2, bench.BATCH_VALU_FILE) is not None, cfg
