"""Split tiles on the host: the queue simulation, the choice of pieces and the
table layout the split kernel variant reads (include/forcepho_amd.h FpxTile).
No GPU needed."""

import numpy as np
import pytest

from forcepho_amd import tile_split
from forcepho_amd._lib import TILE_DTYPE, MAX_TILE_PARTS


def _tiles(n, ppt=256):
    t = np.zeros(n, dtype=TILE_DTYPE)
    t["exposure"] = np.arange(n) // 8
    t["band"] = np.arange(n) // 16
    t["pix_start"] = np.arange(n) * ppt
    t["pix_count"] = ppt
    return t


def test_list_schedule_known_answers():
    # one worker: the sum; as many workers as items: the largest
    assert tile_split.list_schedule([3, 1, 2], 1) == 6
    assert tile_split.list_schedule([3, 1, 2], 3) == 3
    assert tile_split.list_schedule([], 4) == 0
    # queue order matters: 2 workers, [1, 1, 2] -> 3, [2, 1, 1] -> 2
    assert tile_split.list_schedule([1, 1, 2], 2) == 3
    assert tile_split.list_schedule([2, 1, 1], 2) == 2


def test_list_schedule_matches_brute_force_simulation():
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = rng.uniform(0, 10, rng.integers(1, 40))
        g = int(rng.integers(1, 6))
        # the first free worker takes the next item (lowest index on ties)
        free = np.zeros(g)
        end = 0.0
        for x in c:
            w = int(np.argmin(free))
            free[w] += x
            end = max(end, free[w])
        assert tile_split.list_schedule(c, g) == pytest.approx(end)


def test_plan_parts_leaves_even_work_whole():
    cost = np.full(1024, 5e4)
    assert np.all(tile_split.plan_parts(cost, 512, 5000) == 1)
    assert np.all(tile_split.plan_parts(cost[:3], 1, 5000) == 1)   # one worker: no gain


def test_plan_parts_splits_the_tiles_that_hold_the_launch_back():
    # 1024 tiles on 512 workers, 64 of them 4x the others: whole, the launch
    # takes the heavy tile's time; split, they spread over the idle workers
    cost = np.full(1024, 5e4)
    cost[:64] *= 4
    parts = tile_split.plan_parts(cost, 512, 5000)
    assert np.all(parts[:64] > 1) and np.all(parts <= 8)
    whole = tile_split.list_schedule(np.sort(cost)[::-1], 512)
    order, piece = tile_split._queue(cost, parts, 5000)
    split = tile_split.list_schedule(np.repeat(piece[order], parts[order]), 512)
    assert split < 0.8 * whole


def test_plan_parts_respects_overhead():
    # when a piece's fixed cost exceeds what splitting saves, nothing is split
    cost = np.full(1024, 5e4)
    cost[:64] *= 4
    assert np.all(tile_split.plan_parts(cost, 512, 1e6) == 1)


def test_split_table_layout():
    tiles = _tiles(10)
    cost = np.arange(10, dtype=float) * 1e4 + 1e3
    parts = np.array([1, 1, 1, 1, 1, 1, 1, 2, 3, 5])
    out = tile_split.split_table(tiles, cost, parts, 1000)
    assert len(out) == parts.sum()
    # every tile's pixels appear in exactly `parts` entries, pieces in a row
    k = 0
    seen = {}
    while k < len(out):
        p = max(int(out["parts"][k]), 1)
        run = out[k:k + p]
        assert np.array_equal(run["part"], np.arange(p))
        assert np.all(run["pix_start"] == run["pix_start"][0])
        assert np.all(run["exposure"] == run["exposure"][0])
        assert np.all(run["band"] == run["band"][0])
        seen[int(run["pix_start"][0])] = p
        k += p
    assert seen == {int(t["pix_start"]): int(p) for t, p in zip(tiles, parts)}
    # whole tiles keep parts 0 (what fpx_plan_tiles writes)
    assert np.all(out["parts"][out["parts"] != 0] >= 2)
    # queue order: decreasing piece cost (cost / parts + overhead when split)
    piece = {int(t["pix_start"]): c / p + (1000 if p > 1 else 0)
             for t, c, p in zip(tiles, cost, parts)}
    firsts = [piece[int(s)] for s, q in zip(out["pix_start"], out["part"]) if q == 0]
    assert firsts == sorted(firsts, reverse=True)


def test_split_table_rejects_bad_parts():
    tiles = _tiles(3)
    with pytest.raises(ValueError):
        tile_split.split_table(tiles, np.ones(3), [1, 0, 1], 0)
    with pytest.raises(ValueError):
        tile_split.split_table(tiles, np.ones(3), [1, MAX_TILE_PARTS + 1, 1], 0)


def test_check_table_accepts_planned_tables_and_refuses_broken_runs():
    tiles = _tiles(6)
    out = tile_split.split_table(tiles, np.arange(6.0), [1, 3, 1, 2, 1, 1], 0.0)
    tile_split.check_table(out)
    tile_split.check_table(tiles)                       # whole tiles only
    first = int(np.argmax(out["parts"] > 1))
    broken = {
        "a piece missing": out[np.arange(len(out)) != first + 1],
        "pieces out of order": out.copy(),
        "a piece of other pixels": out.copy(),
        "a whole tile with a piece index": out.copy(),
        "too many pieces": out.copy(),
    }
    broken["pieces out of order"]["part"][first] = 1
    broken["pieces out of order"]["part"][first + 1] = 0
    broken["a piece of other pixels"]["pix_start"][first + 1] += 256
    whole = int(np.argmax(out["parts"] == 0))
    broken["a whole tile with a piece index"]["part"][whole] = 1
    broken["too many pieces"]["parts"][first:first + 3] = MAX_TILE_PARTS + 1
    for why, t in broken.items():
        with pytest.raises(ValueError, match="split tile"):
            tile_split.check_table(t)
