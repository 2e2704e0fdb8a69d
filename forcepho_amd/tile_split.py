"""Splitting the heaviest tiles of a persistent launch.

A persistent launch hands tiles out heaviest first to a resident grid of
``groups`` workgroups, so it lasts at least as long as its heaviest tile.
When a few tiles carry several times the mean work (sources crowded near a
few pixels: the central tiles of a stamp), they hold the launch back while
the other workgroups are idle.  A split tile (FpxTile.parts >= 2) is
evaluated by ``parts`` workgroups, each over its share of the tile's
Gaussians; each piece still loads the pixels and culls the Gaussians, so it
costs its share plus a fixed overhead.

This module picks the number of pieces per tile by simulating the queue (a
list schedule in table order onto ``groups`` workers) for a few thresholds
and keeping the one with the shortest simulated launch, and writes the table
in queue order.  Costs are in the units the caller gives (pairs inside the
cutoff plus a per-tile overhead, from ``fpx_count_pairs``).
"""

import heapq

import numpy as np

from ._lib import TILE_DTYPE, MAX_TILE_PARTS

__all__ = ["list_schedule", "plan_parts", "split_table", "check_table"]


def list_schedule(costs, groups):
    """Finish time of a queue of items with ``costs`` (in queue order) taken
    one at a time by the first free of ``groups`` workers."""
    costs = np.asarray(costs, dtype=np.float64)
    if len(costs) == 0:
        return 0.0
    free = [0.0] * min(int(groups), len(costs))
    heapq.heapify(free)
    end = 0.0
    for c in costs:
        t = heapq.heappop(free) + c
        end = max(end, t)
        heapq.heappush(free, t)
    return end


def _queue(cost, parts, piece_overhead):
    """Piece costs in queue order: tiles by decreasing piece cost, each
    tile's pieces in a row.  Returns (tile order, piece cost per tile)."""
    piece = cost / parts + np.where(parts > 1, piece_overhead, 0.0)
    order = np.argsort(-piece, kind="stable")
    return order, piece


def plan_parts(cost, groups, piece_overhead, max_parts=8,
               thresholds=(1.0, 0.75, 0.5, 0.35, 0.25)):
    """Pieces per tile (int array, 1 = whole) for tiles of the given costs
    on ``groups`` resident workgroups.  A tile is cut into
    ceil(cost / (f * L)) pieces (at most ``max_parts``), L being the mean
    work per workgroup and f one of ``thresholds``; the choice with the
    shortest simulated launch wins, no split if none beats the whole
    tiles."""
    cost = np.asarray(cost, dtype=np.float64)
    n = len(cost)
    best = np.ones(n, dtype=np.int64)
    if n == 0 or groups < 2:
        return best
    max_parts = int(min(max_parts, MAX_TILE_PARTS, groups))
    order, piece = _queue(cost, best, piece_overhead)
    best_time = list_schedule(piece[order], groups)
    load = cost.sum() / groups
    for f in thresholds:
        parts = np.clip(np.ceil(cost / (f * load)), 1, max_parts).astype(np.int64)
        if np.all(parts == 1):
            continue
        order, piece = _queue(cost, parts, piece_overhead)
        t = list_schedule(np.repeat(piece[order], parts[order]), groups)
        if t < best_time * 0.999:
            best, best_time = parts, t
    return best


def split_table(tiles, cost, parts, piece_overhead):
    """Tile table in queue order (numpy TILE_DTYPE): tiles by decreasing
    piece cost, each split tile as ``parts`` entries with part 0..parts-1."""
    tiles = np.asarray(tiles, dtype=TILE_DTYPE)
    parts = np.asarray(parts, dtype=np.int64)
    if np.any(parts < 1) or np.any(parts > MAX_TILE_PARTS):
        raise ValueError(f"parts must lie in [1, {MAX_TILE_PARTS}]")
    order, _ = _queue(np.asarray(cost, dtype=np.float64), parts, piece_overhead)
    reps = parts[order]
    out = np.repeat(tiles[order], reps)
    first = np.repeat(np.cumsum(reps) - reps, reps)
    out["part"] = np.where(np.repeat(reps, reps) > 1, np.arange(len(out)) - first, 0)
    out["parts"] = np.where(np.repeat(reps, reps) > 1, np.repeat(reps, reps), 0)
    return out


def check_table(tiles):
    """The table as a contiguous TILE_DTYPE array, checked the way the split
    kernel variant reads it: every split tile (parts >= 2, at most
    MAX_TILE_PARTS) appears as parts entries in a row, with part 0..parts-1
    and the same exposure, band and pixels; whole tiles have part 0.
    Raises ValueError otherwise."""
    tiles = np.ascontiguousarray(tiles, dtype=TILE_DTYPE)
    parts = np.maximum(tiles["parts"].astype(np.int64), 1)
    if np.any(parts > MAX_TILE_PARTS):
        raise ValueError(f"a split tile has more than {MAX_TILE_PARTS} pieces")
    k = 0
    while k < len(tiles):
        p = int(parts[k])
        run = tiles[k:k + p]
        same = all(np.all(run[f] == run[f][0]) for f in ("exposure", "band", "pix_start",
                                                           "pix_count", "parts"))
        if len(run) != p or not np.array_equal(run["part"], np.arange(p)) or not same:
            raise ValueError(f"tile table entry {k}: the pieces of a split tile must follow "
                             "each other with part 0..parts-1 and the same pixels")
        k += p
    return tiles
