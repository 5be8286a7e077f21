"""Autotuner of the batched solve's launch shape.

The step kernel's shape is the number of wavefronts per workgroup and the
SELL slices (64 block rows) each wavefront streams per launch (rows per
wavefront; a workgroup's tile is their product in slices), the number of
launches captured per hipGraph replay, the tile
count at or below which a launch runs the kernel's tail form (the batch's
last, latency-bound launches, once only the stiffest QPs are left), and the
placement of the tiles on the 8 XCDs.  Every candidate changes only the
schedule (and, through the tile grouping of partial sums, the last bits of
the reductions; the tail switch point and the placement not even those),
never the algorithm.  ``tune`` times each candidate on the caller's own
batch from the same warm start and keeps the fastest: first the
wavefront/graph shapes, then the tail switch points at the best of them,
then the placements.
``tuned_params`` looks the answer up in a JSON cache keyed by device, batch
size class and precision, tuning on a miss.
"""
from __future__ import annotations

import json
import math
import os
import time

import mprgp_amd as M

DEFAULT_CANDIDATES = [
    {"tiles": 4, "graph_steps": 16},
    {"tiles": 2, "graph_steps": 16},
    {"tiles": 2, "graph_steps": 32},
    {"tiles": 2, "graph_steps": 8},
    {"tiles": 4, "graph_steps": 32},
    {"tiles": 8, "graph_steps": 16},
    {"tiles": 1, "graph_steps": 16},
]

# Slices (64 rows) each wavefront streams per launch, tried at the best shape
# (1: one slice per wavefront; a tile then holds tiles x this many slices).
SPW_CANDIDATES = [1, 2]

# Tail switch points tried at the best shape (0: the library's default, 256
# tiles, one workgroup per CU).
TAIL_CANDIDATES = [0, 64, 128, 512, 1024, 2048]

# Placements of the tiles on the 8 XCDs tried last (0: runs of 4 tiles
# rotating over the XCDs, 1: one contiguous chunk of each QP per XCD).
LAYOUT_CANDIDATES = [0, 1]

# Bumped whenever the step kernel changes in a way that moves the best shape.
CACHE_VERSION = 7


def cache_path():
    return os.environ.get("MPRGP_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "mprgp_amd", "tune.json")


def batch_key(qps, precision, device=0):
    """Device, precision and the batch's size class (log2 of total slices and
    of the QP count): batches in one class share a launch shape."""
    slices = sum(-(-q.n // 3 // 64) for q in qps)
    return f"v{CACHE_VERSION}/dev{device}/{'mixed' if precision == M.MIXED else 'fp64'}/" \
           f"s{int(math.log2(max(slices, 1)))}/q{int(math.log2(max(len(qps), 1)))}"


def apply(p: M.Params, shape):
    p.tiles = int(shape["tiles"])
    p.graph_steps = int(shape["graph_steps"])
    p.tail_tiles = int(shape.get("tail_tiles", 0))
    p.xcd_layout = int(shape.get("xcd_layout", 0))
    p.slices_per_wave = int(shape.get("slices_per_wave", 1))
    return p


def tune(ctx, qps, per_qp, candidates=None, spws=None, tails=None, layouts=None, repeats=2, log=None):
    """Times every candidate shape on the batch (each QP restarted from its
    warm start with reset_x), keeps each candidate's best of `repeats` timed
    runs (single runs differ by about 1%, as much as the candidates do), then
    the slices per wavefront `spws` at the best shape, the tail switch points
    `tails` at the best so far, then the XCD layouts `layouts`; returns
    (best_shape, [(shape, ms)])."""
    candidates = candidates or DEFAULT_CANDIDATES
    spws = SPW_CANDIDATES if spws is None else spws
    tails = TAIL_CANDIDATES if tails is None else tails
    layouts = LAYOUT_CANDIDATES if layouts is None else layouts
    results = []
    timed = _timer(ctx, qps, per_qp, repeats, log, results)
    for shape in candidates:
        timed(dict(shape))
    best = min(results, key=lambda r: r[1])[0]
    for k in spws:
        if k != best.get("slices_per_wave", 1):
            timed(dict(best, slices_per_wave=k))
    best = min(results, key=lambda r: r[1])[0]
    for t in tails:
        if t != best.get("tail_tiles", 0):
            timed(dict(best, tail_tiles=t))
    best = min(results, key=lambda r: r[1])[0]
    for lay in layouts:
        if lay != best.get("xcd_layout", 0):
            timed(dict(best, xcd_layout=lay))
    for q in qps:
        q.reset_x()
    return min(results, key=lambda r: r[1])[0], results


def _timer(ctx, qps, per_qp, repeats, log, results):
    def timed(shape):
        ps = [apply(M.Params.from_buffer_copy(p), shape) for p in per_qp]
        best = math.inf
        for rep in range(repeats + 1):  # the first run captures the graphs
            for q in qps:
                q.reset_x()
            t0 = time.perf_counter()
            M.solve_batch(ctx, qps, per_qp=ps)
            ms = 1e3 * (time.perf_counter() - t0)
            if rep > 0:
                best = min(best, ms)
        results.append((shape, best))
        if log:
            log(f"tune {shape}: {best:.2f} ms")
    return timed


def tuned_params(ctx, qps, per_qp, precision, device=0, path=None, retune=False, log=None):
    """per_qp with the cached (or freshly tuned) shape applied; returns
    (params list, shape, from_cache)."""
    path = path or cache_path()
    key = batch_key(qps, precision, device)
    cache = {}
    if os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    hit = key in cache and not retune
    if hit:
        shape = cache[key]["shape"]
    else:
        shape, results = tune(ctx, qps, per_qp, log=log)
        cache[key] = {"shape": shape, "ms": [[s, round(ms, 3)] for s, ms in results]}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path + ".tmp", "w") as f:
                json.dump(cache, f, indent=1)
            os.replace(path + ".tmp", path)
        except OSError:
            pass  # read-only home: tuned for this process only
    return [apply(M.Params.from_buffer_copy(p), shape) for p in per_qp], shape, hit
