"""Launch-shape autotuner for the persistent walk kernels.

The walk kernels take six knobs that never change a result (each walk's
random stream is keyed by its (point, walk) index, not by the lane that runs
it, and both record widths visit the same nodes in the same order): threads
per block (64 / 128 / 256), a cap on resident blocks per CU (the occupancy:
fewer waves per SIMD means less L2 pressure per wave), whether the points are
sorted by Morton cell and split into one range per XCD, the lane scheduling
(interleaved traversals or step-synchronous walks; the gate threshold of the
interleaved loop is left at its default, flat within 2% over 16-32 on the
icosphere), the node width of the closest-point tree records (2: one binary
level per dependent fetch, 4: two) and the walks per lane of the fused solve
(walk pairs of one point a lane runs back to back per job it takes).  ``tune``
times every combination of the first four on a sample of the workload, then
each node width with the fastest of them, then each walks-per-lane value with
the fastest so far (fused solves), checks that all runs return the same bits,
and stores the fastest configuration on the solver.
"""

from __future__ import annotations

import itertools
import time
from dataclasses import replace

import torch

from .shapes import _as_points, _Scene
from .solvers import Solver

BLOCK_SIZES = (64, 128, 256)
BLOCKS_PER_CU = (0, 1, 2)  # 0: as many as the occupancy allows
ORDERS = (1, 2)  # spatial order on / off
SCHEDULES = (1, 2)  # interleaved / step-synchronous
NODE_WIDTHS = (2, 4)
WALKS_PER_LANE = (1, 2, 4)

_cache: dict = {}


def _key(solver: Solver, scene: _Scene, n: int):
    return (scene.dim, scene.n_prims, solver.nwalks, solver.exact_cones, max(1, n).bit_length())


def tune(solver: Solver, scene: _Scene, points, seed: int = 0, sample: int = 8192, reps: int = 2,
         fused: bool = True, use_cache: bool = True) -> dict:
    """Times all launch shapes on the first ``sample`` points (full walks per
    point) and sets the fastest on ``solver``.  Returns the timings."""
    p = _as_points(points, scene.dim)
    key = _key(solver, scene, p.shape[0])
    if use_cache and key in _cache:
        best = _cache[key]
        _set(solver, best["config"])
        return best
    q = p[:sample].contiguous()
    results, ref = [], None

    def time_config(cfg):
        nonlocal ref
        s = replace(solver)
        _set(s, cfg)
        run = (lambda: s.solve_grad(q, scene, seed=seed)["u"]) if fused else (lambda: s.solve(q, scene, seed=seed))
        out = run()  # warm-up (allocations, first launch)
        torch.cuda.synchronize()
        if ref is None:
            ref = out.clone()
        elif not torch.equal(out.view(torch.int32), ref.view(torch.int32)):
            raise RuntimeError(f"launch configuration {cfg} changed the result")
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        results.append((cfg, min(times)))

    for bs, pc, order, sched in itertools.product(BLOCK_SIZES, BLOCKS_PER_CU, ORDERS, SCHEDULES):
        time_config((bs, pc, order, sched, NODE_WIDTHS[0], WALKS_PER_LANE[0]))
    lead = min(results, key=lambda r: r[1])[0]
    for nw in NODE_WIDTHS[1:]:
        time_config(lead[:4] + (nw, WALKS_PER_LANE[0]))
    if fused:
        lead = min(results, key=lambda r: r[1])[0]
        for wl in WALKS_PER_LANE[1:]:
            time_config(lead[:5] + (wl,))
    results.sort(key=lambda r: r[1])
    default = (256, 0, 1, 1 if scene.dim == 3 else 2, 2, 1)
    best = dict(config=results[0][0], seconds=results[0][1], seconds_default=dict(results)[default], table=results)
    _set(solver, best["config"])
    _cache[key] = best
    return best


def _set(solver: Solver, cfg) -> None:
    (solver.block_size, solver.blocks_per_cu, solver.spatial_order, solver.schedule, solver.node_width,
     solver.walks_per_lane) = cfg
