"""Many scenes and query sets in one persistent launch.

The reference solves one scene per call (``Solver.solve``,
``diff_wost/solvers/solver.py:36-71``); a workload of many small scenes then
leaves the GPU mostly idle between and inside launches.  ``BatchSolver`` puts
every (solver, scene, points, seed) item into one job space served by one
persistent fused u + grad u launch, optionally captured into a hipGraph and
replayed.  Each item's results are bit-identical to ``solver.solve_grad`` on
its own (walk streams are keyed by (point, pair) per item).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, lib, ptr, stream_handle
from .shapes import BoundaryCondition, _as_points, _Scene
from .solvers import Solver


@dataclass
class BatchItem:
    """One (solver, scene, points, seed) solve of a batch.  ``boundary``
    (optional) replaces the scene's boundary data for this item only, so items
    can share one uploaded scene with different g / h (e.g. perturbed per-vertex
    values)."""

    solver: Solver
    scene: _Scene
    points: torch.Tensor
    seed: int = 0
    point_base: int = 0
    boundary: BoundaryCondition | None = None


class BatchSolver:
    """Prepared multi-scene solve.  ``points`` must be contiguous float32
    (N, dim) CUDA tensors on one device; they are captured by address: update
    them in place between ``solve()`` calls, never reallocate them.
    ``graph=True`` records the launch sequence into a hipGraph on the first
    solve and replays it afterwards."""

    def __init__(self, items: list[BatchItem], graph: bool = False):
        if not items:
            raise ValueError("a batch needs at least one item")
        dims = {it.scene.dim for it in items}
        if len(dims) != 1:
            raise ValueError("all scenes of a batch need the same dimension")
        self.dim = dims.pop()
        dev = None
        self._keep = []  # tensors whose addresses the batch holds
        self.outputs = []
        arr = (_lib.BatchItem * len(items))()
        for k, it in enumerate(items):
            it.solver._check(it.scene)
            p = _as_points(it.points, self.dim)
            if not isinstance(it.points, torch.Tensor) or p.data_ptr() != it.points.data_ptr() or \
                    p.shape != it.points.shape:
                # the batch reads the points by address on every solve: a converted
                # copy would keep solving the values it held at creation
                raise ValueError(f"batch item {k}: points must be a contiguous float32 (N, {self.dim}) CUDA tensor "
                                 "(the batch solves whatever it holds at each solve)")
            if dev is None:
                dev = p.device
            elif p.device != dev:
                raise ValueError("all points of a batch must live on one device")
            n = p.shape[0]
            u = torch.empty(n, device=p.device)
            g = torch.empty(n, self.dim, device=p.device)
            bnd = it.boundary if it.boundary is not None else it.scene.boundary
            bc = bnd.to_c(it.scene.n_vertices)
            arr[k] = _lib.BatchItem(it.scene.handle, ptr(p), n, it.solver._params(it.seed, it.point_base), bc,
                                    ptr(u), ptr(g))
            self._keep.append((p, it.scene, bnd))
            self.outputs.append(dict(u=u, grad=g))
        self.device = dev
        self.steps = torch.zeros(1, dtype=torch.int64, device=dev)
        h = C.c_void_p()
        check(lib().dw_batch_create(arr, len(items), ptr(self.steps), _lib.BATCH_GRAPH if graph else 0, C.byref(h)))
        self._h = h

    def solve(self) -> list[dict]:
        """Enqueue one solve of every item on the current stream; returns the
        per-item output dicts (``u``, ``grad``), overwritten by each solve."""
        check(lib().dw_batch_solve(self._h, stream_handle(self.device)))
        return self.outputs

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            torch.cuda.synchronize(self.device)
            lib().dw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
