"""Walk-on-Stars solvers with the diff_wost solver API, running on libdwost.

``WoSt`` / ``WoSt3D`` keep the reference dataclass fields
(``diff_wost/solvers/solver.py:30-34``, ``wost.py:17-19``, ``wost_3d.py:13-15``)
and methods:

* ``solve(p, scene, seed)``         -- Solver.solve + WoSt.u (solver.py:36-71);
                                      one PCG32 stream per (point, walk).
* ``solve_dude(p, e, scene, seed)`` -- WoSt.solve_dude (wost.py:21-102); one
                                      stream per point, walks serial.
* ``solve_grad(p, scene, seed)``    -- fused u and grad u from the same walks
                                      (the throughput form; antithetic pairs,
                                      one stream per (point, pair)).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import ClassVar

import torch

from . import _lib, ops
from ._lib import check, lib, ptr, stream_handle
from .shapes import BC_PER_VERTEX, _as_points, _Scene


@dataclass
class Solver:
    eps: float = 1e-3
    r_min: float = 1e-4  # unused by the reference walks (kept for API parity)
    nwalks: int = 1000
    nsubwalks: int = 1
    nsteps: int = 128
    control_variates: bool = True
    antithetic_variates: bool = True
    min_R: float = 1e-3
    dim: int = 2
    # True: evaluate the SNCH cone test literally (acos/asin/atan2, bit-for-bit
    # the reference arithmetic).  False: the transcendental-free conservative
    # form, identical except where the literal test's rounding drops a silhouette.
    exact_cones: bool = False
    # Launch shape of the walk kernels (never changes results; 0 = library
    # default).  dwost_amd.autotune picks them per workload.
    block_size: int = 0
    blocks_per_cu: int = 0
    spatial_order: int = 0  # 0 auto, 1 on, 2 off
    schedule: int = 0  # lane scheduling: 0 auto, 1 interleaved, 2 step-synchronous
    gate: int = 0  # interleaved schedule: waiting lanes per walk-logic pass (0: 24)
    # closest-point tree records: 2 (binary, one level per fetch) or 4 (two
    # levels per fetch); the same nodes in the same order (0: 2)
    node_width: int = 0
    # fused solve: walk pairs of one point a lane runs back to back per job it
    # takes from the queue (1, 2, 4 or 8; 0: 1); scheduling only
    walks_per_lane: int = 0

    _walk_flag: ClassVar[int] = 0  # walk family (WoSt; WoS sets DW_SOLVE_WOS)

    def _params(self, seed: int, point_base: int = 0) -> _lib.SolverParams:
        flags = (_lib.SOLVE_EXACT_CONES if self.exact_cones else 0) | self._walk_flag
        return _lib.SolverParams(int(self.nwalks), int(self.nsteps), int(self.nsubwalks), int(self.control_variates),
                                 int(self.antithetic_variates), int(seed) & 0xFFFFFFFF, float(self.eps),
                                 float(self.min_R), flags,
                                 int(point_base), int(self.block_size), int(self.blocks_per_cu),
                                 int(self.spatial_order), int(self.schedule), int(self.gate),
                                 int(self.node_width), int(self.walks_per_lane))

    def _op_params(self, seed: int, point_base: int = 0) -> tuple[list[int], list[float]]:
        """The int and float parameter lists of the torch.ops.dwost solvers."""
        prm = self._params(seed, point_base)
        return [getattr(prm, f) for f in ops.IP_FIELDS], [float(getattr(prm, f)) for f in ops.FP_FIELDS]

    def _check(self, scene: _Scene):
        if scene.dim != self.dim:
            raise ValueError(f"{type(self).__name__} needs a {self.dim}D scene")

    def solve(self, p, scene: _Scene, seed: int = 0, per_walk: bool = False, count_steps: bool = False,
              walk_flags: torch.Tensor | None = None, point_base: int = 0):
        """u at each point: mean over ``nwalks`` WoSt walks (Solver.solve).

        ``walk_flags`` (uint8, n * nwalks, optional) receives per walk a bit per
        subwalk that survived the outlier guard; the adjoint replays need it.
        ``point_base`` is the global index of ``p[0]``: streams are keyed by the
        global (point, walk) index, so slices of a point set reproduce the
        whole-set result."""
        self._check(scene)
        p = _as_points(p, scene.dim)
        n, dev = p.shape[0], p.device
        u = torch.empty(n, device=dev)
        walk_u = walk_steps = walk_term = None
        if per_walk:
            walk_u = torch.empty(n * self.nwalks, device=dev)
            walk_steps = torch.empty(n * self.nwalks, dtype=torch.int32, device=dev)
            walk_term = torch.empty(n * self.nwalks, dtype=torch.int32, device=dev)
        steps = torch.zeros(1, dtype=torch.int64, device=dev) if count_steps else None
        prm = self._params(seed, point_base)
        bc = scene.boundary.to_c(scene.n_vertices)
        if walk_flags is not None:
            _check_flags(walk_flags, n * self.nwalks, dev)
        check(lib().dw_solve_u(scene.handle, ptr(p), n, C.byref(prm), C.byref(bc), ptr(u), ptr(walk_u),
                               ptr(walk_steps), ptr(walk_term), ptr(walk_flags), ptr(steps), stream_handle(dev)))
        if not (per_walk or count_steps):
            return u
        out = dict(u=u)
        if per_walk:
            out.update(walk_u=walk_u, walk_steps=walk_steps, walk_term=walk_term)
        if count_steps:
            out["steps"] = steps
        return out

    def solve_dude(self, p, e, scene: _Scene, seed: int = 0, count_steps: bool = False, point_base: int = 0):
        """Directional derivative du/de (WoSt.solve_dude / WoSt3D.solve_dude)."""
        self._check(scene)
        p = _as_points(p, scene.dim)
        e = _as_points(e, scene.dim)
        n, dev = p.shape[0], p.device
        out = torch.empty(n, device=dev)
        steps = torch.zeros(1, dtype=torch.int64, device=dev) if count_steps else None
        prm = self._params(seed, point_base)
        bc = scene.boundary.to_c(scene.n_vertices)
        check(lib().dw_solve_dude(scene.handle, ptr(p), ptr(e), n, C.byref(prm), C.byref(bc), ptr(out), ptr(steps),
                                  stream_handle(dev)))
        return (out, steps) if count_steps else out

    def solve_grad(self, p, scene: _Scene, seed: int = 0, per_pair: bool = False, count_steps: bool = False,
                   pair_flags: torch.Tensor | None = None, point_base: int = 0):
        """Fused (u, grad u): both estimates come from the same antithetic walk pairs.

        ``pair_flags`` (uint8, n * ceil(nwalks / 2), optional) receives per pair
        bit 0 / bit 1 set when walk + / walk - survived the outlier guard."""
        self._check(scene)
        p = _as_points(p, scene.dim)
        n, dev = p.shape[0], p.device
        D = scene.dim
        npairs = (self.nwalks + 1) // 2
        u = torch.empty(n, device=dev)
        g = torch.empty(n, D, device=dev)
        pv = torch.empty(n * npairs, D + 1, device=dev) if per_pair else None
        ps = torch.empty(n * npairs, dtype=torch.int32, device=dev) if per_pair else None
        steps = torch.zeros(1, dtype=torch.int64, device=dev) if count_steps else None
        prm = self._params(seed, point_base)
        bc = scene.boundary.to_c(scene.n_vertices)
        if pair_flags is not None:
            _check_flags(pair_flags, n * npairs, dev)
        check(lib().dw_solve_fused(scene.handle, ptr(p), n, C.byref(prm), C.byref(bc), ptr(u), ptr(g), ptr(pv),
                                   ptr(ps), ptr(pair_flags), ptr(steps), stream_handle(dev)))
        out = dict(u=u, grad=g)
        if per_pair:
            out.update(pair_vals=pv, pair_steps=ps)
        if count_steps:
            out["steps"] = steps
        return out


def _check_flags(t: torch.Tensor, numel: int, dev) -> None:
    if t.dtype != torch.uint8 or not t.is_contiguous() or t.numel() != numel or t.device != dev:
        raise ValueError(f"flags must be a contiguous uint8 tensor of {numel} elements on {dev}")


def _same_walks(a: Solver, b: Solver) -> bool:
    """True when two solvers draw identical walks for the same seed."""
    return (a._walk_flag == b._walk_flag and a.dim == b.dim and a.nwalks == b.nwalks and a.nsubwalks == b.nsubwalks and a.nsteps == b.nsteps
            and a.control_variates == b.control_variates and a.antithetic_variates == b.antithetic_variates
            and a.eps == b.eps and a.min_R == b.min_R and a.exact_cones == b.exact_cones)


def _per_vertex(scene: _Scene):
    bc = scene.boundary
    if bc is None or bc.code != BC_PER_VERTEX:
        raise ValueError("CustomSolver differentiates per-vertex boundary data: use BoundaryCondition.per_vertex(g, h)")
    return bc.g, bc.h


@dataclass
class CustomSolver:
    """Differentiable solve (diff_wost/solvers/solver.py:255-293).

    The forward value comes from ``fwd_solver``; the backward pass replays
    ``bwd_solver``'s walks from the same seed and scatter-adds dL/dg and dL/dh
    into the scene's per-vertex boundary tensors (``BoundaryCondition.per_vertex``
    with ``requires_grad``).  Both run as the autograd-registered operators
    ``torch.ops.dwost.solve_u`` / ``solve_fused``; nothing is stored per walk
    but one guard-flag byte."""

    fwd_solver: Solver = None
    bwd_solver: Solver = None

    def __post_init__(self):
        if self.bwd_solver is None:
            self.bwd_solver = self.fwd_solver

    def _prep(self, p, scene, seed):
        if self.fwd_solver is None:
            raise ValueError("CustomSolver needs a fwd_solver")
        self.fwd_solver._check(scene)
        self.bwd_solver._check(scene)
        _per_vertex(scene)
        ip, fp = self.fwd_solver._op_params(seed)
        bip, bfp = self.bwd_solver._op_params(seed)
        if _same_walks(self.fwd_solver, self.bwd_solver):
            bip, bfp = ip, fp  # identical walks: the backward reuses the forward's guard flags
        return (scene.handle.value, _as_points(p, scene.dim), ip, fp, bip, bfp) + scene.boundary.op_args()

    def solve(self, p, scene: _Scene, seed: int = 0) -> torch.Tensor:
        """u at each point, differentiable w.r.t. the per-vertex g and h."""
        return torch.ops.dwost.solve_u(*self._prep(p, scene, seed))[0]

    def solve_grad(self, p, scene: _Scene, seed: int = 0):
        """(u, grad u) from the fused estimator, both differentiable w.r.t. g and h."""
        out = torch.ops.dwost.solve_fused(*self._prep(p, scene, seed))
        return out[0], out[1]


@dataclass
class WoSt(Solver):
    """2D Walk on Stars (diff_wost/solvers/wost.py:15-221)."""

    control_variates: bool = True
    antithetic_variates: bool = True
    min_R: float = 1e-3
    dim: int = 2


@dataclass
class WoSt3D(Solver):
    """3D Walk on Stars (diff_wost/solvers/wost_3d.py:11-175)."""

    control_variates: bool = False
    antithetic_variates: bool = False
    min_R: float = 1e-3
    dim: int = 3


@dataclass
class WoS(Solver):
    """2D Walk on Spheres (diff_wost/solvers/wos.py:10-105): every boundary
    primitive is Dirichlet; ``solve`` and ``solve_dude`` only."""

    control_variates: bool = True
    antithetic_variates: bool = True
    use_bvh: bool = False  # reference field; the scene is always a BVH here
    dim: int = 2
    _walk_flag: ClassVar[int] = _lib.SOLVE_WOS


@dataclass
class WoS3D(Solver):
    """3D Walk on Spheres (diff_wost/solvers/wos_3d.py:10-102)."""

    control_variates: bool = True
    antithetic_variates: bool = True
    dim: int = 3
    _walk_flag: ClassVar[int] = _lib.SOLVE_WOS


@dataclass
class WoStGrad(Solver):
    """2D WoStGrad, the robust derivative walk (diff_wost/solvers/wost_grad.py:17-254).
    ``WoStGrad3D`` is its 3D form (all-Neumann scenes).

    ``solve_dude(p, e, scene, seed)`` estimates du/de with walks that reflect off
    Neumann boundaries (Solver.solve_dude, solver.py:113-166: one stream per
    (point, walk), ``nsubwalks`` dudes per walk).  The reference defines no
    ``u`` for it.  Boundary data must be an analytic field: the estimator uses
    the tangential derivatives of g and h."""

    ignore_neumann: bool = False  # reference field, unused by its walk
    ignore_boundary: bool = False
    ignore_silhouette: bool = False
    ignore_source: bool = False
    control_variates: bool = True
    russian_roulette: bool = False  # reference field, unused by its walk (the roulette always runs)
    min_R: float = 1e-2
    rho_max: float = 1.2
    dim: int = 2

    def _grad_params(self) -> _lib.GradParams:
        return _lib.GradParams(float(self.rho_max), int(self.ignore_boundary), int(self.ignore_silhouette),
                               int(self.ignore_source))

    def solve(self, *args, **kwargs):
        raise NotImplementedError("WoStGrad estimates derivatives only (solve_dude)")

    def solve_grad(self, *args, **kwargs):
        raise NotImplementedError("WoStGrad estimates derivatives only (solve_dude)")

    def solve_dude(self, p, e, scene: _Scene, seed: int = 0, per_walk: bool = False, count_steps: bool = False,
                   point_base: int = 0):
        """du/de at each point: mean of ``nwalks`` WoStGrad walks."""
        self._check(scene)
        p = _as_points(p, scene.dim)
        e = _as_points(e, scene.dim)
        n, dev = p.shape[0], p.device
        out = torch.empty(n, device=dev)
        wv = torch.empty(n * self.nwalks, device=dev) if per_walk else None
        ws = torch.empty(n * self.nwalks, dtype=torch.int32, device=dev) if per_walk else None
        steps = torch.zeros(1, dtype=torch.int64, device=dev) if count_steps else None
        prm = self._params(seed, point_base)
        gp = self._grad_params()
        bc = scene.boundary.to_c(scene.n_vertices)
        check(lib().dw_grad_solve_dude(scene.handle, ptr(p), ptr(e), n, C.byref(prm), C.byref(gp), C.byref(bc),
                                       ptr(out), ptr(wv), ptr(ws), ptr(steps), stream_handle(dev)))
        if not (per_walk or count_steps):
            return out
        r = dict(dude=out)
        if per_walk:
            r.update(walk_dude=wv, walk_steps=ws)
        if count_steps:
            r["steps"] = steps
        return r


@dataclass
class WoStGrad3D(WoStGrad):
    """3D WoStGrad (diff_wost/solvers/wost_grad_3d.py:13-206) on all-Neumann
    scenes: the reference's 3D walk has no normal-derivative estimate at
    Dirichlet points.  Its ignore_boundary / ignore_silhouette fields are unused
    by the reference walk (kept for API parity)."""

    min_R: float = 1e-1
    dim: int = 3
