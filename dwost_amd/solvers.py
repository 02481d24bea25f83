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

import torch

from . import _lib
from ._lib import check, lib, ptr, stream_handle
from .shapes import _as_points, _Scene


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

    def _params(self, seed: int) -> _lib.SolverParams:
        return _lib.SolverParams(int(self.nwalks), int(self.nsteps), int(self.nsubwalks), int(self.control_variates),
                                 int(self.antithetic_variates), int(seed) & 0xFFFFFFFF, float(self.eps),
                                 float(self.min_R), _lib.SOLVE_EXACT_CONES if self.exact_cones else 0)

    def _check(self, scene: _Scene):
        if scene.dim != self.dim:
            raise ValueError(f"{type(self).__name__} needs a {self.dim}D scene")

    def solve(self, p, scene: _Scene, seed: int = 0, per_walk: bool = False, count_steps: bool = False):
        """u at each point: mean over ``nwalks`` WoSt walks (Solver.solve)."""
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
        prm = self._params(seed)
        bc = scene.boundary.to_c(scene.n_vertices)
        check(lib().dw_solve_u(scene.handle, ptr(p), n, C.byref(prm), C.byref(bc), ptr(u), ptr(walk_u),
                               ptr(walk_steps), ptr(walk_term), ptr(steps), stream_handle(dev)))
        if not (per_walk or count_steps):
            return u
        out = dict(u=u)
        if per_walk:
            out.update(walk_u=walk_u, walk_steps=walk_steps, walk_term=walk_term)
        if count_steps:
            out["steps"] = steps
        return out

    def solve_dude(self, p, e, scene: _Scene, seed: int = 0, count_steps: bool = False):
        """Directional derivative du/de (WoSt.solve_dude / WoSt3D.solve_dude)."""
        self._check(scene)
        p = _as_points(p, scene.dim)
        e = _as_points(e, scene.dim)
        n, dev = p.shape[0], p.device
        out = torch.empty(n, device=dev)
        steps = torch.zeros(1, dtype=torch.int64, device=dev) if count_steps else None
        prm = self._params(seed)
        bc = scene.boundary.to_c(scene.n_vertices)
        check(lib().dw_solve_dude(scene.handle, ptr(p), ptr(e), n, C.byref(prm), C.byref(bc), ptr(out), ptr(steps),
                                  stream_handle(dev)))
        return (out, steps) if count_steps else out

    def solve_grad(self, p, scene: _Scene, seed: int = 0, per_pair: bool = False, count_steps: bool = False):
        """Fused (u, grad u): both estimates come from the same antithetic walk pairs."""
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
        prm = self._params(seed)
        bc = scene.boundary.to_c(scene.n_vertices)
        check(lib().dw_solve_fused(scene.handle, ptr(p), n, C.byref(prm), C.byref(bc), ptr(u), ptr(g), ptr(pv),
                                   ptr(ps), ptr(steps), stream_handle(dev)))
        out = dict(u=u, grad=g)
        if per_pair:
            out.update(pair_vals=pv, pair_steps=ps)
        if count_steps:
            out["steps"] = steps
        return out


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
