"""Streamline tracer over the walk-based gradient.

``StreamlineTracer`` follows ``examples/example_2d/trace.py:9-68`` and
``examples/example_3d/trace.py:9-60``: each call of ``next_position`` moves every
point one step along the estimated gradient, averaging ``nwalks`` independent
samples per point.  Here the solver is a ``WoSt`` / ``WoSt3D`` (its
``walk_dude``, ``solver.py:148-166``, along each axis gives the gradient); the
whole step (samples, walks, sdf tests, averaging) is one GPU launch
(``dw_trace_step``).  2D keeps the reference's step rule: only points inside
move, and a step whose end is inside too keeps half its length (the corrector
term ``fy_tilde = current_p - y_tilde`` is zero as written).  3D takes the full
Euler step.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from ._lib import check, lib, ptr, stream_handle
from .shapes import _as_points, _Scene
from .solvers import WoS, WoS3D, Solver


@dataclass
class StreamlineTracer:
    scene: _Scene
    solver: Solver
    step_size: float = 0.1
    nwalks: int = 100  # samples per point and step

    def next_position(self, p, seed: int = 0, count_steps: bool = False):
        """One streamline step from p (n x dim); ``seed`` is the iteration index
        in the reference example.  Returns the new positions."""
        if isinstance(self.solver, (WoS, WoS3D)):
            raise ValueError("the tracer runs WoSt walks (WoSt / WoSt3D solvers)")
        if self.nwalks < 1:
            raise ValueError("nwalks must be >= 1")
        self.solver._check(self.scene)
        p = _as_points(p, self.scene.dim)
        n, dev = p.shape[0], p.device
        out = torch.empty(n, self.scene.dim, device=dev)
        steps = torch.zeros(1, dtype=torch.int64, device=dev) if count_steps else None
        prm = self.solver._params(seed)
        bc = self.scene.boundary.to_c(self.scene.n_vertices)
        check(lib().dw_trace_step(self.scene.handle, ptr(p), n, C.byref(prm), C.byref(bc), int(self.nwalks),
                                  float(self.step_size), ptr(out), ptr(steps), stream_handle(dev)))
        return (out, steps) if count_steps else out

    def trace(self, p, niter: int) -> list[torch.Tensor]:
        """``niter`` steps with seeds 0 .. niter-1 (the reference example's loop);
        returns the positions after every step."""
        path = []
        for i in range(niter):
            p = self.next_position(p, i)
            path.append(p)
        return path
