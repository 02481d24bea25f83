"""Streamline tracer over the walk-based gradient.

``StreamlineTracer`` follows ``examples/example_2d/trace.py:9-68`` and
``examples/example_3d/trace.py:9-60``: each call of ``next_position`` moves every
point one step along the estimated gradient, averaging ``nwalks`` independent
samples per point.  The solver is the reference example's ``WoStGrad`` (2D,
``dw_grad_trace_step``) or a ``WoSt`` / ``WoSt3D`` (``dw_trace_step``); its
``walk_dude`` (``solver.py:146-166``) along each axis gives the gradient, and
the whole step (samples, walks, sdf tests, averaging) is one GPU launch.  2D keeps the reference's step rule: only points inside
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
from .solvers import WoS, WoS3D, Solver, WoStGrad


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
        if isinstance(self.solver, WoStGrad):
            gp = self.solver._grad_params()
            check(lib().dw_grad_trace_step(self.scene.handle, ptr(p), n, C.byref(prm), C.byref(gp), C.byref(bc),
                                           int(self.nwalks), float(self.step_size), ptr(out), ptr(steps),
                                           stream_handle(dev)))
        else:
            check(lib().dw_trace_step(self.scene.handle, ptr(p), n, C.byref(prm), C.byref(bc), int(self.nwalks),
                                      float(self.step_size), ptr(out), ptr(steps), stream_handle(dev)))
        return (out, steps) if count_steps else out

    def trace_graph(self, p, niter: int) -> "TraceGraph":
        """The reference example's loop (``niter`` steps, seeds 0 .. niter-1,
        examples/example_2d/run.py:85-127) captured once into a HIP graph:
        ``replay()`` reruns all ``niter`` launches with one graph launch and
        returns the positions after every step, the bits of ``trace``.  The
        start positions are read from ``p`` (a CUDA tensor; update it in place
        between replays)."""
        return TraceGraph(self, p, niter)

    def trace(self, p, niter: int) -> list[torch.Tensor]:
        """``niter`` steps with seeds 0 .. niter-1 (the reference example's loop);
        returns the positions after every step."""
        path = []
        for i in range(niter):
            p = self.next_position(p, i)
            path.append(p)
        return path


class TraceGraph:
    """A streamline trace of fixed point count and length captured into a HIP
    graph (torch.cuda.CUDAGraph).  One eager trace on the capture stream first
    sizes the library's per-stream workspace, so the captured launches
    allocate nothing; each step's seed is its iteration index, fixed at capture."""

    def __init__(self, tracer: StreamlineTracer, p, niter: int):
        if not isinstance(p, torch.Tensor) or not p.is_cuda or p.dtype != torch.float32 or not p.is_contiguous():
            raise ValueError("start points must be a contiguous float32 CUDA tensor (read by address on replay)")
        self.tracer, self.start, self.niter = tracer, p, int(niter)
        self.stream = torch.cuda.Stream()
        self.stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.stream):
            tracer.trace(self.start, self.niter)  # warm-up: workspace and queue counters of this stream
        self.stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.path = tracer.trace(self.start, self.niter)

    def replay(self) -> list[torch.Tensor]:
        """Runs the captured trace; returns the positions after every step
        (tensors owned by the graph, overwritten by the next replay)."""
        self.graph.replay()
        return self.path
