"""PyTorch custom operators over the C ABI: ``torch.ops.dwost.*``.

Each operator is a functional wrapper of one libdwost entry point
(include/dwost.h): it allocates its outputs on the input's device, calls the
entry point on torch's current stream, and has a fake (meta) implementation,
so the operators trace under ``torch.compile`` / ``torch.export`` and report
shapes without running.  A scene crosses the operator boundary as the int64
address of its opaque ``dw_scene`` (``Scene2D.handle``); solver fields as an
int list and a float list (``Solver._op_params``); boundary data as its code,
coefficients, optional per-vertex ``g`` / ``h`` tensors and the source as
[code, four coefficients] (``BoundaryCondition.op_args``).

``solve_u`` and ``solve_fused`` are differentiable with respect to the
per-vertex ``g`` and ``h``: the backward pass replays the walks of the
backward solver's parameters from the same seed and scatter-adds dL/dg, dL/dh
(the reference's CustomSolver, ``diff_wost/solvers/solver.py:202-293``).
"""

from __future__ import annotations

import ctypes as C

import torch
from torch.library import custom_op, register_autograd

from . import _lib
from ._lib import check, lib, ptr, stream_handle

# int parameter list: Solver fields and launch shape (dw_solver_params)
IP_FIELDS = ("nwalks", "nsteps", "nsubwalks", "control_variates", "antithetic_variates", "seed", "flags",
             "point_base", "block_size", "blocks_per_cu", "spatial_order", "schedule", "gate", "node_width",
             "walks_per_lane")
FP_FIELDS = ("eps", "min_R")


def params_c(ip: list[int], fp: list[float]) -> _lib.SolverParams:
    if len(ip) != len(IP_FIELDS) or len(fp) != len(FP_FIELDS):
        raise ValueError(f"solver parameters: {len(IP_FIELDS)} ints and {len(FP_FIELDS)} floats")
    d = dict(zip(IP_FIELDS, ip))
    return _lib.SolverParams(int(d["nwalks"]), int(d["nsteps"]), int(d["nsubwalks"]), int(d["control_variates"]),
                             int(d["antithetic_variates"]), int(d["seed"]) & 0xFFFFFFFF, float(fp[0]), float(fp[1]),
                             int(d["flags"]), int(d["point_base"]), int(d["block_size"]), int(d["blocks_per_cu"]),
                             int(d["spatial_order"]), int(d["schedule"]), int(d["gate"]), int(d["node_width"]),
                             int(d["walks_per_lane"]))


def boundary_c(bc_code: int, coeffs: list[float], g, h, source: list[float]) -> _lib.Boundary:
    b = _lib.Boundary()
    b.bc_code = int(bc_code)
    for k in range(4):
        b.coeffs[k] = float(coeffs[k]) if k < len(coeffs) else 0.0
    if bc_code == 3:  # DW_BC_PER_VERTEX
        if g is None or h is None:
            raise ValueError("per-vertex boundary data needs g and h")
        b.g_vertex = g.data_ptr()
        b.h_vertex = h.data_ptr()
    b.src_code = int(source[0])  # [code, then the four coefficients (constant: the value first)]
    b.src_value = float(source[1])
    for k in range(4):
        b.src_coeffs[k] = float(source[1 + k])
    return b


def _scene(handle: int) -> C.c_void_p:
    return C.c_void_p(int(handle))


def _npairs(ip):
    return (int(ip[0]) + 1) // 2


# -- queries ---------------------------------------------------------------
@custom_op("dwost::closest_point", mutates_args=())
def closest_point(scene: int, sub: int, pts: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor,
                                                                     torch.Tensor, torch.Tensor]:
    """BVH.closest_point / SNCH3D.closest_point: (prim, d, p, n, tuv)."""
    n, D = pts.shape
    dev = pts.device
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    d, p, nr, tuv = (torch.empty(n, device=dev), torch.empty(n, D, device=dev), torch.empty(n, D, device=dev),
                     torch.empty(n, 2, device=dev))
    check(lib().dw_closest_point(_scene(scene), sub, ptr(pts), n, ptr(prim), ptr(d), ptr(p), ptr(nr), ptr(tuv),
                                 stream_handle(dev)))
    return prim, d, p, nr, tuv


@closest_point.register_fake
def _(scene, sub, pts):
    n, D = pts.shape
    return (pts.new_empty(n, dtype=torch.int32), pts.new_empty(n), pts.new_empty(n, D), pts.new_empty(n, D),
            pts.new_empty(n, 2))


@custom_op("dwost::intersect", mutates_args=())
def intersect(scene: int, sub: int, origin: torch.Tensor, direction: torch.Tensor, rmax: torch.Tensor,
              mode: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """BVH.intersect / SNCH3D.intersect (mode DW_RAY_*): (prim, d, p, n, tuv); prim -1 on a miss."""
    n, D = origin.shape
    dev = origin.device
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    d, p, nr, tuv = (torch.empty(n, device=dev), torch.empty(n, D, device=dev), torch.empty(n, D, device=dev),
                     torch.empty(n, 2, device=dev))
    check(lib().dw_intersect(_scene(scene), sub, ptr(origin), ptr(direction), ptr(rmax), n, mode, ptr(prim), ptr(d),
                             ptr(p), ptr(nr), ptr(tuv), stream_handle(dev)))
    return prim, d, p, nr, tuv


@intersect.register_fake
def _(scene, sub, origin, direction, rmax, mode):
    n, D = origin.shape
    return (origin.new_empty(n, dtype=torch.int32), origin.new_empty(n), origin.new_empty(n, D),
            origin.new_empty(n, D), origin.new_empty(n, 2))


@custom_op("dwost::closest_silhouette", mutates_args=())
def closest_silhouette(scene: int, sub: int, pts: torch.Tensor, rmax: torch.Tensor) -> tuple[torch.Tensor,
                                                                                            torch.Tensor,
                                                                                            torch.Tensor]:
    """SNCH.closest_silhouette: (sil, d, p); sil -1 when none lies within rmax."""
    n, D = pts.shape
    dev = pts.device
    sil = torch.empty(n, dtype=torch.int32, device=dev)
    d, p = torch.empty(n, device=dev), torch.empty(n, D, device=dev)
    check(lib().dw_closest_silhouette(_scene(scene), sub, ptr(pts), ptr(rmax), n, ptr(sil), ptr(d), ptr(p),
                                      stream_handle(dev)))
    return sil, d, p


@closest_silhouette.register_fake
def _(scene, sub, pts, rmax):
    n, D = pts.shape
    return pts.new_empty(n, dtype=torch.int32), pts.new_empty(n), pts.new_empty(n, D)


@custom_op("dwost::star_radius", mutates_args=())
def star_radius(scene: int, sub: int, pts: torch.Tensor, rmax: torch.Tensor) -> torch.Tensor:
    """Polyline.star_radius / Mesh.star_radius."""
    out = torch.empty(pts.shape[0], device=pts.device)
    check(lib().dw_star_radius(_scene(scene), sub, ptr(pts), ptr(rmax), pts.shape[0], ptr(out),
                               stream_handle(pts.device)))
    return out


@star_radius.register_fake
def _(scene, sub, pts, rmax):
    return pts.new_empty(pts.shape[0])


@custom_op("dwost::sdf", mutates_args=())
def sdf(scene: int, pts: torch.Tensor) -> torch.Tensor:
    """Polyline.sdf / Mesh.sdf on the full scene (negative inside)."""
    out = torch.empty(pts.shape[0], device=pts.device)
    check(lib().dw_sdf(_scene(scene), ptr(pts), pts.shape[0], ptr(out), stream_handle(pts.device)))
    return out


@sdf.register_fake
def _(scene, pts):
    return pts.new_empty(pts.shape[0])


@custom_op("dwost::sample_boundary", mutates_args=())
def sample_boundary(scene: int, sub: int, dim: int, u: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor,
                                                                              torch.Tensor, torch.Tensor,
                                                                              torch.Tensor]:
    """Polyline / Mesh sample_boundary from given draws: (p, n, tuv, pdf, prim)."""
    n = u.shape[0]
    dev = u.device
    p, nr, tuv, pdf = (torch.empty(n, dim, device=dev), torch.empty(n, dim, device=dev),
                       torch.empty(n, 2, device=dev), torch.empty(n, device=dev))
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    check(lib().dw_sample_boundary(_scene(scene), sub, ptr(u), n, ptr(p), ptr(nr), ptr(tuv), ptr(pdf), ptr(prim),
                                   stream_handle(dev)))
    return p, nr, tuv, pdf, prim


@sample_boundary.register_fake
def _(scene, sub, dim, u):
    n = u.shape[0]
    return (u.new_empty(n, dim), u.new_empty(n, dim), u.new_empty(n, 2), u.new_empty(n),
            u.new_empty(n, dtype=torch.int32))


# -- solvers ---------------------------------------------------------------
@custom_op("dwost::solve_u", mutates_args=())
def solve_u(scene: int, pts: torch.Tensor, iparams: list[int], fparams: list[float], bwd_iparams: list[int],
            bwd_fparams: list[float], bc_code: int, bc_coeffs: list[float], g: torch.Tensor | None,
            h: torch.Tensor | None, source: list[float]) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Solver.solve with WoSt.u / WoSt3D.u (or WoS): (u, walk_flags, steps).
    walk_flags (uint8, n * nwalks) marks the subwalks the outlier guard kept;
    steps (int64[1]) is the walk-step total.  bwd_* are the backward solver's
    parameters (the replay of the backward pass)."""
    n = pts.shape[0]
    dev = pts.device
    u = torch.empty(n, device=dev)
    flags = torch.empty(n * int(iparams[0]), dtype=torch.uint8, device=dev)
    steps = torch.zeros(1, dtype=torch.int64, device=dev)
    prm, bc = params_c(iparams, fparams), boundary_c(bc_code, bc_coeffs, g, h, source)
    check(lib().dw_solve_u(_scene(scene), ptr(pts), n, C.byref(prm), C.byref(bc), ptr(u), None, None, None,
                           ptr(flags), ptr(steps), stream_handle(dev)))
    return u, flags, steps


@solve_u.register_fake
def _(scene, pts, iparams, fparams, bwd_iparams, bwd_fparams, bc_code, bc_coeffs, g, h, source):
    n = pts.shape[0]
    return (pts.new_empty(n), pts.new_empty(n * int(iparams[0]), dtype=torch.uint8),
            pts.new_empty(1, dtype=torch.int64))


@custom_op("dwost::solve_dude", mutates_args=())
def solve_dude(scene: int, pts: torch.Tensor, e: torch.Tensor, iparams: list[int], fparams: list[float],
               bc_code: int, bc_coeffs: list[float], g: torch.Tensor | None, h: torch.Tensor | None,
               source: list[float]) -> torch.Tensor:
    """WoSt.solve_dude / WoSt3D.solve_dude: du/de per point."""
    n = pts.shape[0]
    out = torch.empty(n, device=pts.device)
    prm, bc = params_c(iparams, fparams), boundary_c(bc_code, bc_coeffs, g, h, source)
    check(lib().dw_solve_dude(_scene(scene), ptr(pts), ptr(e), n, C.byref(prm), C.byref(bc), ptr(out), None,
                              stream_handle(pts.device)))
    return out


@solve_dude.register_fake
def _(scene, pts, e, iparams, fparams, bc_code, bc_coeffs, g, h, source):
    return pts.new_empty(pts.shape[0])


@custom_op("dwost::solve_fused", mutates_args=())
def solve_fused(scene: int, pts: torch.Tensor, iparams: list[int], fparams: list[float], bwd_iparams: list[int],
                bwd_fparams: list[float], bc_code: int, bc_coeffs: list[float], g: torch.Tensor | None,
                h: torch.Tensor | None, source: list[float]) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor,
                                                                torch.Tensor]:
    """Fused u + grad u: (u, grad, pair_flags, steps).  pair_flags (uint8,
    n * ceil(nwalks / 2)): bit 0 / 1 set when walk + / - survived the outlier
    guard.  bwd_* are the backward solver's parameters."""
    n, D = pts.shape
    dev = pts.device
    u = torch.empty(n, device=dev)
    gr = torch.empty(n, D, device=dev)
    flags = torch.empty(n * _npairs(iparams), dtype=torch.uint8, device=dev)
    steps = torch.zeros(1, dtype=torch.int64, device=dev)
    prm, bc = params_c(iparams, fparams), boundary_c(bc_code, bc_coeffs, g, h, source)
    check(lib().dw_solve_fused(_scene(scene), ptr(pts), n, C.byref(prm), C.byref(bc), ptr(u), ptr(gr), None, None,
                               ptr(flags), ptr(steps), stream_handle(dev)))
    return u, gr, flags, steps


@solve_fused.register_fake
def _(scene, pts, iparams, fparams, bwd_iparams, bwd_fparams, bc_code, bc_coeffs, g, h, source):
    n, D = pts.shape
    return (pts.new_empty(n), pts.new_empty(n, D), pts.new_empty(n * _npairs(iparams), dtype=torch.uint8),
            pts.new_empty(1, dtype=torch.int64))


@custom_op("dwost::adjoint_u", mutates_args=())
def adjoint_u(scene: int, pts: torch.Tensor, iparams: list[int], fparams: list[float], bc_code: int,
              bc_coeffs: list[float], g: torch.Tensor, h: torch.Tensor, source: list[float], dl_du: torch.Tensor,
              walk_flags: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Adjoint of solve_u for per-vertex g, h: (dL/dg, dL/dh)."""
    dev = pts.device
    dg, dh = torch.zeros_like(g), torch.zeros_like(h)
    prm, bc = params_c(iparams, fparams), boundary_c(bc_code, bc_coeffs, g, h, source)
    check(lib().dw_adjoint_u(_scene(scene), ptr(pts), pts.shape[0], C.byref(prm), C.byref(bc), ptr(dl_du),
                             ptr(walk_flags), ptr(dg), ptr(dh), stream_handle(dev)))
    return dg, dh


@adjoint_u.register_fake
def _(scene, pts, iparams, fparams, bc_code, bc_coeffs, g, h, source, dl_du, walk_flags):
    return torch.empty_like(g), torch.empty_like(h)


@custom_op("dwost::adjoint_fused", mutates_args=())
def adjoint_fused(scene: int, pts: torch.Tensor, iparams: list[int], fparams: list[float], bc_code: int,
                  bc_coeffs: list[float], g: torch.Tensor, h: torch.Tensor, source: list[float], dl_du: torch.Tensor,
                  dl_dgrad: torch.Tensor, pair_flags: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Adjoint of solve_fused for per-vertex g, h: (dL/dg, dL/dh)."""
    dev = pts.device
    dg, dh = torch.zeros_like(g), torch.zeros_like(h)
    prm, bc = params_c(iparams, fparams), boundary_c(bc_code, bc_coeffs, g, h, source)
    check(lib().dw_adjoint_fused(_scene(scene), ptr(pts), pts.shape[0], C.byref(prm), C.byref(bc), ptr(dl_du),
                                 ptr(dl_dgrad), ptr(pair_flags), ptr(dg), ptr(dh), stream_handle(dev)))
    return dg, dh


@adjoint_fused.register_fake
def _(scene, pts, iparams, fparams, bc_code, bc_coeffs, g, h, source, dl_du, dl_dgrad, pair_flags):
    return torch.empty_like(g), torch.empty_like(h)


# -- autograd: replay from seeds ------------------------------------------
def _setup(ctx, inputs, output):
    scene, pts, ip, fp, bip, bfp, bc_code, coeffs, g, h, source = inputs
    ctx.args = (scene, list(ip), list(fp), list(bip), list(bfp), bc_code, list(coeffs), list(source))
    ctx.save_for_backward(pts, g, h, output[-2])  # the guard flags of the forward walks


def _replay_flags(op, ctx, pts, g, h, flags):
    """The backward solver's guard flags: the forward's when it drew the same
    walks, else a forward pass of the backward solver."""
    scene, ip, fp, bip, bfp, bc_code, coeffs, source = ctx.args
    if (ip, fp) == (bip, bfp):
        return flags
    return op(scene, pts, bip, bfp, bip, bfp, bc_code, coeffs, g, h, source)[-2]


def _grads(ctx, dg, dh):
    return (None, None, None, None, None, None, None, None, dg if ctx.needs_input_grad[8] else None,
            dh if ctx.needs_input_grad[9] else None, None)


def _backward_u(ctx, dl_du, _flags, _steps):
    pts, g, h, flags = ctx.saved_tensors
    if g is None or h is None or dl_du is None:
        return _grads(ctx, None, None)
    scene, ip, fp, bip, bfp, bc_code, coeffs, source = ctx.args
    fl = _replay_flags(torch.ops.dwost.solve_u, ctx, pts, g, h, flags)
    dg, dh = torch.ops.dwost.adjoint_u(scene, pts, bip, bfp, bc_code, coeffs, g, h, source,
                                       dl_du.float().contiguous(), fl)
    return _grads(ctx, dg, dh)


def _backward_fused(ctx, dl_du, dl_dgrad, _flags, _steps):
    pts, g, h, flags = ctx.saved_tensors
    if g is None or h is None or (dl_du is None and dl_dgrad is None):
        return _grads(ctx, None, None)
    scene, ip, fp, bip, bfp, bc_code, coeffs, source = ctx.args
    n, D = pts.shape
    dl_du = torch.zeros(n, device=pts.device) if dl_du is None else dl_du.float().contiguous()
    dl_dgrad = torch.zeros(n, D, device=pts.device) if dl_dgrad is None else dl_dgrad.float().contiguous()
    fl = _replay_flags(torch.ops.dwost.solve_fused, ctx, pts, g, h, flags)
    dg, dh = torch.ops.dwost.adjoint_fused(scene, pts, bip, bfp, bc_code, coeffs, g, h, source, dl_du, dl_dgrad, fl)
    return _grads(ctx, dg, dh)


register_autograd("dwost::solve_u", _backward_u, setup_context=_setup)
register_autograd("dwost::solve_fused", _backward_fused, setup_context=_setup)


@custom_op("dwost::star_radius_2", mutates_args=())
def star_radius_2(scene: int, sub: int, pts: torch.Tensor, e: torch.Tensor, rho_max: float) -> torch.Tensor:
    """Polyline.star_radius_2 / Mesh.star_radius_2 (rho-bounded star radius)."""
    out = torch.empty(pts.shape[0], device=pts.device)
    check(lib().dw_star_radius_rho(_scene(scene), sub, ptr(pts), ptr(e), float(rho_max), pts.shape[0], ptr(out),
                                   stream_handle(pts.device)))
    return out


@star_radius_2.register_fake
def _(scene, sub, pts, e, rho_max):
    return pts.new_empty(pts.shape[0])


def math_check(device=None) -> dict:
    """Largest float-step distance between acosf / asinf / atanf / cosf and
    their once-rounded fp64 forms over each whole domain, and the bound the
    bracketed cone tests assume (dw_math_check)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    out = torch.empty(5, dtype=torch.int32, device=dev)
    check(lib().dw_math_check(ptr(out), stream_handle(dev)))
    v = out.cpu().tolist()
    return dict(acos=v[0], asin=v[1], atan=v[2], cos=v[3], bound=v[4])
