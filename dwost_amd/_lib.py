"""ctypes binding of libdwost.so (the C ABI declared in include/dwost.h).

torch is imported first so that the process already holds torch's HIP
runtime (soname libamdhip64.so.7) when the library is loaded; the library then
binds to that same runtime instead of loading a second copy.
"""

from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (must precede the library load, see above)

from .build import LIB

DW_OK = 0
SUB_FULL, SUB_DIRICHLET, SUB_NEUMANN = 0, 1, 2

_lib = None


class DwostError(RuntimeError):
    pass


class SceneDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n_vertices", C.c_int64), ("n_prims", C.c_int64),
                ("vertices", C.c_void_p), ("indices", C.c_void_p), ("types", C.c_void_p), ("flags", C.c_int32)]

SCENE_HOST_ONLY = 1


class Boundary(C.Structure):
    _fields_ = [("bc_code", C.c_int32), ("coeffs", C.c_float * 4), ("g_vertex", C.c_void_p),
                ("h_vertex", C.c_void_p), ("src_code", C.c_int32), ("src_value", C.c_float),
                ("src_coeffs", C.c_float * 4)]


class SolverParams(C.Structure):
    _fields_ = [("nwalks", C.c_int32), ("nsteps", C.c_int32), ("nsubwalks", C.c_int32),
                ("control_variates", C.c_int32), ("antithetic_variates", C.c_int32), ("seed", C.c_uint32),
                ("eps", C.c_float), ("min_R", C.c_float), ("flags", C.c_int32), ("point_base", C.c_int64),
                ("block_size", C.c_int32), ("blocks_per_cu", C.c_int32), ("spatial_order", C.c_int32),
                ("schedule", C.c_int32), ("gate", C.c_int32), ("node_width", C.c_int32),
                ("walks_per_lane", C.c_int32)]

NOALLOC_SCRATCH_BYTES = 64 << 20  # DW_NOALLOC_SCRATCH_BYTES
class GradParams(C.Structure):
    _fields_ = [("rho_max", C.c_float), ("ignore_boundary", C.c_int32), ("ignore_silhouette", C.c_int32),
                ("ignore_source", C.c_int32)]


SOLVE_EXACT_CONES = 1
SOLVE_WOS = 2
RAY_REFERENCE, RAY_NEAREST = 0, 1  # dw_intersect modes
BASE_CLOSEST, BASE_RAY, BASE_SILHOUETTE = 0, 1, 2  # dw_query_baseline kinds


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("pts", C.c_void_p), ("n", C.c_int64), ("params", SolverParams),
                ("boundary", Boundary), ("u", C.c_void_p), ("grad", C.c_void_p)]

BATCH_GRAPH = 1


def lib():
    """Load libdwost.so; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("DWOST_LIB", LIB)
    if not os.path.exists(path):
        raise DwostError(f"{path} is missing: build it with `python -m dwost_amd.build` (no CPU fallback exists)")
    L = C.CDLL(path)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.dw_last_error.restype = C.c_char_p
    L.dw_version.restype = C.c_int
    L.dw_scene_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(vp)]
    L.dw_scene_destroy.argtypes = [vp]
    L.dw_scene_info.argtypes = [vp, i32, vp]
    L.dw_scene_export.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.dw_closest_point.argtypes = [vp, i32, vp, i64, vp, vp, vp, vp, vp, vp]
    L.dw_intersect.argtypes = [vp, i32, vp, vp, vp, i64, i32, vp, vp, vp, vp, vp, vp]
    L.dw_closest_silhouette.argtypes = [vp, i32, vp, vp, i64, vp, vp, vp, vp]
    L.dw_star_radius.argtypes = [vp, i32, vp, vp, i64, vp, vp]
    L.dw_star_radius_rho.argtypes = [vp, i32, vp, vp, C.c_float, i64, vp, vp]
    L.dw_math_check.argtypes = [vp, vp]
    L.dw_sdf.argtypes = [vp, vp, i64, vp, vp]
    L.dw_sample_boundary.argtypes = [vp, i32, vp, i64, vp, vp, vp, vp, vp, vp]
    L.dw_query_baseline.argtypes = [vp, i32, i32, vp, vp, vp, i64, vp, vp, vp]
    L.dw_scene_prim_map.argtypes = [vp, i32, vp]
    L.dw_solve_u.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), vp, vp, vp, vp, vp, vp, vp]
    L.dw_solve_dude.argtypes = [vp, vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), vp, vp, vp]
    L.dw_grad_solve_dude.argtypes = [vp, vp, vp, i64, C.POINTER(SolverParams), C.POINTER(GradParams),
                                     C.POINTER(Boundary), vp, vp, vp, vp, vp]
    L.dw_grad_trace_step.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(GradParams), C.POINTER(Boundary),
                                     C.c_int32, C.c_float, vp, vp, vp]
    L.dw_trace_step.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), C.c_int32, C.c_float, vp,
                                vp, vp]
    L.dw_solve_fused.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), vp, vp, vp, vp, vp, vp,
                                 vp]
    L.dw_solve_fused_noalloc.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), vp, vp, vp, vp]
    L.dw_batch_create.argtypes = [C.POINTER(BatchItem), i32, vp, i32, C.POINTER(vp)]
    L.dw_batch_solve.argtypes = [vp, vp]
    L.dw_batch_destroy.argtypes = [vp]
    L.dw_batch_destroy.restype = None
    L.dw_adjoint_u.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), vp, vp, vp, vp, vp]
    L.dw_adjoint_fused.argtypes = [vp, vp, i64, C.POINTER(SolverParams), C.POINTER(Boundary), vp, vp, vp, vp, vp,
                                   vp]
    for name in ("dw_scene_create", "dw_scene_destroy", "dw_scene_info", "dw_scene_export", "dw_closest_point",
                 "dw_intersect", "dw_closest_silhouette", "dw_star_radius", "dw_star_radius_rho", "dw_math_check",
                 "dw_sdf", "dw_sample_boundary",
                 "dw_query_baseline", "dw_scene_prim_map", "dw_solve_u", "dw_solve_dude", "dw_grad_solve_dude", "dw_grad_trace_step",
                 "dw_trace_step",
                 "dw_solve_fused", "dw_solve_fused_noalloc", "dw_adjoint_u", "dw_adjoint_fused", "dw_batch_create", "dw_batch_solve"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != DW_OK:
        msg = lib().dw_last_error().decode(errors="replace")
        raise DwostError(f"libdwost error {rc}: {msg}")


def ptr(t) -> C.c_void_p | None:
    """Device (or host numpy) pointer of a contiguous tensor / array; None passes NULL."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        if not t.is_contiguous():
            raise DwostError("tensor arguments must be contiguous")
        return C.c_void_p(t.data_ptr())
    return t.ctypes.data_as(C.c_void_p)


def stream_handle(device: torch.device | None = None):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)
