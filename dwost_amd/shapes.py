"""Scene2D / Scene3D: the diff_wost scene API over device-resident BVH + SNCH.

Mirrors ``diff_wost/shapes/polyline.py`` (``Scene2D(vertices, indices, types)``,
:623-692) and ``diff_wost/shapes/mesh.py`` (``Scene3D``, :592-635).  The
reference scene derives ``g`` and ``h`` from Python callables ``u`` / ``dudx``
defined on a Scene subclass (polyline.py:52-58); a HIP kernel cannot trace
Python, so boundary data is a :class:`BoundaryCondition` -- an analytic
harmonic field code or per-vertex ``g`` / ``h`` tensors.

Queries take ``(N, dim)`` CUDA tensors (the reference's ``(dim, N)`` layout is
accepted too) and return CUDA tensors.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum

import numpy as np
import torch

from . import _lib
from ._lib import SUB_DIRICHLET, SUB_FULL, SUB_NEUMANN, check, lib, ptr, stream_handle

BC_EXPCOS, BC_QUAD, BC_LINEAR, BC_PER_VERTEX = 0, 1, 2, 3
SRC_ZERO, SRC_CONST, SRC_LINEAR = 0, 1, 2


class BoundaryType(Enum):
    """diff_wost/shapes/primitive.py:10-19."""

    Neumann = 0
    Dirichlet = 1


@dataclass
class BoundaryCondition:
    """g = u(p) on Dirichlet boundaries, h = n . grad u(p) on Neumann ones.

    ``code`` selects u: BC_EXPCOS (e^x cos y), BC_QUAD (x^2 - y^2),
    BC_LINEAR (c0 + c1 x + c2 y + c3 z) or BC_PER_VERTEX (``g`` / ``h`` float32
    CUDA tensors with one value per scene vertex, interpolated linearly along
    segments and barycentrically over triangles).  ``source`` is the
    reference's ``scene.f``: a constant, or (s0, s1, s2[, s3]) for the linear
    f = s0 + s1 x + s2 y + s3 z."""

    code: int = BC_EXPCOS
    coeffs: tuple = (0.0, 0.0, 0.0, 0.0)
    g: torch.Tensor | None = None
    h: torch.Tensor | None = None
    source: float | tuple = 0.0

    @staticmethod
    def exp_cos() -> "BoundaryCondition":
        return BoundaryCondition(BC_EXPCOS)

    @staticmethod
    def quadratic() -> "BoundaryCondition":
        return BoundaryCondition(BC_QUAD)

    @staticmethod
    def linear(c0=0.0, c1=0.0, c2=0.0, c3=0.0) -> "BoundaryCondition":
        return BoundaryCondition(BC_LINEAR, (c0, c1, c2, c3))

    @staticmethod
    def per_vertex(g: torch.Tensor, h: torch.Tensor) -> "BoundaryCondition":
        return BoundaryCondition(BC_PER_VERTEX, g=g, h=h)

    def op_args(self) -> tuple:
        """(bc_code, coeffs, g, h, source): the boundary arguments of the
        torch.ops.dwost solvers."""
        c = [float(self.coeffs[k]) if k < len(self.coeffs) else 0.0 for k in range(4)]
        pv = self.code == BC_PER_VERTEX
        return int(self.code), c, self.g if pv else None, self.h if pv else None, self.source_list()

    def source_list(self) -> list[float]:
        """[source code, its four coefficients] (constant: the value first)."""
        if isinstance(self.source, (tuple, list)):
            s = [float(v) for v in self.source] + [0.0] * (4 - len(self.source))
            return [float(SRC_LINEAR)] + s[:4]
        v = float(self.source)
        return [float(SRC_CONST if v != 0.0 else SRC_ZERO), v, 0.0, 0.0, 0.0]

    def to_c(self, n_vertices: int | None = None) -> _lib.Boundary:
        b = _lib.Boundary()
        b.bc_code = int(self.code)
        for k in range(4):
            b.coeffs[k] = float(self.coeffs[k]) if k < len(self.coeffs) else 0.0
        if self.code == BC_PER_VERTEX:
            for name, t in (("g", self.g), ("h", self.h)):
                if t is None or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
                    raise ValueError(f"per-vertex {name} must be a contiguous float32 CUDA tensor")
                if n_vertices is not None and t.numel() != n_vertices:
                    raise ValueError(f"per-vertex {name} needs {n_vertices} values, got {t.numel()}")
            b.g_vertex = self.g.data_ptr()
            b.h_vertex = self.h.data_ptr()
        src = self.source_list()
        b.src_code = int(src[0])
        b.src_value = src[1]
        for k in range(4):
            b.src_coeffs[k] = src[1 + k]
        return b


def _as_points(p, dim: int) -> torch.Tensor:
    if not isinstance(p, torch.Tensor):
        p = torch.as_tensor(np.asarray(p, dtype=np.float32))
    if p.dim() != 2:
        raise ValueError("points must be a 2-D tensor")
    if p.shape[1] != dim and p.shape[0] == dim:
        p = p.t()
    if p.shape[1] != dim:
        raise ValueError(f"points must have shape (N, {dim})")
    return p.to(device="cuda", dtype=torch.float32).contiguous()


class _Scene:
    dim = 2

    def __init__(self, vertices, indices, types, boundary: BoundaryCondition | None = None, host_only: bool = False):
        V = np.ascontiguousarray(np.asarray(_np(vertices), dtype=np.float32))
        I = np.ascontiguousarray(np.asarray(_np(indices), dtype=np.int32))
        T = np.ascontiguousarray(np.asarray(_np(types), dtype=np.int32).reshape(-1))
        d = self.dim
        if V.ndim == 2 and V.shape[1] != d and V.shape[0] == d:
            V = np.ascontiguousarray(V.T)
        if I.ndim == 2 and I.shape[1] != d and I.shape[0] == d:
            I = np.ascontiguousarray(I.T)
        if V.ndim != 2 or V.shape[1] != d or I.ndim != 2 or I.shape[1] != d or len(T) != len(I):
            raise ValueError(f"expected vertices (V,{d}), indices (P,{d}), types (P,)")
        self.vertices, self.indices, self.types = V, I, T
        self.boundary = boundary if boundary is not None else BoundaryCondition()
        desc = _lib.SceneDesc(d, len(V), len(I), V.ctypes.data, I.ctypes.data, T.ctypes.data,
                              _lib.SCENE_HOST_ONLY if host_only else 0)
        h = C.c_void_p()
        self.host_only = host_only
        if not host_only:
            torch.cuda.init()
        check(lib().dw_scene_create(C.byref(desc), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            if not self.host_only:
                torch.cuda.synchronize()
            lib().dw_scene_destroy(h)
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def n_vertices(self) -> int:
        return len(self.vertices)

    @property
    def n_prims(self) -> int:
        return len(self.indices)

    # -- structure -------------------------------------------------------
    def info(self, sub: int = SUB_FULL) -> dict:
        out = np.zeros(8, np.int64)
        check(lib().dw_scene_info(self._h, sub, out.ctypes.data_as(C.c_void_p)))
        keys = ["present", "nverts", "nprims", "nnodes", "nrefs", "max_depth", "py_sils", "nsil_topology"]
        return {k: int(v) for k, v in zip(keys, out)}

    def export(self, sub: int = SUB_FULL) -> dict:
        inf = self.info(sub)
        D = self.dim
        nn, npr, nr = inf["nnodes"], inf["nprims"], inf["nrefs"]
        box = np.zeros((nn, 2 * D), np.float32)
        ints = np.zeros((nn, 4), np.int32)
        cone = np.zeros((nn, D + 2), np.float32)
        pidx = np.zeros(npr, np.int32)
        refs = np.zeros(max(nr, 1), np.int32)
        check(lib().dw_scene_export(self._h, sub, *(a.ctypes.data_as(C.c_void_p) for a in (box, ints, cone, pidx, refs))))
        return dict(box=box, ints=ints, cone=cone, prim_index=pidx, refs=refs[:nr], **inf)

    @property
    def dirichlet_present(self) -> bool:
        return bool(self.info(SUB_DIRICHLET)["present"])

    @property
    def neumann_present(self) -> bool:
        return bool(self.info(SUB_NEUMANN)["present"])

    # -- queries ---------------------------------------------------------
    def closest_point(self, p, sub: int = SUB_FULL) -> dict:
        """BVH.closest_point (bvh.py:111-188) / SNCH3D.closest_point (snch3d.py:334-409)."""
        p = _as_points(p, self.dim)
        n = p.shape[0]
        dev = p.device
        out = dict(prim=torch.empty(n, dtype=torch.int32, device=dev), d=torch.empty(n, device=dev),
                   p=torch.empty(n, self.dim, device=dev), n=torch.empty(n, self.dim, device=dev),
                   tuv=torch.empty(n, 2, device=dev))
        check(lib().dw_closest_point(self._h, sub, ptr(p), n, ptr(out["prim"]), ptr(out["d"]), ptr(out["p"]),
                                     ptr(out["n"]), ptr(out["tuv"]), stream_handle(dev)))
        return out

    def distance(self, p, sub: int = SUB_FULL) -> torch.Tensor:
        return self.closest_point(p, sub)["d"]

    def intersect(self, origin, direction, r_max, sub: int = SUB_FULL, nearest: bool = False) -> dict:
        """BVH.intersect (bvh.py:190-257) / SNCH3D.intersect (snch3d.py:411-476); prim = -1 on a miss.

        2D rays return the first valid hit in traversal order, as the reference's
        BVH does (bvh.py:215-217); ``nearest=True`` returns the nearest hit
        (Polyline.intersect_baseline, polyline.py:468-496).  3D rays are nearest
        either way."""
        o = _as_points(origin, self.dim)
        v = _as_points(direction, self.dim)
        n = o.shape[0]
        dev = o.device
        r = torch.as_tensor(r_max, dtype=torch.float32, device=dev).expand(n).contiguous()
        out = dict(prim=torch.empty(n, dtype=torch.int32, device=dev), d=torch.empty(n, device=dev),
                   p=torch.empty(n, self.dim, device=dev), n=torch.empty(n, self.dim, device=dev),
                   tuv=torch.empty(n, 2, device=dev))
        mode = _lib.RAY_NEAREST if nearest else _lib.RAY_REFERENCE
        check(lib().dw_intersect(self._h, sub, ptr(o), ptr(v), ptr(r), n, mode, ptr(out["prim"]), ptr(out["d"]),
                                 ptr(out["p"]), ptr(out["n"]), ptr(out["tuv"]), stream_handle(dev)))
        return out

    def closest_silhouette(self, p, r_max=float("inf"), sub: int = SUB_FULL) -> dict:
        """SNCH closest silhouette (snch.py:312-378, snch3d.py:594-663); sil = -1 if none."""
        p = _as_points(p, self.dim)
        n = p.shape[0]
        dev = p.device
        r = torch.as_tensor(r_max, dtype=torch.float32, device=dev).expand(n).contiguous()
        out = dict(sil=torch.empty(n, dtype=torch.int32, device=dev), d=torch.empty(n, device=dev),
                   p=torch.empty(n, self.dim, device=dev))
        check(lib().dw_closest_silhouette(self._h, sub, ptr(p), ptr(r), n, ptr(out["sil"]), ptr(out["d"]),
                                          ptr(out["p"]), stream_handle(dev)))
        return out

    def sdf(self, p) -> torch.Tensor:
        """Polyline.sdf / Mesh.sdf (polyline.py:188-194, mesh.py:290-293): the
        full-scene distance, negative inside."""
        p = _as_points(p, self.dim)
        out = torch.empty(p.shape[0], device=p.device)
        check(lib().dw_sdf(self._h, ptr(p), p.shape[0], ptr(out), stream_handle(p.device)))
        return out

    def inside(self, p) -> torch.Tensor:
        """Polyline.inside / Mesh.inside (polyline.py:232-234, mesh.py:295-297): sdf < 0."""
        return self.sdf(p) < 0

    def sample_boundary(self, u, sub: int = SUB_FULL) -> dict:
        """Polyline.sample_boundary / Mesh.sample_boundary (polyline.py:498-512,
        mesh.py:474-490) with the sampler's uniform draws given: ``u`` of shape
        (N,) in 2D (the edge CDF draw) or (N, 3) in 3D (face CDF draw, r1, r2),
        values in [0, 1).  Returns p, n, t (2D) / uv (3D), pdf and prim (the
        sub-scene's primitive id)."""
        u = torch.as_tensor(u, dtype=torch.float32, device="cuda").contiguous()
        want = (u.shape[0],) if self.dim == 2 else (u.shape[0], 3)
        if tuple(u.shape) != want:
            raise ValueError(f"draws must have shape {'(N,)' if self.dim == 2 else '(N, 3)'}")
        n, dev, D = u.shape[0], u.device, self.dim
        out = dict(p=torch.empty(n, D, device=dev), n=torch.empty(n, D, device=dev), tuv=torch.empty(n, 2, device=dev),
                   pdf=torch.empty(n, device=dev), prim=torch.empty(n, dtype=torch.int32, device=dev))
        check(lib().dw_sample_boundary(self._h, sub, ptr(u), n, ptr(out["p"]), ptr(out["n"]), ptr(out["tuv"]),
                                       ptr(out["pdf"]), ptr(out["prim"]), stream_handle(dev)))
        if D == 2:
            out["t"] = out["tuv"][:, 0]
        else:
            out["uv"] = out["tuv"]
        return out

    def _baseline(self, kind, p, v=None, r_max=None, sub=SUB_FULL) -> dict:
        p = _as_points(p, self.dim)
        n, dev = p.shape[0], p.device
        v = None if v is None else _as_points(v, self.dim)
        r = None if r_max is None else torch.as_tensor(r_max, dtype=torch.float32, device=dev).expand(n).contiguous()
        out = dict(prim=torch.empty(n, dtype=torch.int32, device=dev), d=torch.empty(n, device=dev))
        check(lib().dw_query_baseline(self._h, sub, kind, ptr(p), ptr(v), ptr(r), n, ptr(out["prim"]),
                                      ptr(out["d"]), stream_handle(dev)))
        return out

    def closest_point_baseline(self, p, sub: int = SUB_FULL) -> dict:
        """Brute-force closest point (polyline.py:246-279, mesh.py:312-335): prim, d."""
        return self._baseline(_lib.BASE_CLOSEST, p, sub=sub)

    def intersect_baseline(self, origin, direction, r_max=float("inf"), sub: int = SUB_FULL) -> dict:
        """Brute-force nearest ray hit (polyline.py:468-496, mesh.py:440-472): prim (-1: miss), d."""
        return self._baseline(_lib.BASE_RAY, origin, direction, r_max, sub)

    def closest_silhouette_baseline(self, p, r_max=float("inf"), sub: int = SUB_FULL) -> dict:
        """Brute-force closest silhouette (polyline.py:289-302, mesh.py:558-570)
        over the topology silhouettes: prim = topology id (-1: none), d."""
        return self._baseline(_lib.BASE_SILHOUETTE, p, None, r_max, sub)

    def global_prim_ids(self, sub: int) -> np.ndarray:
        """Sub-scene primitive id -> scene primitive id (polyline.py:636-645)."""
        inf = self.info(sub)
        out = np.zeros(max(inf["nprims"], 1), np.int32)
        check(lib().dw_scene_prim_map(self._h, sub, out.ctypes.data_as(C.c_void_p)))
        return out[: inf["nprims"]]

    @property
    def dirichlet_2_global_indices(self) -> np.ndarray:
        return self.global_prim_ids(SUB_DIRICHLET)

    @property
    def neumann_2_global_indices(self) -> np.ndarray:
        return self.global_prim_ids(SUB_NEUMANN)

    def star_radius(self, p, r_max=float("inf"), sub: int = SUB_FULL) -> torch.Tensor:
        """Polyline.star_radius (polyline.py:304-314) / Mesh.star_radius (mesh.py:337-346)."""
        p = _as_points(p, self.dim)
        n = p.shape[0]
        r = torch.as_tensor(r_max, dtype=torch.float32, device=p.device).expand(n).contiguous()
        out = torch.empty(n, device=p.device)
        check(lib().dw_star_radius(self._h, sub, ptr(p), ptr(r), n, ptr(out), stream_handle(p.device)))
        return out

    def star_radius_2(self, p, e, rho_max: float | None = None, sub: int = SUB_FULL) -> torch.Tensor:
        """Polyline.star_radius_2 (polyline.py:317-330) / Mesh.star_radius_2
        (mesh.py:349-361, r_max = inf): the star radius along e bounded by the
        Poisson-kernel ratio rho_max (reference defaults 1.0 in 2D, 1.01 in 3D)."""
        p = _as_points(p, self.dim)
        e = _as_points(e, self.dim).to(p.device)
        if e.shape != p.shape:
            raise ValueError("e must have one direction per point")
        if rho_max is None:
            rho_max = 1.0 if self.dim == 2 else 1.01
        out = torch.empty(p.shape[0], device=p.device)
        check(lib().dw_star_radius_rho(self._h, sub, ptr(p), ptr(e), float(rho_max), p.shape[0], ptr(out),
                                       stream_handle(p.device)))
        return out


def _np(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return a


class Scene2D(_Scene):
    """Polyline scene with Dirichlet / Neumann segments (polyline.py:623-692)."""

    dim = 2


class Scene3D(_Scene):
    """Triangle-mesh scene with Dirichlet / Neumann faces (mesh.py:592-635)."""

    dim = 3
