"""Seeded synthetic scenes and query sets for the Walk-on-Stars workloads.

All geometry is generated in float64 with ``numpy.random.default_rng(seed)``
and cast to float32.  Boundary conventions follow the reference examples
(``examples/example_2d/scene.py:61-64,90``): outer loops are counter-clockwise
with edges ``(i, i+1)`` so that ``n = (d.y, -d.x)`` points outward; holes are
clockwise.  Types use ``BoundaryType`` (``diff_wost/shapes/primitive.py:18-19``):
Neumann = 0, Dirichlet = 1.

Workloads:
  * ``star_scene_2d``   -- one seeded star polygon (256 segments, mixed D/N),
                           g = e^x cos y, 1024 interior queries.
  * ``holes_scene_2d``  -- 4096-segment star with 8 Neumann holes, g = x^2-y^2,
                           65,536 queries on a square lattice.
  * ``icosphere_scene`` -- noisy icosphere (81,920 triangles), y>0 Dirichlet,
                           g = e^x cos y, 262,144 interior queries.
  * ``batch_scenes_2d`` -- 64 uneven star polygons with 0-4 holes.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NEUMANN = 0
DIRICHLET = 1

# Boundary-value codes understood by the kernels (g = u(p) on Dirichlet,
# h = n . grad u on Neumann, from an analytic harmonic u).
BC_EXPCOS = 0  # u = e^x cos y
BC_QUAD = 1  # u = x^2 - y^2
BC_LINEAR = 2  # u = c0 + c1 x + c2 y + c3 z
BC_PER_VERTEX = 3  # piecewise-linear per-vertex g and h buffers


@dataclass
class SceneSpec:
    dim: int
    vertices: np.ndarray  # (V, dim) float32
    indices: np.ndarray  # (P, dim) int32
    types: np.ndarray  # (P,) int32
    bc_code: int
    bc_coeffs: tuple = (0.0, 0.0, 0.0, 0.0)
    queries: np.ndarray | None = None  # (N, dim) float32
    meta: dict = field(default_factory=dict)


# ---------------------------------------------------------------------------
# analytic harmonic fields (float64) used as the ground truth
# ---------------------------------------------------------------------------
def analytic_u(code: int, p: np.ndarray, coeffs=(0.0, 0.0, 0.0, 0.0)) -> np.ndarray:
    p = np.asarray(p, dtype=np.float64)
    x, y = p[:, 0], p[:, 1]
    if code == BC_EXPCOS:
        return np.exp(x) * np.cos(y)
    if code == BC_QUAD:
        return x * x - y * y
    z = p[:, 2] if p.shape[1] > 2 else 0.0
    return coeffs[0] + coeffs[1] * x + coeffs[2] * y + coeffs[3] * z


def analytic_grad(code: int, p: np.ndarray, coeffs=(0.0, 0.0, 0.0, 0.0)) -> np.ndarray:
    p = np.asarray(p, dtype=np.float64)
    x, y = p[:, 0], p[:, 1]
    g = np.zeros_like(p)
    if code == BC_EXPCOS:
        g[:, 0] = np.exp(x) * np.cos(y)
        g[:, 1] = -np.exp(x) * np.sin(y)
    elif code == BC_QUAD:
        g[:, 0] = 2 * x
        g[:, 1] = -2 * y
    else:
        g[:, 0] = coeffs[1]
        g[:, 1] = coeffs[2]
        if p.shape[1] > 2:
            g[:, 2] = coeffs[3]
    return g


# ---------------------------------------------------------------------------
# 2D geometry helpers (float64)
# ---------------------------------------------------------------------------
def _star_loop(rng: np.random.Generator, n: int, k: int = 5, amp: float = 0.3, jitter: float = 0.02):
    phi = rng.uniform(0.0, 2.0 * math.pi)
    theta = 2.0 * math.pi * np.arange(n) / n
    r = 1.0 + amp * np.sin(k * theta + phi) + jitter * rng.uniform(-1.0, 1.0, size=n)
    return np.stack([r * np.cos(theta), r * np.sin(theta)], axis=1)


def _circle_loop(c, radius, n):
    theta = 2.0 * math.pi * np.arange(n) / n
    return np.stack([c[0] + radius * np.cos(theta), c[1] + radius * np.sin(theta)], axis=1)


def segment_distance(pts: np.ndarray, a: np.ndarray, b: np.ndarray, chunk: int = 4096) -> np.ndarray:
    """Unsigned distance from each point to the nearest segment (float64)."""
    out = np.empty(len(pts))
    ba = b - a
    bb = np.maximum((ba * ba).sum(1), 1e-300)
    for s in range(0, len(pts), chunk):
        p = pts[s : s + chunk, None, :]
        pa = p - a[None]
        h = np.clip((pa * ba[None]).sum(2) / bb[None], 0.0, 1.0)
        d = pa - ba[None] * h[..., None]
        out[s : s + chunk] = np.sqrt((d * d).sum(2)).min(1)
    return out


def winding_inside(pts: np.ndarray, a: np.ndarray, b: np.ndarray, chunk: int = 4096) -> np.ndarray:
    """Even-odd crossing test against a set of closed loops (float64)."""
    out = np.empty(len(pts), dtype=bool)
    for s in range(0, len(pts), chunk):
        p = pts[s : s + chunk, None, :]
        ay, by = a[None, :, 1], b[None, :, 1]
        ax, bx = a[None, :, 0], b[None, :, 0]
        py, px = p[..., 1], p[..., 0]
        cond = (ay > py) != (by > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = ax + (py - ay) * (bx - ax) / (by - ay)
        cross = cond & (px < xi)
        out[s : s + chunk] = (cross.sum(1) % 2) == 1
    return out


def _loops_to_mesh(loops, loop_types):
    verts, idx, types = [], [], []
    base = 0
    for loop, t in zip(loops, loop_types):
        n = len(loop)
        verts.append(loop)
        e = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1) + base
        idx.append(e)
        types.append(np.full(n, t, dtype=np.int32) if np.isscalar(t) else t)
        base += n
    return np.concatenate(verts), np.concatenate(idx).astype(np.int32), np.concatenate(types).astype(np.int32)


def _interior_2d(pts, V, I, clearance):
    a, b = V[I[:, 0]], V[I[:, 1]]
    ins = winding_inside(pts, a, b)
    d = segment_distance(pts, a, b)
    return ins & (d > clearance)


def _interior_binned_2d(pts, V, I, clearance, nbins: int = 256, chunk: int = 2048):
    """_interior_2d with the points sorted into horizontal bins: each bin tests
    only the segments whose y range, widened by the clearance (plus 1e-9), meets
    the bin (one bin of slack each side), which holds every segment a point's
    horizontal ray can cross and every segment within the clearance of it.  The
    per-segment float64 arithmetic is the dense test's, so the mask is the same."""
    out = np.zeros(len(pts), dtype=bool)
    if len(pts) == 0:
        return out
    a, b = V[I[:, 0]], V[I[:, 1]]
    ba = b - a
    bb = np.maximum((ba * ba).sum(1), 1e-300)
    band = clearance + 1e-9
    ylo = np.minimum(a[:, 1], b[:, 1]) - band
    yhi = np.maximum(a[:, 1], b[:, 1]) + band
    y0, y1 = pts[:, 1].min(), pts[:, 1].max()
    h = (y1 - y0) / nbins if y1 > y0 else 1.0
    pbin = np.clip(np.floor((pts[:, 1] - y0) / h), 0, nbins - 1).astype(np.int64)
    s0 = np.clip(np.floor((ylo - y0) / h) - 1, 0, nbins - 1).astype(np.int64)
    s1 = np.clip(np.floor((yhi - y0) / h) + 1, 0, nbins - 1).astype(np.int64)
    order = np.argsort(pbin, kind="stable")
    starts = np.searchsorted(pbin[order], np.arange(nbins + 1))
    for k in range(nbins):
        idx_all = order[starts[k] : starts[k + 1]]
        if len(idx_all) == 0:
            continue
        segs = np.nonzero((s0 <= k) & (k <= s1))[0]
        if len(segs) == 0:
            continue  # no crossing: outside
        sa, sbb, sba, sbn = a[segs], b[segs], ba[segs], bb[segs]
        for c in range(0, len(idx_all), chunk):
            idx = idx_all[c : c + chunk]
            p = pts[idx, None, :]
            py, px = p[..., 1], p[..., 0]
            ay, by, ax, bx = sa[None, :, 1], sbb[None, :, 1], sa[None, :, 0], sbb[None, :, 0]
            cond = (ay > py) != (by > py)
            with np.errstate(divide="ignore", invalid="ignore"):
                xi = ax + (py - ay) * (bx - ax) / (by - ay)
            inside = ((cond & (px < xi)).sum(1) % 2) == 1
            pa = p - sa[None]
            hh = np.clip((pa * sba[None]).sum(2) / sbn[None], 0.0, 1.0)
            d = pa - sba[None] * hh[..., None]
            out[idx] = inside & (np.sqrt((d * d).sum(2)).min(1) > clearance)
    return out


def _interior_lattice_2d(xs, ys, V, I, clearance):
    """_interior_2d for the lattice xs x ys (float64 values already rounded to
    float32), row by row in meshgrid "xy" order, with the same per-segment float64
    arithmetic: a row's crossings come from the segments that straddle its y
    (counted by a sorted search), and its distances from the segments whose y
    range lies within the clearance (plus 1e-9) of it, the only ones that can be
    that near.  Returns exactly what _interior_2d returns on those points."""
    a, b = V[I[:, 0]], V[I[:, 1]]
    ba = b - a
    bb = np.maximum((ba * ba).sum(1), 1e-300)
    ylo, yhi = np.minimum(a[:, 1], b[:, 1]), np.maximum(a[:, 1], b[:, 1])
    band = clearance + 1e-9
    out = np.empty((len(ys), len(xs)), dtype=bool)
    for r, py in enumerate(ys):
        ay, by, ax, bx = a[:, 1], b[:, 1], a[:, 0], b[:, 0]
        cond = (ay > py) != (by > py)
        xi = np.sort(ax[cond] + (py - ay[cond]) * (bx[cond] - ax[cond]) / (by[cond] - ay[cond]))
        inside = ((len(xi) - np.searchsorted(xi, xs, side="right")) % 2) == 1
        near = (ylo - band <= py) & (py <= yhi + band)
        far = np.ones(len(xs), dtype=bool)
        if near.any():
            p = np.stack([xs, np.full(len(xs), py)], 1)[:, None, :]
            pa = p - a[near][None]
            h = np.clip((pa * ba[near][None]).sum(2) / bb[near][None], 0.0, 1.0)
            d = pa - ba[near][None] * h[..., None]
            far = np.sqrt((d * d).sum(2)).min(1) > clearance
        out[r] = inside & far
    return out.ravel()


def star_scene_2d(n_seg: int = 256, n_query: int = 1024, eps: float = 1e-3, scene_seed: int = 1,
                  query_seed: int = 2) -> SceneSpec:
    """Star polygon r = 1 + 0.3 sin(5 theta + phi) + 0.02 jitter; x<0 midpoints Dirichlet."""
    rng = np.random.default_rng(scene_seed)
    loop = _star_loop(rng, n_seg)
    V, I, _ = _loops_to_mesh([loop], [0])
    mid = 0.5 * (V[I[:, 0]] + V[I[:, 1]])
    types = np.where(mid[:, 0] < 0.0, DIRICHLET, NEUMANN).astype(np.int32)
    Vf = V.astype(np.float32).astype(np.float64)
    qrng = np.random.default_rng(query_seed)
    lo, hi = Vf.min(0), Vf.max(0)
    pts = np.zeros((0, 2))
    while len(pts) < n_query:
        cand = qrng.uniform(lo, hi, size=(4 * n_query, 2))
        cand = cand.astype(np.float32).astype(np.float64)
        pts = np.concatenate([pts, cand[_interior_binned_2d(cand, Vf, I, 2 * eps)]])
    return SceneSpec(2, V.astype(np.float32), I, types, BC_EXPCOS, queries=pts[:n_query].astype(np.float32),
                     meta=dict(eps=eps, name=f"star{n_seg}"))


def holes_scene_2d(n_outer: int = 4096, n_holes: int = 8, hole_seg: int = 64, n_query: int = 65536,
                   eps: float = 1e-4, scene_seed: int = 1, lattice: bool = True, query_seed: int = 2,
                   bc_code: int = BC_QUAD) -> SceneSpec:
    """Star polygon (Dirichlet) with circular Neumann holes (clockwise loops)."""
    rng = np.random.default_rng(scene_seed)
    outer = _star_loop(rng, n_outer)
    oa = outer
    ob = np.roll(outer, -1, axis=0)
    holes = []
    tries = 0
    while len(holes) < n_holes:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place holes")
        rad = rng.uniform(0.05, 0.15)
        c = rng.uniform(-1.0, 1.0, size=2)
        if not winding_inside(c[None], oa, ob)[0]:
            continue
        if segment_distance(c[None], oa, ob)[0] < rad + 0.02:
            continue
        if any(np.hypot(*(c - hc)) < rad + hr + 0.02 for hc, hr in holes):
            continue
        holes.append((c, rad))
    loops = [outer] + [_circle_loop(c, r, hole_seg)[::-1] for c, r in holes]
    loop_types = [DIRICHLET] + [NEUMANN] * n_holes
    V, I, types = _loops_to_mesh(loops, loop_types)
    Vf = V.astype(np.float32).astype(np.float64)
    a, b = Vf[I[:, 0]], Vf[I[:, 1]]
    lo, hi = Vf.min(0), Vf.max(0)
    if lattice:
        k = int(math.ceil(math.sqrt(n_query)))
        while True:
            xs = np.linspace(lo[0], hi[0], k + 2)[1:-1]
            ys = np.linspace(lo[1], hi[1], k + 2)[1:-1]
            X, Y = np.meshgrid(xs, ys, indexing="xy")
            cand = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32).astype(np.float64)
            ok = _interior_lattice_2d(xs.astype(np.float32).astype(np.float64),
                                      ys.astype(np.float32).astype(np.float64), Vf, I, 2 * eps)
            if ok.sum() >= n_query:
                pts = cand[ok][:n_query]
                break
            k = int(math.ceil(k * math.sqrt(n_query / max(ok.sum(), 1)) * 1.01)) + 1
    else:
        qrng = np.random.default_rng(query_seed)
        pts = np.zeros((0, 2))
        while len(pts) < n_query:
            cand = qrng.uniform(lo, hi, size=(4 * n_query, 2)).astype(np.float32).astype(np.float64)
            pts = np.concatenate([pts, cand[_interior_binned_2d(cand, Vf, I, 2 * eps)]])
        pts = pts[:n_query]
    return SceneSpec(2, V.astype(np.float32), I, types, bc_code, queries=pts.astype(np.float32),
                     meta=dict(eps=eps, name=f"holes{n_outer}x{n_holes}", holes=holes))


# ---------------------------------------------------------------------------
# 3D: noisy icosphere
# ---------------------------------------------------------------------------
def icosphere(level: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = V[i] + V[j]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    return np.array(V), np.array(F, dtype=np.int32)


def _real_sh_basis(dirs: np.ndarray, lmax: int) -> np.ndarray:
    from scipy.special import lpmv

    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    th = np.arccos(np.clip(z, -1, 1))
    ph = np.arctan2(y, x)
    cols = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            P = lpmv(am, l, np.cos(th))
            if m == 0:
                cols.append(norm * P)
            elif m > 0:
                cols.append(math.sqrt(2) * norm * P * np.cos(am * ph))
            else:
                cols.append(math.sqrt(2) * norm * P * np.sin(am * ph))
    return np.stack(cols, 1)


def _noise_coeffs(seed: int, lmax: int = 6):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((lmax + 1) ** 2)


def icosphere_scene(level: int = 6, n_query: int = 262144, eps: float = 1e-3, scene_seed: int = 3,
                    query_seed: int = 4, amp: float = 0.15) -> SceneSpec:
    """Icosphere with seeded low-frequency radial noise; centroid y>0 faces are Dirichlet."""
    V0, F = icosphere(level)
    coeff = _noise_coeffs(scene_seed)
    noise = _real_sh_basis(V0, 6) @ coeff
    scale = 1.0 / np.abs(noise).max()
    r = 1.0 + amp * noise * scale
    V = V0 * r[:, None]
    cen = V[F].mean(1)
    types = np.where(cen[:, 1] > 0.0, DIRICHLET, NEUMANN).astype(np.int32)
    qrng = np.random.default_rng(query_seed)
    pts = np.zeros((0, 3))
    lo, hi = V.min(0), V.max(0)
    margin = 2 * eps + 2e-3  # exceeds the piecewise-linear vs smooth-surface gap
    while len(pts) < n_query:
        cand = qrng.uniform(lo, hi, size=(2 * n_query, 3))
        rad = np.linalg.norm(cand, axis=1)
        dirs = cand / np.maximum(rad, 1e-30)[:, None]
        rs = 1.0 + amp * (_real_sh_basis(dirs, 6) @ coeff) * scale
        pts = np.concatenate([pts, cand[rad < rs - margin]])
    return SceneSpec(3, V.astype(np.float32), F, types, BC_EXPCOS, queries=pts[:n_query].astype(np.float32),
                     meta=dict(eps=eps, name=f"icosphere{len(F)}", coeff=coeff, scale=scale, amp=amp))


def batch_scenes_2d(n_scenes: int = 64, seed: int = 5, eps: float = 1e-3, max_queries: int | None = None):
    """Uneven batch: segments log-U[64,16384], 0-4 holes x 64 segments, queries log-U[100,1e5]."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        nseg = int(round(math.exp(rng.uniform(math.log(64), math.log(16384)))))
        nh = int(rng.integers(0, 5))
        code = BC_EXPCOS if rng.integers(0, 2) == 0 else BC_QUAD
        nq = int(round(math.exp(rng.uniform(math.log(100), math.log(1e5)))))
        W = int(rng.choice([64, 256, 1024]))
        if max_queries is not None:
            nq = min(nq, max_queries)
        sub_seed = int(rng.integers(0, 2**31 - 1))
        spec = holes_scene_2d(n_outer=nseg, n_holes=nh, hole_seg=64, n_query=nq, eps=eps, scene_seed=sub_seed,
                              lattice=False, query_seed=sub_seed + 1, bc_code=code)
        spec.meta["nwalks"] = W
        out.append(spec)
    return out


def smooth_vertex_field(values: np.ndarray, faces: np.ndarray, n_vertices: int, rounds: int = 8) -> np.ndarray:
    """``rounds`` passes of v <- (v + mean of edge neighbours) / 2 over the mesh graph."""
    import scipy.sparse as sp

    f = np.asarray(faces, np.int64)
    k = f.shape[1]
    rows = np.concatenate([f[:, i] for i in range(k) for j in range(k) if i != j])
    cols = np.concatenate([f[:, j] for i in range(k) for j in range(k) if i != j])
    A = sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n_vertices, n_vertices)).tocsr()
    A.data[:] = 1.0  # duplicate edges from the two faces sharing them count once
    deg = np.maximum(np.asarray(A.sum(1)).ravel(), 1.0)
    v = np.asarray(values, np.float64)
    for _ in range(rounds):
        v = 0.5 * v + 0.5 * (A @ v) / deg
    return v


@dataclass
class AdjointProblem:
    spec: SceneSpec
    g0: np.ndarray  # (V,) float32 initial per-vertex Dirichlet values
    h0: np.ndarray  # (V,) float32 initial per-vertex Neumann values
    u_target: np.ndarray  # (N,) float32 e^x cos y at the queries
    grad_target: np.ndarray  # (N, dim) float32 its gradient


def adjoint_problem_3d(level: int = 6, n_query: int = 16384, eps: float = 1e-3, g_seed: int = 6, h_seed: int = 7,
                       query_seed: int = 8, smooth_rounds: int = 10) -> AdjointProblem:
    """Noisy icosphere (level 6: 40,962 vertices) with trainable per-vertex g and h
    (N(0,1) from seeds 6 / 7, 10 rounds of uniform-Laplacian smoothing with
    lambda = 0.5) and the target field e^x cos y at interior points (seed 8);
    the loss is mean(|u - u*|^2 + |grad u - grad u*|^2)."""
    spec = icosphere_scene(level=level, n_query=n_query, eps=eps, query_seed=query_seed)
    V = len(spec.vertices)
    g0 = smooth_vertex_field(np.random.default_rng(g_seed).standard_normal(V), spec.indices, V, smooth_rounds)
    h0 = smooth_vertex_field(np.random.default_rng(h_seed).standard_normal(V), spec.indices, V, smooth_rounds)
    g0, h0 = g0.astype(np.float32), h0.astype(np.float32)
    spec.bc_code = BC_PER_VERTEX
    q = spec.queries
    return AdjointProblem(spec, g0, h0, analytic_u(BC_EXPCOS, q).astype(np.float32),
                          analytic_grad(BC_EXPCOS, q).astype(np.float32))
