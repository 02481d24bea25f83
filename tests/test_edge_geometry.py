"""Geometric and numeric edge cases on the GPU against the oracle:
* a polygon with a zero-length segment (a repeated vertex: 0/0 in the segment
  parameter, a NaN normal, zero sampling weight);
* queries exactly on shared vertices (closest-point ties between the two
  segments or several triangles meeting there), on silhouette vertices, and
  just inside and outside them along the vertex normal;
* per-vertex boundary data holding NaN and +Inf.
Queries must agree bit for bit; walks per walk (value bits, or NaN for NaN,
step count, terminal primitive)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import dwost_amd

    dwost_amd._lib.lib()
    return dwost_amd


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _eq(a, b):
    """Bitwise equality, NaN equal to NaN (payloads aside)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype != np.float32:
        return np.array_equal(a, b)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def _degenerate_polygon():
    """A 64-segment star polygon whose vertex 10 is repeated (segment 10 has
    zero length); x < 0 Dirichlet, the rest Neumann."""
    rng = np.random.default_rng(12)
    th = 2 * np.pi * np.arange(64) / 64
    r = 1 + 0.25 * np.sin(5 * th + 0.3) + 0.02 * rng.uniform(-1, 1, 64)
    P = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    P = np.insert(P, 11, P[10], axis=0).astype(np.float32)  # vertex 11 == vertex 10
    n = len(P)
    I = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    mid = 0.5 * (P[I[:, 0]] + P[I[:, 1]])
    T = (mid[:, 0] < 0).astype(np.int32)
    return P, I, T


def _vertex_queries(V, I, dim):
    """Every vertex exactly, and +-1e-4 / +-1e-2 along the averaged normal of
    its primitives."""
    nrm = np.zeros_like(V, dtype=np.float64)
    Vd = V.astype(np.float64)
    for f in I:
        if dim == 2:
            d = Vd[f[1]] - Vd[f[0]]
            n = np.array([d[1], -d[0]])
        else:
            n = np.cross(Vd[f[1]] - Vd[f[0]], Vd[f[2]] - Vd[f[0]])
        for v in f:
            nrm[v] += n
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    q = [V] + [(Vd + s * nrm).astype(np.float32) for s in (1e-4, -1e-4, 1e-2, -1e-2)]
    return np.concatenate(q).astype(np.float32)


def _query_parity(gs, os_, oracle, q, dim):
    for sub in range(3):
        if not os_.info(sub)["present"]:
            continue
        g = gs.closest_point(_t(q), sub=sub)
        o = os_.query(sub, oracle.KIND_CLOSEST, q)
        for k in ("prim", "d", "p", "n"):
            assert _eq(_np(g[k]), o[k]), (sub, k)
        for rm in (np.inf, 0.05):
            r = np.full(len(q), rm, np.float32)
            g = gs.closest_silhouette(_t(q), _t(r), sub=sub)
            o = os_.query(sub, oracle.KIND_SIL, q, rmax=r)
            assert _eq(_np(g["sil"]), o["prim"]), (sub, "sil", rm)
            assert _eq(_np(gs.star_radius(_t(q), _t(r), sub=sub)), os_.query(sub, oracle.KIND_STAR, q, rmax=r)["d"])
    rng = np.random.default_rng(5)
    v = rng.standard_normal(q.shape)
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    r = np.full(len(q), 0.3, np.float32)
    g = gs.intersect(_t(q), _t(v), _t(r))
    o = os_.query(0, oracle.KIND_RAY, q, v, r)
    assert _eq(_np(g["prim"]), o["prim"])
    hit = o["prim"] >= 0
    assert _eq(_np(g["d"])[hit], o["d"][hit])
    assert _eq(_np(gs.sdf(_t(q))), os_.sdf(q))


def _walk_parity(gpu, gs, os_, q, spec_bc, dim, gv=None, hv=None, W=16):
    Solver = gpu.WoSt if dim == 2 else gpu.WoSt3D
    sol = Solver(eps=1e-3, nwalks=W, nsteps=1024, exact_cones=True)
    g = sol.solve(_t(q), gs, seed=3, per_walk=True)
    o = os_.solve_u(q, spec_bc, W, nsteps=1024, eps=1e-3, min_R=sol.min_R, cv=sol.control_variates,
                    av=sol.antithetic_variates, seed=3, per_walk=True, gv=gv, hv=hv)
    assert _eq(_np(g["walk_u"]), o["walk_u"])
    assert np.array_equal(_np(g["walk_steps"]), o["walk_steps"])
    assert np.array_equal(_np(g["walk_term"]), o["walk_term"])
    f = sol.solve_grad(_t(q), gs, seed=4, per_pair=True)
    of = os_.solve_fused(q, spec_bc, W, nsteps=1024, eps=1e-3, min_R=sol.min_R, seed=4, per_pair=True,
                         cv=sol.control_variates, gv=gv, hv=hv)
    pv = _np(f["pair_vals"])
    assert _eq(pv[:, 0], of["pair_u"]) and _eq(pv[:, 1:], of["pair_g"])
    assert np.array_equal(_np(f["pair_steps"]), of["pair_steps"])
    return o


def test_zero_length_segment_and_vertex_queries_2d(gpu, oracle):
    V, I, T = _degenerate_polygon()
    gs = gpu.Scene2D(V, I, T, gpu.BoundaryCondition.exp_cos())
    os_ = oracle.OracleScene(V, I, T)
    q = _vertex_queries(V, I, 2)
    _query_parity(gs, os_, oracle, q, 2)
    # sampling never picks the zero-length segment (zero weight)
    u = np.random.default_rng(1).random(20000, dtype=np.float32)
    for sub in range(3):
        if os_.info(sub)["present"]:
            s = gs.sample_boundary(_t(u), sub=sub)
            o = os_.sample_boundary(sub, u)
            assert _eq(_np(s["prim"]), o["prim"]) and _eq(_np(s["p"]), o["p"]) and _eq(_np(s["pdf"]), o["pdf"])
    pr = _np(gs.sample_boundary(_t(u), sub=0)["prim"])
    assert 10 not in set(pr.tolist())
    # walks from the vertices (on the boundary: Dirichlet ones terminate at once)
    # and from just inside
    inside = _np(gs.sdf(_t(q))) < -1e-3
    _walk_parity(gpu, gs, os_, np.concatenate([q[: len(V)], q[inside][:64]]), 0, 2)


def test_vertex_queries_3d(gpu, oracle):
    from dwost_amd import scenes

    spec = scenes.icosphere_scene(level=2, n_query=8)
    gs = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = _vertex_queries(spec.vertices, spec.indices, 3)
    _query_parity(gs, os_, oracle, q, 3)
    inside = _np(gs.sdf(_t(q))) < -1e-3
    _walk_parity(gpu, gs, os_, np.concatenate([q[: len(spec.vertices)], q[inside][:64]]), spec.bc_code, 3)


@pytest.mark.parametrize("dim", [2, 3])
def test_nan_and_inf_boundary_data(gpu, oracle, dim):
    """Per-vertex g with a NaN and a +Inf at Dirichlet vertices, h with a NaN at
    a Neumann vertex: the walks that meet them carry NaN / Inf exactly as the
    oracle's (the 2D value guard zeroes them, 3D keeps them), the rest stay
    finite and bit-identical."""
    from dwost_amd import scenes

    spec = scenes.star_scene_2d(n_query=64) if dim == 2 else scenes.icosphere_scene(level=2, n_query=64)
    V = len(spec.vertices)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(V).astype(np.float32)
    h = (0.1 * rng.standard_normal(V)).astype(np.float32)
    dv = np.unique(spec.indices[spec.types == 1].ravel())
    nv = np.setdiff1d(np.unique(spec.indices[spec.types == 0].ravel()), dv)
    g[dv[::29]] = np.nan
    g[dv[13::29]] = np.inf
    h[nv[::41]] = np.nan
    cls = gpu.Scene2D if dim == 2 else gpu.Scene3D
    gs = cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(_t(g), _t(h)))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    o = _walk_parity(gpu, gs, os_, spec.queries[:48], 3, dim, gv=g, hv=h, W=32)
    wu = o["walk_u"]
    assert np.isfinite(wu).any()
    if dim == 3:  # no value guard in 3D: the bad data reaches some walks' values
        assert not np.isfinite(wu).all()
