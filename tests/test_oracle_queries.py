"""The restated traversals (bvh.py, snch.py, snch3d.py) against the reference's
own brute-force baselines (polyline.py:246-302,468-496; mesh.py:312-335,
440-472,558-570).  Tie rule: distances within 1e-6 relative are ties; any of
the tied primitive indices is accepted.  CPU only."""

import numpy as np
import pytest

from dwost_amd import scenes


def _seg_dist64(V, I, p, prim):
    a, b = V[I[prim, 0]].astype(np.float64), V[I[prim, 1]].astype(np.float64)
    ba = b - a
    h = np.clip(((p - a) * ba).sum(1) / (ba * ba).sum(1), 0, 1)
    return np.linalg.norm(a + ba * h[:, None] - p, axis=1)


def _sub(spec, sub):
    T = spec.types
    sel = np.arange(len(T)) if sub == 0 else np.nonzero(T == (1 if sub == 1 else 0))[0]
    used = np.unique(spec.indices[sel])
    remap = -np.ones(len(spec.vertices), int)
    remap[used] = np.arange(len(used))
    return spec.vertices[used], remap[spec.indices[sel]]


@pytest.mark.parametrize("sub", [0, 1, 2])
def test_closest_point_2d_vs_brute_force(oracle, star_spec, sub):
    sc = oracle.OracleScene(star_spec.vertices, star_spec.indices, star_spec.types)
    rng = np.random.default_rng(3)
    q = np.concatenate([star_spec.queries, rng.uniform(-2, 2, (512, 2)).astype(np.float32)])
    r = sc.query(sub, oracle.KIND_CLOSEST, q)
    b = sc.query(sub, oracle.KIND_BRUTE_CLOSEST, q)
    V, I = _sub(star_spec, sub)
    d_bvh = _seg_dist64(V, I, q.astype(np.float64), r["prim"])
    d_min = _seg_dist64(V, I, q.astype(np.float64), b["prim"])
    # the baseline evaluates |pa - ba*h| instead of |a + ba*h - p|: the two differ
    # by a few ulps of the coordinates (|x| <= 2 here), i.e. an absolute 4e-7
    assert np.all(np.abs(r["d"] - b["d"]) <= 1e-6 * b["d"] + 4e-7)
    assert np.all(d_bvh <= d_min * (1 + 1e-6) + 4e-7)  # tie rule
    # the closest point lies on the returned segment at parameter t
    a, bb = V[I[r["prim"], 0]], V[I[r["prim"], 1]]
    pt = a + (bb - a) * r["tuv"][:, :1]
    assert np.allclose(pt, r["p"], atol=1e-6)


def test_closest_point_3d_vs_brute_force(oracle, ico_spec):
    sc = oracle.OracleScene(ico_spec.vertices, ico_spec.indices, ico_spec.types)
    rng = np.random.default_rng(4)
    q = np.concatenate([ico_spec.queries[:512], rng.uniform(-1.5, 1.5, (256, 3)).astype(np.float32)])
    for sub in (0, 1):
        r = sc.query(sub, oracle.KIND_CLOSEST, q)
        b = sc.query(sub, oracle.KIND_BRUTE_CLOSEST, q)
        assert np.all(np.abs(r["d"] - b["d"]) <= 1e-6 * b["d"] + 1e-12)  # same formula in 3D
        same = r["prim"] == b["prim"]
        # different triangles only at exact ties (shared vertex / edge nearest)
        assert np.array_equal(r["d"][~same], b["d"][~same])


def test_ray_2d_vs_brute_force(oracle, star_spec):
    """Inside a WoSt ball (r_max <= star radius) a ray meets the boundary at most
    once, so the first hit in traversal order (bvh.py:215-217) is the nearest."""
    sc = oracle.OracleScene(star_spec.vertices, star_spec.indices, star_spec.types)
    q = star_spec.queries
    rng = np.random.default_rng(5)
    th = rng.uniform(0, 2 * np.pi, len(q))
    v = np.stack([np.cos(th), np.sin(th)], 1).astype(np.float32)
    R = sc.query(0, oracle.KIND_STAR, q, rmax=np.full(len(q), np.inf, np.float32))["d"]
    r = sc.query(0, oracle.KIND_RAY, q, v, R)
    b = sc.query(0, oracle.KIND_BRUTE_RAY, q, v, R)
    hit = b["prim"] >= 0
    assert np.array_equal(r["prim"] >= 0, hit)
    assert np.array_equal(r["prim"][hit], b["prim"][hit])
    assert np.allclose(r["d"][hit], b["d"][hit], rtol=1e-5)
    # with a long ray the first-hit quirk returns some valid hit, not necessarily the nearest
    r2 = sc.query(0, oracle.KIND_RAY, q, v, np.full(len(q), 10.0, np.float32))
    b2 = sc.query(0, oracle.KIND_BRUTE_RAY, q, v, np.full(len(q), 10.0, np.float32))
    assert np.all(r2["d"] >= b2["d"] * (1 - 1e-5))


def test_ray_3d_vs_brute_force(oracle, ico_spec):
    sc = oracle.OracleScene(ico_spec.vertices, ico_spec.indices, ico_spec.types)
    q = ico_spec.queries[:512]
    rng = np.random.default_rng(6)
    v = rng.standard_normal((len(q), 3))
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    r = sc.query(0, oracle.KIND_RAY, q, v, np.full(len(q), 10.0, np.float32))
    b = sc.query(0, oracle.KIND_BRUTE_RAY, q, v, np.full(len(q), 10.0, np.float32))
    assert np.array_equal(r["prim"] >= 0, b["prim"] >= 0)
    assert np.allclose(r["d"], b["d"], rtol=1e-6)


@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_closest_silhouette_vs_brute_force(oracle, which, star_spec, holes_spec, ico_spec):
    spec = dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which]
    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = spec.queries[:1024]
    for sub in (0, 2):
        if not sc.info(sub)["present"]:
            continue
        for rmax in (np.inf, 0.1):
            rm = np.full(len(q), rmax, np.float32)
            r = sc.query(sub, oracle.KIND_SIL, q, rmax=rm)
            b = sc.query(sub, oracle.KIND_BRUTE_SIL, q, rmax=rm)
            assert np.array_equal(r["prim"] >= 0, b["prim"] >= 0)
            ok = r["prim"] >= 0
            assert np.array_equal(r["d"][ok], b["d"][ok])
            topo = np.array([sc.ref_topology(sub, int(x)) for x in r["prim"][ok]])
            same = topo == b["prim"][ok]
            if spec.dim == 2:  # 3D nearest points are often shared vertices of several edges (exact ties)
                assert same.mean() > 0.99


def test_star_radius_bounds(oracle, star_spec):
    """star radius <= bbox bound and <= r_max; equals the closest silhouette distance when one exists."""
    sc = oracle.OracleScene(star_spec.vertices, star_spec.indices, star_spec.types)
    q = star_spec.queries
    rm = np.full(len(q), 0.3, np.float32)
    R = sc.query(2, oracle.KIND_STAR, q, rmax=rm)["d"]
    s = sc.query(2, oracle.KIND_SIL, q, rmax=rm)
    assert np.all(R <= 0.3)
    has = s["prim"] >= 0
    assert np.array_equal(R[has], s["d"][has])
    assert np.all(R[~has] == np.float32(0.3))


@pytest.mark.parametrize("which", ["star", "holes"])
def test_cv_closest_point_bounded_by_dirichlet_distance(oracle, which, star_spec, holes_spec):
    """The 2D walk's control-variate closest point (full scene, wost.py:123-126)
    is no farther than the Dirichlet distance found at the same point.  Started
    with its pruning radius just above that distance (relative 2^-16 plus 1e-6
    of the coordinate scale) the traversal only skips subtrees farther than the
    answer, so the record, including which of equally near segments wins, is
    the unbounded query's bit for bit: checked on interior points, points on and
    beside the boundary, and points around the vertices where segments meet."""
    spec = dict(star=star_spec, holes=holes_spec)[which]
    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    rng = np.random.default_rng(17)
    V, I = spec.vertices.astype(np.float64), spec.indices
    f = rng.integers(0, len(I), 4096)
    t = rng.choice([0.0, 1.0, 0.5, 1e-6, 1 - 1e-6], 4096)[:, None] * np.ones((1, 1))
    t = np.where(rng.random((4096, 1)) < 0.5, t, rng.random((4096, 1)))
    a, b = V[I[f, 0]], V[I[f, 1]]
    n = np.stack([(b - a)[:, 1], -(b - a)[:, 0]], 1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    off = rng.choice([0.0, 1e-6, -1e-6, 1e-4, -1e-4, 1e-2, -1e-2], 4096)[:, None]
    near = a + (b - a) * t + off * n
    corner = V[rng.integers(0, len(V), 2048)] + rng.normal(0, 1e-3, (2048, 2))
    pts = np.concatenate([spec.queries[:4096].astype(np.float64), near, corner]).astype(np.float32)
    same, n_full, n_bound = sc.cv_bound_check(pts)
    print(f"{which}: {same.mean():.6f} identical; nodes visited {n_full} -> {n_bound} "
          f"({n_bound / n_full:.3f})")
    assert same.all()
    assert n_bound < n_full


@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_dirichlet_query_bounded_by_previous_step(oracle, which, star_spec, holes_spec, ico_spec):
    """Each walk step's Dirichlet closest point starts its pruning radius at the
    bound the previous step gives: the distance to a set is 1-Lipschitz and a
    step moves the walk by at most R + 1e-5 (R <= the previous Dirichlet
    distance), so d(x_new) <= d(x_prev) + R + 1e-5, widened by relative 2^-16
    plus 1e-6 of the coordinate scale.  The bounded query returns the unbounded
    query's record bit for bit (equally near segments or triangles included) on
    moves of every length up to the bound, from interior points, points beside
    the boundary and points around vertices."""
    spec = dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which]
    D = spec.dim
    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    rng = np.random.default_rng(23)
    V, I = spec.vertices.astype(np.float64), spec.indices
    f = rng.integers(0, len(I), 3000)
    w = rng.dirichlet(np.ones(D), 3000)
    on = np.einsum("nk,nkd->nd", w, V[I[f]])
    corner = V[rng.integers(0, len(V), 1500)] + rng.normal(0, 1e-3, (1500, D))
    prev = np.concatenate([spec.queries[:3000].astype(np.float64), on + rng.normal(0, 1e-3, on.shape), corner])
    d_prev = sc.query(1, oracle.KIND_CLOSEST, prev.astype(np.float32))["d"].astype(np.float64)
    keep = np.isfinite(d_prev)
    prev, d_prev = prev[keep], d_prev[keep]
    R = d_prev * rng.choice([1.0, 0.98, 0.5, 1e-3], len(prev))
    dirs = rng.normal(size=prev.shape)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    s = (R + 1e-5) * rng.choice([1.0, 0.999, 0.5, 0.0], len(prev))
    nxt = prev + dirs * s[:, None]
    same, n_full, n_bound = sc.step_bound_check(prev.astype(np.float32), nxt.astype(np.float32), R.astype(np.float32))
    print(f"{which}: {same.mean():.6f} identical; nodes visited {n_full} -> {n_bound} ({n_bound / n_full:.3f})")
    assert same.all()
    assert n_bound < n_full
    # planted mistake: a bound that forgets the move (R = 0) loses the answer
    bad, _, _ = sc.step_bound_check(prev.astype(np.float32), nxt.astype(np.float32), np.zeros(len(prev), np.float32))
    assert not bad.all()
