"""Bracketed cone tests.  The SNCH cone tests (SNCHNode.overlaps, snch.py:46-97;
SNCHNode.overlaps_2, snch.py:100-174; SNCHNode3D.visit, snch3d.py:196-271)
decide from fp32 acosf / asinf / atanf / cosf bracketed by a step bound to the
fp64-evaluated, once-rounded values, and run the fp64 form only when a bracket
straddles a threshold.  Checked here: the bound over every float of each
domain, and the rho-bounded star radius (Polyline.star_radius_2,
polyline.py:317-330; Mesh.star_radius_2, mesh.py:349-361) bit for bit against
the oracle on many points and directions, near-boundary points included."""

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


def test_math_bound_holds(gpu):
    r = gpu.ops.math_check()
    print(r)
    assert r["bound"] >= 1
    assert max(r["acos"], r["asin"], r["atan"], r["cos"]) <= r["bound"], r


def _points(spec, n_near, seed):
    """The scene's query points, points within 1e-3 of boundary vertices and
    points spread over the bounding box (outside the domain too)."""
    rng = np.random.default_rng(seed)
    D = spec.dim
    V = spec.vertices.astype(np.float32)
    near = V[rng.integers(0, len(V), n_near)] + rng.uniform(-1e-3, 1e-3, (n_near, D)).astype(np.float32)
    lo, hi = V.min(0), V.max(0)
    box = rng.uniform(lo - 0.1, hi + 0.1, (n_near, D)).astype(np.float32)
    p = np.concatenate([spec.queries.astype(np.float32), near, box])
    e = rng.standard_normal(p.shape)
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    return np.ascontiguousarray(p), np.ascontiguousarray(e)


@pytest.mark.parametrize("which,types,sub,rho", [
    ("star", "mixed", 2, 1.0),
    ("star", "allN", 0, 1.2),
    ("holes", "mixed", 2, 1.2),
    ("holes", "mixed", 0, 2.0),
    ("ico", "allN", 0, 1.01),
    ("ico", "allN", 0, 1.2),
])
def test_star_radius_2_matches_oracle(gpu, oracle, which, types, sub, rho):
    from dwost_amd import scenes

    if which == "star":
        spec = scenes.star_scene_2d(n_query=4096)
    elif which == "holes":
        spec = scenes.holes_scene_2d(n_outer=1024, n_holes=4, n_query=4096, eps=1e-3)
    else:
        spec = scenes.icosphere_scene(level=4, n_query=2048)
    T = np.zeros_like(spec.types) if types == "allN" else spec.types
    Scene = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    sc = Scene(spec.vertices, spec.indices, T)
    osc = oracle.OracleScene(spec.vertices, spec.indices, T)
    p, e = _points(spec, 2048, seed=len(which) * 7 + sub)
    got = sc.star_radius_2(torch.as_tensor(p).cuda(), torch.as_tensor(e).cuda(), rho_max=rho, sub=sub).cpu().numpy()
    want, _ = osc.star_radius2_rho(sub, p, e, rho_max=rho)
    same = got.view(np.uint32) == want.view(np.uint32)
    print(f"{which} sub {sub} rho {rho}: {same.sum()} / {len(same)} identical")
    assert same.all(), np.flatnonzero(~same)[:20].tolist()
    # the torch operator is the same entry point
    op = torch.ops.dwost.star_radius_2(int(sc.handle.value), sub, torch.as_tensor(p).cuda(), torch.as_tensor(e).cuda(), rho)
    assert np.array_equal(op.cpu().numpy().view(np.uint32), got.view(np.uint32))


def test_star_radius_2_arguments(gpu, star_spec):
    sc = gpu.Scene2D(star_spec.vertices, star_spec.indices, star_spec.types)
    q = torch.as_tensor(star_spec.queries[:8]).cuda()
    e = torch.zeros_like(q)
    e[:, 0] = 1.0
    with pytest.raises(gpu.DwostError):
        sc.star_radius_2(q, e, rho_max=0.0)
    with pytest.raises(ValueError):
        sc.star_radius_2(q, e[:4])
    assert sc.star_radius_2(q[:0], e[:0]).shape == (0,)
    # the reference default (1.0 in 2D) is used when rho_max is not given
    assert torch.equal(sc.star_radius_2(q, e), sc.star_radius_2(q, e, rho_max=1.0))


@pytest.mark.parametrize("dim", [2, 3])
def test_star_radius_2_on_vertices_and_edges(gpu, oracle, dim, star_spec):
    """Points exactly on boundary vertices and segment / edge midpoints, with
    directions along the adjacent segments and edges and their normals (where
    cone and plane tests sit on their thresholds): bit for bit against the
    oracle."""
    from dwost_amd import scenes

    spec = star_spec if dim == 2 else scenes.icosphere_scene(level=3, n_query=16)
    V = spec.vertices.astype(np.float32)
    I = spec.indices
    a, b = V[I[:, 0]], V[I[:, 1]]
    t = b - a
    if dim == 2:
        nrm = np.stack([t[:, 1], -t[:, 0]], 1)
        mids = (a + b) * np.float32(0.5)
    else:
        c = V[I[:, 2]]
        nrm = np.cross(t, c - a)
        mids = (a + b + c) / np.float32(3.0)
    dirs = np.concatenate([t, -t, nrm, -nrm]).astype(np.float64)
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    base = np.concatenate([a, mids, b, a]).astype(np.float32)
    T = np.zeros_like(spec.types)
    Scene = gpu.Scene2D if dim == 2 else gpu.Scene3D
    sc = Scene(spec.vertices, spec.indices, T)
    osc = oracle.OracleScene(spec.vertices, spec.indices, T)
    p, e = np.ascontiguousarray(base), np.ascontiguousarray(dirs)
    for rho in (1.0, 1.01, 1.2, 3.0):
        got = sc.star_radius_2(torch.as_tensor(p).cuda(), torch.as_tensor(e).cuda(), rho_max=rho).cpu().numpy()
        want, _ = osc.star_radius2_rho(0, p, e, rho_max=rho)
        same = got.view(np.uint32) == want.view(np.uint32)
        print(f"dim {dim} rho {rho}: {same.sum()} / {len(same)} identical")
        assert same.all(), np.flatnonzero(~same)[:20].tolist()
