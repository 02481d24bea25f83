"""The Scene2D / Scene3D query surface beyond the walk's own queries: sdf and
inside, boundary sampling from given draws, the brute-force baselines and the
sub-scene -> scene primitive map, each against the oracle (GPU) or the
reference's definition (CPU)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _gpu():
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


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _spec(which, star_spec, holes_spec, ico_spec):
    return dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which]


def _points(spec, rng, m=1024):
    lo, hi = spec.vertices.min(0), spec.vertices.max(0)
    return np.concatenate([spec.queries[:m], rng.uniform(lo - 0.3, hi + 0.3, (m, spec.dim))]).astype(np.float32)


def test_prim_map_is_the_compressed_type_mask(star_spec):
    """Scene2D._dirichlet_2_global_indices / _neumann_2_global_indices
    (polyline.py:636-645): dr.compress of the type masks, i.e. the global ids
    of each type in ascending order."""
    from dwost_amd.shapes import Scene2D, Scene3D
    from dwost_amd import scenes

    for spec, cls in ((star_spec, Scene2D), (scenes.icosphere_scene(level=2, n_query=4), Scene3D)):
        s = cls(spec.vertices, spec.indices, spec.types, host_only=True)
        assert np.array_equal(s.dirichlet_2_global_indices, np.flatnonzero(spec.types == 1))
        assert np.array_equal(s.neumann_2_global_indices, np.flatnonzero(spec.types == 0))
        assert np.array_equal(s.global_prim_ids(0), np.arange(len(spec.types)))
        # the sub-scene's primitive k is the scene primitive map[k]: same vertices
        ex = s.export(1)
        assert ex["nprims"] == int((spec.types == 1).sum())


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_sdf_and_inside(oracle, which, star_spec, holes_spec, ico_spec):
    gpu = _gpu()
    spec = _spec(which, star_spec, holes_spec, ico_spec)
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    gs = cls(spec.vertices, spec.indices, spec.types)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = _points(spec, np.random.default_rng(4))
    g = _np(gs.sdf(_t(q)))
    assert np.array_equal(_bits(g), _bits(os_.sdf(q)))
    assert np.array_equal(_np(gs.inside(_t(q))), g < 0)
    assert np.all(g[: min(1024, len(spec.queries))] < 0)  # the query sets are interior points


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_sample_boundary(oracle, which, star_spec, holes_spec, ico_spec):
    gpu = _gpu()
    spec = _spec(which, star_spec, holes_spec, ico_spec)
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    gs = cls(spec.vertices, spec.indices, spec.types)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    rng = np.random.default_rng(8)
    n = 4096
    for sub in range(3):
        if not os_.info(sub)["present"]:
            continue
        u = rng.random((n,) if spec.dim == 2 else (n, 3), dtype=np.float32)
        u.flat[:4] = [0.0, np.nextafter(np.float32(1), np.float32(0)), 0.5, 0.25]
        g = gs.sample_boundary(_t(u), sub=sub)
        o = os_.sample_boundary(sub, u)
        for k in ("p", "n", "tuv", "pdf", "prim"):
            assert np.array_equal(_bits(_np(g[k])), _bits(o[k])), (sub, k)
        # draws outside [0, 1) are refused per sample
        bad = u[:3].copy()
        bad[0, ...] = 1.0  # sample 0: draws of 1, sample 1: a negative draw
        bad[1, ...] = -0.5
        b = gs.sample_boundary(_t(bad), sub=sub)
        pr = _np(b["prim"])
        assert pr[0] == -1 and pr[1] == -1 and np.isnan(_np(b["pdf"])[:2]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_baselines_match_oracle_brute_force(oracle, which, star_spec, holes_spec, ico_spec):
    """closest_point_baseline / intersect_baseline / closest_silhouette_baseline:
    distances and primitives bit for bit against the oracle's restated loops."""
    gpu = _gpu()
    spec = _spec(which, star_spec, holes_spec, ico_spec)
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    gs = cls(spec.vertices, spec.indices, spec.types)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    rng = np.random.default_rng(9)
    q = _points(spec, rng, 512)
    v = rng.standard_normal(q.shape)
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    rm = np.full(len(q), 0.4, np.float32)
    for sub in range(3):
        if not os_.info(sub)["present"]:
            continue
        cases = ((gs.closest_point_baseline(_t(q), sub=sub), os_.query(sub, oracle.KIND_BRUTE_CLOSEST, q)),
                 (gs.intersect_baseline(_t(q), _t(v), 10.0, sub=sub),
                  os_.query(sub, oracle.KIND_BRUTE_RAY, q, v, np.full(len(q), 10.0, np.float32))),
                 (gs.closest_silhouette_baseline(_t(q), _t(rm), sub=sub),
                  os_.query(sub, oracle.KIND_BRUTE_SIL, q, rmax=rm)))
        for kind, (g, o) in enumerate(cases):
            gp, gd = _np(g["prim"]), _np(g["d"])
            hit = o["prim"] >= 0
            assert np.array_equal(gp >= 0, hit), (sub, kind)
            assert np.array_equal(_bits(gd[hit]), _bits(o["d"][hit])), (sub, kind)
            # both loop in the reference's index order and keep the first of
            # exact ties (silhouettes: the lowest topology id)
            assert np.array_equal(gp, o["prim"]), (sub, kind, np.flatnonzero(gp != o["prim"])[:10])
            assert hit.sum() > 0
        if spec.dim == 3:  # the tree query and the loop minimise the same per-triangle distance
            bvh = gs.closest_point(_t(q), sub=sub)
            assert np.array_equal(_bits(_np(bvh["d"])), _bits(_np(cases[0][0]["d"])))
