"""The benchmark workloads (bench.py CONFIGS) have the frozen sizes and rules:
segment / triangle counts, the Dirichlet / Neumann split, query counts, and
queries inside the domain with their clearance.  The row-wise lattice and the
binned interior tests that build the query sets return exactly the dense test's
mask."""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def bench_mod():
    import bench

    return bench


def test_lattice_interior_matches_dense_test():
    from dwost_amd import scenes as S

    for spec in (S.holes_scene_2d(n_query=512), S.star_scene_2d(n_query=64)):
        V = spec.vertices.astype(np.float64)
        lo, hi = V.min(0), V.max(0)
        for k in (37, 96):
            xs = np.linspace(lo[0], hi[0], k + 2)[1:-1].astype(np.float32).astype(np.float64)
            ys = np.linspace(lo[1], hi[1], k + 2)[1:-1].astype(np.float32).astype(np.float64)
            X, Y = np.meshgrid(xs, ys, indexing="xy")
            cand = np.stack([X.ravel(), Y.ravel()], 1)
            for clearance in (2e-4, 2e-2):
                dense = S._interior_2d(cand, V, spec.indices, clearance)
                rows = S._interior_lattice_2d(xs, ys, V, spec.indices, clearance)
                assert np.array_equal(dense, rows), (spec.meta["name"], k, clearance)


def test_binned_interior_matches_dense_test():
    """Random points, points on and beside segments and at vertices: the binned
    test used for random query sets returns the dense test's mask."""
    from dwost_amd import scenes as S

    rng = np.random.default_rng(4)
    for spec in (S.star_scene_2d(n_query=64), S.holes_scene_2d(n_outer=1024, n_holes=3, n_query=256)):
        V = spec.vertices.astype(np.float64)
        I = spec.indices
        lo, hi = V.min(0), V.max(0)
        for clearance in (2e-4, 2e-3, 5e-2):
            f = rng.integers(0, len(I), 3000)
            a, b = V[I[f, 0]], V[I[f, 1]]
            t = rng.choice([0.0, 1.0, 0.5, 0.25], 3000)[:, None]
            pts = np.concatenate([rng.uniform(lo - 0.05, hi + 0.05, (8000, 2)),
                                  a + (b - a) * t + rng.normal(0, clearance, (3000, 2)), V])
            pts = pts.astype(np.float32).astype(np.float64)
            assert np.array_equal(S._interior_2d(pts, V, I, clearance), S._interior_binned_2d(pts, V, I, clearance))


def test_star_workload(bench_mod):
    from dwost_amd import scenes as S

    s = bench_mod.make_spec("star")
    assert s.indices.shape == (256, 2) and s.queries.shape == (1024, 2)
    mid = 0.5 * (s.vertices[s.indices[:, 0]] + s.vertices[s.indices[:, 1]])
    assert np.array_equal(s.types == S.DIRICHLET, mid[:, 0] < 0)
    assert S._interior_2d(s.queries.astype(np.float64), s.vertices.astype(np.float64), s.indices, 2e-3).all()


def test_holes_workload(bench_mod):
    from dwost_amd import scenes as S

    s = bench_mod.make_spec("holes")
    assert s.indices.shape == (4096 + 8 * 64, 2) and s.queries.shape == (65536, 2)
    assert (s.types[:4096] == S.DIRICHLET).all() and (s.types[4096:] == S.NEUMANN).all()
    assert s.bc_code == S.BC_QUAD and s.meta["eps"] == 1e-4
    # a sample of the lattice queries: inside, outside the holes, 2 eps from the boundary
    q = s.queries[np.random.default_rng(0).choice(len(s.queries), 2048, replace=False)].astype(np.float64)
    assert S._interior_2d(q, s.vertices.astype(np.float64), s.indices, 2e-4).all()
    # row-major lattice order: y does not decrease
    assert np.all(np.diff(s.queries[:, 1]) >= 0)


def test_icosphere_workload(bench_mod):
    from dwost_amd import scenes as S

    s = bench_mod.make_spec("ico")
    assert s.vertices.shape == (40962, 3) and s.indices.shape == (81920, 3) and s.queries.shape == (262144, 3)
    cen = s.vertices[s.indices].mean(1)
    assert np.array_equal(s.types == S.DIRICHLET, cen[:, 1] > 0)
