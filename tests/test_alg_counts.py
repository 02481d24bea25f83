"""Algorithmic bytes and flops per walk step (the roofline inputs of bench.py):
the oracle's counters of the fused estimator, in the reference traversal order
with the reference record sizes.  The frozen figures in bench.py come from
`python bench.py --alg-counts --config <name>` on 1024 points; a smaller sample
must land near them, and the byte count must lie between what the counted
nodes, primitives and silhouettes cost at the smallest and the largest record
sizes."""

import numpy as np
import pytest


def _sample(oracle, spec, n, W):
    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    idx = np.sort(np.random.default_rng(1).choice(len(spec.queries), n, replace=False))
    return sc.solve_fused(spec.queries[idx], spec.bc_code, W, nsteps=1024, eps=spec.meta["eps"], seed=3, count=True)


@pytest.mark.parametrize("name", ["star", "holes"])
def test_counts_near_frozen_figures(oracle, name):
    import bench
    from dwost_amd import scenes

    spec = scenes.star_scene_2d() if name == "star" else scenes.holes_scene_2d(n_query=4096, eps=1e-4)
    o = _sample(oracle, spec, 96, 64)
    b, f = o["bytes"] / o["steps"], o["flops"] / o["steps"]
    ref = bench.ALG[name]
    print(f"{name}: {b:.1f} B/step (frozen {ref['b_alg']}), {f:.1f} flops/step (frozen {ref['f_alg']})")
    assert abs(b / ref["b_alg"] - 1) < 0.1
    assert abs(f / ref["f_alg"] - 1) < 0.1
    # record sizes: nodes 16-48 B (2D ray boxes 16, BVH 24, SNCH 48), segments 20 B,
    # silhouette vertices 40 B; CDF gathers 8 B plus 4 B per halving
    assert 16 * o["nodes"] + 20 * o["prims"] + 40 * o["sils"] <= o["bytes"]
    assert o["bytes"] <= 48 * o["nodes"] + 20 * o["prims"] + 40 * o["sils"] + 60 * o["steps"]
    assert o["ff"] == 0


def test_counts_3d(oracle):
    import bench
    from dwost_amd import scenes

    spec = scenes.icosphere_scene(level=6, n_query=2048)
    o = _sample(oracle, spec, 48, 32)
    b = o["bytes"] / o["steps"]
    print(f"ico: {b:.1f} B/step (frozen {bench.ALG['ico']['b_alg']})")
    assert abs(b / bench.ALG["ico"]["b_alg"] - 1) < 0.15
    # 3D: SNCH nodes 60 B (ray boxes 24 B), triangles 40 B, silhouette edges 68 B
    assert 24 * o["nodes"] + 40 * o["prims"] + 68 * o["sils"] <= o["bytes"]
    assert o["bytes"] <= 60 * o["nodes"] + 40 * o["prims"] + 68 * o["sils"] + 80 * o["steps"]


def test_counting_leaves_results_unchanged(oracle, star_spec):
    sc = oracle.OracleScene(star_spec.vertices, star_spec.indices, star_spec.types)
    q = star_spec.queries[:16]
    a = sc.solve_fused(q, star_spec.bc_code, 16, nsteps=1024, eps=1e-3, seed=2, per_pair=True)
    b = sc.solve_fused(q, star_spec.bc_code, 16, nsteps=1024, eps=1e-3, seed=2, per_pair=True, count=True)
    assert np.array_equal(a["pair_u"].view(np.uint32), b["pair_u"].view(np.uint32))
    assert a["steps"] == b["steps"]
