"""The source term of the Poisson problem -Lap u = f (the reference's scene.f:
WoSt.handle_source, wost.py:128-143, and the source-gradient term of
WoSt.dude, wost.py:59-79), pinned on the CPU for a linear f = a . x.
u* = x - (a . x) |x|^2 / (2 dim + 4) solves -Lap u* = a . x; on an all-Dirichlet
scene whose per-vertex g is u* at the vertices, the oracle's walk means for u
(Solver.solve), for the fused u and grad u, and for the serial dude lie within
their standard errors of u* and grad u* (the piecewise-linear g is exact to
~1e-4 here).  Dropping the source or flipping its sign moves the means far away."""

import numpy as np
import pytest

A2 = np.array([1.5, -2.0])
A3 = np.array([1.5, -2.0, 1.0])


def _truth(q, a):
    d = q.shape[1]
    q = q.astype(np.float64)
    ax = q @ a
    r2 = np.sum(q * q, 1)
    u = q[:, 0] - ax * r2 / (2 * d + 4)
    g = -(np.outer(ax, 2 * np.ones(d)) * q + np.outer(r2, a)) / (2 * d + 4)
    g[:, 0] += 1.0
    return u, g


def _setup(dim, oracle):
    from dwost_amd import scenes

    spec = scenes.star_scene_2d(n_query=64) if dim == 2 else scenes.icosphere_scene(level=3, n_query=64)
    types = np.ones_like(spec.types)  # all Dirichlet
    a = A2 if dim == 2 else A3
    gv = _truth(spec.vertices, a)[0].astype(np.float32)
    hv = np.zeros_like(gv)
    sc = oracle.OracleScene(spec.vertices, spec.indices, types)
    return spec, sc, a, gv, hv


def _z(est_walks, truth, nb=32):
    bm = est_walks.reshape(len(truth), nb, -1).astype(np.float64).mean(2)
    return (bm.mean(1) - truth) / (bm.std(1, ddof=1) / np.sqrt(nb))


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_source_u_and_grad(oracle, dim):
    spec, sc, a, gv, hv = _setup(dim, oracle)
    q = spec.queries[:16]
    u_t, g_t = _truth(q, a)
    src = [0.0] + [float(v) for v in a] + ([0.0] if dim == 2 else [])
    meds = {}
    for label, s, code in (("ref", src, 2), ("dropped", src, 0), ("flipped", [-v for v in src], 2)):
        o = oracle.OracleScene.solve_u(sc, q, 3, 2048, nsteps=1024, eps=1e-3, seed=1, gv=gv, hv=hv, src_code=code,
                                       src_c=s, per_walk=True)
        zu = _z(o["walk_u"], u_t)
        f = sc.solve_fused(q, 3, 2048, nsteps=1024, eps=1e-3, seed=2, gv=gv, hv=hv, src_code=code, src_c=s,
                           per_pair=True)
        pg = f["pair_g"].reshape(len(q), 32, -1, dim).astype(np.float64).sum(2) / (2 * 1024 / 32)
        zg = (pg.mean(1) - g_t) / (pg.std(1, ddof=1) / np.sqrt(32))
        meds[label] = (float(np.median(np.abs(zu))), float(np.median(np.abs(zg))))
        print(dim, label, meds[label])
    assert meds["ref"][0] < 1.5 and meds["ref"][1] < 1.5
    assert meds["dropped"][0] > 3 and meds["flipped"][0] > 3


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_source_dude(oracle, dim):
    """The serial dude (one stream per point, its walks in order) carries the
    source-gradient term with its running-mean control variate."""
    spec, sc, a, gv, hv = _setup(dim, oracle)
    q = spec.queries[:16]
    _, g_t = _truth(q, a)
    src = [0.0] + [float(v) for v in a] + ([0.0] if dim == 2 else [])
    e = np.random.default_rng(3).standard_normal((len(q), dim))
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    truth = np.sum(g_t * e, 1)
    res = []
    for seed in range(24):  # independent dudes of 256 walks
        o = sc.solve_dude(q, e, 3, 256, nsteps=1024, eps=1e-3, seed=seed, gv=gv, hv=hv, src_code=2, src_c=src)
        res.append(o["dude"])
    res = np.array(res, np.float64)
    z = (res.mean(0) - truth) / (res.std(0, ddof=1) / np.sqrt(len(res)))
    print(dim, "dude median |z|", np.median(np.abs(z)))
    assert np.median(np.abs(z)) < 1.5
