"""The oracle's WoSt estimators against FP64 analytic harmonic fields (CPU).

e^x cos y (and x^2 - y^2 on the holes scene) is harmonic, so with f = 0 and the
scene's Dirichlet data g and Neumann data h taken from the field, the walk
estimators of u (wost.py:145-221, wost_3d.py:43-106), the fused antithetic
u + grad u pair estimator, and the serial dude (wost.py:21-102,
wost_3d.py:108-175) are unbiased up to the eps shell.  Per-point z-scores of the
walk means must lie within 3 sigma for all but a few points.  Planted mistakes
show that the checks have power: dropping the Neumann term (h = 0 through the
per-vertex data) and flipping the sign of the gradient both fail them.  The GPU
kernels agree with the oracle walk for walk (tests/test_gpu_parity.py), so these
checks pin both.
"""

import numpy as np
import pytest

EPS = 1e-3
NSTEPS = 1024


def _scene(oracle, name):
    from dwost_amd import scenes

    if name == "star":
        spec, code = scenes.star_scene_2d(n_query=256), scenes.BC_EXPCOS
    elif name == "holes":
        spec, code = scenes.holes_scene_2d(n_outer=512, n_holes=4, hole_seg=32, n_query=256), scenes.BC_QUAD
    else:
        spec, code = scenes.icosphere_scene(level=3, n_query=256), scenes.BC_EXPCOS
    return spec, code, oracle.OracleScene(spec.vertices, spec.indices, spec.types)


def _z(samples, exact):
    """z-scores of per-point sample means (samples: points x draws [x dim]); the
    4 eps term covers the eps-shell bias of points that start within the shell
    (every walk then ends at once on the same closest point)."""
    n = samples.shape[1]
    sd = np.sqrt(samples.var(1) / n + (4 * EPS) ** 2)
    return (samples.mean(1) - exact) / sd


@pytest.fixture(scope="module", params=["star", "holes", "ico"])
def case(request, oracle):
    return (request.param,) + _scene(oracle, request.param)


def test_u_unbiased(oracle, case):
    from dwost_amd import scenes

    name, spec, code, sc = case
    q, W = spec.queries[:96], 512
    o = sc.solve_u(q, code, W, nsteps=NSTEPS, eps=EPS, seed=3, per_walk=True)
    wu = o["walk_u"].reshape(len(q), W).astype(np.float64)
    z = _z(wu, scenes.analytic_u(code, q))
    print(f"{name}: |z|<3 for {np.mean(np.abs(z) < 3):.3f} of the points, max |z| {np.abs(z).max():.2f}")
    assert np.mean(np.abs(z) < 3) >= 0.93
    # the mean output is the mean of the walks
    assert np.allclose(o["u"], wu.mean(1), rtol=1e-5, atol=1e-5)


def test_fused_unbiased(oracle, case):
    from dwost_amd import scenes

    name, spec, code, sc = case
    q, W = spec.queries[:96], 512
    P = (W + 1) // 2
    f = sc.solve_fused(q, code, W, nsteps=NSTEPS, eps=EPS, seed=3, per_pair=True)
    # a pair's outputs are the sums over its two antithetic walks
    pu = f["pair_u"].reshape(len(q), P).astype(np.float64) / 2
    pg = f["pair_g"].reshape(len(q), P, spec.dim).astype(np.float64) / 2
    zu = _z(pu, scenes.analytic_u(code, q))
    ga = scenes.analytic_grad(code, q)
    # grad u: (dim / R) u dir has variance ~ 1/R^2; score the points off the shell
    far = sc.query(oracle.SUB_FULL, oracle.KIND_CLOSEST, q)["d"] > 0.05
    # a point passes when every component of its gradient lies within 3 sigma
    ok = np.all(np.abs(_z(pg[far], ga[far])) < 3, axis=1)
    flipped = np.all(np.abs(_z(-pg[far], ga[far])) < 3, axis=1)
    print(f"{name}: u |z|<3 {np.mean(np.abs(zu) < 3):.3f}; grad {np.mean(ok):.3f} over {far.sum()} points, "
          f"sign-flipped grad {np.mean(flipped):.3f}")
    assert far.sum() >= 48
    assert np.mean(np.abs(zu) < 3) >= 0.93
    assert np.mean(ok) >= 0.9
    # planted mistake: the gradient with its sign flipped fails the same check
    assert np.mean(flipped) < 0.5
    assert np.allclose(f["u"], pu.mean(1), rtol=1e-5, atol=1e-5)
    assert np.allclose(f["grad"], pg.mean(1), rtol=1e-5, atol=1e-5)


def test_dude_unbiased(oracle, case):
    """Serial dude along a random unit direction per point; the spread comes from
    8 independent seeds of 64 walks (t with 7 degrees of freedom: P(|t| < 3) ~ 0.98)."""
    from dwost_amd import scenes

    name, spec, code, sc = case
    q = spec.queries[:96]
    rng = np.random.default_rng(11)
    e = rng.standard_normal((len(q), spec.dim))
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    ds = np.stack([sc.solve_dude(q, e, code, 64, nsteps=NSTEPS, eps=EPS, seed=s)["dude"] for s in range(8)],
                  1).astype(np.float64)
    exact = np.sum(scenes.analytic_grad(code, q) * e, axis=1)
    t = (ds.mean(1) - exact) / np.sqrt(ds.var(1, ddof=1) / 8 + (4 * EPS) ** 2)
    print(f"{name}: dude |t|<3 for {np.mean(np.abs(t) < 3):.3f} of the points")
    assert np.mean(np.abs(t) < 3) >= 0.9


def test_dropped_neumann_term_fails(oracle):
    """Planted mistake: the same walks with h = 0 (g exact at the vertices) no
    longer estimate e^x cos y on the star, whose boundary is half Neumann."""
    from dwost_amd import scenes

    spec, code, sc = _scene(oracle, "star")
    q, W = spec.queries[:96], 512
    gv = scenes.analytic_u(code, spec.vertices.astype(np.float64)).astype(np.float32)
    o = sc.solve_u(q, scenes.BC_PER_VERTEX, W, nsteps=NSTEPS, eps=EPS, seed=3, per_walk=True, gv=gv,
                   hv=np.zeros_like(gv))
    z = _z(o["walk_u"].reshape(len(q), W).astype(np.float64), scenes.analytic_u(code, q))
    assert np.mean(np.abs(z) < 3) < 0.5


@pytest.mark.parametrize("name", ["star", "ico"])
def test_linear_in_vertex_data(oracle, name):
    """For fixed seeds the walks depend on the geometry only, so u and grad u
    are linear in the per-vertex g and h: u(g1 + 2 g2) = u(g1) + 2 u(g2)."""
    from dwost_amd import scenes

    spec, _, sc = _scene(oracle, name)
    q = spec.queries[:32]
    V = len(spec.vertices)
    rng = np.random.default_rng(5)
    g1, h1, g2, h2 = (rng.standard_normal(V).astype(np.float32) for _ in range(4))

    def run(g, h):
        u = sc.solve_u(q, scenes.BC_PER_VERTEX, 64, nsteps=NSTEPS, eps=EPS, seed=9, gv=g, hv=h)["u"]
        f = sc.solve_fused(q, scenes.BC_PER_VERTEX, 64, nsteps=NSTEPS, eps=EPS, seed=9, gv=g, hv=h)
        return u.astype(np.float64), f["u"].astype(np.float64), f["grad"].astype(np.float64)

    a, b, c = run(g1, h1), run(g2, h2), run(g1 + 2 * g2, h1 + 2 * h2)
    for x, y, z in zip(a, b, c):
        scale = np.abs(x).max() + 2 * np.abs(y).max() + 1.0
        assert np.abs(z - (x + 2 * y)).max() <= 1e-4 * scale
