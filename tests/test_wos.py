"""Plain Walk on Spheres (wos.py, wos_3d.py): the oracle's restatement is
pinned by the FP64 analytic harmonic field on all-Dirichlet scenes (per-point
z-scores, unbiased estimator), and the GPU walks match the oracle walk by walk."""

import numpy as np
import pytest


def _dirichlet(spec):
    return np.ones_like(spec.types)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_wos_unbiased(oracle, star_spec, dim):
    from dwost_amd import scenes

    spec = star_spec if dim == 2 else scenes.icosphere_scene(level=3, n_query=64)
    sc = oracle.OracleScene(spec.vertices, spec.indices, _dirichlet(spec))
    q = spec.queries[:64]
    W = 256
    o = sc.solve_u(q, scenes.BC_EXPCOS, W, nsteps=1024, eps=1e-4, seed=5, per_walk=True, wos=True)
    wu = o["walk_u"].reshape(len(q), W).astype(np.float64)
    ua = scenes.analytic_u(scenes.BC_EXPCOS, q)
    z = (wu.mean(1) - ua) / (wu.std(1) / np.sqrt(W) + 1e-12)
    # the eps-shell bias is O(eps); with eps = 1e-4 it is far below the noise
    assert np.mean(np.abs(z) < 3) > 0.95
    # WoS never takes more than the ball-shrinking steps a walk needs, and it
    # consumes one direction per non-terminal step: steps are finite
    assert o["walk_steps"].max() < 1024




@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_wos_gpu_matches_oracle(oracle, star_spec, dim):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import dwost_amd
    from dwost_amd import scenes

    spec = star_spec if dim == 2 else scenes.icosphere_scene(level=4, n_query=64)
    T = _dirichlet(spec)
    cls = dwost_amd.Scene2D if dim == 2 else dwost_amd.Scene3D
    gs = cls(spec.vertices, spec.indices, T, dwost_amd.BoundaryCondition.exp_cos())
    os_ = oracle.OracleScene(spec.vertices, spec.indices, T)
    q = spec.queries[:48]
    sol = (dwost_amd.WoS if dim == 2 else dwost_amd.WoS3D)(eps=1e-3, nwalks=64, nsteps=1024, nsubwalks=2)
    qt = torch.as_tensor(q).cuda()
    g = sol.solve(qt, gs, seed=9, per_walk=True)
    o = os_.solve_u(q, scenes.BC_EXPCOS, 64, nsteps=1024, eps=1e-3, seed=9, nsubwalks=2, per_walk=True, wos=True)
    torch.cuda.synchronize()
    gu = g["walk_u"].cpu().numpy()
    assert np.array_equal(gu.view(np.uint32), o["walk_u"].view(np.uint32))
    assert np.array_equal(g["walk_steps"].cpu().numpy(), o["walk_steps"])
    assert np.array_equal(g["walk_term"].cpu().numpy(), o["walk_term"])
    e = np.random.default_rng(4).standard_normal((len(q), dim)).astype(np.float32)
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    sol1 = (dwost_amd.WoS if dim == 2 else dwost_amd.WoS3D)(eps=1e-3, nwalks=64, nsteps=1024)
    gd = sol1.solve_dude(qt, torch.as_tensor(e).cuda(), gs, seed=2).cpu().numpy()
    od = os_.solve_dude(q, e, scenes.BC_EXPCOS, 64, nsteps=1024, eps=1e-3, seed=2, wos=True)["dude"]
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    with pytest.raises(dwost_amd.DwostError):
        sol1.solve_grad(qt, gs, seed=2)
