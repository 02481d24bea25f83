"""A linear source f = s0 + s1 x + s2 y + s3 z (the reference's scene.f and
scene.dfdx) on the GPU against the oracle, bit for bit: the WoSt walks of
Solver.solve (f at each sampled source point), the fused pairs (f at the
sample and at its mirror), the serial dudes (the mirrored sample, the
running-mean control variate), and WoStGrad (f at the Neumann sample, grad f in
its source term, f in the normal-derivative estimate)."""

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


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_source_walks_match_oracle(gpu, oracle, dim):
    from dwost_amd import scenes

    spec = scenes.star_scene_2d(n_query=64) if dim == 2 else scenes.icosphere_scene(level=3, n_query=64)
    src = (0.3, 1.5, -2.0, 0.0) if dim == 2 else (0.3, 1.5, -2.0, 1.0)
    bc = gpu.BoundaryCondition(spec.bc_code, source=src)
    cls = gpu.Scene2D if dim == 2 else gpu.Scene3D
    sc = cls(spec.vertices, spec.indices, spec.types, bc)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = np.ascontiguousarray(spec.queries[:24])
    qt = torch.as_tensor(q).cuda()
    sol = (gpu.WoSt if dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=32, nsteps=1024, exact_cones=True)
    kw = dict(nsteps=1024, eps=1e-3, src_code=2, src_c=list(src))
    g = sol.solve(qt, sc, seed=4, per_walk=True)
    o = os_.solve_u(q, spec.bc_code, 32, seed=4, per_walk=True, min_R=sol.min_R, cv=sol.control_variates,
                    av=sol.antithetic_variates, **kw)
    assert np.array_equal(_bits(g["walk_u"].cpu().numpy()), _bits(o["walk_u"]))
    f = sol.solve_grad(qt, sc, seed=5, per_pair=True)
    of = os_.solve_fused(q, spec.bc_code, 32, seed=5, per_pair=True, min_R=sol.min_R, cv=sol.control_variates, **kw)
    pv = f["pair_vals"].cpu().numpy()
    assert np.array_equal(_bits(pv[:, 0]), _bits(of["pair_u"])) and np.array_equal(_bits(pv[:, 1:]), _bits(of["pair_g"]))
    e = np.random.default_rng(6).standard_normal((len(q), dim))
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    d = sol.solve_dude(qt, torch.as_tensor(e).cuda(), sc, seed=6)
    od = os_.solve_dude(q, e, spec.bc_code, 32, seed=6, min_R=sol.min_R, cv=sol.control_variates,
                        av=sol.antithetic_variates, **kw)["dude"]
    assert np.array_equal(_bits(d.cpu().numpy()), _bits(od))


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_source_wostgrad_matches_oracle(gpu, oracle, dim):
    from dwost_amd import scenes

    if dim == 2:
        spec = scenes.star_scene_2d(n_query=64)
        types, src = spec.types, (0.3, 1.5, -2.0, 0.0)
    else:
        spec = scenes.icosphere_scene(level=2, n_query=64)
        types, src = np.zeros_like(spec.types), (0.3, 1.5, -2.0, 1.0)
    cls = gpu.Scene2D if dim == 2 else gpu.Scene3D
    sc = cls(spec.vertices, spec.indices, types, gpu.BoundaryCondition(0, source=src))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, types)
    q = np.ascontiguousarray(spec.queries[:16])
    e = np.random.default_rng(7).standard_normal((len(q), dim))
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    sol = (gpu.WoStGrad if dim == 2 else gpu.WoStGrad3D)(eps=1e-3, nwalks=16, nsteps=1024)
    g = sol.solve_dude(torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda(), sc, seed=8, per_walk=True)
    o = os_.grad_dude(q, e, 0, 16, nsteps=1024, eps=1e-3, min_R=sol.min_R, cv=sol.control_variates, seed=8,
                      rho_max=sol.rho_max, src_code=2, src_c=list(src), per_walk=True)
    assert np.array_equal(_bits(g["walk_dude"].cpu().numpy()), _bits(o["walk_dude"]))
