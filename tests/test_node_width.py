"""Node width of the closest-point tree records.  Width 4 reads one record per
two binary levels (a node's children and the pair of the child descended
into) and visits the same nodes in the same order with the same arithmetic as
width 2, so every estimator returns the same bits: against the oracle walk by
walk (fused pairs, Solver.solve walks, dudes, WoS) on the 2D star and holes
scenes and the 3D icosphere, against width 2 under both lane schedules and in
the adjoint, and in a batch."""

from dataclasses import replace

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


def _bits(t):
    return t.detach().cpu().numpy().view(np.uint32)


def _scene(gpu, spec):
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    return cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))


@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_width4_fused_and_walks_match_oracle(gpu, oracle, which, star_spec, holes_spec, ico_spec):
    spec = dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which]
    sc = _scene(gpu, spec)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    eps = spec.meta["eps"]
    n, W = 48, 64
    q = np.ascontiguousarray(spec.queries[:n])
    Sol = gpu.WoSt if spec.dim == 2 else gpu.WoSt3D
    sol = Sol(eps=eps, nwalks=W, nsteps=1024, exact_cones=True, node_width=4)
    g = sol.solve_grad(torch.as_tensor(q).cuda(), sc, seed=5, per_pair=True, count_steps=True)
    o = os_.solve_fused(q, spec.bc_code, W, nsteps=1024, eps=eps, seed=5, per_pair=True)
    pv = g["pair_vals"].cpu().numpy()
    assert np.array_equal(pv[:, 0].view(np.uint32), o["pair_u"].view(np.uint32))
    assert np.array_equal(pv[:, 1:].view(np.uint32), o["pair_g"].view(np.uint32))
    assert np.array_equal(g["pair_steps"].cpu().numpy(), o["pair_steps"])
    u = sol.solve(torch.as_tensor(q).cuda(), sc, seed=7, per_walk=True, count_steps=True)
    ou = os_.solve_u(q, spec.bc_code, W, nsteps=1024, eps=eps, min_R=sol.min_R, cv=sol.control_variates,
                     av=sol.antithetic_variates, seed=7, per_walk=True)
    assert np.array_equal(_bits(u["walk_u"]), ou["walk_u"].view(np.uint32))
    assert np.array_equal(u["walk_steps"].cpu().numpy(), ou["walk_steps"])
    assert np.array_equal(u["walk_term"].cpu().numpy(), ou["walk_term"])


@pytest.mark.parametrize("which", ["star", "ico"])
def test_width4_dude_and_wos_match_oracle(gpu, oracle, which, star_spec, ico_spec):
    spec = dict(star=star_spec, ico=ico_spec)[which]
    sc = _scene(gpu, spec)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    n, W = 32, 32
    q = np.ascontiguousarray(spec.queries[:n])
    rng = np.random.default_rng(13)
    e = rng.normal(size=(n, spec.dim)).astype(np.float32)
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    Sol = gpu.WoSt if spec.dim == 2 else gpu.WoSt3D
    sol = Sol(eps=1e-3, nwalks=W, nsteps=1024, exact_cones=True, node_width=4)
    g = sol.solve_dude(torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda(), sc, seed=3)
    o = os_.solve_dude(q, e, spec.bc_code, W, nsteps=1024, eps=1e-3, cv=sol.control_variates,
                       av=sol.antithetic_variates, seed=3)["dude"]
    assert np.array_equal(_bits(g), o.view(np.uint32))
    # WoS walks need an all-Dirichlet boundary
    T = np.ones_like(spec.types)
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    sd = cls(spec.vertices, spec.indices, T, gpu.BoundaryCondition(spec.bc_code))
    od_ = oracle.OracleScene(spec.vertices, spec.indices, T)
    Wos = gpu.WoS if spec.dim == 2 else gpu.WoS3D
    w = Wos(eps=1e-3, nwalks=W, nsteps=1024, node_width=4).solve(torch.as_tensor(q).cuda(), sd, seed=4,
                                                                 per_walk=True)
    ow = od_.solve_u(q, spec.bc_code, W, nsteps=1024, eps=1e-3, seed=4, per_walk=True, wos=True)
    assert np.array_equal(_bits(w["walk_u"]), ow["walk_u"].view(np.uint32))
    assert np.array_equal(w["walk_steps"].cpu().numpy(), ow["walk_steps"])


@pytest.mark.parametrize("dim", [2, 3])
def test_width4_equals_width2_both_schedules(gpu, dim):
    from dwost_amd import scenes

    if dim == 2:
        spec = scenes.holes_scene_2d(n_query=400, eps=1e-3)
        sol = gpu.WoSt(eps=1e-3, nwalks=64, nsteps=1024)
    else:
        spec = scenes.icosphere_scene(level=4, n_query=300, eps=1e-3)
        sol = gpu.WoSt3D(eps=1e-3, nwalks=32, nsteps=1024)
    sc = _scene(gpu, spec)
    q = torch.as_tensor(spec.queries).cuda()
    ref = sol.solve_grad(q, sc, seed=7, per_pair=True, count_steps=True)
    for sched in (1, 2):
        g = replace(sol, schedule=sched, node_width=4).solve_grad(q, sc, seed=7, per_pair=True, count_steps=True)
        for k in ("u", "grad", "pair_vals", "pair_steps"):
            assert np.array_equal(_bits(g[k]), _bits(ref[k])), (sched, k)


def test_width4_adjoint_and_batch(gpu, holes_spec, star_spec):
    from dwost_amd import scenes

    prob = scenes.adjoint_problem_3d(level=3, n_query=128)
    spec = prob.spec
    q = torch.as_tensor(spec.queries).cuda()
    ut, gt = torch.as_tensor(prob.u_target).cuda(), torch.as_tensor(prob.grad_target).cuda()
    grads = []
    for nw in (2, 4):
        g = torch.as_tensor(prob.g0).cuda().requires_grad_(True)
        h = torch.as_tensor(prob.h0).cuda().requires_grad_(True)
        scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
        cs = gpu.CustomSolver(gpu.WoSt3D(eps=1e-3, nwalks=32, nsteps=1024, node_width=nw))
        u, gr = cs.solve_grad(q, scene, seed=11)
        (torch.mean((u - ut) ** 2) + torch.mean(torch.sum((gr - gt) ** 2, 1))).backward()
        grads.append((g.grad.cpu().numpy(), h.grad.cpu().numpy(), u.detach().cpu().numpy()))
    (g1, h1, u1), (g2, h2, u2) = grads
    assert np.array_equal(u1.view(np.uint32), u2.view(np.uint32))
    for a, b in ((g1, g2), (h1, h2)):  # same scatter-adds, in another order
        assert np.allclose(a, b, rtol=1e-4, atol=1e-6 * float(np.abs(a).max()))
    items = []
    for spec, nw in ((holes_spec, 4), (star_spec, 2), (star_spec, 4)):
        sc = _scene(gpu, spec)
        items.append(gpu.BatchItem(gpu.WoSt(eps=1e-3, nwalks=64, nsteps=1024, node_width=nw), sc,
                                   torch.as_tensor(spec.queries[:200]).cuda(), seed=3))
    outs = gpu.BatchSolver(items).solve()
    for it, out in zip(items, outs):
        ref = replace(it.solver, node_width=2).solve_grad(it.points, it.scene, seed=3)
        assert np.array_equal(_bits(out["u"]), _bits(ref["u"]))
        assert np.array_equal(_bits(out["grad"]), _bits(ref["grad"]))


def test_node_width_rejects_other_values(gpu, star_spec):
    sc = _scene(gpu, star_spec)
    q = torch.as_tensor(star_spec.queries[:4]).cuda()
    for nw in (1, 3, 8):
        with pytest.raises(Exception):
            gpu.WoSt(nwalks=4, node_width=nw).solve(q, sc)
