"""Walks per lane of the fused solve: a lane runs m walk pairs of one point back
to back per job it takes from the queue.  Each pair keeps its own stream, so
every m returns the bits of m = 1 (u, grad u, per-pair values, step counts and
guard flags), also when m does not divide the pair count, for points outside
the domain, in batches mixing m, in the adjoint, and against the oracle."""

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
def test_walks_per_lane_keeps_results(gpu, which, star_spec, holes_spec, ico_spec):
    spec = dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which]
    sc = _scene(gpu, spec)
    eps = spec.meta["eps"]
    # 50 walks = 25 pairs: no m > 1 divides it; the last point lies outside
    q = np.concatenate([spec.queries[:95], np.full((1, spec.dim), 50.0, np.float32)])
    qt = torch.as_tensor(q).cuda()
    Sol = gpu.WoSt if spec.dim == 2 else gpu.WoSt3D
    base = Sol(eps=eps, nwalks=50, nsteps=1024)
    ref = base.solve_grad(qt, sc, seed=5, per_pair=True, count_steps=True)
    for m in (2, 4, 8):
        for sched in (1, 2):
            g = replace(base, walks_per_lane=m, schedule=sched).solve_grad(qt, sc, seed=5, per_pair=True,
                                                                            count_steps=True)
            for k in ("u", "grad", "pair_vals", "pair_steps"):
                assert np.array_equal(_bits(g[k]), _bits(ref[k])), (m, sched, k)
            assert int(g["steps"][0]) == int(ref["steps"][0])
    assert float(ref["u"][-1]) == 0.0


def test_walks_per_lane_matches_oracle(gpu, oracle, holes_spec):
    spec = holes_spec
    sc = _scene(gpu, spec)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = np.ascontiguousarray(spec.queries[:48])
    sol = gpu.WoSt(eps=1e-3, nwalks=64, nsteps=1024, exact_cones=True, walks_per_lane=4)
    g = sol.solve_grad(torch.as_tensor(q).cuda(), sc, seed=5, per_pair=True, count_steps=True)
    o = os_.solve_fused(q, spec.bc_code, 64, nsteps=1024, eps=1e-3, seed=5, per_pair=True)
    pv = g["pair_vals"].cpu().numpy()
    assert np.array_equal(pv[:, 0].view(np.uint32), o["pair_u"].view(np.uint32))
    assert np.array_equal(pv[:, 1:].view(np.uint32), o["pair_g"].view(np.uint32))
    assert np.array_equal(g["pair_steps"].cpu().numpy(), o["pair_steps"])


def test_walks_per_lane_batch_and_adjoint(gpu, holes_spec, star_spec):
    from dwost_amd import scenes

    items = []
    for spec, m in ((holes_spec, 4), (star_spec, 1), (star_spec, 8)):
        sc = _scene(gpu, spec)
        items.append(gpu.BatchItem(gpu.WoSt(eps=1e-3, nwalks=42, nsteps=1024, walks_per_lane=m), sc,
                                   torch.as_tensor(spec.queries[:150]).cuda(), seed=3))
    for graph in (False, True):
        b = gpu.BatchSolver(items, graph=graph)
        for _ in range(2):
            outs = b.solve()
            torch.cuda.synchronize()
            for it, out in zip(items, outs):
                ref = replace(it.solver, walks_per_lane=1).solve_grad(it.points, it.scene, seed=3)
                assert np.array_equal(_bits(out["u"]), _bits(ref["u"]))
                assert np.array_equal(_bits(out["grad"]), _bits(ref["grad"]))
        b.close()
    prob = scenes.adjoint_problem_3d(level=3, n_query=128)
    spec = prob.spec
    q = torch.as_tensor(spec.queries).cuda()
    ut, gt = torch.as_tensor(prob.u_target).cuda(), torch.as_tensor(prob.grad_target).cuda()
    grads = []
    for m in (1, 4):
        g = torch.as_tensor(prob.g0).cuda().requires_grad_(True)
        h = torch.as_tensor(prob.h0).cuda().requires_grad_(True)
        scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
        cs = gpu.CustomSolver(gpu.WoSt3D(eps=1e-3, nwalks=32, nsteps=1024, walks_per_lane=m))
        u, gr = cs.solve_grad(q, scene, seed=11)
        (torch.mean((u - ut) ** 2) + torch.mean(torch.sum((gr - gt) ** 2, 1))).backward()
        grads.append((g.grad.cpu().numpy(), h.grad.cpu().numpy(), u.detach().cpu().numpy()))
    (g1, h1, u1), (g2, h2, u2) = grads
    assert np.array_equal(u1.view(np.uint32), u2.view(np.uint32))
    for a, b_ in ((g1, g2), (h1, h2)):  # same scatter-adds, in another order
        assert np.allclose(a, b_, rtol=1e-4, atol=1e-6 * float(np.abs(a).max()))


def test_walks_per_lane_rejects_other_values(gpu, star_spec):
    sc = _scene(gpu, star_spec)
    q = torch.as_tensor(star_spec.queries[:4]).cuda()
    for m in (-1, 3, 16):
        with pytest.raises(Exception):
            gpu.WoSt(nwalks=4, walks_per_lane=m).solve_grad(q, sc)
