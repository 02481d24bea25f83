"""Lane scheduling of the persistent walk kernels never changes a result: the
interleaved and the step-synchronous loops (and any gate threshold) return the
same bits for u, grad u and per-walk values in 2D and 3D, and the same adjoint
up to the order of its atomic scatter-adds.  Each
schedule is also checked against the oracle walk by walk in the parity tests
(AUTO picks interleaved in 3D, step-synchronous in 2D); here the opposite
schedule of each dimension is compared with the default."""

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


def _case(gpu, dim):
    from dwost_amd import scenes

    if dim == 2:
        spec = scenes.holes_scene_2d(n_query=400, eps=1e-3)
        sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoSt(eps=1e-3, nwalks=64, nsteps=1024)
    else:
        spec = scenes.icosphere_scene(level=3, n_query=300, eps=1e-3)
        sc = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoSt3D(eps=1e-3, nwalks=32, nsteps=1024)
    return sc, sol, torch.as_tensor(spec.queries).cuda()


@pytest.mark.parametrize("dim", [2, 3])
def test_schedules_agree(gpu, dim):
    sc, sol, q = _case(gpu, dim)
    ref_g = sol.solve_grad(q, sc, seed=7, per_pair=True, count_steps=True)
    ref_u = sol.solve(q, sc, seed=7, per_walk=True)
    for sched, gate in ((1, 0), (2, 0), (1, 1), (1, 64)):
        s = replace(sol, schedule=sched, gate=gate)
        g = s.solve_grad(q, sc, seed=7, per_pair=True, count_steps=True)
        u = s.solve(q, sc, seed=7, per_walk=True)
        for k in ("u", "grad", "pair_vals", "pair_steps"):
            assert np.array_equal(_bits(g[k]), _bits(ref_g[k])), (sched, gate, k)
        assert int(g["steps"][0]) == int(ref_g["steps"][0])
        for k in ("u", "walk_u", "walk_steps", "walk_term"):
            assert np.array_equal(_bits(u[k]), _bits(ref_u[k])), (sched, gate, k)


def test_adjoint_schedules_agree(gpu):
    from dwost_amd import scenes

    prob = scenes.adjoint_problem_3d(level=3, n_query=128)
    spec = prob.spec
    q = torch.as_tensor(spec.queries).cuda()
    ut, gt = torch.as_tensor(prob.u_target).cuda(), torch.as_tensor(prob.grad_target).cuda()
    grads = []
    for sched in (1, 2):
        g = torch.as_tensor(prob.g0).cuda().requires_grad_(True)
        h = torch.as_tensor(prob.h0).cuda().requires_grad_(True)
        scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
        cs = gpu.CustomSolver(gpu.WoSt3D(eps=1e-3, nwalks=32, nsteps=1024, schedule=sched))
        u, gr = cs.solve_grad(q, scene, seed=11)
        (torch.mean((u - ut) ** 2) + torch.mean(torch.sum((gr - gt) ** 2, 1))).backward()
        grads.append((g.grad.cpu().numpy(), h.grad.cpu().numpy(), u.detach().cpu().numpy()))
    (g1, h1, u1), (g2, h2, u2) = grads
    assert np.array_equal(u1.view(np.uint32), u2.view(np.uint32))
    for a, b in ((g1, g2), (h1, h2)):
        assert np.allclose(a, b, rtol=1e-4, atol=1e-6 * float(np.abs(a).max()))


def test_schedule_rejects_bad_values(gpu):
    sc, sol, q = _case(gpu, 2)
    with pytest.raises(Exception):
        replace(sol, schedule=3).solve(q[:4], sc)
    with pytest.raises(Exception):
        replace(sol, gate=65).solve(q[:4], sc)


@pytest.mark.parametrize("dim", [2, 3])
def test_short_walks_match_oracle(gpu, oracle, dim):
    """Step limits 0, 1, 2 and 3 subwalks (the step machine's edge transitions:
    a walk that ends before its first query, at its first step boundary, and
    walks restarting inside one job) under both schedules, walk by walk against
    the oracle."""
    from dwost_amd import scenes

    if dim == 2:
        spec = scenes.holes_scene_2d(n_query=64, eps=1e-3)
        sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        Sol = gpu.WoSt
    else:
        spec = scenes.icosphere_scene(level=3, n_query=64, eps=1e-3)
        sc = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        Sol = gpu.WoSt3D
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = np.ascontiguousarray(spec.queries)
    for nsteps in (0, 1, 2):
        o = os_.solve_u(q, spec.bc_code, 8, nsteps=nsteps, eps=1e-3, seed=9, nsubwalks=3, per_walk=True)
        for sched in (1, 2):
            sol = Sol(eps=1e-3, nwalks=8, nsteps=nsteps, nsubwalks=3, schedule=sched)
            g = sol.solve(torch.as_tensor(q).cuda(), sc, seed=9, per_walk=True)
            assert np.array_equal(_bits(g["walk_u"]), o["walk_u"].view(np.uint32)), (nsteps, sched)
            assert np.array_equal(g["walk_steps"].cpu().numpy(), o["walk_steps"]), (nsteps, sched)


@pytest.mark.parametrize("wos", [False, True])
@pytest.mark.parametrize("dim", [2, 3])
def test_persistent_dude_matches_simple(gpu, dim, wos, monkeypatch):
    """solve_dude runs its dudes (WoSt and WoS) as persistent point jobs; with the
    literal cone test both schedules return the bits of the one-point-per-thread
    kernel."""
    sc, sol, q = _case(gpu, dim)
    if wos:
        sol = (gpu.WoS if dim == 2 else gpu.WoS3D)(eps=1e-3, nwalks=12, nsteps=1024)
    sol = replace(sol, exact_cones=True, nwalks=12)
    rng = np.random.default_rng(3)
    e = rng.standard_normal((q.shape[0], dim)).astype(np.float32)
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    e = torch.as_tensor(e).cuda()
    monkeypatch.setenv("DWOST_KERNEL", "simple")
    ref, ref_steps = sol.solve_dude(q, e, sc, seed=4, count_steps=True)
    torch.cuda.synchronize()
    monkeypatch.delenv("DWOST_KERNEL")
    for sched in (1, 2):
        got, steps = replace(sol, schedule=sched).solve_dude(q, e, sc, seed=4, count_steps=True)
        assert np.array_equal(_bits(got), _bits(ref)), sched
        assert int(steps[0]) == int(ref_steps[0])


@pytest.mark.parametrize("dim", [2, 3])
def test_simple_switch_fused_and_flags(gpu, dim, monkeypatch):
    """DWOST_KERNEL=simple: an exact-cone fused solve without guard flags runs the
    one-pair-per-thread kernel and returns the persistent kernel's bits; a solve
    that asks for guard flags (the CustomSolver forward) or for the fast cone
    test keeps the persistent kernel, so the flags are written and the cone mode
    is the one asked for."""
    sc, sol, q = _case(gpu, dim)
    sol = replace(sol, nwalks=16)
    n, npairs = q.shape[0], 8
    exact = replace(sol, exact_cones=True)
    ref = exact.solve_grad(q, sc, seed=9, per_pair=True)
    fast_ref = sol.solve_grad(q, sc, seed=9, per_pair=True)
    flags_ref = torch.full((n * npairs,), 0xEE, dtype=torch.uint8, device="cuda")
    exact.solve_grad(q, sc, seed=9, pair_flags=flags_ref)
    torch.cuda.synchronize()
    monkeypatch.setenv("DWOST_KERNEL", "simple")
    got = exact.solve_grad(q, sc, seed=9, per_pair=True)
    got_fast = sol.solve_grad(q, sc, seed=9, per_pair=True)
    flags = torch.full((n * npairs,), 0xEE, dtype=torch.uint8, device="cuda")
    with_flags = exact.solve_grad(q, sc, seed=9, per_pair=True, pair_flags=flags)
    torch.cuda.synchronize()
    monkeypatch.delenv("DWOST_KERNEL")
    for k in ("u", "grad", "pair_vals", "pair_steps"):
        assert np.array_equal(_bits(got[k]), _bits(ref[k])), k
        assert np.array_equal(_bits(got_fast[k]), _bits(fast_ref[k])), k
        assert np.array_equal(_bits(with_flags[k]), _bits(ref[k])), k
    f = flags.cpu().numpy()
    assert np.array_equal(f, flags_ref.cpu().numpy())
    assert np.all(f <= 3)  # written: two guard bits per pair, never the fill value
