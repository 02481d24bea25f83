"""Adjoint (replay-from-seed) gradients of the walk estimators with respect to
per-vertex boundary data, through the C ABI on the GPU.

For fixed seeds every walk estimate is linear in the per-vertex g and h (the
walks themselves depend on geometry only), so the exact gradient of
L = sum_i w_i u_i (+ sum_i c_i . grad u_i) is J^T w with J the Jacobian of the
estimator.  Column v of J is the oracle's estimate with g = e_v, h = 0 (or the
reverse), computed walk by walk on the CPU: that pins the adjoint without
restating it.  Central finite differences of a quadratic loss are exact for
a linear estimator, so the CustomSolver check uses wide steps and a 1e-2
relative bound (observed agreement is at fp32 rounding level)."""

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


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _per_vertex_scene(gpu, spec, g, h):
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    bc = gpu.BoundaryCondition.per_vertex(_t(g.astype(np.float32)), _t(h.astype(np.float32)))
    return cls(spec.vertices, spec.indices, spec.types, bc)


def _boundary_vertices(spec, kind, rng, k):
    """k random vertices touched by primitives of the given type (1 Dirichlet, 0 Neumann)."""
    vs = np.unique(spec.indices[spec.types == kind].ravel())
    return rng.choice(vs, size=min(k, len(vs)), replace=False)


def _oracle_column_u(os_, spec, q, sol, seed, vertex, which):
    V = len(spec.vertices)
    gv = np.zeros(V, np.float32)
    hv = np.zeros(V, np.float32)
    (gv if which == "g" else hv)[vertex] = 1.0
    o = os_.solve_u(q, spec.bc_code_pv, sol.nwalks, nsteps=sol.nsteps, eps=sol.eps, min_R=sol.min_R,
                    cv=sol.control_variates, av=sol.antithetic_variates, seed=seed, nsubwalks=sol.nsubwalks,
                    gv=gv, hv=hv, per_walk=True)
    return o["walk_u"].astype(np.float64).reshape(len(q), sol.nwalks).mean(1)  # walk_u: mean over subwalks


def _oracle_column_fused(os_, spec, q, sol, seed, vertex, which):
    V = len(spec.vertices)
    gv = np.zeros(V, np.float32)
    hv = np.zeros(V, np.float32)
    (gv if which == "g" else hv)[vertex] = 1.0
    o = os_.solve_fused(q, spec.bc_code_pv, sol.nwalks, nsteps=sol.nsteps, eps=sol.eps, min_R=sol.min_R, seed=seed,
                        gv=gv, hv=hv, per_pair=True, cv=sol.control_variates)
    npairs = (sol.nwalks + 1) // 2
    # per pair: the sum over its two walks (u) and the two-walk gradient sum;
    # the estimates divide by 2 * npairs
    pu = o["pair_u"].astype(np.float64).reshape(len(q), npairs).sum(1) / (2 * npairs)
    pg = o["pair_g"].astype(np.float64).reshape(len(q), npairs, spec.dim).sum(1) / (2 * npairs)
    return pu, pg


def _spec(which, star_spec, ico_small):
    spec = dict(star=star_spec, ico=ico_small)[which]
    spec.bc_code_pv = 3
    return spec


@pytest.fixture(scope="module")
def ico_small():
    from dwost_amd import scenes

    return scenes.icosphere_scene(level=3, n_query=256)


def _fields(spec, seed):
    from dwost_amd import scenes

    rng = np.random.default_rng(seed)
    V = len(spec.vertices)
    g = scenes.analytic_u(scenes.BC_EXPCOS, spec.vertices.astype(np.float64)) + 0.1 * rng.standard_normal(V)
    h = 0.2 * rng.standard_normal(V)
    return g.astype(np.float32), h.astype(np.float32)


@pytest.mark.parametrize("nsub", [1, 2])
@pytest.mark.parametrize("which", ["star", "ico"])
def test_adjoint_u_matches_oracle_jacobian(gpu, oracle, which, nsub, star_spec, ico_small):
    spec = _spec(which, star_spec, ico_small)
    g, h = _fields(spec, 21)
    scene = _per_vertex_scene(gpu, spec, g, h)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    Solver = gpu.WoSt if spec.dim == 2 else gpu.WoSt3D
    sol = Solver(eps=1e-3, nwalks=32, nsteps=1024, nsubwalks=nsub)
    q = spec.queries[:16]
    n = len(q)
    flags = torch.empty(n * sol.nwalks, dtype=torch.uint8, device="cuda")
    sol.solve(_t(q), scene, seed=9, walk_flags=flags)
    fl = _np(flags)
    assert np.all(fl == (1 << nsub) - 1), "an outlier guard fired; the Jacobian check assumes none"
    rng = np.random.default_rng(5)
    w = rng.standard_normal(n).astype(np.float32)
    V = len(spec.vertices)
    dg = torch.zeros(V, device="cuda")
    dh = torch.zeros(V, device="cuda")
    bc = scene.boundary.to_c(V)
    import ctypes as C

    L = gpu._lib
    prm = sol._params(9)
    qt, wt = _t(q), _t(w)  # keep the device copies alive until the kernel has run
    L.check(L.lib().dw_adjoint_u(scene.handle, L.ptr(qt), n, C.byref(prm), C.byref(bc), L.ptr(wt), L.ptr(flags),
                                 L.ptr(dg), L.ptr(dh), None))
    dg, dh = _np(dg), _np(dh)
    checked = 0
    for kind, which_f, grad in ((1, "g", dg), (0, "h", dh)):
        if not np.any(spec.types == kind):
            continue
        for v in _boundary_vertices(spec, kind, rng, 12):
            col = _oracle_column_u(os_, spec, q, sol, 9, v, which_f)
            ref = float(np.dot(w.astype(np.float64), col))
            scale = float(np.dot(np.abs(w), np.abs(col))) + 1e-6
            assert abs(grad[v] - ref) <= 1e-4 * scale, (which_f, v, grad[v], ref)
            checked += 1
    # vertices touched by no primitive of the kind get nothing
    only_n = np.setdiff1d(np.arange(V), np.unique(spec.indices[spec.types == 1].ravel()))
    assert np.all(dg[only_n] == 0)
    assert checked >= 12
    print(f"adjoint u: {checked} Jacobian columns match")


@pytest.mark.parametrize("which", ["star", "ico"])
def test_adjoint_fused_matches_oracle_jacobian(gpu, oracle, which, star_spec, ico_small):
    spec = _spec(which, star_spec, ico_small)
    g, h = _fields(spec, 22)
    scene = _per_vertex_scene(gpu, spec, g, h)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    Solver = gpu.WoSt if spec.dim == 2 else gpu.WoSt3D
    sol = Solver(eps=1e-3, nwalks=32, nsteps=1024)
    q = spec.queries[:16]
    n, D = len(q), spec.dim
    npairs = (sol.nwalks + 1) // 2
    flags = torch.empty(n * npairs, dtype=torch.uint8, device="cuda")
    sol.solve_grad(_t(q), scene, seed=4, pair_flags=flags)
    assert np.all(_np(flags) == 3), "an outlier guard fired; the Jacobian check assumes none"
    rng = np.random.default_rng(6)
    wu = rng.standard_normal(n).astype(np.float32)
    wg = rng.standard_normal((n, D)).astype(np.float32)
    V = len(spec.vertices)
    dg = torch.zeros(V, device="cuda")
    dh = torch.zeros(V, device="cuda")
    import ctypes as C

    L = gpu._lib
    prm = sol._params(4)
    bc = scene.boundary.to_c(V)
    qt, wut, wgt = _t(q), _t(wu), _t(wg)  # alive until the kernel has run
    L.check(L.lib().dw_adjoint_fused(scene.handle, L.ptr(qt), n, C.byref(prm), C.byref(bc), L.ptr(wut), L.ptr(wgt),
                                     L.ptr(flags), L.ptr(dg), L.ptr(dh), None))
    dg, dh = _np(dg), _np(dh)
    checked = 0
    for kind, which_f, grad in ((1, "g", dg), (0, "h", dh)):
        if not np.any(spec.types == kind):
            continue
        for v in _boundary_vertices(spec, kind, rng, 12):
            pu, pg = _oracle_column_fused(os_, spec, q, sol, 4, v, which_f)
            ref = float(np.dot(wu, pu) + np.sum(wg * pg))
            scale = float(np.dot(np.abs(wu), np.abs(pu)) + np.sum(np.abs(wg * pg))) + 1e-6
            assert abs(grad[v] - ref) <= 1e-4 * scale, (which_f, v, grad[v], ref)
            checked += 1
    assert checked >= 12
    print(f"adjoint fused: {checked} Jacobian columns match")


def test_custom_solver_fd_3d(gpu):
    """The 3D adjoint workload in miniature: per-vertex g, h trainable,
    loss = MSE(u) + MSE(grad u) vs e^x cos y; central differences on 32
    random vertices (16 g, 16 h) with common random numbers."""
    from dwost_amd import scenes

    prob = scenes.adjoint_problem_3d(level=3, n_query=256)
    spec = prob.spec
    g = _t(prob.g0).requires_grad_(True)
    h = _t(prob.h0).requires_grad_(True)
    scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
    cs = gpu.CustomSolver(gpu.WoSt3D(eps=1e-3, nwalks=64, nsteps=1024))
    q = _t(spec.queries)
    ut, gt = _t(prob.u_target), _t(prob.grad_target)

    def loss_of(gv, hv):
        scene.boundary = gpu.BoundaryCondition.per_vertex(gv, hv)
        u, gr = cs.solve_grad(q, scene, seed=17)
        return torch.mean((u - ut) ** 2) + torch.mean(torch.sum((gr - gt) ** 2, 1))

    loss = loss_of(g, h)
    loss.backward()
    dg, dh = _np(g.grad).astype(np.float64), _np(h.grad).astype(np.float64)
    assert np.isfinite(dg).all() and np.isfinite(dh).all()
    rng = np.random.default_rng(8)
    g0, h0 = prob.g0.astype(np.float64), prob.h0.astype(np.float64)
    errs = []
    with torch.no_grad():
        for kind, grad in ((1, dg), (0, dh)):
            for v in _boundary_vertices(spec, kind, rng, 16):
                d = 0.25
                vals = []
                for s in (1.0, -1.0):
                    gp, hp = g0.copy(), h0.copy()
                    (gp if kind == 1 else hp)[v] += s * d
                    vals.append(float(loss_of(_t(gp.astype(np.float32)), _t(hp.astype(np.float32)))))
                fd = (vals[0] - vals[1]) / (2 * d)
                errs.append(abs(fd - grad[v]) / max(abs(fd), 1e-3 * float(np.abs(grad).max())))
    print(f"CustomSolver FD: max rel err {max(errs):.2e} over {len(errs)} vertices")
    assert len(errs) == 32
    assert max(errs) <= 1e-2


def test_custom_solver_u_and_replay_solver(gpu, star_spec):
    """CustomSolver.solve (u only) on the 2D star: autograd gradient equals the
    direct adjoint, and a different bwd_solver replays its own walks."""
    spec = star_spec
    g, h = _fields(spec, 23)
    gt, ht = _t(g).requires_grad_(True), _t(h).requires_grad_(True)
    scene = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(gt, ht))
    q = _t(spec.queries[:64])
    w = _t(np.random.default_rng(3).standard_normal(64).astype(np.float32))
    fwd = gpu.WoSt(eps=1e-3, nwalks=32, nsteps=1024)
    bwd = gpu.WoSt(eps=1e-3, nwalks=48, nsteps=1024)

    def grads(cs):
        gt.grad = None
        ht.grad = None
        u = cs.solve(q, scene, seed=2)
        (u * w).sum().backward()
        return _np(gt.grad).copy(), _np(ht.grad).copy(), _np(u)

    g1, h1, u1 = grads(gpu.CustomSolver(fwd, bwd))
    g2, h2, u2 = grads(gpu.CustomSolver(bwd, bwd))
    g3, h3, u3 = grads(gpu.CustomSolver(fwd, fwd))
    assert np.array_equal(u1, u3)  # the forward value is fwd_solver's
    assert np.allclose(g1, g2, rtol=1e-5, atol=1e-6) and np.allclose(h1, h2, rtol=1e-5, atol=1e-6)
    assert not np.allclose(g1, g3)  # other walks, other (unbiased) gradient
    # gradient of sum w*u is linear in w: the adjoint of 2w is twice the adjoint of w
    gt.grad = None
    ht.grad = None
    (gpu.CustomSolver(fwd).solve(q, scene, seed=2) * (2 * w)).sum().backward()
    assert np.allclose(_np(gt.grad), 2 * g3, rtol=1e-5, atol=1e-6)


def test_custom_solver_rejects_analytic_bc(gpu, star_spec):
    spec = star_spec
    scene = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.exp_cos())
    with pytest.raises(ValueError):
        gpu.CustomSolver(gpu.WoSt()).solve(_t(spec.queries[:4]), scene)
