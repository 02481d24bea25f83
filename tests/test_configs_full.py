"""The benchmark workloads at full size on the GPU, against the oracle.

* noisy icosphere (81,920 triangles, level 6): per-walk and per-pair parity on
  256 points spread over the query set, 64 walks each;
* 64 uneven 2D scenes (segments log-U[64, 16384], 0-4 holes, queries
  log-U[100, 1e5] uncapped, 64/256/1024 walks): the one persistent batched
  launch equals the 64 per-scene solves bit for bit, eager and graph-replayed,
  and the per-scene walk pairs equal the oracle's on points of 8 scenes of
  different sizes;
* the 3D adjoint problem (40,962 vertices, per-vertex g and h, 16,384 points x
  256 walks): adjoint Jacobian columns against the oracle's forward columns,
  and central finite differences of the loss on 32 vertices over all 16,384
  points.
"""

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


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def ico6():
    from dwost_amd import scenes

    return scenes.icosphere_scene(level=6, n_query=4096)


@pytest.mark.parametrize("exact", [True, False])
def test_icosphere_full_walk_parity(gpu, oracle, ico6, exact):
    """WoSt3D.u walks (Solver.solve) on the 81,920-triangle icosphere: every walk's
    value, step count and terminal primitive equal the oracle's."""
    spec = ico6
    gs = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = np.ascontiguousarray(spec.queries[::16])  # 256 points over the whole interior
    assert len(q) == 256
    sol = gpu.WoSt3D(eps=1e-3, nwalks=64, nsteps=1024, exact_cones=exact)
    g = sol.solve(_t(q), gs, seed=21, per_walk=True, count_steps=True)
    o = os_.solve_u(q, spec.bc_code, 64, nsteps=1024, eps=1e-3, min_R=sol.min_R, cv=sol.control_variates,
                    av=sol.antithetic_variates, seed=21, per_walk=True)
    same = ((_bits(_np(g["walk_u"])) == _bits(o["walk_u"])) & (_np(g["walk_steps"]) == o["walk_steps"])
            & (_np(g["walk_term"]) == o["walk_term"]))
    bad = np.flatnonzero(~same)
    print(f"icosphere walks (exact_cones={exact}): {same.sum()} of {same.size} identical")
    assert bad.size <= (0 if exact else int(1e-3 * same.size)), bad[:20].tolist()
    assert int(_np(g["steps"])[0]) == int(_np(g["walk_steps"]).astype(np.int64).sum())
    assert np.allclose(_np(g["u"]), o["u"], rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("exact", [True, False])
def test_icosphere_full_fused_parity(gpu, oracle, ico6, exact):
    """The fused u + grad u estimator (the benchmark's) on the same 256 points:
    every antithetic pair's value bits and step count equal the oracle's."""
    spec = ico6
    gs = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = np.ascontiguousarray(spec.queries[8::16])
    sol = gpu.WoSt3D(eps=1e-3, nwalks=64, nsteps=1024, exact_cones=exact)
    g = sol.solve_grad(_t(q), gs, seed=22, per_pair=True, count_steps=True)
    o = os_.solve_fused(q, spec.bc_code, 64, nsteps=1024, eps=1e-3, seed=22, per_pair=True)
    pv = _np(g["pair_vals"])
    same = ((_bits(pv[:, 0]) == _bits(o["pair_u"])) & np.all(_bits(pv[:, 1:]) == _bits(o["pair_g"]), axis=1)
            & (_np(g["pair_steps"]) == o["pair_steps"]))
    bad = np.flatnonzero(~same)
    print(f"icosphere pairs (exact_cones={exact}): {same.sum()} of {same.size} identical")
    assert bad.size <= (0 if exact else int(1e-3 * same.size)), bad[:20].tolist()
    assert int(_np(g["steps"])[0]) == o["steps"]
    assert np.allclose(_np(g["u"]), o["u"], rtol=1e-5, atol=1e-5)
    assert np.allclose(_np(g["grad"]), o["grad"], rtol=1e-4, atol=1e-4)


def test_batch_64_scenes_uncapped(gpu, oracle):
    """All 64 scenes of the uneven batch at their drawn query counts: the
    batched launch equals every per-scene solve bit for bit (eager and
    graph-replayed), and per-scene walk pairs equal the oracle's on slices of
    8 scenes spanning the size range."""
    from dwost_amd import scenes

    specs = scenes.batch_scenes_2d(n_scenes=64, seed=5)
    n_pts = sum(len(s.queries) for s in specs)
    print(f"64 scenes, {n_pts} points, {sum(s.vertices.shape[0] for s in specs)} vertices")
    items = []
    for k, spec in enumerate(specs):
        sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoSt(eps=spec.meta["eps"], nwalks=spec.meta["nwalks"], nsteps=1024)
        items.append(gpu.BatchItem(sol, sc, _t(spec.queries), seed=k))
    refs = []
    for it in items:
        r = it.solver.solve_grad(it.points, it.scene, seed=it.seed)
        refs.append((r["u"], r["grad"]))
    torch.cuda.synchronize()
    refs = [(_np(u), _np(g)) for u, g in refs]
    for graph in (False, True):
        b = gpu.BatchSolver(items, graph=graph)
        for rep in range(2 if graph else 1):
            outs = b.solve()
            torch.cuda.synchronize()
            for k, (o, (u, g)) in enumerate(zip(outs, refs)):
                assert np.array_equal(_bits(_np(o["u"])), _bits(u)), (graph, rep, k)
                assert np.array_equal(_bits(_np(o["grad"])), _bits(g)), (graph, rep, k)
        b.close()
    # oracle on 8 scenes of increasing segment count, 4 points each (a slice
    # solved with point_base = its offset draws the whole set's streams)
    order = np.argsort([len(s.indices) for s in specs])
    picks = order[np.linspace(0, len(order) - 1, 8).astype(int)]
    for k in picks:
        spec, it = specs[k], items[k]
        j = len(spec.queries) // 2
        q = np.ascontiguousarray(spec.queries[j:j + 4])
        W = spec.meta["nwalks"]
        g = it.solver.solve_grad(_t(q), it.scene, seed=it.seed, per_pair=True, point_base=j)
        os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
        o = os_.solve_fused(q, spec.bc_code, W, nsteps=1024, eps=spec.meta["eps"], seed=it.seed, per_pair=True,
                            base=j)
        pv = _np(g["pair_vals"])
        same = ((_bits(pv[:, 0]) == _bits(o["pair_u"])) & np.all(_bits(pv[:, 1:]) == _bits(o["pair_g"]), axis=1)
                & (_np(g["pair_steps"]) == o["pair_steps"]))
        print(f"scene {k}: {len(spec.indices)} segments, W={W}: {same.sum()} of {same.size} pairs identical")
        assert same.sum() >= same.size - int(1e-3 * same.size), np.flatnonzero(~same)[:20].tolist()
        assert np.allclose(refs[k][0][j:j + 4], o["u"], rtol=1e-5, atol=1e-5)


@pytest.fixture(scope="module")
def adjoint6():
    from dwost_amd import scenes

    return scenes.adjoint_problem_3d(level=6, n_query=16384)


def test_adjoint_full_jacobian_columns(gpu, oracle, adjoint6):
    """dw_adjoint_fused on the 40,962-vertex problem (64 of its points, 256
    walks): dL/dg and dL/dh at 12 + 12 vertices touched by the walks equal
    w^T J columns, each column the oracle's forward estimate with g = e_v
    (or h = e_v) and everything else zero."""
    import ctypes as C

    prob = adjoint6
    spec = prob.spec
    V = len(spec.vertices)
    g, h = _t(prob.g0), _t(prob.h0)
    scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    sol = gpu.WoSt3D(eps=1e-3, nwalks=256, nsteps=1024)
    q = np.ascontiguousarray(spec.queries[::256])
    n, D, npairs = len(q), 3, 128
    flags = torch.empty(n * npairs, dtype=torch.uint8, device="cuda")
    sol.solve_grad(_t(q), scene, seed=31, pair_flags=flags)
    fl = _np(flags)
    rng = np.random.default_rng(17)
    wu = rng.standard_normal(n).astype(np.float32)
    wg = rng.standard_normal((n, D)).astype(np.float32)
    dg = torch.zeros(V, device="cuda")
    dh = torch.zeros(V, device="cuda")
    L = gpu._lib
    prm = sol._params(31)
    bc = scene.boundary.to_c(V)
    qt, wut, wgt = _t(q), _t(wu), _t(wg)
    L.check(L.lib().dw_adjoint_fused(scene.handle, L.ptr(qt), n, C.byref(prm), C.byref(bc), L.ptr(wut), L.ptr(wgt),
                                     L.ptr(flags), L.ptr(dg), L.ptr(dh), None))
    dg, dh = _np(dg), _np(dh)
    # the oracle applies no guard flags: weight the columns by the same survivors
    keep = fl.reshape(n, npairs)
    checked = 0
    for which, grad in (("g", dg), ("h", dh)):
        touched = np.flatnonzero(grad != 0)
        assert len(touched) >= 12, (which, len(touched))
        big = touched[np.argsort(-np.abs(grad[touched]))[:6]]
        rest = rng.choice(np.setdiff1d(touched, big), 6, replace=False)
        for v in np.concatenate([big, rest]):
            gv = np.zeros(V, np.float32)
            hv = np.zeros(V, np.float32)
            (gv if which == "g" else hv)[v] = 1.0
            o = os_.solve_fused(q, 3, 256, nsteps=1024, eps=1e-3, seed=31, gv=gv, hv=hv, per_pair=True,
                                cv=sol.control_variates)
            pu = o["pair_u"].astype(np.float64).reshape(n, npairs)
            pg = o["pair_g"].astype(np.float64).reshape(n, npairs, D)
            assert np.all(keep == 3), "an outlier guard fired; the column check assumes none"
            ref = float(np.dot(wu, pu.sum(1)) + np.sum(wg * pg.sum(1))) / (2 * npairs)
            scale = float(np.dot(np.abs(wu), np.abs(pu).sum(1)) + np.sum(np.abs(wg * pg.sum(1)))) / (2 * npairs)
            assert abs(grad[v] - ref) <= 1e-4 * scale + 1e-12, (which, v, grad[v], ref)
            checked += 1
    print(f"adjoint on 40,962 vertices: {checked} Jacobian columns match the oracle")
    assert checked == 24


def _progress(msg):
    """A line on the process's real stderr (pytest captures sys.stderr), so a
    long test shows it is alive."""
    import sys

    sys.__stderr__.write(f"[progress] {msg}\n")
    sys.__stderr__.flush()


@pytest.mark.timeout(900)
def test_adjoint_full_fd(gpu, adjoint6):
    """The adjoint workload at full size: CustomSolver.solve_grad on 16,384 points
    x 256 walks, loss = mean(|u - u*|^2 + |grad u - grad u*|^2) in float64;
    central differences (step 0.25, common random numbers: the same seed) on
    16 Dirichlet and 16 Neumann vertices agree with the autograd gradient
    within 1e-2 relative (the estimator is linear in g and h for fixed seeds,
    so central differences of the quadratic loss are exact up to rounding).
    The 64 perturbed forward solves run as batched launches of 16 items over
    the one uploaded scene (BatchItem.boundary), which return the per-call
    solve's bits."""
    import time

    prob = adjoint6
    spec = prob.spec
    g = _t(prob.g0).requires_grad_(True)
    h = _t(prob.h0).requires_grad_(True)
    scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
    sol = gpu.WoSt3D(eps=1e-3, nwalks=256, nsteps=1024)
    cs = gpu.CustomSolver(sol)
    q = _t(spec.queries)
    assert q.shape == (16384, 3)
    ut, gt = _t(prob.u_target).double(), _t(prob.grad_target).double()

    def loss(u, gr):
        return torch.mean((u.double() - ut) ** 2) + torch.mean(torch.sum((gr.double() - gt) ** 2, 1))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u0, gr0 = cs.solve_grad(q, scene, seed=5)
    l0 = loss(u0, gr0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    l0.backward()
    torch.cuda.synchronize()
    _progress(f"adjoint fd: forward {t1 - t0:.2f} s, backward {time.perf_counter() - t1:.2f} s, loss {float(l0):.4f}")
    dg, dh = _np(g.grad).astype(np.float64), _np(h.grad).astype(np.float64)
    assert np.isfinite(dg).all() and np.isfinite(dh).all()
    rng = np.random.default_rng(9)
    g0, h0 = prob.g0.astype(np.float64), prob.h0.astype(np.float64)
    d = 0.25
    cases = []  # (kind, vertex, sign, boundary)
    for kind in (1, 0):
        vs = np.unique(spec.indices[spec.types == kind].ravel())
        for v in rng.choice(vs, 16, replace=False):
            for sgn in (1.0, -1.0):
                gp, hp = g0.copy(), h0.copy()
                (gp if kind == 1 else hp)[v] += sgn * d
                cases.append((kind, int(v), sgn,
                              gpu.BoundaryCondition.per_vertex(_t(gp.astype(np.float32)), _t(hp.astype(np.float32)))))
    vals = {}
    t0 = time.perf_counter()
    with torch.no_grad():
        for c0 in range(0, len(cases), 16):
            chunk = cases[c0:c0 + 16]
            b = gpu.BatchSolver([gpu.BatchItem(sol, scene, q, seed=5, boundary=c[3]) for c in chunk])
            outs = b.solve()
            torch.cuda.synchronize()
            for c, o in zip(chunk, outs):
                vals[(c[0], c[1], c[2])] = float(loss(o["u"], o["grad"]))
            b.close()
            _progress(f"adjoint fd: {c0 + len(chunk)} of {len(cases)} perturbed solves, "
                      f"{time.perf_counter() - t0:.1f} s")
    errs = []
    for kind, grad in ((1, dg), (0, dh)):
        for v in sorted({c[1] for c in cases if c[0] == kind}):
            fd = (vals[(kind, v, 1.0)] - vals[(kind, v, -1.0)]) / (2 * d)
            errs.append(abs(fd - grad[v]) / max(abs(fd), 1e-3 * float(np.abs(grad).max())))
    print(f"FD on 32 vertices ({time.perf_counter() - t0:.1f} s): max rel err {max(errs):.2e}")
    _progress(f"adjoint fd: max rel err {max(errs):.2e}")
    assert len(errs) == 32
    assert max(errs) <= 1e-2
