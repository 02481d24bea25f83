"""WoStGrad (diff_wost/solvers/wost_grad.py) on the GPU against the oracle, walk
for walk: the per-walk du/de bits and step counts must be the oracle's on a
mixed star polygon (e^x cos y), an all-Neumann one with u = x (the field of the
reference's streamline example, examples/example_2d/scene.py), the holes scene
(x^2 - y^2), with two subwalks, with the ignore switches and a small min_R;
the mean is the Kahan-reduced walk mean."""

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


def _dirs(n, seed):
    e = np.random.default_rng(seed).standard_normal((n, 2))
    return (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)


def _case(name):
    from dwost_amd import scenes

    if name == "holes":
        spec = scenes.holes_scene_2d(n_outer=1024, n_holes=4, n_query=64, eps=1e-3)
        return spec, spec.types, spec.bc_code, (0, 0, 0, 0)
    spec = scenes.star_scene_2d(n_query=64)
    if name == "allN_linear":
        return spec, np.zeros_like(spec.types), 2, (0, 1, 0, 0)
    if name == "allD_expcos":
        return spec, np.ones_like(spec.types), 0, (0, 0, 0, 0)
    return spec, spec.types, 0, (0, 0, 0, 0)


@pytest.mark.parametrize("name,W,kw", [
    ("mixed", 64, {}),
    ("allN_linear", 64, {}),
    ("holes", 32, {}),
    ("allD_expcos", 32, {}),
    ("mixed", 16, dict(nsubwalks=2)),
    ("mixed", 32, dict(min_R=1e-4)),
    ("allN_linear", 32, dict(ignore_silhouette=True, ignore_boundary=True)),
    ("mixed", 32, dict(ignore_source=True, control_variates=False, rho_max=2.0)),
])
def test_wostgrad_matches_oracle(gpu, oracle, name, W, kw):
    spec, types, code, coeffs = _case(name)
    bc = gpu.BoundaryCondition(code, coeffs)
    sc = gpu.Scene2D(spec.vertices, spec.indices, types, bc)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, types)
    q = np.ascontiguousarray(spec.queries[:32])
    e = _dirs(len(q), 4)
    sol = gpu.WoStGrad(eps=1e-3, nwalks=W, nsteps=1024, **kw)
    g = sol.solve_dude(torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda(), sc, seed=9, per_walk=True,
                       count_steps=True)
    torch.cuda.synchronize()
    okw = dict(nsteps=1024, eps=1e-3, min_R=sol.min_R, cv=sol.control_variates, seed=9, nsubwalks=sol.nsubwalks,
               rho_max=sol.rho_max, ignore_boundary=sol.ignore_boundary, ignore_silhouette=sol.ignore_silhouette,
               ignore_source=sol.ignore_source, coeffs=coeffs)
    o = os_.grad_dude(q, e, code, W, per_walk=True, **okw)
    gw, gs = g["walk_dude"].cpu().numpy(), g["walk_steps"].cpu().numpy()
    same = (gw.view(np.uint32) == o["walk_dude"].view(np.uint32)) & (gs == o["walk_steps"])
    bad = np.flatnonzero(~same)
    print(f"{name} {kw}: {same.sum()} of {same.size} walks identical; mean steps {gs.mean():.1f}")
    assert bad.size == 0, bad[:20].tolist()
    assert int(g["steps"][0]) == int(gs.astype(np.int64).sum())
    assert np.allclose(g["dude"].cpu().numpy(), o["dude"], rtol=1e-5, atol=1e-5)


def test_wostgrad_rejects_per_vertex_data_and_3d(gpu, ico_spec, star_spec):
    V = len(star_spec.vertices)
    g = torch.zeros(V, device="cuda")
    sc = gpu.Scene2D(star_spec.vertices, star_spec.indices, star_spec.types, gpu.BoundaryCondition.per_vertex(g, g))
    q = torch.as_tensor(star_spec.queries[:4]).cuda()
    with pytest.raises(gpu.DwostError):
        gpu.WoStGrad(nwalks=4).solve_dude(q, q, sc)
    with pytest.raises(NotImplementedError):
        gpu.WoStGrad(nwalks=4).solve(q, sc)


def test_tracer_with_wostgrad_matches_oracle(gpu, oracle):
    """StreamlineTracer driven by WoStGrad, as in the reference example
    (all-Neumann scene, u = x): two chained steps bit for bit against the oracle."""
    spec, types, code, coeffs = _case("allN_linear")
    sc = gpu.Scene2D(spec.vertices, spec.indices, types, gpu.BoundaryCondition(code, coeffs))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, types)
    sol = gpu.WoStGrad(eps=1e-3, nwalks=1, nsteps=1024)
    tr = gpu.StreamlineTracer(sc, sol, step_size=0.02, nwalks=8)
    q = np.ascontiguousarray(spec.queries[:24])
    p = torch.as_tensor(q).cuda()
    o = q
    for it in range(2):
        p, st = tr.next_position(p, seed=it, count_steps=True)
        r = os_.grad_trace_step(o, code, 8, 0.02, nsteps=1024, eps=1e-3, min_R=sol.min_R, seed=it, coeffs=coeffs)
        g = p.cpu().numpy()
        assert np.array_equal(g.view(np.uint32), r["pos"].view(np.uint32)), it
        assert int(st[0]) == r["steps"]
        o = g
    # u = x: the streamlines move along +x on average
    assert np.mean(g[:, 0] - q[:, 0]) > 0


@pytest.mark.parametrize("solver", ["wostgrad", "wost"])
def test_trace_graph_replays_match_eager(gpu, solver):
    """The whole trace loop captured into one HIP graph: every replay (with eager
    work on the default stream between replays, and new start points written in
    place) returns the eager trace's bits at every step."""
    spec, types, code, coeffs = _case("allN_linear" if solver == "wostgrad" else "mixed")
    sc = gpu.Scene2D(spec.vertices, spec.indices, types, gpu.BoundaryCondition(code, coeffs))
    sol = gpu.WoStGrad(eps=1e-3, nwalks=1, nsteps=1024) if solver == "wostgrad" else \
        gpu.WoSt(eps=1e-3, nwalks=8, nsteps=1024)
    tr = gpu.StreamlineTracer(sc, sol, step_size=0.02, nwalks=4)
    q = torch.as_tensor(np.ascontiguousarray(spec.queries[:40])).cuda()
    start = q.clone()
    tg = tr.trace_graph(start, 3)
    for rep in range(3):
        if rep == 2:
            start.copy_(q.flip(0))
        ref = tr.trace(start.clone(), 3)
        torch.cuda.synchronize()
        out = tg.replay()
        torch.cuda.synchronize()
        for k in range(3):
            assert torch.equal(out[k].view(torch.int32), ref[k].view(torch.int32)), (rep, k)
        sol.solve_dude(q[:8], q[:8], sc, seed=rep) if solver == "wostgrad" else sol.solve(q[:8], sc, seed=rep)
        torch.cuda.synchronize()


@pytest.mark.parametrize("code,coeffs,kw", [(2, (0, 1, 0, 0), {}), (1, (0, 0, 0, 0), {}),
                                            (0, (0, 0, 0, 0), dict(min_R=1e-2, nsubwalks=2)),
                                            (2, (0, 0.3, -0.5, 0.8), dict(ignore_source=True, rho_max=2.0))])
def test_wostgrad3d_matches_oracle(gpu, oracle, code, coeffs, kw):
    """WoStGrad3D on an all-Neumann noisy icosphere (the reference example's
    setting): per-walk bits and step counts equal the oracle's."""
    from dwost_amd import scenes

    spec = scenes.icosphere_scene(level=2, n_query=64)
    types = np.zeros_like(spec.types)
    sc = gpu.Scene3D(spec.vertices, spec.indices, types, gpu.BoundaryCondition(code, coeffs))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, types)
    q = np.ascontiguousarray(spec.queries[:16])
    e = np.random.default_rng(7).standard_normal((16, 3))
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    sol = gpu.WoStGrad3D(eps=1e-3, nwalks=16, nsteps=1024, **kw)
    g = sol.solve_dude(torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda(), sc, seed=3, per_walk=True,
                       count_steps=True)
    torch.cuda.synchronize()
    o = os_.grad_dude(q, e, code, 16, nsteps=1024, eps=1e-3, min_R=sol.min_R, cv=sol.control_variates, seed=3,
                      nsubwalks=sol.nsubwalks, rho_max=sol.rho_max, ignore_source=sol.ignore_source, coeffs=coeffs,
                      per_walk=True)
    gw, gs = g["walk_dude"].cpu().numpy(), g["walk_steps"].cpu().numpy()
    same = (gw.view(np.uint32) == o["walk_dude"].view(np.uint32)) & (gs == o["walk_steps"])
    print(f"3D code {code} {kw}: {same.sum()} of {same.size} walks identical; mean steps {gs.mean():.1f}")
    assert same.all(), np.flatnonzero(~same)[:20].tolist()
    assert np.allclose(g["dude"].cpu().numpy(), o["dude"], rtol=1e-5, atol=1e-5)


def test_wostgrad3d_needs_all_neumann(gpu, ico_spec):
    sc = gpu.Scene3D(ico_spec.vertices, ico_spec.indices, ico_spec.types, gpu.BoundaryCondition(0))
    q = torch.as_tensor(ico_spec.queries[:4]).cuda()
    with pytest.raises(gpu.DwostError):
        gpu.WoStGrad3D(nwalks=4).solve_dude(q, q, sc)


@pytest.mark.parametrize("dim", [2, 3])
def test_wostgrad_persistent_matches_simple(gpu, dim, monkeypatch):
    """WoStGrad walks run as persistent jobs, one step per loop iteration; every
    walk (value and step count) equals the one-walk-per-thread kernel's, with
    several sub-walks per job and points outside the domain among the jobs."""
    from dwost_amd import scenes

    if dim == 2:
        spec = scenes.holes_scene_2d(n_query=300, eps=1e-3)
        sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoStGrad(eps=1e-3, nwalks=24, nsteps=1024, nsubwalks=2)
    else:
        spec = scenes.icosphere_scene(level=3, n_query=200)
        types = np.zeros_like(spec.types)
        sc = gpu.Scene3D(spec.vertices, spec.indices, types, gpu.BoundaryCondition(2, (0, 1, 0, 0)))
        sol = gpu.WoStGrad3D(eps=1e-3, nwalks=16, nsteps=1024, nsubwalks=2)
    q = np.concatenate([spec.queries, np.full((3, dim), 9.0, np.float32)])
    e = np.random.default_rng(5).standard_normal(q.shape)
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    qt, et = torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda()
    monkeypatch.setenv("DWOST_KERNEL", "simple")
    ref = sol.solve_dude(qt, et, sc, seed=6, per_walk=True, count_steps=True)
    torch.cuda.synchronize()
    monkeypatch.delenv("DWOST_KERNEL")
    got = sol.solve_dude(qt, et, sc, seed=6, per_walk=True, count_steps=True)
    torch.cuda.synchronize()
    for k in ("dude", "walk_dude", "walk_steps"):
        assert np.array_equal(got[k].cpu().numpy().view(np.uint32), ref[k].cpu().numpy().view(np.uint32)), k
    assert int(got["steps"][0]) == int(ref["steps"][0])
    assert float(got["dude"][-1]) == 0.0


def test_wostgrad_tracer_persistent_matches_simple(gpu, monkeypatch):
    """The WoStGrad tracer's samples as persistent jobs (x then y walks, nested
    dudn walks a step at a time) equal the one-sample-per-thread kernel's bit
    for bit on a mixed scene with two sub-walks and points outside the domain."""
    from dwost_amd import scenes

    spec = scenes.holes_scene_2d(n_outer=1024, n_holes=4, n_query=200, eps=1e-3)
    sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    sol = gpu.WoStGrad(eps=1e-3, nwalks=1, nsteps=1024, nsubwalks=2)
    tr = gpu.StreamlineTracer(sc, sol, step_size=0.01, nwalks=6)
    q = np.concatenate([spec.queries, np.full((3, 2), 9.0, np.float32)])
    p = torch.as_tensor(q).cuda()
    monkeypatch.setenv("DWOST_KERNEL", "simple")
    ref, ref_st = tr.next_position(p, seed=3, count_steps=True)
    torch.cuda.synchronize()
    monkeypatch.delenv("DWOST_KERNEL")
    got, got_st = tr.next_position(p, seed=3, count_steps=True)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), ref.cpu().numpy().view(np.uint32))
    assert int(got_st[0]) == int(ref_st[0]) > 0
    assert np.array_equal(got[-3:].cpu().numpy(), q[-3:])


@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_wostgrad_bench_scenes_match_oracle(gpu, oracle, which):
    """The scenes of the WoStGrad measurement at their full size (the 256-segment
    star, the 4,096-segment holes scene, the 20,480-triangle icosphere, all
    Neumann), on points spread over each query set: every walk's bits and step
    count equal the oracle's."""
    from dwost_amd import scenes

    if which == "ico":
        spec = scenes.icosphere_scene(level=5, n_query=4096)
        types, code, coeffs = np.zeros_like(spec.types), 2, (0.0, 1.0, 0.0, 0.0)
        sc = gpu.Scene3D(spec.vertices, spec.indices, types, gpu.BoundaryCondition(code, coeffs))
        sol = gpu.WoStGrad3D(eps=1e-3, nwalks=8, nsteps=1024)
        pick = np.arange(0, 4096, 64)
    else:
        spec = scenes.star_scene_2d(n_query=65536) if which == "star" else scenes.holes_scene_2d(n_query=8192, eps=1e-3)
        types, code, coeffs = spec.types, spec.bc_code, (0.0, 0.0, 0.0, 0.0)
        sc = gpu.Scene2D(spec.vertices, spec.indices, types, gpu.BoundaryCondition(code))
        sol = gpu.WoStGrad(eps=1e-3, nwalks=16, nsteps=1024)
        pick = np.arange(0, len(spec.queries), len(spec.queries) // 256)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, types)
    q = np.ascontiguousarray(spec.queries[pick])
    e = np.zeros_like(q)
    e[:, 0] = 1.0
    g = sol.solve_dude(torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda(), sc, seed=1, per_walk=True,
                       count_steps=True)
    torch.cuda.synchronize()
    o = os_.grad_dude(q, e, code, sol.nwalks, nsteps=1024, eps=1e-3, min_R=sol.min_R, seed=1, rho_max=sol.rho_max,
                      coeffs=coeffs, per_walk=True)
    gw, gs = g["walk_dude"].cpu().numpy(), g["walk_steps"].cpu().numpy()
    same = (gw.view(np.uint32) == o["walk_dude"].view(np.uint32)) & (gs == o["walk_steps"])
    print(f"{which}: {same.sum()} of {same.size} walks identical; mean steps {gs.mean():.1f}, max {gs.max()}")
    assert same.all(), np.flatnonzero(~same)[:20].tolist()


def test_wostgrad_tracer_mixed_scene_matches_oracle(gpu, oracle):
    """The WoStGrad tracer on the holes scene (Dirichlet outer loop: every
    sample's walks end in a nested dudn walk), two sub-walks per sample and a
    point outside the domain: two chained steps bit for bit against the oracle."""
    from dwost_amd import scenes

    spec = scenes.holes_scene_2d(n_outer=1024, n_holes=4, n_query=64, eps=1e-3)
    code = spec.bc_code
    sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(code))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    sol = gpu.WoStGrad(eps=1e-3, nwalks=1, nsteps=1024, nsubwalks=2)
    tr = gpu.StreamlineTracer(sc, sol, step_size=0.01, nwalks=6)
    q = np.ascontiguousarray(np.concatenate([spec.queries[:24], np.full((1, 2), 9.0, np.float32)]))
    p, o = torch.as_tensor(q).cuda(), q
    for it in range(2):
        p, st = tr.next_position(p, seed=it, count_steps=True)
        r = os_.grad_trace_step(o, code, 6, 0.01, nsubwalks=2, nsteps=1024, eps=1e-3, min_R=sol.min_R, seed=it)
        g = p.cpu().numpy()
        assert np.array_equal(g.view(np.uint32), r["pos"].view(np.uint32)), it
        assert int(st[0]) == r["steps"]
        o = g


def test_zero_radius_walks_match_oracle(gpu, oracle):
    """Points on Neumann segments of the mixed star polygon where the
    rho-bounded star radius along +x is 0 (and along +y is not): a du/dx walk
    from there takes R = 0 at every step until its step cap, so each step's
    source sample fails all 1,000 rejection rounds (the GPU jumps over 999 of
    them).  In a tracer sample the y walk continues on the x walk's stream, so
    its bits pin the number of draws the stuck x walk consumed."""
    from dwost_amd import scenes

    spec = scenes.star_scene_2d(n_query=64)
    code = spec.bc_code
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    V = spec.vertices.astype(np.float32)
    a, b = V[spec.indices[:, 0]], V[spec.indices[:, 1]]
    t = (np.arange(1, 64) / 64).astype(np.float32)
    P = np.ascontiguousarray((a[:, None, :] + t[None, :, None] * (b - a)[:, None, :]).reshape(-1, 2))
    ex, ey = np.zeros_like(P), np.zeros_like(P)
    ex[:, 0] = 1.0
    ey[:, 1] = 1.0
    rx, _ = os_.star_radius2_rho(2, P, ex, rho_max=1.2)
    ry, _ = os_.star_radius2_rho(2, P, ey, rho_max=1.2)
    cand = np.flatnonzero((rx == 0) & (ry > 0))
    ox = os_.grad_dude(P[cand], ex[cand], code, 1, nsteps=1024, eps=1e-3, min_R=1e-2, seed=1, per_walk=True)
    q = np.ascontiguousarray(P[cand[ox["walk_steps"] == 1024]])
    print(f"{len(cand)} candidates, {len(q)} stuck du/dx walks")
    assert len(q) >= 16

    sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(code))
    sol = gpu.WoStGrad(eps=1e-3, nwalks=4, nsteps=1024)
    e = np.zeros_like(q)
    e[:, 0] = 1.0
    got = sol.solve_dude(torch.as_tensor(q).cuda(), torch.as_tensor(e).cuda(), sc, seed=2, count_steps=True, per_walk=True)
    want = os_.grad_dude(q, e, code, 4, nsteps=1024, eps=1e-3, min_R=1e-2, seed=2, per_walk=True)
    assert (want["walk_steps"] == 1024).all()
    assert np.array_equal(got["walk_steps"].cpu().numpy(), want["walk_steps"])
    assert np.array_equal(got["walk_dude"].cpu().numpy().view(np.uint32), want["walk_dude"].view(np.uint32))

    tr = gpu.StreamlineTracer(sc, gpu.WoStGrad(eps=1e-3, nwalks=1, nsteps=1024), step_size=0.01, nwalks=4)
    p, st = tr.next_position(torch.as_tensor(q).cuda(), seed=3, count_steps=True)
    r = os_.grad_trace_step(q, code, 4, 0.01, nsteps=1024, eps=1e-3, min_R=1e-2, seed=3)
    g = p.cpu().numpy()
    assert np.array_equal(g.view(np.uint32), r["pos"].view(np.uint32))
    assert int(st[0]) == r["steps"]
    # the y walks move the points: their values depend on the stream position
    assert (g[:, 1] != q[:, 1]).sum() >= len(q) // 4
