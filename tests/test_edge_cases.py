"""Edge cases through the C ABI: empty point sets, ragged sizes (point counts
that fill no wave, odd walk counts, a single walk) and a batch item without
points.  Ragged runs are compared with the oracle walk for walk."""

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


def _spec(which):
    from dwost_amd import scenes

    if which == "star":
        return scenes.star_scene_2d(n_query=64)
    return scenes.icosphere_scene(level=3, n_query=64)


def _gpu_scene(gpu, spec):
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    return cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))


def _solver(gpu, spec, W, **kw):
    return (gpu.WoSt if spec.dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=W, nsteps=1024, **kw)


@pytest.mark.parametrize("which", ["star", "ico"])
def test_empty_point_sets(gpu, which):
    spec = _spec(which)
    D = spec.dim
    sc = _gpu_scene(gpu, spec)
    p = torch.empty(0, D, device="cuda")
    assert sc.closest_point(p)["d"].shape == (0,)
    assert sc.intersect(p, p, 1.0)["prim"].shape == (0,)
    assert sc.closest_silhouette(p, 1.0)["sil"].shape == (0,)
    assert sc.star_radius(p, 1.0).shape == (0,)
    sol = _solver(gpu, spec, 8)
    r = sol.solve(p, sc, per_walk=True, count_steps=True)
    assert r["u"].shape == (0,) and r["walk_u"].shape == (0,) and int(r["steps"][0]) == 0
    r = sol.solve_grad(p, sc, per_pair=True, count_steps=True)
    assert r["grad"].shape == (0, D) and int(r["steps"][0]) == 0
    du, st = sol.solve_dude(p, p, sc, count_steps=True)
    assert du.shape == (0,) and int(st[0]) == 0
    out = gpu.StreamlineTracer(sc, sol, step_size=0.05, nwalks=4).next_position(p)
    assert out.shape == (0, D)
    torch.cuda.synchronize()


@pytest.mark.parametrize("which,n,W,cvav", [("star", 37, 33, True), ("star", 5, 1, True), ("ico", 13, 7, False),
                                             ("ico", 13, 7, True)])
def test_ragged_sizes_match_oracle(gpu, oracle, which, n, W, cvav):
    spec = _spec(which)
    sc = _gpu_scene(gpu, spec)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = spec.queries[:n]
    qt = torch.as_tensor(q).cuda()
    sol = _solver(gpu, spec, W, control_variates=cvav, antithetic_variates=cvav)
    g = sol.solve(qt, sc, seed=11, per_walk=True, count_steps=True)
    flags = dict(cv=sol.control_variates, av=sol.antithetic_variates)
    o = os_.solve_u(q, spec.bc_code, W, nsteps=1024, eps=1e-3, seed=11, per_walk=True, **flags)
    torch.cuda.synchronize()
    assert np.array_equal(g["walk_u"].cpu().numpy().view(np.uint32), o["walk_u"].view(np.uint32))
    assert np.array_equal(g["walk_steps"].cpu().numpy(), o["walk_steps"])
    assert int(g["steps"][0]) == int(o["walk_steps"].astype(np.int64).sum())
    f = sol.solve_grad(qt, sc, seed=11, per_pair=True, count_steps=True)
    of = os_.solve_fused(q, spec.bc_code, W, nsteps=1024, eps=1e-3, seed=11, per_pair=True, cv=sol.control_variates)
    torch.cuda.synchronize()
    pv = f["pair_vals"].cpu().numpy()
    assert pv.shape == (n * ((W + 1) // 2), spec.dim + 1)
    assert np.array_equal(pv[:, 0].view(np.uint32), of["pair_u"].view(np.uint32))
    assert np.array_equal(pv[:, 1:].view(np.uint32), of["pair_g"].view(np.uint32))
    assert int(f["steps"][0]) == of["steps"]
    rng = np.random.default_rng(3)
    e = rng.standard_normal((n, spec.dim))
    e = (e / np.linalg.norm(e, axis=1, keepdims=True)).astype(np.float32)
    du = sol.solve_dude(qt, torch.as_tensor(e).cuda(), sc, seed=11)
    od = os_.solve_dude(q, e, spec.bc_code, W, nsteps=1024, eps=1e-3, seed=11, **flags)["dude"]
    torch.cuda.synchronize()
    assert np.array_equal(du.cpu().numpy().view(np.uint32), od.view(np.uint32))


def test_batch_with_an_empty_item(gpu):
    spec = _spec("star")
    sc = _gpu_scene(gpu, spec)
    sol = _solver(gpu, spec, 16)
    q = torch.as_tensor(spec.queries[:21]).cuda()
    items = [gpu.BatchItem(sol, sc, q[:0], seed=1), gpu.BatchItem(sol, sc, q, seed=2)]
    b = gpu.BatchSolver(items)
    outs = b.solve()
    ref = sol.solve_grad(q, sc, seed=2)
    torch.cuda.synchronize()
    assert outs[0]["u"].shape == (0,)
    assert np.array_equal(outs[1]["u"].cpu().numpy().view(np.uint32), ref["u"].cpu().numpy().view(np.uint32))
    assert np.array_equal(outs[1]["grad"].cpu().numpy().view(np.uint32), ref["grad"].cpu().numpy().view(np.uint32))
    b.close()
