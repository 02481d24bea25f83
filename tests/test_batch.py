"""Batched multi-scene solves: one persistent launch over many scenes of
uneven size must reproduce every per-scene solve bit for bit (walk streams are
keyed by (point, pair) per item), eagerly and when replayed from a captured
hipGraph with eager work between the replays."""

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


@pytest.fixture(scope="module")
def batch_items(gpu):
    from dwost_amd import scenes

    specs = scenes.batch_scenes_2d(n_scenes=7, seed=5, max_queries=300)
    items, refs = [], []
    for k, spec in enumerate(specs):
        sc = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoSt(eps=1e-3, nwalks=min(spec.meta["nwalks"], 256), nsteps=1024)
        q = torch.as_tensor(spec.queries).cuda()
        items.append(gpu.BatchItem(sol, sc, q, seed=100 + k))
    for it in items:
        refs.append(it.solver.solve_grad(it.points, it.scene, seed=it.seed))
    torch.cuda.synchronize()
    return items, [(r["u"].cpu().numpy(), r["grad"].cpu().numpy()) for r in refs]


def _check(outs, refs):
    for k, (o, (u, g)) in enumerate(zip(outs, refs)):
        assert np.array_equal(o["u"].cpu().numpy().view(np.uint32), u.view(np.uint32)), k
        assert np.array_equal(o["grad"].cpu().numpy().view(np.uint32), g.view(np.uint32)), k


def test_batch_matches_per_scene(gpu, batch_items):
    items, refs = batch_items
    b = gpu.BatchSolver(items)
    outs = b.solve()
    torch.cuda.synchronize()
    _check(outs, refs)
    total = sum(int(it.solver.solve_grad(it.points, it.scene, seed=it.seed, count_steps=True)["steps"][0])
                for it in items)
    assert int(b.steps[0]) == total
    b.close()


def test_batch_graph_replays(gpu, batch_items):
    items, refs = batch_items
    b = gpu.BatchSolver(items, graph=True)
    for rep in range(3):
        outs = b.solve()
        torch.cuda.synchronize()
        _check(outs, refs)
        # eager work between replays, on the same stream
        it = items[rep % len(items)]
        it.solver.solve_grad(it.points, it.scene, seed=it.seed + 1000)
        torch.cuda.synchronize()
    b.close()


def test_batch_3d_matches_per_scene(gpu):
    """3D items (two noisy icospheres of different resolution and noise, and
    different walk counts) in one interleaved persistent launch, eager and
    graph-replayed, against their own solves."""
    from dwost_amd import scenes

    items = []
    for k, (level, n, W) in enumerate(((3, 150, 32), (4, 90, 16))):
        spec = scenes.icosphere_scene(level=level, n_query=n, scene_seed=3 + k)
        sc = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        items.append(gpu.BatchItem(gpu.WoSt3D(eps=1e-3, nwalks=W, nsteps=1024), sc,
                                   torch.as_tensor(spec.queries).cuda(), seed=40 + k))
    refs = [it.solver.solve_grad(it.points, it.scene, seed=it.seed) for it in items]
    torch.cuda.synchronize()
    refs = [(r["u"].cpu().numpy(), r["grad"].cpu().numpy()) for r in refs]
    for graph in (False, True):
        b = gpu.BatchSolver(items, graph=graph)
        for _ in range(2 if graph else 1):
            outs = b.solve()
            torch.cuda.synchronize()
            _check(outs, refs)
        b.close()


def test_batch_rejects_mixed_dims(gpu, batch_items, ico_spec):
    items, _ = batch_items
    sc3 = gpu.Scene3D(ico_spec.vertices, ico_spec.indices, ico_spec.types)
    bad = items[:1] + [gpu.BatchItem(gpu.WoSt3D(), sc3, torch.as_tensor(ico_spec.queries[:8]).cuda())]
    with pytest.raises(ValueError):
        gpu.BatchSolver(bad)


def test_autotune_keeps_results(gpu, holes_spec):
    """Every launch shape the autotuner tries returns the same bits (it checks
    this itself and raises otherwise); the chosen shape is stored on the solver."""
    from dwost_amd import autotune

    sc = gpu.Scene2D(holes_spec.vertices, holes_spec.indices, holes_spec.types,
                     gpu.BoundaryCondition(holes_spec.bc_code))
    sol = gpu.WoSt(eps=1e-4, nwalks=64, nsteps=1024)
    q = torch.as_tensor(holes_spec.queries[:1024]).cuda()
    ref = sol.solve_grad(q, sc, seed=3)["u"].clone()
    rep = autotune.tune(sol, sc, q, seed=3, sample=1024, reps=1, use_cache=False)
    assert (sol.block_size, sol.blocks_per_cu, sol.spatial_order, sol.schedule, sol.node_width,
            sol.walks_per_lane) == rep["config"]
    # 36 launch shapes at node width 2 and one walk pair per lane, then node
    # width 4 at the fastest, then 2 and 4 walk pairs per lane at the fastest
    assert len(rep["table"]) == 39
    assert torch.equal(sol.solve_grad(q, sc, seed=3)["u"].view(torch.int32), ref.view(torch.int32))


def test_batch_points_updated_in_place(gpu, batch_items):
    """The batch reads each item's points by address: new values written in place
    between solves (eager and graph-replayed) give the per-scene results on the
    new points, bit for bit."""
    items, _ = batch_items
    keep = [it.points.clone() for it in items]
    try:
        for graph in (False, True):
            b = gpu.BatchSolver(items, graph=graph)
            b.solve()
            torch.cuda.synchronize()
            for k, it in enumerate(items):  # reversed point order: other values at every address
                it.points.copy_(keep[k].flip(0))
            outs = b.solve()
            torch.cuda.synchronize()
            refs = [it.solver.solve_grad(it.points, it.scene, seed=it.seed) for it in items]
            torch.cuda.synchronize()
            _check(outs, [(r["u"].cpu().numpy(), r["grad"].cpu().numpy()) for r in refs])
            for k, it in enumerate(items):
                it.points.copy_(keep[k])
            b.close()
    finally:
        for k, it in enumerate(items):
            it.points.copy_(keep[k])
        torch.cuda.synchronize()


def test_batch_rejects_converted_points(gpu, batch_items):
    """Points the batch would have to copy (host arrays, (dim, N) layout, float64,
    non-contiguous views) are refused: the copy would never see in-place updates."""
    items, _ = batch_items
    it = items[0]
    p = it.points
    for bad in (p.cpu().numpy(), p.t(), p.double(), torch.cat([p, p], 1)[:, :2]):
        with pytest.raises(ValueError):
            gpu.BatchSolver([gpu.BatchItem(it.solver, it.scene, bad, seed=it.seed)])
