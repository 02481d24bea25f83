"""Streamline tracer (examples/example_2d/trace.py, examples/example_3d/trace.py
with a WoSt / WoSt3D solver).

CPU: the oracle's tracer is pinned by the harmonic field: per-sample
displacements / step estimate grad u without bias (3D Euler step; 2D half step
for interior points whose step stays inside, the reference's corrector term
being zero as written), points outside do not move, and a point's position is
its samples' mean summed in sample order.  GPU: dw_trace_step against the oracle
on the same seeds, over two chained steps."""

import numpy as np
import pytest

from dwost_amd.scenes import analytic_grad


def _interior(os_, spec, n, dmin):
    """The first n queries at least dmin from the boundary."""
    from oracle import oracle as O

    q = spec.queries[: 16 * n]
    d = np.abs(os_.query(0, O.KIND_CLOSEST, q)["d"])
    return np.ascontiguousarray(q[d > dmin][:n])


def _unbiased(samples, q, scale, exact, sigmas=5.0):
    g = (samples - q[:, None, :]) / scale
    mean, sd = g.mean(1), g.std(1, ddof=1)
    err = np.abs(mean - exact)
    lim = sigmas * sd / np.sqrt(g.shape[1]) + 2e-3
    print("max err / limit", float((err / lim).max()))
    return bool((err <= lim).all())


def _ordered_mean(samples):
    acc = np.zeros(samples[:, 0].shape, np.float32)
    for k in range(samples.shape[1]):
        acc = acc + samples[:, k]
    return acc / np.float32(samples.shape[1])


def test_oracle_trace_3d_unbiased(oracle):
    from dwost_amd import scenes

    spec = scenes.icosphere_scene(level=3, n_query=512)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = _interior(os_, spec, 6, 0.2)
    tn, step = 24, 0.05
    r = os_.trace_step(q, spec.bc_code, nwalks=16, samples_per_point=tn, step_size=step, nsubwalks=2, nsteps=1024,
                       eps=1e-3, seed=3)
    smp = r["samples"].reshape(len(q), tn, 3)
    assert _unbiased(smp.astype(np.float64), q.astype(np.float64), step, analytic_grad(spec.bc_code, q))
    assert np.array_equal(r["pos"].view(np.uint32), _ordered_mean(smp).view(np.uint32))


def test_oracle_trace_2d_half_step_and_outside(oracle):
    from dwost_amd import scenes

    spec = scenes.star_scene_2d(n_query=512)
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    q = _interior(os_, spec, 6, 0.25)
    q = np.concatenate([q, np.array([[4.0, 4.0]], np.float32)])  # outside: never moves
    tn, step = 24, 0.02
    r = os_.trace_step(q, spec.bc_code, nwalks=32, samples_per_point=tn, step_size=step, nsteps=1024, eps=1e-3,
                       seed=5)
    smp = r["samples"].reshape(len(q), tn, 2)
    assert np.array_equal(smp[-1], np.repeat(q[-1:], tn, 0))
    assert np.array_equal(r["pos"][-1], q[-1])
    assert _unbiased(smp[:-1].astype(np.float64), q[:-1].astype(np.float64), 0.5 * step,
                     analytic_grad(spec.bc_code, q[:-1]))
    assert np.array_equal(r["pos"].view(np.uint32), _ordered_mean(smp).view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("which", ["star", "ico"])
def test_trace_parity(oracle, which, exact, star_spec, ico_spec):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import dwost_amd as gpu

    spec = dict(star=star_spec, ico=ico_spec)[which]
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    scene = cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    os_ = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    sol = (gpu.WoSt if spec.dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=16, nsteps=1024, exact_cones=exact)
    tr = gpu.StreamlineTracer(scene, sol, step_size=0.02, nwalks=4)
    q = np.ascontiguousarray(spec.queries[:24])
    p = torch.as_tensor(q).cuda()
    o = q
    for it in range(2):
        p = tr.next_position(p, seed=it)
        o = os_.trace_step(o, spec.bc_code, 16, 4, 0.02, nsteps=1024, eps=1e-3, cv=sol.control_variates,
                           av=sol.antithetic_variates, seed=it)["pos"]
        g = p.cpu().numpy()
        same = np.all(g.view(np.uint32) == o.view(np.uint32), axis=1)
        bad = np.flatnonzero(~same)
        print(f"trace step {it} (exact_cones={exact}): {same.sum()} of {same.size} points bit-identical")
        # exact cones: every point; fast cones: a branch flip in one of the
        # 4 x 16 x dim walks behind a point moves it, at most one point allowed
        assert bad.size <= (0 if exact else 1), bad.tolist()
        assert np.allclose(g, o, rtol=1e-3, atol=1e-3)
        o = g  # continue both from the GPU positions
    with pytest.raises(Exception):
        gpu.StreamlineTracer(scene, gpu.WoS(eps=1e-3) if spec.dim == 2 else gpu.WoS3D(eps=1e-3)).next_position(p)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["star", "ico"])
def test_persistent_trace_matches_simple(which, star_spec, ico_spec, monkeypatch):
    """The tracer runs its samples as persistent jobs; with the literal cone test
    both schedules return the bits of the one-sample-per-thread kernel."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from dataclasses import replace

    import dwost_amd as gpu

    spec = dict(star=star_spec, ico=ico_spec)[which]
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    scene = cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    sol = (gpu.WoSt if spec.dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=8, nsubwalks=2, nsteps=1024,
                                                      exact_cones=True)
    q = torch.as_tensor(np.ascontiguousarray(spec.queries[:40])).cuda()
    monkeypatch.setenv("DWOST_KERNEL", "simple")
    ref, ref_steps = gpu.StreamlineTracer(scene, sol, step_size=0.03, nwalks=3).next_position(q, seed=6,
                                                                                            count_steps=True)
    torch.cuda.synchronize()
    monkeypatch.delenv("DWOST_KERNEL")
    for sched in (1, 2):
        tr = gpu.StreamlineTracer(scene, replace(sol, schedule=sched), step_size=0.03, nwalks=3)
        got, steps = tr.next_position(q, seed=6, count_steps=True)
        assert np.array_equal(got.cpu().numpy().view(np.uint32), ref.cpu().numpy().view(np.uint32)), sched
        assert int(steps[0]) == int(ref_steps[0])
