"""dw_solve_fused_noalloc: the fused u + grad u solve without any allocation
(its scratch is one static device array per GPU), called through the C ABI.
Two calls in a row, the second on fewer points (with a point_base), must return
the bits of dw_solve_fused with the same arguments, in 2D (with the spatial
order of the points, which lives in the scratch too) and 3D; a call whose
scratch would exceed the fixed size fails before any launch."""

import ctypes as C

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


def _solve(gpu, fn, scene, sol, q, seed, point_base):
    L = gpu._lib
    n, D = q.shape
    u = torch.full((n,), float("nan"), device="cuda")
    g = torch.full((n, D), float("nan"), device="cuda")
    steps = torch.zeros(1, dtype=torch.int64, device="cuda")
    prm = sol._params(seed, point_base)
    bc = scene.boundary.to_c(scene.n_vertices)
    st = L.stream_handle()
    if fn == "noalloc":
        rc = L.lib().dw_solve_fused_noalloc(scene.handle, L.ptr(q), n, C.byref(prm), C.byref(bc), L.ptr(u), L.ptr(g),
                                            L.ptr(steps), st)
    else:
        rc = L.lib().dw_solve_fused(scene.handle, L.ptr(q), n, C.byref(prm), C.byref(bc), L.ptr(u), L.ptr(g), None,
                                    None, None, L.ptr(steps), st)
    L.check(rc)
    torch.cuda.synchronize()
    return u.cpu().numpy(), g.cpu().numpy(), int(steps[0])


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_matches_fused(gpu, dim):
    from dwost_amd import scenes

    if dim == 2:
        spec = scenes.holes_scene_2d(n_query=9000, eps=1e-3)
        scene = gpu.Scene2D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoSt(eps=1e-3, nwalks=64, nsteps=1024)
        sizes = ((0, 9000), (1234, 1534))  # 9,000 points use the spatial order
    else:
        spec = scenes.icosphere_scene(level=3, n_query=2000)
        scene = gpu.Scene3D(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
        sol = gpu.WoSt3D(eps=1e-3, nwalks=32, nsteps=1024)
        sizes = ((0, 2000), (77, 177))
    qa = torch.as_tensor(spec.queries).cuda()
    for seed, (lo, hi) in zip((3, 4), sizes):  # second call: fewer points, other point_base
        q = qa[lo:hi].contiguous()
        ref = _solve(gpu, "fused", scene, sol, q, seed, lo)
        got = _solve(gpu, "noalloc", scene, sol, q, seed, lo)
        assert np.array_equal(_bits(got[0]), _bits(ref[0])), (dim, lo, "u")
        assert np.array_equal(_bits(got[1]), _bits(ref[1])), (dim, lo, "grad")
        assert got[2] == ref[2] > 0
        print(f"{dim}D: {hi - lo} points x {sol.nwalks} walks identical, {got[2]} walk steps")


def test_noalloc_rejects_oversize(gpu, star_spec):
    L = gpu._lib
    scene = gpu.Scene2D(star_spec.vertices, star_spec.indices, star_spec.types, gpu.BoundaryCondition.exp_cos())
    sol = gpu.WoSt(eps=1e-3, nwalks=1024, nsteps=1024)
    q = torch.as_tensor(star_spec.queries[:8]).cuda()
    u = torch.empty(8, device="cuda")
    g = torch.empty(8, 2, device="cuda")
    prm = sol._params(0)
    bc = scene.boundary.to_c(scene.n_vertices)
    n = L.NOALLOC_SCRATCH_BYTES // (512 * 16) + 1  # more (point, pair) slots than the scratch holds
    rc = L.lib().dw_solve_fused_noalloc(scene.handle, L.ptr(q), n, C.byref(prm), C.byref(bc), L.ptr(u), L.ptr(g),
                                        None, L.stream_handle())
    assert rc == -1
    assert b"scratch" in L.lib().dw_last_error()
    torch.cuda.synchronize()
