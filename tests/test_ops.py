"""torch.ops.dwost: the C ABI as PyTorch custom operators.  CPU: every operator
is registered with a fake implementation of the right output shapes.  GPU:
each operator returns the bits of the ctypes path it wraps, passes
torch.library.opcheck (schema, fake tensors, autograd registration, AOT
dispatch), and composes with torch.compile (the aot_eager backend) and the
CustomSolver autograd."""

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

OPS = ("closest_point", "intersect", "closest_silhouette", "star_radius", "sdf", "sample_boundary", "solve_u",
       "solve_dude", "solve_fused", "adjoint_u", "adjoint_fused")


def test_ops_registered_with_fake_shapes():
    import dwost_amd

    for name in OPS:
        assert hasattr(torch.ops.dwost, name), name
    ip, fp = dwost_amd.WoSt3D(nwalks=7)._op_params(3)
    p = torch.empty(5, 3, device="meta")
    bc = (0, [0.0] * 4, None, None, [0.0] * 5)
    u, g, fl, st = torch.ops.dwost.solve_fused(0, p, ip, fp, ip, fp, *bc)
    assert (u.shape, g.shape, fl.shape, fl.dtype, st.shape) == ((5,), (5, 3), (20,), torch.uint8, (1,))
    u, fl, st = torch.ops.dwost.solve_u(0, p, ip, fp, ip, fp, *bc)
    assert (u.shape, fl.shape) == ((5,), (35,))
    assert [t.shape for t in torch.ops.dwost.closest_point(0, 0, p)] == [(5,), (5,), (5, 3), (5, 3), (5, 2)]
    assert torch.ops.dwost.sample_boundary(0, 2, 3, torch.empty(4, 3, device="meta"))[0].shape == (4, 3)
    gv = torch.empty(9, device="meta")
    dg, dh = torch.ops.dwost.adjoint_fused(0, p, ip, fp, 3, [0.0] * 4, gv, gv, [0.0] * 5, torch.empty(5, device="meta"),
                                           torch.empty(5, 3, device="meta"), fl)
    assert dg.shape == (9,) and dh.shape == (9,)


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import dwost_amd

    dwost_amd._lib.lib()
    return dwost_amd


def _bits(t):
    a = t.detach().cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _same(a, b):
    return np.array_equal(_bits(a), _bits(b))


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_ops_match_ctypes_path(dim, star_spec, ico_spec):
    gpu = _gpu()
    spec = star_spec if dim == 2 else ico_spec
    cls = gpu.Scene2D if dim == 2 else gpu.Scene3D
    sc = cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition(spec.bc_code))
    H = sc.handle.value
    q = torch.as_tensor(spec.queries[:200]).cuda()
    rng = np.random.default_rng(2)
    v = rng.standard_normal((200, dim))
    v = torch.as_tensor((v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)).cuda()
    r = torch.full((200,), 0.3, device="cuda")
    a = torch.ops.dwost.closest_point(H, 0, q)
    b = sc.closest_point(q)
    for x, k in zip(a, ("prim", "d", "p", "n", "tuv")):
        assert _same(x, b[k]), k
    a = torch.ops.dwost.intersect(H, 0, q, v, r, 0)
    b = sc.intersect(q, v, r)
    for x, k in zip(a, ("prim", "d", "p", "n", "tuv")):
        assert _same(x, b[k]), k
    a = torch.ops.dwost.closest_silhouette(H, 0, q, r)
    b = sc.closest_silhouette(q, r)
    assert _same(a[0], b["sil"]) and _same(a[1], b["d"])
    assert _same(torch.ops.dwost.star_radius(H, 0, q, r), sc.star_radius(q, r))
    assert _same(torch.ops.dwost.star_radius_2(H, 0, q, v, 1.2), sc.star_radius_2(q, v, rho_max=1.2))
    assert _same(torch.ops.dwost.sdf(H, q), sc.sdf(q))
    u = torch.as_tensor(rng.random((300,) if dim == 2 else (300, 3), dtype=np.float32)).cuda()
    a = torch.ops.dwost.sample_boundary(H, 0, dim, u)
    b = sc.sample_boundary(u)
    for x, k in zip(a, ("p", "n", "tuv", "pdf", "prim")):
        assert _same(x, b[k]), k
    sol = (gpu.WoSt if dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=16, nsteps=1024)
    ip, fp = sol._op_params(5)
    bc = sc.boundary.op_args()
    uu, fl, st = torch.ops.dwost.solve_u(H, q, ip, fp, ip, fp, *bc)
    flags = torch.empty(200 * 16, dtype=torch.uint8, device="cuda")
    ref = sol.solve(q, sc, seed=5, walk_flags=flags, count_steps=True)
    assert _same(uu, ref["u"]) and _same(fl, flags) and int(st[0]) == int(ref["steps"][0])
    e = v
    assert _same(torch.ops.dwost.solve_dude(H, q, e, ip, fp, *bc), sol.solve_dude(q, e, sc, seed=5))
    uf, gf, pf, st = torch.ops.dwost.solve_fused(H, q, ip, fp, ip, fp, *bc)
    pflags = torch.empty(200 * 8, dtype=torch.uint8, device="cuda")
    ref = sol.solve_grad(q, sc, seed=5, pair_flags=pflags, count_steps=True)
    assert _same(uf, ref["u"]) and _same(gf, ref["grad"]) and _same(pf, pflags)
    assert int(st[0]) == int(ref["steps"][0])


def _adjoint_case(gpu, spec, n=24):
    from dwost_amd import scenes

    rng = np.random.default_rng(4)
    V = len(spec.vertices)
    g = torch.as_tensor((scenes.analytic_u(0, spec.vertices.astype(np.float64)) + 0.1 * rng.standard_normal(V))
                        .astype(np.float32)).cuda()
    h = torch.as_tensor((0.2 * rng.standard_normal(V)).astype(np.float32)).cuda()
    cls = gpu.Scene2D if spec.dim == 2 else gpu.Scene3D
    sc = cls(spec.vertices, spec.indices, spec.types, gpu.BoundaryCondition.per_vertex(g, h))
    return sc, g, h, torch.as_tensor(spec.queries[:n]).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_adjoint_ops_match_ctypes_path(dim, star_spec, ico_spec):
    gpu = _gpu()
    L = gpu._lib
    spec = star_spec if dim == 2 else ico_spec
    sc, g, h, q = _adjoint_case(gpu, spec)
    n = q.shape[0]
    sol = (gpu.WoSt if dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=16, nsteps=1024)
    ip, fp = sol._op_params(7)
    bc = sc.boundary.op_args()
    wu = torch.randn(n, device="cuda")
    wg = torch.randn(n, dim, device="cuda")
    _, _, pf, _ = torch.ops.dwost.solve_fused(sc.handle.value, q, ip, fp, ip, fp, *bc)
    dg, dh = torch.ops.dwost.adjoint_fused(sc.handle.value, q, ip, fp, *bc[:2], g, h, bc[4], wu, wg, pf)
    rg, rh = torch.zeros_like(g), torch.zeros_like(h)
    prm, cb = sol._params(7), sc.boundary.to_c(sc.n_vertices)
    L.check(L.lib().dw_adjoint_fused(sc.handle, L.ptr(q), n, C.byref(prm), C.byref(cb), L.ptr(wu), L.ptr(wg),
                                     L.ptr(pf), L.ptr(rg), L.ptr(rh), None))
    # atomic scatter-adds: the same sums in another order
    assert torch.allclose(dg, rg, rtol=1e-5, atol=1e-6) and torch.allclose(dh, rh, rtol=1e-5, atol=1e-6)
    _, fl, _ = torch.ops.dwost.solve_u(sc.handle.value, q, ip, fp, ip, fp, *bc)
    dg, dh = torch.ops.dwost.adjoint_u(sc.handle.value, q, ip, fp, *bc[:2], g, h, bc[4], wu, fl)
    rg.zero_(), rh.zero_()
    L.check(L.lib().dw_adjoint_u(sc.handle, L.ptr(q), n, C.byref(prm), C.byref(cb), L.ptr(wu), L.ptr(fl),
                                 L.ptr(rg), L.ptr(rh), None))
    assert torch.allclose(dg, rg, rtol=1e-5, atol=1e-6) and torch.allclose(dh, rh, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_opcheck(star_spec, ico_spec):
    gpu = _gpu()
    for spec in (star_spec, ico_spec):
        dim = spec.dim
        sc, g, h, q = _adjoint_case(gpu, spec, 16)
        H = sc.handle.value
        sol = (gpu.WoSt if dim == 2 else gpu.WoSt3D)(eps=1e-3, nwalks=8, nsteps=1024)
        ip, fp = sol._op_params(1)
        bc = sc.boundary.op_args()
        gq, hq = g.clone().requires_grad_(True), h.clone().requires_grad_(True)
        r = torch.full((16,), 0.3, device="cuda")
        v = torch.nn.functional.normalize(torch.randn(16, dim, device="cuda"), dim=1)
        cases = [
            (torch.ops.dwost.closest_point.default, (H, 0, q)),
            (torch.ops.dwost.intersect.default, (H, 0, q, v, r, 0)),
            (torch.ops.dwost.closest_silhouette.default, (H, 0, q, r)),
            (torch.ops.dwost.star_radius.default, (H, 0, q, r)),
            (torch.ops.dwost.star_radius_2.default, (H, 0, q, v, 1.2)),
            (torch.ops.dwost.sdf.default, (H, q)),
            (torch.ops.dwost.sample_boundary.default,
             (H, 0, dim, torch.rand(16, device="cuda") if dim == 2 else torch.rand(16, 3, device="cuda"))),
            (torch.ops.dwost.solve_dude.default, (H, q, v, ip, fp, *bc)),
            (torch.ops.dwost.solve_u.default, (H, q, ip, fp, ip, fp, bc[0], bc[1], gq, hq, bc[4])),
            (torch.ops.dwost.solve_fused.default, (H, q, ip, fp, ip, fp, bc[0], bc[1], gq, hq, bc[4])),
        ]
        for op, args in cases:
            torch.library.opcheck(op, args)


@pytest.mark.gpu
def test_compile_and_custom_solver_through_ops(ico_spec):
    """torch.compile (aot_eager, full graph) of a loss over solve_fused equals
    eager, values and gradients; CustomSolver's gradients are the adjoint op's."""
    gpu = _gpu()
    sc, g, h, q = _adjoint_case(gpu, ico_spec, 32)
    sol = gpu.WoSt3D(eps=1e-3, nwalks=16, nsteps=1024)
    ip, fp = sol._op_params(11)
    H = sc.handle.value
    code, coeffs, _, _, src = sc.boundary.op_args()

    def loss(gv, hv):
        u, gr, _, _ = torch.ops.dwost.solve_fused(H, q, ip, fp, ip, fp, code, coeffs, gv, hv, src)
        return (u ** 2).sum() + (gr ** 2).sum()

    ge, he = g.clone().requires_grad_(True), h.clone().requires_grad_(True)
    le = loss(ge, he)
    le.backward()
    gc, hc = g.clone().requires_grad_(True), h.clone().requires_grad_(True)
    lc = torch.compile(loss, backend="aot_eager", fullgraph=True)(gc, hc)
    lc.backward()
    assert _same(le, lc)
    assert torch.allclose(ge.grad, gc.grad, rtol=1e-5, atol=1e-6)
    assert torch.allclose(he.grad, hc.grad, rtol=1e-5, atol=1e-6)
    # CustomSolver runs the same operator
    gs, hs = g.clone().requires_grad_(True), h.clone().requires_grad_(True)
    sc.boundary = gpu.BoundaryCondition.per_vertex(gs, hs)
    u, gr = gpu.CustomSolver(sol).solve_grad(q, sc, seed=11)
    ((u ** 2).sum() + (gr ** 2).sum()).backward()
    assert torch.allclose(gs.grad, ge.grad, rtol=1e-5, atol=1e-6)
