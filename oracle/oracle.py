"""ctypes wrapper of the CPU restatement (oracle/wost_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (dwost_amd/) never
imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

KIND_CLOSEST, KIND_RAY, KIND_SIL, KIND_STAR = 0, 1, 2, 3
KIND_BRUTE_CLOSEST, KIND_BRUTE_RAY, KIND_BRUTE_SIL = 10, 11, 12
SUB_FULL, SUB_DIRICHLET, SUB_NEUMANN = 0, 1, 2


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        L = _lib
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_sub_info.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_sub_export.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 5
        L.orc_query.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64] + [
            C.c_void_p] * 5
        L.orc_ref_topology.restype = C.c_int32
        L.orc_ref_topology.argtypes = [C.c_void_p, C.c_int, C.c_int32]
        common = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_solve_u.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + common + [C.c_void_p] * 5
        L.orc_solve_dude.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64] + common + [C.c_void_p] * 2
        L.orc_solve_fused.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + common + [C.c_void_p] * 7
        L.orc_trace_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + common + [C.c_int32, C.c_float] + [
            C.c_void_p] * 3
        L.orc_tea.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_pcg_seq.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]
        L.orc_pcg_lane_floats.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_void_p]
        L.orc_math.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_num_threads.restype = C.c_int
        L.orc_cv_bound_check.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.orc_far_check.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_sdf.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_closest_iters.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_step_bound_check.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                           C.c_void_p]
        L.orc_grad_trace_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                          C.c_float] + [C.c_void_p] * 3
        L.orc_star_radius2_rho.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                                           C.c_void_p, C.c_void_p]
        L.orc_grad_dude.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_void_p] * 4
        L.orc_sample_boundary.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64] + [C.c_void_p] * 5
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _src4(src_c):
    """The source's coefficients: a constant (SRC_CONST) or (s0, s1, s2, s3) of
    the linear f = s0 + s1 x + s2 y + s3 z (SRC_LINEAR)."""
    v = np.zeros(4, np.float32)
    a = np.atleast_1d(np.asarray(src_c, np.float32))
    v[: len(a)] = a
    return v


def tea(v0: int, v1: int, rounds: int = 4, key=None):
    out = np.zeros(2, np.uint32)
    k = None if key is None else np.ascontiguousarray(key, dtype=np.uint32)
    lib().orc_tea(v0, v1, rounds, _p(k), _p(out))
    return int(out[0]), int(out[1])


def pcg_sequence(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_pcg_seq(initstate, initseq, n, _p(out))
    return out


def pcg_lane_floats(seed: int, point: int, lane: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().orc_pcg_lane_floats(seed, point, lane, n, _p(out))
    return out


MATH_FUNCS = {"log": 0, "exp": 1, "sin": 2, "cos": 3, "acos": 4, "asin": 5, "atan2": 6, "cbrt": 7}


def math_fn(name: str, x, y=None) -> np.ndarray:
    x = _f32(x)
    y = _f32(np.zeros_like(x) if y is None else y)
    out = np.zeros_like(x)
    lib().orc_math(MATH_FUNCS[name], _p(x), _p(y), x.size, _p(out))
    return out


class OracleScene:
    """CPU restatement of Scene2D / Scene3D with its sub-scenes."""

    def __init__(self, vertices, indices, types):
        self.V = _f32(vertices)
        self.I = np.ascontiguousarray(indices, dtype=np.int32)
        self.T = np.ascontiguousarray(types, dtype=np.int32)
        self.dim = self.V.shape[1]
        self.h = lib().orc_scene_create(self.dim, _p(self.V), len(self.V), _p(self.I), _p(self.T), len(self.I))

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_destroy(self.h)
            self.h = None

    def info(self, sub: int) -> dict:
        out = np.zeros(8, np.int64)
        lib().orc_sub_info(self.h, sub, _p(out))
        keys = ["present", "nverts", "nprims", "nnodes", "nrefs", "max_depth", "py_sils", "nsil_topology"]
        return {k: int(v) for k, v in zip(keys, out)}

    def export(self, sub: int) -> dict:
        inf = self.info(sub)
        D = self.dim
        nn, npr, nr = inf["nnodes"], inf["nprims"], inf["nrefs"]
        box = np.zeros((nn, 2 * D), np.float32)
        ints = np.zeros((nn, 4), np.int32)
        cone = np.zeros((nn, D + 2), np.float32)
        pidx = np.zeros(npr, np.int32)
        refs = np.zeros(max(nr, 1), np.int32)
        lib().orc_sub_export(self.h, sub, _p(box), _p(ints), _p(cone), _p(pidx), _p(refs))
        return dict(box=box, ints=ints, cone=cone, prim_index=pidx, refs=refs[:nr], **inf)

    def query(self, sub: int, kind: int, pts, dirs=None, rmax=None):
        pts = _f32(pts)
        n = len(pts)
        D = self.dim
        dirs = _f32(np.zeros_like(pts) if dirs is None else dirs)
        rmax = _f32(np.full(n, np.inf) if rmax is None else rmax)
        prim = np.zeros(n, np.int32)
        d = np.zeros(n, np.float32)
        p = np.zeros((n, D), np.float32)
        nrm = np.zeros((n, D), np.float32)
        tuv = np.zeros((n, 2), np.float32)
        lib().orc_query(self.h, sub, kind, _p(pts), _p(dirs), _p(rmax), n, _p(prim), _p(d), _p(p), _p(nrm), _p(tuv))
        return dict(prim=prim, d=d, p=p, n=nrm, tuv=tuv)

    def closest_iters(self, sub: int, pts) -> np.ndarray:
        """Per closest-point query on sub-scene `sub`: (iterations at node width
        2, at width 4, interior-record fetches at width 2, at width 4) of the
        GPU loop, replayed from the reference traversal."""
        pts = _f32(pts)
        out = np.zeros((len(pts), 4), np.int64)
        lib().orc_closest_iters(self.h, sub, _p(pts), len(pts), _p(out))
        return out

    def step_bound_check(self, prev, nxt, R):
        """The Dirichlet closest point at nxt queried with its pruning radius
        started at the bound the walk derives from the step prev -> nxt of
        radius R (|nxt - prev| <= R + 1e-5), against the unbounded query.
        Returns (per-point bitwise agreement, nodes visited unbounded, bounded)."""
        prev, nxt, R = _f32(prev), _f32(nxt), _f32(R)
        same = np.zeros(len(prev), np.int32)
        nodes = np.zeros(2, np.uint64)
        lib().orc_step_bound_check(self.h, _p(prev), _p(nxt), _p(R), len(prev), _p(same), _p(nodes))
        return same.astype(bool), int(nodes[0]), int(nodes[1])

    def cv_bound_check(self, pts):
        """2D: the full-scene closest point queried with its pruning radius
        started just above the Dirichlet distance, against the unbounded query.
        Returns (per-point bitwise agreement, nodes visited unbounded, bounded)."""
        pts = _f32(pts)
        same = np.zeros(len(pts), np.int32)
        nodes = np.zeros(2, np.uint64)
        lib().orc_cv_bound_check(self.h, _p(pts), len(pts), _p(same), _p(nodes))
        return same.astype(bool), int(nodes[0]), int(nodes[1])

    def sdf(self, pts):
        pts = _f32(pts)
        out = np.zeros(len(pts), np.float32)
        lib().orc_sdf(self.h, _p(pts), len(pts), _p(out))
        return out

    def sample_boundary(self, sub: int, u):
        """Polyline / Mesh sample_boundary with the sampler's draws given
        (2D: u[n]; 3D: u[n, 3])."""
        u = _f32(u)
        n, D = len(u), self.dim
        out = dict(p=np.zeros((n, D), np.float32), n=np.zeros((n, D), np.float32), tuv=np.zeros((n, 2), np.float32),
                   pdf=np.zeros(n, np.float32), prim=np.zeros(n, np.int32))
        lib().orc_sample_boundary(self.h, sub, _p(u), n, _p(out["p"]), _p(out["n"]), _p(out["tuv"]), _p(out["pdf"]),
                                  _p(out["prim"]))
        return out

    def far_check(self, pts):
        """3D: the distance bounds the walk kernels use far from the scene,
        checked against the reference queries (see orc_far_check)."""
        pts = _f32(pts)
        out = np.zeros((len(pts), 4), np.int32)
        lib().orc_far_check(self.h, _p(pts), len(pts), _p(out))
        return out

    def ref_topology(self, sub: int, ref: int) -> int:
        return lib().orc_ref_topology(self.h, sub, ref)

    @staticmethod
    def _params(nwalks, nsteps, nsubwalks, cv, av, seed, eps, min_R, base=0, wos=False):
        ip = np.array([nwalks, nsteps, nsubwalks, int(cv), int(av), np.uint32(seed).astype(np.int32), int(base),
                       int(wos)], np.int32)
        fp = np.array([eps, min_R], np.float32)
        return ip, fp

    @staticmethod
    def _bc(bc_code, coeffs, gv, hv):
        c = _f32(np.asarray(coeffs if coeffs is not None else (0, 0, 0, 0)))
        gv = None if gv is None else _f32(gv)
        hv = None if hv is None else _f32(hv)
        return c, gv, hv

    def solve_u(self, pts, bc_code, nwalks, nsteps=128, eps=1e-3, min_R=1e-3, cv=True, av=True, seed=0,
                nsubwalks=1, coeffs=None, gv=None, hv=None, src_code=0, src_c=0.0, per_walk=False,
                count_bytes=False, base=0, wos=False):
        pts = _f32(pts)
        n = len(pts)
        ip, fp = self._params(nwalks, nsteps, nsubwalks, cv, av, seed, eps, min_R, base, wos)
        c, gv, hv = self._bc(bc_code, coeffs, gv, hv)
        mean = np.zeros(n, np.float32)
        wv = np.zeros(n * nwalks, np.float32) if per_walk else None
        ws = np.zeros(n * nwalks, np.int32) if per_walk else None
        wt = np.zeros(n * nwalks, np.int32) if per_walk else None
        bt = np.zeros(1, np.uint64) if count_bytes else None
        lib().orc_solve_u(self.h, _p(pts), n, _p(ip), _p(fp), bc_code, _p(c), _p(gv), _p(hv), src_code, _p(_src4(src_c)),
                          _p(mean), _p(wv), _p(ws), _p(wt), _p(bt))
        out = dict(u=mean)
        if per_walk:
            out.update(walk_u=wv, walk_steps=ws, walk_term=wt)
        if count_bytes:
            out["bytes"] = int(bt[0])
        return out

    def solve_dude(self, pts, e, bc_code, nwalks, nsteps=128, eps=1e-3, min_R=1e-3, cv=True, av=True, seed=0,
                   coeffs=None, gv=None, hv=None, src_code=0, src_c=0.0, base=0, wos=False):
        pts = _f32(pts)
        e = _f32(e)
        n = len(pts)
        ip, fp = self._params(nwalks, nsteps, 1, cv, av, seed, eps, min_R, base, wos)
        c, gv, hv = self._bc(bc_code, coeffs, gv, hv)
        out = np.zeros(n, np.float32)
        st = np.zeros(1, np.uint64)
        lib().orc_solve_dude(self.h, _p(pts), _p(e), n, _p(ip), _p(fp), bc_code, _p(c), _p(gv), _p(hv), src_code,
                             _p(_src4(src_c)), _p(out), _p(st))
        return dict(dude=out, steps=int(st[0]))

    def star_radius2_rho(self, sub: int, pts, e, rho_max=1.2):
        """(through the SNCH, by the loop over all segments) of Polyline.star_radius_2."""
        pts, e = _f32(pts), _f32(e)
        t = np.zeros(len(pts), np.float32)
        b = np.zeros(len(pts), np.float32)
        lib().orc_star_radius2_rho(self.h, sub, _p(pts), _p(e), len(pts), rho_max, _p(t), _p(b))
        return t, b

    def grad_dude(self, pts, e, bc_code, nwalks, nsteps=128, eps=1e-3, min_R=1e-2, cv=True, seed=0, nsubwalks=1,
                  rho_max=1.2, ignore_boundary=False, ignore_silhouette=False, ignore_source=False, coeffs=None,
                  src_code=0, src_c=0.0, per_walk=False, base=0, plant=0):
        """WoStGrad.solve_dude (wost_grad.py, solver.py:113-166), 2D scenes."""
        pts, e = _f32(pts), _f32(e)
        n = len(pts)
        ip, fp = self._params(nwalks, nsteps, nsubwalks, cv, True, seed, eps, min_R, base)
        c = _f32(np.asarray(coeffs if coeffs is not None else (0, 0, 0, 0)))
        gp = np.array([rho_max], np.float32)
        gi = np.array([int(ignore_boundary), int(ignore_silhouette), int(ignore_source), int(plant)], np.int32)
        out = np.zeros(n, np.float32)
        wv = np.zeros(n * nwalks, np.float32) if per_walk else None
        ws = np.zeros(n * nwalks, np.int32) if per_walk else None
        st = np.zeros(1, np.uint64)
        lib().orc_grad_dude(self.h, _p(pts), _p(e), n, _p(ip), _p(fp), bc_code, _p(c), src_code, _p(_src4(src_c)), _p(gp),
                            _p(gi), _p(out), _p(wv), _p(ws), _p(st))
        r = dict(dude=out, steps=int(st[0]))
        if per_walk:
            r.update(walk_dude=wv, walk_steps=ws)
        return r

    def grad_trace_step(self, pts, bc_code, samples_per_point, step_size, nsubwalks=1, nsteps=128, eps=1e-3,
                        min_R=1e-2, cv=True, seed=0, rho_max=1.2, coeffs=None, src_code=0, src_c=0.0, base=0):
        """StreamlineTracer.next_position with a WoStGrad solver (2D)."""
        pts = _f32(pts)
        n = len(pts)
        ip, fp = self._params(1, nsteps, nsubwalks, cv, True, seed, eps, min_R, base)
        c = _f32(np.asarray(coeffs if coeffs is not None else (0, 0, 0, 0)))
        gp = np.array([rho_max], np.float32)
        gi = np.zeros(4, np.int32)
        out = np.zeros((n, 2), np.float32)
        smp = np.zeros((n * samples_per_point, 2), np.float32)
        st = np.zeros(1, np.uint64)
        lib().orc_grad_trace_step(self.h, _p(pts), n, _p(ip), _p(fp), bc_code, _p(c), src_code, _p(_src4(src_c)), _p(gp),
                                  _p(gi), int(samples_per_point), float(step_size), _p(out), _p(smp), _p(st))
        return dict(pos=out, samples=smp, steps=int(st[0]))

    def trace_step(self, pts, bc_code, nwalks, samples_per_point, step_size, nsubwalks=1, nsteps=128, eps=1e-3,
                   min_R=1e-3, cv=True, av=True, seed=0, coeffs=None, gv=None, hv=None, src_code=0, src_c=0.0, base=0):
        """StreamlineTracer.next_position with a WoSt solver of nwalks walks per
        dude and nsubwalks dudes per walk_dude; returns the mean new positions
        and the per-sample ones."""
        pts = _f32(pts)
        n, D = len(pts), self.dim
        ip, fp = self._params(nwalks, nsteps, nsubwalks, cv, av, seed, eps, min_R, base)
        c, gv, hv = self._bc(bc_code, coeffs, gv, hv)
        out = np.zeros((n, D), np.float32)
        smp = np.zeros((n * samples_per_point, D), np.float32)
        st = np.zeros(1, np.uint64)
        lib().orc_trace_step(self.h, _p(pts), n, _p(ip), _p(fp), bc_code, _p(c), _p(gv), _p(hv), src_code, _p(_src4(src_c)),
                             int(samples_per_point), float(step_size), _p(out), _p(smp), _p(st))
        return dict(pos=out, samples=smp, steps=int(st[0]))

    def solve_fused(self, pts, bc_code, nwalks, nsteps=128, eps=1e-3, min_R=1e-3, seed=0, coeffs=None, gv=None,
                    hv=None, src_code=0, src_c=0.0, per_pair=False, cv=True, base=0, count=False):
        """count=True adds the algorithmic counters of the walks (reference
        binary-tree order, reference record sizes): bytes, flops, nodes,
        prims, sils, ff (steps from a NaN position, which fetch nothing), and
        for the closest-point queries (closest_queries) the GPU loop's
        iterations and interior-record fetches at node width 2 and 4 (iters2,
        iters4, fetch2, fetch4)."""
        pts = _f32(pts)
        n = len(pts)
        D = self.dim
        ip, fp = self._params(nwalks, nsteps, 1, cv, True, seed, eps, min_R, base)
        c, gv, hv = self._bc(bc_code, coeffs, gv, hv)
        npairs = (nwalks + 1) // 2
        u = np.zeros(n, np.float32)
        g = np.zeros((n, D), np.float32)
        pu = np.zeros(n * npairs, np.float32) if per_pair else None
        pg = np.zeros((n * npairs, D), np.float32) if per_pair else None
        ps = np.zeros(n * npairs, np.int32) if per_pair else None
        st = np.zeros(1, np.uint64)
        cnt = np.zeros(11, np.uint64) if count else None
        lib().orc_solve_fused(self.h, _p(pts), n, _p(ip), _p(fp), bc_code, _p(c), _p(gv), _p(hv), src_code, _p(_src4(src_c)),
                              _p(u), _p(g), _p(pu), _p(pg), _p(ps), _p(st), _p(cnt))
        out = dict(u=u, grad=g, steps=int(st[0]))
        if count:
            keys = ("bytes", "flops", "nodes", "prims", "sils", "ff", "closest_queries", "iters2", "iters4", "fetch2",
                    "fetch4")
            out.update({k: int(v) for k, v in zip(keys, cnt)})
        if per_pair:
            out.update(pair_u=pu, pair_g=pg, pair_steps=ps)
        return out


def num_threads() -> int:
    return lib().orc_num_threads()
