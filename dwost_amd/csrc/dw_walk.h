// dw_walk.h -- boundary data, sampling and the Walk-on-Stars estimators.
#pragma once
#include "dw_device.h"

namespace dwd {

// --------------------------------------------------------------------------
// Boundary data (diff_wost/shapes/polyline.py:52-58, mesh.py:59-63): g = u(p),
// h = n . grad u(p) for an analytic harmonic u, or per-vertex g/h buffers
// interpolated linearly (2D: lerp by t; 3D: barycentric (u, v) as in
// math.py:306-307).
// --------------------------------------------------------------------------
DWD float bc_u(const BCDev& bc, F3 p) {
    if (bc.code == DW_BC_EXPCOS) return cr_exp(p.x) * cr_cos(p.y);
    if (bc.code == DW_BC_QUAD) return p.x * p.x - p.y * p.y;
    return ((bc.c[0] + bc.c[1] * p.x) + bc.c[2] * p.y) + bc.c[3] * p.z;
}
DWD F3 bc_grad(const BCDev& bc, F3 p) {
    if (bc.code == DW_BC_EXPCOS) {
        float ex = cr_exp(p.x);
        return f3(ex * cr_cos(p.y), -ex * cr_sin(p.y), 0.0f);
    }
    if (bc.code == DW_BC_QUAD) return f3(2.0f * p.x, -2.0f * p.y, 0.0f);
    return f3(bc.c[1], bc.c[2], bc.c[3]);
}
// the source f of the Poisson problem (the reference's scene.f / scene.dfdx):
// zero, a constant, or linear f = s0 + s1 x + s2 y + s3 z (2D: z = 0)
DWD float src_f(const BCDev& bc, F3 p) {
    if (bc.src_code == DW_SRC_CONST) return bc.src_c;
    if (bc.src_code == DW_SRC_LINEAR) return ((bc.sc[0] + bc.sc[1] * p.x) + bc.sc[2] * p.y) + bc.sc[3] * p.z;
    return 0.0f;
}
DWD float src_f(const BCDev& bc, F2 p) { return src_f(bc, f3(p.x, p.y, 0.0f)); }
DWD F3 src_grad(const BCDev& bc) {
    if (bc.src_code == DW_SRC_LINEAR) return f3(bc.sc[1], bc.sc[2], bc.sc[3]);
    return f3(0.0f, 0.0f, 0.0f);
}

DWD float g2(const BCDev& bc, const SubDev& S, const CP& c) {
    if (bc.code != DW_BC_PER_VERTEX) return bc_u(bc, c.p);
    return lerpf(bc.gv[S.local_vglobal[2 * c.prim]], bc.gv[S.local_vglobal[2 * c.prim + 1]], c.t);
}
DWD float g3(const BCDev& bc, const SubDev& S, const CP& c) {
    if (bc.code != DW_BC_PER_VERTEX) return bc_u(bc, c.p);
    float ga = bc.gv[S.local_vglobal[3 * c.prim]], gb = bc.gv[S.local_vglobal[3 * c.prim + 1]],
          gc = bc.gv[S.local_vglobal[3 * c.prim + 2]];
    return (gb * c.u + gc * c.v) + ga * ((1.0f - c.u) - c.v);
}
DWD float h2(const BCDev& bc, const SubDev& S, int prim, float t, F2 p, F2 n) {
    if (bc.code != DW_BC_PER_VERTEX) {
        F3 g = bc_grad(bc, f3(p.x, p.y, 0.0f));
        return dot(n, f2(g.x, g.y));
    }
    return lerpf(bc.hv[S.local_vglobal[2 * prim]], bc.hv[S.local_vglobal[2 * prim + 1]], t);
}
DWD float h3(const BCDev& bc, const SubDev& S, int prim, float u, float v, F3 p, F3 n) {
    if (bc.code != DW_BC_PER_VERTEX) return dot(n, bc_grad(bc, p));
    float ha = bc.hv[S.local_vglobal[3 * prim]], hb = bc.hv[S.local_vglobal[3 * prim + 1]],
          hc = bc.hv[S.local_vglobal[3 * prim + 2]];
    return (hb * u + hc * v) + ha * ((1.0f - u) - v);
}

// Random stream of walk (or walk pair) k of point i: TEA(seed, global point
// index) and Dr.Jit PCG32 lane = global point index * per + k (solver.py:55-63),
// so a shard with point_base b draws exactly what one launch over all points does.
DWD void seed_stream(Rng& rng, const ParamsDev& P, int64_t i, int64_t per, int64_t k) {
    int64_t gi = P.point_base + i;
    uint32_t a, b;
    tea4(P.seed, (uint32_t)gi, a, b);
    rng.seed_lane(a, b, (uint64_t)(gi * per + k));
}

// Adjoint of the per-vertex boundary data: scatter w * d(value)/d(vertex value).
// g2/h2: lerp(a, b, t) = a (1 - t) + b t;  g3/h3: b u + c v + a (1 - u - v).
struct Adj {
    float w;    // upstream dL/d(walk value)
    float* dg;  // dL/dg per global vertex (accumulated)
    float* dh;  // dL/dh per global vertex
};
DWD void adj_lin2(float* buf, const SubDev& S, int prim, float t, float w) {
    atomicAdd(buf + S.local_vglobal[2 * prim], w * (1.0f - t));
    atomicAdd(buf + S.local_vglobal[2 * prim + 1], w * t);
}
DWD void adj_lin3(float* buf, const SubDev& S, int prim, float u, float v, float w) {
    atomicAdd(buf + S.local_vglobal[3 * prim], w * ((1.0f - u) - v));
    atomicAdd(buf + S.local_vglobal[3 * prim + 1], w * u);
    atomicAdd(buf + S.local_vglobal[3 * prim + 2], w * v);
}

// DiscreteDistribution.sample_reuse_pmf (distr.py:52-107): Dr.Jit's
// binary_search (the first index whose predicate fails, or n - 1), run inside
// the bracket the host's guide table gives for the draw: the predicate is
// monotone in the index and the draw, so the bracket holds that same index.
DWD int cdf_sample(const SubDev& S, float value, float& reuse, float& pmf) {
    float sample = value * S.pmf_sum;
    int k = (int)(value * (float)(1 << S.guide_shift));  // exact scaling, value < 1
    uint32_t start = (uint32_t)S.cdf_guide[k], end = (uint32_t)S.cdf_guide[k + 1];
    int iters = start < end ? (32 - __builtin_clz(end - start)) : 0;
    for (int it = 0; it < iters; ++it) {
        uint32_t mid = (start + end) >> 1;
        float cv = S.cdf[mid];
        bool cond = ((cv < sample) || (cv == 0.0f)) && (cv != S.pmf_sum);
        if (cond)
            start = (mid + 1 < end) ? mid + 1 : end;
        else
            end = mid;
    }
    int idx = (int)start;
    pmf = S.pmf[idx] * S.pmf_norm;
    float cdf = idx > 0 ? S.cdf[idx - 1] * S.pmf_norm : 0.0f;
    reuse = (value - cdf) / pmf;
    return idx;
}

// GreensBall (greens_fn.py:14-287) and GreensBall3D (greens_fn_3d.py:5-56)
DWD float greens2_G(F2 c, float R, F2 x, F2 y) {
    x = (x - c) / R;
    y = (y - c) / R;
    float v = norm(x - y);
    return kNegInv2Pi * cr_log((1.0f - dot(x, y)) / v);
}
DWD void greens2_sample(F2 c, float R, F2 dir, Rng& rng, F2& p, float& r, float& pdf) {
    float bound = 1.5f / R;
    int iteration = 0;
    bool active = true;
    r = 0.0f;
    pdf = 0.0f;
    if (R == 0.0f) {
        // Every proposal is rejected (r = u R = 0 and R / r = 0 / 0 make the
        // test u < NaN), so the first 999 rounds only consume their two draws
        // each: jump over them and run the last round, whose values the loop
        // leaves behind.  R is 0 where the star radius is 0; such a step's ray
        // has length 0, so a WoStGrad walk there mostly stays put and repeats
        // the step until its step cap.
        rng.advance(2 * 999);
        iteration = 999;
    }
    while ((iteration < 1000) && active) {  // rejection sampling of r ~ r ln(R/r)
        float u = rng.next();
        r = rng.next() * R;
        pdf = (cr_log(R / r) / kTwoPi) / (R * R / 4.0f);
        float pdf_radius = pdf / (1.0f / (kTwoPi * r));
        iteration += 1;
        if (u < pdf_radius / bound) active = false;
    }
    r = fmx(r, 1e-6f);
    if (r > R) r = R / 2.0f;
    p = c + dir * r;
}
DWD F2 greens2_gradient(F2 c, float R, float r, F2 y) {
    float gn = (1.0f / (r * r) - 1.0f / (R * R)) / kTwoPi;
    return (y - c) * gn;
}
DWD float greens3_G(F3 c, float R, F3 x, F3 y) {
    float r = norm(y - x);
    return -(1.0f / r - R / (R * R - dot(x - c, y - c))) / kFourPi;
}
DWD void greens3_sample(F3 c, float R, F3 dir, Rng& rng, F3& p, float& r, float& pdf) {
    float u1 = rng.next();
    float u2 = rng.next();
    float phi = kTwoPi * u2;  // Ulrich's polar method
    r = (1.0f + __builtin_sqrtf(1.0f - cr_cbrt(u1 * u1)) * cr_cos(phi)) * R / 2.0f;
    r = fmx(r, 1e-6f);
    if (r > R) r = R / 2.0f;
    pdf = ((1.0f / r - 1.0f / R) / kFourPi) / (R * R / 6.0f);
    p = c + dir * r;
}
DWD F3 greens3_gradient(F3 c, float R, float r, F3 y) {
    float r3 = r * r * r;
    float gn = (1.0f / r3 - 1.0f / (R * R * R)) / kFourPi;
    return (y - c) * gn;
}
DWD F2 uniform_dir2(Rng& rng) {  // uniform_angle (math.py:275-277)
    float theta = (rng.next() * 2.0f) * kPi;
    return f2(cr_cos(theta), cr_sin(theta));
}
DWD F3 uniform_sphere(Rng& rng) {  // math.py:298-303
    float u0 = rng.next();
    float u1 = rng.next();
    float z = 1.0f - 2.0f * u0;
    float r = __builtin_sqrtf(fmx(0.0f, 1.0f - z * z));
    float theta = kTwoPi * u1;
    return f3(r * cr_cos(theta), r * cr_sin(theta), z);
}

// Polyline.sdf / Mesh.sdf and inside (polyline.py:188-234, mesh.py:290-297) on
// the full scene: distance signed by the side of the interpolated vertex normal
// (2D) or the face normal (3D), negative inside; inside is sdf < 0.
DWD float sdf2(const SceneDev& sc, F2 x) {
    const SubDev& F = sc.sub[0];
    CP c = closest2(F, x);
    const float2* vn = (const float2*)F.vnormal;
    float2 a = vn[F.local_vidx[2 * c.prim]], b = vn[F.local_vidx[2 * c.prim + 1]];
    F2 n = lerp(f2(a.x, a.y), f2(b.x, b.y), c.t);
    float s = dot(n, x - f2(c.p.x, c.p.y)) >= 0.0f ? 1.0f : -1.0f;
    return c.d * s;
}
DWD float sdf3(const SceneDev& sc, F3 x) {
    CP c = closest3(sc.sub[0], x);
    float s = dot(c.n, x - c.p) >= 0.0f ? 1.0f : -1.0f;
    return c.d * s;
}
DWD bool inside2(const SceneDev& sc, F2 x) { return sdf2(sc, x) < 0.0f; }
DWD bool inside3(const SceneDev& sc, F3 x) { return sdf3(sc, x) < 0.0f; }

// --------------------------------------------------------------------------
// WoSt.u (wost.py:145-221) / WoSt3D.u (wost_3d.py:43-106) and WoS.u (wos.py:16-36)
// / WoS3D.u (wos_3d.py:15-34) as a step machine.  walk_init, then walk_advance
// repeatedly: it runs the walk's logic from its current stage until it has
// started the next boundary query in `t` (ADV_QUERY; the caller runs the
// traversal to its end and calls walk_advance again), finished a step
// (ADV_STEP) or finished the walk (ADV_END); then walk_value.
// Stages of one WoSt step: Dirichlet closest point -> star radius (silhouette
// query) -> [2D: Neumann sample + control-variate closest point] -> ray
// intersection -> source term.  Two exact shortcuts against the literal loop:
// the terminating step reuses the Dirichlet closest-point record (the
// reference queries the same point twice, wost.py:171,183) and skips the star
// radius, which has no side effects there.
// --------------------------------------------------------------------------
enum : int { WS_START = 0, WS_WOS, WS_DIR, WS_STAR, WS_SIL, WS_CV, WS_MOVE, WS_ISECT };
enum : int { ADV_QUERY = 0, ADV_STEP = 1, ADV_END = 2 };

template <int D>
struct WalkAux {};
template <>
struct WalkAux<2> {  // Neumann sample held across the control-variate query
    F2 ref, bp, bn, cv_n;
    float rnd, bpdf, cv_t;
    int idx, cv_prim;
};

template <int D>
struct Walk {
    typename VecT<D>::T ip, in;
    float result;
    int i;
    int term;
    int stage;
    bool ob;
    float R, d_dir;  // ball radius and Dirichlet distance of the current step
    WalkAux<D> aux;
};

DWD void set_zero(F2& v) { v = f2(0.0f, 0.0f); }
DWD void set_zero(F3& v) { v = f3(0.0f, 0.0f, 0.0f); }

template <int D>
DWD void walk_init(Walk<D>& w, typename VecT<D>::T x) {
    w.ip = x;
    set_zero(w.in);
    w.ob = false;
    w.result = 0.0f;
    w.i = 0;
    w.term = -1;
    w.stage = WS_START;
    w.R = 0.0f;
    w.d_dir = kInf;  // no previous step: the first Dirichlet query is unbounded
}

// Largest coordinate magnitude of a sub-scene's box (the absolute part of the
// rounding margins below).
DWD float coord_scale(const SubDev& S, int dim) {
    float sc = 0.0f;
    for (int k = 0; k < dim; ++k) sc = fmx(sc, fmx(__builtin_fabsf(S.root_mn[k]), __builtin_fabsf(S.root_mx[k])));
    return sc;
}
// Bound on the Dirichlet distance at the start of a step, from the step before:
// the distance to a set is 1-Lipschitz and a step moves the walk by at most
// R + 1e-5 (its ray starts at most 1e-5 inside a boundary point and stops
// within R), so d <= d_prev + R + 1e-5, widened by relative 2^-16 plus 1e-6 of
// the coordinate scale for rounding.  Started at this radius, the closest-point
// query only skips subtrees farther than its answer and returns the unbounded
// query's record, tie order included (the same argument as the control-variate
// query's bound).  kInf on a walk's first step.
template <int D>
DWD float dirichlet_bound(const Walk<D>& w, const SubDev& F) {
    float r = w.d_dir + w.R + 1e-5f;
    if (!(r < kInf)) return kInf;
    return r * (1.0f + 0x1p-16f) + 1e-6f * (1.0f + coord_scale(F, D));
}

// step finished: count it, end the walk at nsteps
template <int D>
DWD int step_done(Walk<D>& w, const ParamsDev& P) {
    w.stage = WS_START;
    return w.i < P.nsteps ? ADV_STEP : ADV_END;
}

template <bool X, bool A>
DWD int walk_advance(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, Walk<2>& w, Rng& rng, Trav& t,
                     const Adj& adj) {
    const SubDev& F = sc.sub[0];
    const SubDev& Dn = sc.sub[1];
    const SubDev& Nm = sc.sub[2];
    if (w.stage == WS_START) {
        if (!(w.i < P.nsteps)) return ADV_END;
        if (P.flags & DW_SOLVE_WOS) {
            closest2_begin(t, F, w.ip, kInf, P.node_width == 4);
            w.stage = WS_WOS;
            return ADV_QUERY;
        }
        if (Dn.present) {
            closest2_begin(t, Dn, w.ip, dirichlet_bound(w, F), P.node_width == 4);
            w.stage = WS_DIR;
            return ADV_QUERY;
        }
        w.R = w.d_dir = kInf;
        w.stage = WS_STAR;
    } else if (w.stage == WS_WOS) {  // WoS: terminate in the eps shell or jump to the ball
        CP c = closest2_end(t);
        float d = __builtin_fabsf(c.d);
        w.i += 1;
        if (d < P.eps) {
            w.result += g2(bc, F, c);
            if constexpr (A) adj_lin2(adj.dg, F, c.prim, c.t, adj.w);
            w.term = c.prim;
            return ADV_END;
        }
        F2 dir = uniform_dir2(rng);
        w.ip = w.ip + dir * d;
        return step_done(w, P);
    } else if (w.stage == WS_DIR) {
        CP cpd = closest2_end(t);
        w.R = w.d_dir = __builtin_fabsf(cpd.d);
        if (w.d_dir < P.eps) {
            w.result += g2(bc, Dn, cpd);
            if constexpr (A) adj_lin2(adj.dg, Dn, cpd.prim, cpd.t, adj.w);
            w.term = cpd.prim;
            w.i += 1;
            return ADV_END;
        }
        w.stage = WS_STAR;
    }
    if (w.stage == WS_STAR) {
        w.stage = WS_SIL;
        if (Nm.present) {  // star radius on the Neumann part
            F2 p = w.ob ? w.ip - w.in * 1e-5f : w.ip;
            t.r_max = star_bound(Nm, p, w.d_dir);
            if (Nm.py_sils && sil_begin<2>(t, Nm, p, t.r_max)) return ADV_QUERY;
        }
    }
    if (w.stage == WS_SIL) {
        w.aux.ref = f2(0.0f, 0.0f);
        w.aux.cv_prim = -1;
        w.stage = WS_MOVE;
        if (Nm.present) {
            float sr = t.r_max * 0.98f;
            sr = fmx(sr, P.min_R);
            w.R = fmn(w.R, sr);
            float rnd, pmf;
            int idx = cdf_sample(Nm, rng.next(), rnd, pmf);
            float4 e = ((const float4*)Nm.local_prims)[idx];
            F2 a = f2(e.x, e.y), b = f2(e.z, e.w);
            F2 tg = b - a;
            w.aux.bn = normalize(f2(tg.y, -tg.x));
            w.aux.bp = lerp(a, b, rnd);
            w.aux.bpdf = pmf / norm(tg);
            w.aux.rnd = rnd;
            w.aux.idx = idx;
            // boundary_control_variates (wost.py:123-126): the full-scene closest
            // point is no farther than the Dirichlet one found at the same point,
            // so the query starts its pruning radius just above that distance
            // (relative 2^-16 plus 1e-6 of the coordinate scale); it then only
            // skips subtrees farther than the answer and returns the unbounded
            // query's record, tie order included
            float r0 = kInf;
            if (w.d_dir < kInf) r0 = w.d_dir * (1.0f + 0x1p-16f) + 1e-6f * (1.0f + coord_scale(F, 2));
            closest2_begin(t, F, w.ip, r0, P.node_width == 4);
            w.stage = WS_CV;
            return ADV_QUERY;
        }
    } else if (w.stage == WS_CV) {
        CP cr = closest2_end(t);
        float hc = h2(bc, F, cr.prim, cr.t, f2(cr.p.x, cr.p.y), f2(cr.n.x, cr.n.y));
        F2 ref = f2(hc * cr.n.x, hc * cr.n.y);
        F2 bp = w.aux.bp, bn = w.aux.bn;
        float alpha = w.ob ? 2.0f : 1.0f;  // handle_boundary (wost.py:104-121)
        float hb = 0.0f;
        float G = clampf(greens2_G(w.ip, w.R, w.ip, bp), -10.0f, 10.0f);
        if (norm(w.ip - bp) <= w.R) {
            float h = h2(bc, Nm, w.aux.idx, w.aux.rnd, bp, bn);
            if (P.cv) h = h - dot(ref, bn);
            hb = 0.0f - G * h / w.aux.bpdf;
            if constexpr (A) {
                float k = adj.w * alpha * (G / w.aux.bpdf);  // d/dh of -alpha G h / pdf is -k
                adj_lin2(adj.dh, Nm, w.aux.idx, w.aux.rnd, -k);
                if (P.cv) adj_lin2(adj.dh, F, cr.prim, cr.t, k * dot(f2(cr.n.x, cr.n.y), bn));
            }
        }
        w.result += alpha * hb;
        w.aux.ref = ref;
        if constexpr (A) w.aux.cv_prim = cr.prim, w.aux.cv_t = cr.t, w.aux.cv_n = f2(cr.n.x, cr.n.y);
        w.stage = WS_MOVE;
    }
    if (w.stage == WS_MOVE) {
        F2 dir = uniform_dir2(rng);
        if (w.ob && (dot(dir, w.in) > 0.0f)) dir = -dir;
        F2 o = w.ob ? w.ip - w.in * 1e-5f : w.ip;  // Polyline.intersect (polyline.py:435-466)
        w.stage = WS_ISECT;
        if (isect2_begin(t, F, o, dir, w.R)) return ADV_QUERY;
    }
    // WS_ISECT: move to the hit or the ball, then the source term
    Hit its = isect2_end(t);
    F2 o = f2(t.x.x, t.x.y), dir = f2(t.v.x, t.v.y);
    F2 _p = w.ip;
    if (its.valid) {
        w.ip = f2(its.p.x, its.p.y);
        w.in = f2(its.n.x, its.n.y);
        w.ob = true;
    } else {
        w.ip = o + dir * w.R;
        w.in = dir;
        its.d = w.R;
        w.ob = false;
    }
    if (P.cv) w.result += dot(_p - w.ip, w.aux.ref);
    if constexpr (A) {  // d/dhc of dot(_p - ip, hc n_c)
        if (P.cv && w.aux.cv_prim >= 0)
            adj_lin2(adj.dh, F, w.aux.cv_prim, w.aux.cv_t, adj.w * dot(_p - w.ip, w.aux.cv_n));
    }
    F2 sp;  // handle_source (wost.py:128-143)
    float sr_, spdf;
    greens2_sample(_p, w.R, dir, rng, sp, sr_, spdf);
    float src = 0.0f;
    if (norm(sp - _p) < its.d) src += (w.R * w.R / 4.0f) * src_f(bc, sp);
    w.result += src;
    w.i += 1;
    return step_done(w, P);
}

template <bool X, bool A>
DWD int walk_advance(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, Walk<3>& w, Rng& rng, Trav& t,
                     const Adj& adj) {
    const SubDev& F = sc.sub[0];
    const SubDev& Dn = sc.sub[1];
    const SubDev& Nm = sc.sub[2];
    if (w.stage == WS_START) {
        if (!(w.i < P.nsteps)) return ADV_END;
        if (P.flags & DW_SOLVE_WOS) {
            closest3_begin(t, F, w.ip, P.node_width == 4);
            w.stage = WS_WOS;
            return ADV_QUERY;
        }
        if (__builtin_isnan(w.ip.x) || __builtin_isnan(w.ip.y) || __builtin_isnan(w.ip.z)) {
            // A walk that overflowed to infinity keeps a NaN coordinate from here on
            // (ip' = ip + dir R), and every remaining step is then fixed: no
            // closest-point or hit test can succeed (all distances are NaN), so it
            // never terminates early, the Neumann and source terms are skipped
            // (their |y - x| tests are false) and only `result += 0` remains.  The
            // traversals of such steps visit whole trees (NaN box distances clamp
            // to 0), so they are replaced by the outcome: the step count runs out
            // and the stream advances by the draws of the skipped steps (cdf, two
            // triangle coordinates, two direction and two source draws).
            uint64_t left = (uint64_t)(P.nsteps - w.i);
            rng.advance(left * (Nm.present ? 7u : 4u));
            w.result += 0.0f;
            w.i = P.nsteps;
            return ADV_END;
        }
        if (Dn.present) {
            // Far from the scene every distance to the Dirichlet boundary rounds to
            // one value: when the bounds on all of them coincide, the closest-point
            // query (which would visit the whole tree, every test a tie) returns
            // exactly that distance, and a step that does not terminate uses
            // nothing else of its record.
            float lo, hi;
            if (dist_bounds(Dn, w.ip, lo, hi) && lo == hi && !(lo < P.eps)) {
                w.R = w.d_dir = lo;
                w.stage = WS_STAR;
            } else {
                closest3_begin(t, Dn, w.ip, P.node_width == 4, dirichlet_bound(w, F));
                w.stage = WS_DIR;
                return ADV_QUERY;
            }
        } else {
            w.R = w.d_dir = kInf;
            w.stage = WS_STAR;
        }
    } else if (w.stage == WS_WOS) {
        CP c = closest3_end(t);
        float d = __builtin_fabsf(c.d);
        w.i += 1;
        if (d < P.eps) {
            w.result += g3(bc, F, c);
            if constexpr (A) adj_lin3(adj.dg, F, c.prim, c.u, c.v, adj.w);
            w.term = c.prim;
            return ADV_END;
        }
        F3 dir = uniform_sphere(rng);
        w.ip = w.ip + dir * d;
        return step_done(w, P);
    } else if (w.stage == WS_DIR) {
        CP cpd = closest3_end(t);
        w.R = w.d_dir = __builtin_fabsf(cpd.d);
        if (w.d_dir < P.eps) {
            w.result += g3(bc, Dn, cpd);
            if constexpr (A) adj_lin3(adj.dg, Dn, cpd.prim, cpd.u, cpd.v, adj.w);
            w.term = cpd.prim;
            w.i += 1;
            return ADV_END;
        }
        w.stage = WS_STAR;
    }
    if (w.stage == WS_STAR) {  // star radius on the full mesh
        F3 p = w.ob ? w.ip - w.in * 1e-3f : w.ip;
        t.r_max = star_bound(F, p, w.d_dir);
        w.stage = WS_SIL;
        // a silhouette counts only nearer than r_max: none can when every
        // distance to the scene is at least r_max, and the query then returns
        // r_max unchanged
        float lo, hi;
        bool none = dist_bounds(F, p, lo, hi) && t.r_max <= lo;
        if (F.py_sils && !none && sil_begin<3>(t, F, p, t.r_max)) return ADV_QUERY;
    }
    if (w.stage == WS_SIL) {
        float sr = t.r_max * 0.98f;
        sr = fmx(sr, P.min_R);
        w.R = fmn(w.R, sr);
        if (Nm.present) {  // Mesh.sample_boundary (mesh.py:474-490)
            float rnd, pmf;
            int idx = cdf_sample(Nm, rng.next(), rnd, pmf);
            const float* tp = Nm.local_prims + 9 * idx;
            F3 a = f3(tp[0], tp[1], tp[2]), b = f3(tp[3], tp[4], tp[5]), c = f3(tp[6], tp[7], tp[8]);
            F3 cr = cross(b - a, c - a);
            float area = norm(cr) / 2.0f;
            float r1 = rng.next();
            float r2 = rng.next();
            float su = __builtin_sqrtf(r1) * (1.0f - r2);
            float sv = __builtin_sqrtf(r1) * r2;
            F3 bp = (b * su + c * sv) + a * ((1.0f - su) - sv);
            F3 bn = normalize(cr);
            float bpdf = pmf / area;
            float G = clampf(greens3_G(w.ip, w.R, w.ip, bp), -100.0f, 100.0f);
            float alpha = w.ob ? 2.0f : 1.0f;
            if (norm(w.ip - bp) <= w.R) {
                float h = h3(bc, Nm, idx, su, sv, bp, bn);
                w.result -= G * h * alpha / bpdf;
                if constexpr (A) adj_lin3(adj.dh, Nm, idx, su, sv, -adj.w * ((G * alpha) / bpdf));
            }
        }
        F3 dir = uniform_sphere(rng);
        if (w.ob && (dot(dir, w.in) > 0.0f)) dir = -dir;
        F3 o = w.ob ? w.ip - w.in * 1e-5f : w.ip;  // Mesh.intersect (mesh.py:383-415)
        w.stage = WS_ISECT;
        if (isect3_begin(t, F, o, dir, w.R)) return ADV_QUERY;
    }
    // WS_ISECT: move to the hit or the ball, then the source term
    Hit its = isect3_end(t);
    F3 o = t.x, dir = t.v;
    F3 _p = w.ip;
    if (its.valid) {
        w.ip = its.p;
        w.in = its.n;
        w.ob = true;
    } else {
        w.ip = o + dir * w.R;
        w.in = dir;
        its.d = w.R;
        w.ob = false;
    }
    F3 sp;
    float sr_, spdf;
    greens3_sample(_p, w.R, dir, rng, sp, sr_, spdf);
    float src = 0.0f;
    if (norm(sp - _p) < its.d) src += (w.R * w.R / 6.0f) * src_f(bc, sp);
    w.result += src;
    w.i += 1;
    return step_done(w, P);
}

// One whole step (or the walk's end) with the traversals run back to back.
template <bool X, bool A = false, int D>
DWD bool walk_step(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, Walk<D>& w, Rng& rng,
                   const Adj& adj = Adj{0.0f, nullptr, nullptr}) {
    Trav t;
    Ovf o;
    while (true) {
        int r = walk_advance<X, A>(sc, bc, P, w, rng, t, adj);
        if (r != ADV_QUERY) return r == ADV_STEP;
        while (!trav_iter<D, X>(t, o)) {
        }
    }
}

// final value of a finished walk: 2D guard (wost.py:218-219); 3D has none
DWD float walk_value(const Walk<2>& w, const ParamsDev& P) {
    float r = w.result;
    if (P.flags & DW_SOLVE_WOS) return r;  // WoS.u has no guard
    return (bad(r) || __builtin_fabsf(r) > 100.0f) ? 0.0f : r;
}
DWD float walk_value(const Walk<3>& w, const ParamsDev&) { return w.result; }

template <int D, bool X = true>
DWD float walk_u(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, typename VecT<D>::T x, Rng& rng,
                 int& steps, int& term) {
    Walk<D> w;
    walk_init<D>(w, x);
    while (walk_step<X>(sc, bc, P, w, rng)) {
    }
    steps = w.i;
    term = w.term;
    return walk_value(w, P);
}
DWD float walk_u2(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, F2 x, Rng& rng, int& steps, int& term) {
    return walk_u<2>(sc, bc, P, x, rng, steps, term);
}
DWD float walk_u3(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, F3 x, Rng& rng, int& steps, int& term) {
    return walk_u<3>(sc, bc, P, x, rng, steps, term);
}

}  // namespace dwd
