// dw_grad.h -- WoStGrad / WoStGrad3D (diff_wost/solvers/wost_grad.py,
// wost_grad_3d.py), the robust derivative walk: du/de from a walk that reflects
// off Neumann boundaries,
// with the rho-bounded star radius, BVH sampling of the Neumann boundary and of
// its silhouette vertices, and a nested WoSt walk for the normal derivative at
// a Dirichlet point.  The queries walk the sub-scene's node-major tree
// (DNode2 / DCone2 / DSeg, the reference's flat layout) one lane at a time.
#pragma once
#include "dw_walk.h"

namespace dwd {

struct GradDev {
    float rho_max;
    int ignore_boundary, ignore_silhouette, ignore_source;
};

// Second derivatives of the analytic field (for dh/dt): (u_xx, u_xy, u_yy)
DWD void bc_hess2(const BCDev& bc, F2 p, float& hxx, float& hxy, float& hyy) {
    if (bc.code == DW_BC_EXPCOS) {
        float ex = cr_exp(p.x), c = cr_cos(p.y), sn = cr_sin(p.y);
        hxx = ex * c, hxy = -(ex * sn), hyy = -(ex * c);
    } else if (bc.code == DW_BC_QUAD) {
        hxx = 2.0f, hxy = 0.0f, hyy = -2.0f;
    } else {
        hxx = hxy = hyy = 0.0f;
    }
}
// the scene callbacks of the analytic field: h = n . grad u, and the
// derivatives of g and h along the tangent (-n.y, n.x)
DWD float grad_h2(const BCDev& bc, F2 p, F2 n) {
    F3 g = bc_grad(bc, f3(p.x, p.y, 0.0f));
    return dot(n, f2(g.x, g.y));
}
DWD float grad_dgdt2(const BCDev& bc, F2 p, F2 t) {
    F3 g = bc_grad(bc, f3(p.x, p.y, 0.0f));
    return dot(f2(g.x, g.y), t);
}
DWD float grad_dhdt2(const BCDev& bc, F2 p, F2 n) {
    float hxx, hxy, hyy;
    bc_hess2(bc, p, hxx, hxy, hyy);
    F2 t = f2(-n.y, n.x);
    return dot(n, f2(hxx * t.x + hxy * t.y, hxy * t.x + hyy * t.y));
}

DWD F2 node_mn(const DNode2& n) { return f2(n.mnx, n.mny); }
DWD F2 node_mx(const DNode2& n) { return f2(n.mxx, n.mxy); }

// LineSegment.star_radius_2 (line_segment.py:169-212)
DWD float seg_star_radius2(F2 a, F2 b, F2 x, F2 e, float rho_max, float r_max) {
    float R = r_max;
    F2 tangent = normalize(b - a);
    F2 normal = f2(tangent.y, -tangent.x);
    F2 dir = normalize(a - x);
    if (dot(dir, normal) > 0.0f) {
        float en = dot(e, normal), et = dot(e, tangent), d = dot(a - x, normal);
        float t1 = -kInf, t2 = kInf;
        F2 pt = x + normal * d;
        if (__builtin_fabsf(en) > 1e-4f) {
            t1 = d / en * (rho_max + et);
            t2 = -d / en * (rho_max - et);
        }
        float ta = dot(a - x, tangent), tb = dot(b - x, tangent);
        float tmin = fmn(ta, tb), tmax = fmx(ta, tb);
        if ((t1 > tmin) && (t1 < tmax)) R = fmn(R, norm((pt + tangent * t1) - x));
        if ((t2 > tmin) && (t2 < tmax)) R = fmn(R, norm((pt + tangent * t2) - x));
        if (((t1 > tmax) && (t2 > tmax)) || ((t1 < tmin) && (t2 < tmin))) {
            R = fmn(R, norm(a - x));
            R = fmn(R, norm(b - x));
        }
    } else {
        R = fmn(R, norm(a - x));
        R = fmn(R, norm(b - x));
    }
    return R;
}

// SNCHNode.overlaps_2 (snch.py:100-174): the cone part for a valid cone
// (ha >= 0) of a node whose box is within r_max, with fp64 transcendentals
DWD_COLD bool rho_cone2_fp64(F2 mn, F2 mx, const DCone2& cn, F2 x, F2 e, float rho_bound, float d_min) {
    float ha = cn.half_angle;
    bool ov = (ha > kPi2) && (d_min < kDrEps);
    if (!ov) {
        F2 c = (mn + mx) / 2.0f;
        F2 vca = c - x;
        float length = norm(vca);
        vca = vca / length;
        ov = length < cn.radius;
        if (!ov) {
            F2 axis = f2(cn.ax, cn.ay);
            float axis_angle = cr_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
            float vch = cr_asin(cn.radius / length);
            float hs = ha + vch;
            float amin = axis_angle - hs, amax = axis_angle + hs;
            ov = (hs > kPi2) || ((kPi2 >= amin) && (kPi2 <= amax));
            if (!ov) {
                float c0 = cr_cos(amin), c1 = cr_cos(amax);
                float dn_min = fmn(c0, c1), dn_max = fmx(c0, c1);
                float ed_axis = cr_acos(clampf(dot(e, vca), -1.0f, 1.0f));
                float e0 = cr_cos(ed_axis - ha), e1 = cr_cos(ed_axis + ha);
                float ed_min = fmn(e0, e1), ed_max = fmx(e0, e1);
                float en_axis = cr_acos(clampf(dot(e, axis), -1.0f, 1.0f));
                float n0 = cr_cos(en_axis - ha), n1 = cr_cos(en_axis + ha);
                float en_min = fmn(n0, n1), en_max = fmx(n0, n1);
                float en_dn_max = en_max / dn_min;
                float en_dn_min = en_min < 0.0f ? en_min / dn_min : en_min / dn_max;
                auto rho = [](float ed, float q) { return (1.0f - 2.0f * ed * q) + q * q; };
                float rm = rho(ed_min, en_dn_min);
                rm = fmx(rm, rho(ed_min, en_dn_max));
                rm = fmx(rm, rho(ed_max, en_dn_min));
                rm = fmx(rm, rho(ed_max, en_dn_max));
                ov = rm > rho_bound * rho_bound;
            }
        }
    }
    return ov;
}
// The same decision with the transcendentals bracketed (kU: open)
DWD int rho_cone2_bracketed(F2 mn, F2 mx, const DCone2& cn, F2 x, F2 e, float rho_bound, float d_min) {
    float ha = cn.half_angle;
    if ((ha > kPi2) && (d_min < kDrEps)) return kT;
    F2 c = (mn + mx) / 2.0f;
    F2 vca = c - x;
    float length = norm(vca);
    vca = vca / length;
    if (length < cn.radius) return kT;
    F2 axis = f2(cn.ax, cn.ay);
    FI axis_angle = fi_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
    FI hs = fi_add(fi(ha), fi_asin(cn.radius / length));
    FI amin = fi_sub(axis_angle, hs), amax = fi_add(axis_angle, hs);
    int t = t_or(t_gt(hs, kPi2), t_in(kPi2, amin, amax));
    if (t != kF) return t;
    FI c0 = fi_cos(amin), c1 = fi_cos(amax);
    FI dn_min = fi_min(c0, c1), dn_max = fi_max(c0, c1);
    FI ed_axis = fi_acos(clampf(dot(e, vca), -1.0f, 1.0f));
    FI e0 = fi_cos(fi_sub(ed_axis, fi(ha))), e1 = fi_cos(fi_add(ed_axis, fi(ha)));
    FI ed_min = fi_min(e0, e1), ed_max = fi_max(e0, e1);
    FI en_axis = fi_acos(clampf(dot(e, axis), -1.0f, 1.0f));
    FI n0 = fi_cos(fi_sub(en_axis, fi(ha))), n1 = fi_cos(fi_add(en_axis, fi(ha)));
    FI en_min = fi_min(n0, n1), en_max = fi_max(n0, n1);
    int neg = t_lt(en_min, 0.0f);
    if (neg == kU) return kU;
    FI en_dn_max = fi_div(en_max, dn_min);
    FI en_dn_min = fi_div(en_min, neg == kT ? dn_min : dn_max);
    auto rho = [](FI ed, FI q) { return fi_add(fi_sub(fi(1.0f), fi_mul(fi_mul(fi(2.0f), ed), q)), fi_sq(q)); };
    FI rm = rho(ed_min, en_dn_min);
    rm = fi_max(rm, rho(ed_min, en_dn_max));
    rm = fi_max(rm, rho(ed_max, en_dn_min));
    rm = fi_max(rm, rho(ed_max, en_dn_max));
    return t_gt(rm, rho_bound * rho_bound);
}
DWD bool snch_overlaps2(F2 mn, F2 mx, const DCone2& cn, F2 x, F2 e, float rho_bound, float r_max, float& d_min) {
    float d_max;
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    if ((cn.half_angle >= 0.0f) && ov) {
        int t = rho_cone2_bracketed(mn, mx, cn, x, e, rho_bound, d_min);
        ov = t == kU ? rho_cone2_fp64(mn, mx, cn, x, e, rho_bound, d_min) : t == kT;
    }
    return ov;
}

// The cone fields the rho test reads, from a child-in-parent record
DWD DCone2 rho_cone2(float ax, float ay, float ha, float radius) {
    DCone2 c{};
    c.ax = ax, c.ay = ay, c.half_angle = ha, c.radius = radius;
    return c;
}
// Polyline.star_radius_2 (polyline.py:317-330) -> SNCH.star_radius_2 (snch.py:380-453)
// The traversal reads the child-in-parent records: an inner node's SNCH record
// holds both children's boxes and cones and its BVH record their primitive
// references, so each popped node costs one round of dependent fetches.  Same
// nodes in the same order as a walk over the node arrays.
DWD float star_radius2_rho(const SubDev& S, F2 x, F2 e, float rho_max) {
    float m0 = fmx(__builtin_fabsf(x.x - S.bb_min[0]), __builtin_fabsf(x.x - S.bb_max[0]));
    float m1 = fmx(__builtin_fabsf(x.y - S.bb_min[1]), __builtin_fabsf(x.y - S.bb_max[1]));
    float r_max = norm(f2(m0, m1));
    float dmin, dmax;
    box_dist(f2(S.root_mn[0], S.root_mn[1]), f2(S.root_mx[0], S.root_mx[1]), x, dmin, dmax);
    if (!(dmin <= r_max)) return r_max;
    Ovf o;
    int sp = 0;
    stk_push(sp, o, S.root_ref_bvh, dmin);
    while (sp > 0) {
        int ref;
        float cd;
        stk_pop(sp, o, ref, cd);
        if (!(cd <= r_max)) continue;
        if (ref < 0) {
            int off, cnt;
            leaf_range(S, ref, off, cnt);
            for (int j = 0; j < cnt; ++j) {
                DSeg sg = gload<DSeg>(S.prims, off + j);
                float r = seg_star_radius2(f2(sg.ax, sg.ay), f2(sg.bx, sg.by), x, e, rho_max, r_max);
                if (r < r_max) r_max = r;
            }
        } else {
            const RSnch2 rc = gload<RSnch2>(S.rec_snch, ref);
            const int2 br = gload<int2>(S.rec_bvh, ref * 6 + 4);  // RBvh2 ref_l, ref_r
            float d0, d1;
            bool h0 = snch_overlaps2(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), rho_cone2(rc.lax, rc.lay, rc.lha, rc.lrad),
                                     x, e, rho_max, r_max, d0);
            bool h1 = snch_overlaps2(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), rho_cone2(rc.rax, rc.ray, rc.rha, rc.rrad),
                                     x, e, rho_max, r_max, d1);
            // the builder's depth cap (62) keeps the stack within DW_MAX_STACK
            if (h0 && h1) {
                int closer = br.x, other = br.y;
                if (d1 < d0) {
                    closer = br.y, other = br.x;
                    float tmp = d0;
                    d0 = d1, d1 = tmp;
                }
                stk_push(sp, o, other, d1);
                stk_push(sp, o, closer, d0);
            } else if (h0) {
                stk_push(sp, o, br.x, d0);
            } else if (h1) {
                stk_push(sp, o, br.y, d1);
            }
        }
    }
    return r_max;
}

// One root-to-leaf descent of BVH._sample_boundary (bvh.py:289-369; the point
// BVH's twin): each inner node picks a child overlapping the ball with
// probability proportional to |G(x, R, box centre)|; the leaf draws one of the
// primitives in the ball (all of them when the chosen box lies inside it) by
// reservoir sampling weighted by `weight`.  Returns the sorted index (-1: none).
template <typename Acc, typename Wt>
DWD int bvh2_descend(const DNode2* nodes, F2 x, float R, Rng& rng, bool& hit, float& pdf, Acc accept, Wt weight) {
    hit = false;
    pdf = 1.0f;
    int sorted = -1;
    float r_max = kInf;
    int ni = 0;
    while (true) {
        const DNode2 nd = nodes[ni];
        if (nd.nref > 0) {
            float total = 0.0f;
            for (int j = 0; j < nd.nref; ++j) {
                int k = nd.off + j;
                float sa = weight(k);
                if ((r_max <= R) || accept(k)) {
                    hit = true;
                    total += sa;
                    float sel = sa / total;
                    float u = rng.next();
                    if (u < sel) sorted = k;
                }
            }
            if (total > 0.0f) pdf *= (sorted >= 0 ? weight(sorted) : 0.0f) / total;
            return sorted;
        }
        int L = ni + 1, Rn = ni + nd.off;
        const DNode2 n0 = nodes[L], n1 = nodes[Rn];
        float d0, m0, d1, m1;
        box_dist(node_mn(n0), node_mx(n0), x, d0, m0);
        box_dist(node_mn(n1), node_mx(n1), x, d1, m1);
        float w0 = (d0 <= R) ? 1.0f : 0.0f, w1 = (d1 <= R) ? 1.0f : 0.0f;
        if (w0 > 0.0f) w0 *= __builtin_fabsf(greens2_G(x, R, x, (node_mn(n0) + node_mx(n0)) / 2.0f));
        if (w1 > 0.0f) w1 *= __builtin_fabsf(greens2_G(x, R, x, (node_mn(n1) + node_mx(n1)) / 2.0f));
        float tw = w0 + w1;
        if (!(tw > 0.0f)) return sorted;
        float p0 = w0 / tw, p1 = 1.0f - p0;
        float u = rng.next();
        if (u < p0) {
            ni = L;
            pdf *= p0;
            r_max = m0;
        } else {
            ni = Rn;
            pdf *= p1;
            r_max = m1;
        }
    }
}

struct BSample2 {
    bool hit;
    F2 p, n;
    float t, pdf;
    int prim;
};
// Polyline.sample_boundary_bvh (polyline.py:515-517) -> BVH.sample_boundary (bvh.py:266-287)
DWD BSample2 sample_boundary_bvh2(const SubDev& S, F2 x, float R, Rng& rng) {
    BSample2 b{false, f2(0.0f, 0.0f), f2(0.0f, 0.0f), 0.0f, 0.0f, 0};
    const DNode2* nodes = (const DNode2*)S.nodes;
    const DSeg* prims = (const DSeg*)S.prims;
    float dmin, dmax;
    box_dist(node_mn(nodes[0]), node_mx(nodes[0]), x, dmin, dmax);
    if (!(dmin <= R)) return b;
    auto seg = [&](int k, F2& a, F2& c) {
        if (k < 0) {
            a = c = f2(0.0f, 0.0f);
            return;
        }
        DSeg sg = prims[k];
        a = f2(sg.ax, sg.ay), c = f2(sg.bx, sg.by);
    };
    auto weight = [&](int k) {  // LineSegment.surface_area (line_segment.py:32-33)
        F2 a, c;
        seg(k, a, c);
        return norm(c - a);
    };
    auto accept = [&](int k) {  // LineSegment.sphere_intersect (line_segment.py:43-45)
        F2 a, c;
        seg(k, a, c);
        F2 u = c - a, w = x - a;
        float t = clampf(dot(w, u) / dot(u, u), 0.0f, 1.0f);
        return norm(lerp(a, c, t) - x) <= R;
    };
    bool hit;
    float pdf;
    int k = bvh2_descend(nodes, x, R, rng, hit, pdf, accept, weight);
    if (!hit) return b;
    F2 a, c;
    seg(k, a, c);
    float t = rng.next();  // LineSegment.sample_point (line_segment.py:36-40)
    F2 tg = c - a;
    b.hit = true;
    b.p = a + tg * t;
    b.n = normalize(f2(tg.y, -tg.x));
    b.t = t;
    b.pdf = pdf * (1.0f / norm(tg));
    b.prim = k >= 0 ? S.prim_index[k] : 0;
    return b;
}

struct SSample2 {
    bool hit;
    F2 p, n1, n2, t1, t2;
    float pdf;
    int prim;
};
// Polyline.sample_silhouette_bvh (polyline.py:519-535) -> PointBVH.sample_boundary
DWD SSample2 sample_silhouette_bvh2(const SubDev& S, F2 x, float R, Rng& rng) {
    SSample2 r;
    r.hit = false;
    r.pdf = 0.0f;
    r.prim = 0;
    r.p = r.n1 = r.n2 = r.t1 = r.t2 = f2(0.0f, 0.0f);
    if (S.n_psil == 0) return r;
    const DNode2* nodes = (const DNode2*)S.pnodes;
    float dmin, dmax;
    box_dist(node_mn(nodes[0]), node_mx(nodes[0]), x, dmin, dmax);
    if (!(dmin <= R)) return r;
    auto weight = [](int) { return 1.0f; };
    auto accept = [&](int k) {
        const float* q = S.psil + 6 * S.porder[k];
        return norm(f2(q[2], q[3]) - x) <= R;
    };
    bool hit;
    float pdf;
    int k = bvh2_descend(nodes, x, R, rng, hit, pdf, accept, weight);
    if (!hit) return r;
    int si = k >= 0 ? S.porder[k] : 0;
    const float* q = S.psil + 6 * si;
    F2 a = f2(q[0], q[1]), b = f2(q[2], q[3]), c = f2(q[4], q[5]);
    r.hit = true;
    r.t1 = normalize(b - a);
    r.t2 = normalize(b - c);
    r.n1 = f2(r.t1.y, -r.t1.x);
    r.n2 = f2(-r.t2.y, r.t2.x);
    r.p = b;
    r.pdf = pdf * 1.0f;
    r.prim = si;
    return r;
}

// Polyline.off_centered_ball (polyline.py:583-600)
DWD void off_centered_ball2(const SubDev& F, F2 p, F2 n, F2& c, float& radius) {
    float b = fmn(F.bb_max[0] - F.bb_min[0], F.bb_max[1] - F.bb_min[1]) / 2.0f;
    float a = 1e-3f;
    for (int i = 0; i < 10; ++i) {
        float m = (a + b) / 2.0f;
        float d = __builtin_fabsf(closest2(F, p - n * m).d);
        if (__builtin_fabsf(d - m) < 1e-3f)
            a = m;
        else
            b = m;
    }
    c = p - n * a;
    radius = a;
}

// WoStGrad.dudn (wost_grad.py:207-254): the normal derivative at a Dirichlet
// point from one WoSt walk (nsteps 200, eps 1e-4) started on the off-centred
// ball, in the parts before and after that walk, so a scheduler can run the
// walk a step at a time.  Dudn2 carries the nested walk and the values the
// second part needs (and, for the caller, the tangent frame and dg/dt).
struct Dudn2 {
    Walk<2> w;
    F2 c, p, ip, nrm, tg;
    float R, theta, u_ref, dudt;
};
DWD ParamsDev dudn_params(const ParamsDev& P) {
    ParamsDev W = P;
    W.nwalks = 1, W.nsubwalks = 1, W.nsteps = 200, W.eps = 1e-4f, W.min_R = 1e-3f, W.cv = 1, W.av = 1;
    W.flags = DW_SOLVE_EXACT_CONES;
    return W;
}
DWD void dudn2_begin(const SceneDev& sc, const BCDev& bc, const CP& its, Rng& rng, Dudn2& d) {
    F2 ip = f2(its.p.x, its.p.y), in = f2(its.n.x, its.n.y);
    off_centered_ball2(sc.sub[0], ip, in, d.c, d.R);
    float theta = (rng.next() * 2.0f) * kPi;  // uniform_angle
    theta = clampf(theta, 3e-1f, kTwoPi - 3e-1f);
    F2 pd = f2(-in.y, in.x);
    F2 dir = in * cr_cos(theta) + pd * cr_sin(theta);
    d.p = d.c + dir * d.R;
    d.u_ref = bc_u(bc, its.p);
    d.theta = theta;
    d.ip = ip;
    walk_init<2>(d.w, d.p);
}
DWD float dudn2_end(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, const Dudn2& d, Rng& rng, int& steps) {
    const SubDev& Dn = sc.sub[1];
    const SubDev& Nm = sc.sub[2];
    const F2 c = d.c, p = d.p, ip = d.ip;
    const float R = d.R, theta = d.theta, u_ref = d.u_ref;
    float u = walk_value(d.w, dudn_params(P));
    steps += d.w.i;
    if (Nm.present) {
        float dd = closest2(Dn, p).d;
        float dnm = closest2(Nm, p).d;
        if ((dd < 2e-2f) && (dnm < 2e-2f)) u = u_ref;
    }
    float Pk = 1.0f / (cr_cos(theta) - 1.0f);
    float dudn = 0.0f;
    if (P.cv)
        dudn += Pk * (u - u_ref) / R;
    else
        dudn += Pk * u / R;
    float rr = __builtin_sqrtf(rng.next());  // uniform_in_disk (math.py:285-288)
    float th = kTwoPi * rng.next();
    F2 p_src = c + f2(rr * cr_cos(th), rr * cr_sin(th)) * R;
    float f_ref = src_f(bc, ip), f = src_f(bc, p_src);
    float x2 = dot(ip - c, ip - c);
    float y2 = dot(p_src - c, p_src - c);
    float Pq = (x2 - y2) / fmx(dot(ip - p_src, ip - p_src), 1e-6f) * R / 2.0f;
    if (P.cv)
        dudn -= (f - f_ref) * Pq + f_ref * R / 2.0f;
    else
        dudn -= f * Pq;
    return dudn;
}

// State of one WoStGrad walk between its steps (2D and 3D): position, normal
// and on-boundary flag, the direction e being differentiated along, the
// running estimate, the 2D throughput, the step count and whether it goes on.
template <int D>
struct GradWalk {
    typename VecT<D>::T ip, in, e;
    float result, thr;
    int i;
    bool ob, active;
};
DWD void grad_init(const SceneDev& sc, GradWalk<2>& g, F2 x, F2 e) {
    g.ip = x;
    g.in = f2(0.0f, 0.0f);
    g.e = e;
    g.ob = false;
    g.result = 0.0f;
    g.i = 0;
    g.active = g.ob || inside2(sc, g.ip);
    g.thr = 1.0f;
}
template <int D>
DWD bool grad_running(const GradWalk<D>& g, const ParamsDev& P) {
    return (g.i < P.nsteps) && g.active;
}

// One step of WoStGrad.dude (wost_grad.py:91-205), from an interior point.
// grad_step2_begin returns false when the step is done, true when it has
// reached a Dirichlet point and the normal derivative's WoSt walk (d.w, run
// with dudn_params) is pending; grad_step2_end then completes the step.
DWD bool grad_step2_begin(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, const GradDev& G, GradWalk<2>& g,
                          Rng& rng, Dudn2& d) {
    const SubDev& F = sc.sub[0];
    const SubDev& Dn = sc.sub[1];
    const SubDev& Nm = sc.sub[2];
    F2 ip = g.ip, in = g.in, e = g.e;
    bool ob = g.ob, active = g.active;
    float result = g.result, thr = g.thr;
    int i = g.i;
    {
        bool inflate = false;
        float d_dir = kInf;
        if (Dn.present) d_dir = __builtin_fabsf(closest2(Dn, ip).d);
        if (d_dir < P.eps) {  // a Dirichlet point: dg/dt and the normal derivative's walk
            CP cr = closest2(F, ip);
            d.nrm = f2(cr.n.x, cr.n.y), d.tg = f2(-d.nrm.y, d.nrm.x);
            d.dudt = grad_dgdt2(bc, f2(cr.p.x, cr.p.y), d.tg);
            dudn2_begin(sc, bc, cr, rng, d);
            return true;
        } else {
            float R = d_dir;
            if (Nm.present) R = fmn(R, star_radius2_rho(Nm, ip, e, G.rho_max));
            if (ob) {
                if (R < P.min_R) inflate = true;
                R = fmx(R, P.min_R);
            }
            float alpha = ob ? 2.0f : 1.0f;
            if (Nm.present) {
                if (!G.ignore_boundary) {
                    BSample2 b = sample_boundary_bvh2(Nm, ip, R, rng);
                    if (b.hit) {  // handle_boundary (wost_grad.py:28-51)
                        float Gv = clampf(greens2_G(ip, R, ip, b.p), -100.0f, 100.0f);
                        float res = 0.0f;
                        if (norm(ip - b.p) <= R) {
                            F2 tg = f2(-b.n.y, b.n.x);
                            float e_t = dot(e, tg), e_n = dot(e, b.n);
                            float f_tn = grad_dhdt2(bc, b.p, b.n);
                            res += e_t * f_tn - e_n * src_f(bc, b.p);
                        }
                        result += thr * alpha * (-Gv * res / b.pdf);
                    }
                }
                if (!G.ignore_silhouette) {
                    SSample2 sr = sample_silhouette_bvh2(Nm, ip, R, rng);
                    if (!inflate && sr.hit) {  // handle_silhouette (wost_grad.py:53-72)
                        float Gv = clampf(greens2_G(ip, R, ip, sr.p), -100.0f, 100.0f);
                        float res = 0.0f;
                        if (norm(sr.p - ip) <= R) {
                            float h1 = grad_h2(bc, sr.p, sr.n1), h2 = grad_h2(bc, sr.p, sr.n2);
                            res -= h1 * dot(e, sr.t1);
                            res -= h2 * dot(e, sr.t2);
                        }
                        result += thr * alpha * (-Gv * res / sr.pdf);
                    }
                }
            }
            F2 dir = uniform_dir2(rng);
            if (ob && (dot(dir, in) > 0.0f)) dir = -dir;
            F2 _p = ip;
            F2 o = ob ? ip - in * 1e-5f : ip;  // Polyline.intersect (polyline.py:435-466)
            Hit its = intersect2(F, o, dir, R);
            if (!its.valid) {
                F2 q = o + dir * R;
                its.p = f3(q.x, q.y, 0.0f);
                its.n = f3(dir.x, dir.y, 0.0f);
                its.d = R;
                its.on_boundary = false;
            }
            ip = f2(its.p.x, its.p.y);
            in = f2(its.n.x, its.n.y);
            ob = its.on_boundary;
            if (!G.ignore_source) {  // handle_source (wost_grad.py:74-89)
                F2 sp;
                float sr_, spdf;
                greens2_sample(_p, R, dir, rng, sp, sr_, spdf);
                float res = 0.0f;
                F3 df = src_grad(bc);
                if (norm(sp - _p) < __builtin_fabsf(its.d)) res += (R * R / 4.0f) * dot(f2(df.x, df.y), e);
                result += thr * res;
            }
            if (ob) {
                F2 tg = f2(-in.y, in.x);
                float e_t = dot(e, tg), e_n = dot(e, in);
                float K = dot(dir, tg) / dot(dir, in);
                result += thr * e_n * grad_h2(bc, ip, in);
                thr *= e_t - K * e_n;
                e = tg;
            }
            if (__builtin_fabsf(thr) < 1.0f) {  // Russian roulette
                if (rng.next() > __builtin_fabsf(thr))
                    active = false;
                else
                    thr /= __builtin_fabsf(thr);
            }
        }
        i += 1;
        if (bad(result) || __builtin_fabsf(result) > 1000.0f) result = 0.0f;
    }
    g.ip = ip, g.in = in, g.e = e;
    g.ob = ob, g.active = active;
    g.result = result, g.thr = thr;
    g.i = i;
    return false;
}
DWD void grad_step2_end(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, GradWalk<2>& g, Rng& rng,
                        const Dudn2& d, int& steps) {
    float dudn = dudn2_end(sc, bc, P, d, rng, steps);
    float result = g.result;
    result += g.thr * d.dudt * dot(g.e, d.tg);
    result += g.thr * dudn * dot(g.e, d.nrm);
    g.active = false;
    g.i += 1;
    if (bad(result) || __builtin_fabsf(result) > 1000.0f) result = 0.0f;
    g.result = result;
}
DWD void grad_step2(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, const GradDev& G, GradWalk<2>& g,
                    Rng& rng, int& steps) {
    Dudn2 d;
    if (grad_step2_begin(sc, bc, P, G, g, rng, d)) {
        const ParamsDev W = dudn_params(P);
        while (walk_step<true>(sc, bc, W, d.w, rng)) {
        }
        grad_step2_end(sc, bc, P, g, rng, d, steps);
    }
}
DWD float grad_dude2(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, const GradDev& G, F2 x, F2 e,
                     Rng& rng, int& steps) {
    GradWalk<2> g;
    grad_init(sc, g, x, e);
    while (grad_running(g, P)) grad_step2(sc, bc, P, G, g, rng, steps);
    steps += g.i;
    return g.result;
}

// ---------------------------------------------------------------------------
// WoStGrad3D (diff_wost/solvers/wost_grad_3d.py).  The reference calls a dudn
// it never defines once a walk reaches a Dirichlet point, so the estimator
// runs on all-Neumann scenes.
// ---------------------------------------------------------------------------
// Hessian rows of the analytic field
DWD void bc_hess3(const BCDev& bc, F3 p, F3 H[3]) {
    if (bc.code == DW_BC_EXPCOS) {
        float ex = cr_exp(p.x), c = cr_cos(p.y), sn = cr_sin(p.y);
        H[0] = f3(ex * c, -(ex * sn), 0.0f), H[1] = f3(-(ex * sn), -(ex * c), 0.0f), H[2] = f3(0.0f, 0.0f, 0.0f);
    } else if (bc.code == DW_BC_QUAD) {
        H[0] = f3(2.0f, 0.0f, 0.0f), H[1] = f3(0.0f, -2.0f, 0.0f), H[2] = f3(0.0f, 0.0f, 0.0f);
    } else {
        H[0] = H[1] = H[2] = f3(0.0f, 0.0f, 0.0f);
    }
}

// DiscreteDistribution.sample_reuse_pmf (distr.py:52-107) over the silhouette
// lengths, searched inside the guide table's bracket (as cdf_sample)
DWD int psil3_sample(const SubDev& S, float value, float& reuse, float& pmf) {
    float sample = value * S.ps_sum;
    int k = (int)(value * (float)(1 << S.ps_shift));
    uint32_t start = (uint32_t)S.ps_guide[k], end = (uint32_t)S.ps_guide[k + 1];
    int iters = start < end ? (32 - __builtin_clz(end - start)) : 0;
    for (int it = 0; it < iters; ++it) {
        uint32_t mid = (start + end) >> 1;
        float cv = S.ps_cdf[mid];
        bool cond = ((cv < sample) || (cv == 0.0f)) && (cv != S.ps_sum);
        if (cond)
            start = (mid + 1 < end) ? mid + 1 : end;
        else
            end = mid;
    }
    int idx = (int)start;
    pmf = S.ps_pmf[idx] * S.ps_norm;
    float cdf = idx > 0 ? S.ps_cdf[idx - 1] * S.ps_norm : 0.0f;
    reuse = (value - cdf) / pmf;
    return idx;
}

DWD F3 node3_mn(const DNode3& n) { return f3(n.mnx, n.mny, n.mnz); }
DWD F3 node3_mx(const DNode3& n) { return f3(n.mxx, n.mxy, n.mxz); }

// SNCHNode3D.visit (snch3d.py:196-271): the cone part for a valid cone
// (ha >= 0) of a node whose box is within r_max, with fp64 transcendentals
DWD_COLD bool rho_cone3_fp64(F3 mn, F3 mx, const DCone3& cn, F3 x, F3 v, float rho_max, float d_min) {
    float ha = cn.half_angle;
    bool ov = (ha > kPi2) && (d_min < kDrEps);
    if (!ov) {
        F3 c = (mn + mx) / 2.0f;
        F3 vca = c - x;
        float l = norm(vca);
        vca = vca / l;
        ov = l < cn.radius;
        if (!ov) {
            F3 axis = f3(cn.ax, cn.ay, cn.az);
            float axis_angle = cr_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
            float vch = cr_asin(cn.radius / l);
            float hs = ha + vch;
            float amin = axis_angle - hs, amax = axis_angle + hs;
            ov = (hs > kPi2) || ((kPi2 >= amin) && (kPi2 <= amax));
            if (!ov) {
                float c0 = cr_cos(amin), c1 = cr_cos(amax);
                float cmin_a = fmn(c0, c1), cmax_a = fmx(c0, c1);
                if ((amin <= 0.0f) && (amax >= 0.0f)) cmax_a = fmx(cmax_a, 1.0f);
                float beta = cr_acos(clampf(dot(axis, v), -1.0f, 1.0f));
                float bmin = beta - ha, bmax = beta + ha;
                float b0 = cr_cos(bmin), b1 = cr_cos(bmax);
                float cmin_b = fmn(b0, b1), cmax_b = fmx(b0, b1);
                if ((bmin <= 0.0f) && (bmax >= 0.0f)) cmax_b = fmx(cmax_b, 1.0f);
                float gamma = cr_acos(clampf(dot(v, vca), -1.0f, 1.0f));
                float gmin = gamma - vch, gmax = gamma + vch;
                float g0 = cr_cos(gmin), g1 = cr_cos(gmax);
                float cmin_g = fmn(g0, g1), cmax_g = fmx(g0, g1);
                if ((gmin <= 0.0f) && (gmax >= 0.0f)) cmax_g = fmx(cmax_g, 1.0f);
                float a_min = cmin_b / cmax_a, a_max = cmax_b / cmin_a;
                float b_min = cmin_g, b_max = cmax_g;
                float rho2 = (1.0f - 2.0f * fmn(fmn(a_min * b_min, a_min * b_max), fmn(a_max * b_min, a_max * b_max))) +
                             fmx(a_min * a_min, a_max * a_max);
                ov = rho2 > rho_max * rho_max;
            }
        }
    }
    return ov;
}
// [min, max] of cos over [lo, hi] as the test forms it: cos at both ends,
// raised to 1 when the range holds 0 (kU: open)
DWD int fi_cos_range(FI lo, FI hi, FI& cmin, FI& cmax) {
    FI c0 = fi_cos(lo), c1 = fi_cos(hi);
    cmin = fi_min(c0, c1), cmax = fi_max(c0, c1);
    int z = t_and(t_le(lo, 0.0f), t_ge(hi, 0.0f));
    if (z == kT) cmax = fi_max(cmax, fi(1.0f));
    return z;
}
// The same decision with the transcendentals bracketed (kU: open)
DWD int rho_cone3_bracketed(F3 mn, F3 mx, const DCone3& cn, F3 x, F3 v, float rho_max, float d_min) {
    float ha = cn.half_angle;
    if ((ha > kPi2) && (d_min < kDrEps)) return kT;
    F3 c = (mn + mx) / 2.0f;
    F3 vca = c - x;
    float l = norm(vca);
    vca = vca / l;
    if (l < cn.radius) return kT;
    F3 axis = f3(cn.ax, cn.ay, cn.az);
    FI axis_angle = fi_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
    FI vch = fi_asin(cn.radius / l);
    FI hs = fi_add(fi(ha), vch);
    FI amin = fi_sub(axis_angle, hs), amax = fi_add(axis_angle, hs);
    int t = t_or(t_gt(hs, kPi2), t_in(kPi2, amin, amax));
    if (t != kF) return t;
    FI cmin_a, cmax_a, cmin_b, cmax_b, cmin_g, cmax_g;
    if (fi_cos_range(amin, amax, cmin_a, cmax_a) == kU) return kU;
    FI beta = fi_acos(clampf(dot(axis, v), -1.0f, 1.0f));
    if (fi_cos_range(fi_sub(beta, fi(ha)), fi_add(beta, fi(ha)), cmin_b, cmax_b) == kU) return kU;
    FI gamma = fi_acos(clampf(dot(v, vca), -1.0f, 1.0f));
    if (fi_cos_range(fi_sub(gamma, vch), fi_add(gamma, vch), cmin_g, cmax_g) == kU) return kU;
    FI a_min = fi_div(cmin_b, cmax_a), a_max = fi_div(cmax_b, cmin_a);
    FI m = fi_min(fi_min(fi_mul(a_min, cmin_g), fi_mul(a_min, cmax_g)),
                  fi_min(fi_mul(a_max, cmin_g), fi_mul(a_max, cmax_g)));
    FI rho2 = fi_add(fi_sub(fi(1.0f), fi_mul(fi(2.0f), m)), fi_max(fi_sq(a_min), fi_sq(a_max)));
    return t_gt(rho2, rho_max * rho_max);
}
DWD bool snch_visit3(F3 mn, F3 mx, const DCone3& cn, F3 x, F3 v, float rho_max, float r_max, float& d_min) {
    float d_max;
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    if ((cn.half_angle >= 0.0f) && ov) {
        int t = rho_cone3_bracketed(mn, mx, cn, x, v, rho_max, d_min);
        ov = t == kU ? rho_cone3_fp64(mn, mx, cn, x, v, rho_max, d_min) : t == kT;
    }
    return ov;
}

// inside_triangle (math.py:311-331)
DWD bool inside_triangle3(F3 p, F3 a, F3 b, F3 c) {
    F3 v0 = c - a, v1 = b - a, v2 = p - a;
    float d00 = dot(v0, v0), d01 = dot(v0, v1), d02 = dot(v0, v2), d11 = dot(v1, v1), d12 = dot(v1, v2);
    float denom = d00 * d11 - d01 * d01;
    if (denom == 0.0f) return false;
    float uu = (d11 * d02 - d01 * d12) / denom;
    float vv = (d00 * d12 - d01 * d02) / denom;
    return (uu >= 0.0f) && (vv >= 0.0f) && (uu + vv <= 1.0f);
}

// Triangle.star_radius_2 (triangle.py:109-151) with Plane.bounding_point
// (plane.py:11-34) and line_sphere_intersection (math.py:373-386)
DWD float tri_star_radius2(const DTri& tr, F3 x, F3 e, float rho_max, float r_max) {
    F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), cc = f3(tr.cx, tr.cy, tr.cz);
    float R = r_max;
    F3 n = normalize(cross(b - a, cc - a));
    F3 c = f3(kInf, kInf, kInf), p = f3(kInf, kInf, kInf);
    float r = kInf;
    if (__builtin_fabsf(dot(e, n)) > kDrEps) {
        float d = dot(a - x, n) / dot(e, n);
        c = x + e * d;
        F3 xp = x + n * dot(a - x, n);
        r = rho_max * __builtin_fabsf(d);
        if (norm(xp - c) > kDrEps)
            p = c + normalize(xp - c) * r;
        else if (norm(n - f3(1.0f, 0.0f, 0.0f)) < 1e-4f)
            p = c + cross(n, f3(0.0f, 1.0f, 0.0f)) * r;
        else
            p = c + cross(n, f3(0.0f, 0.0f, 1.0f)) * r;
    }
    F3 qp;
    float qu, qv, qd;
    tri_closest(tr, x, qp, qu, qv, qd);
    bool inside_circle = (norm(a - c) < r) && (norm(b - c) < r) && (norm(cc - c) < r);
    if (inside_circle) {
    } else if (norm(qp - c) > r) {
        R = fmn(R, norm(qp - x));
    } else if (inside_triangle3(p, a, b, cc)) {
        R = fmn(R, norm(p - x));
    } else {
        F3 vs[3] = {a, b, cc};
        for (int k = 0; k < 3; ++k) {
            F3 s0 = vs[k], s1 = vs[(k + 1) % 3];
            F3 dd = s1 - s0, L = s0 - c;
            float qa = dot(dd, dd), qb = 2.0f * dot(dd, L), qc = dot(L, L) - r * r;
            float disc = qb * qb - 4.0f * qa * qc;
            if (disc > 0.0f) {
                float t0 = (-qb - __builtin_sqrtf(disc)) / (2.0f * qa);
                float t1 = (-qb + __builtin_sqrtf(disc)) / (2.0f * qa);
                if ((t0 >= 0.0f) && (t0 <= 1.0f)) R = fmn(R, norm((s0 + dd * t0) - x));
                if ((t1 >= 0.0f) && (t1 <= 1.0f)) R = fmn(R, norm((s0 + dd * t1) - x));
            }
        }
    }
    return R;
}

// The SNCH traversal of SNCH3D.star_radius_3 (snch3d.py:742-811) in
// resumable form: rho3_begin pushes the root when its box is within r0;
// each rho3_iter pops one node (a leaf's triangles bound r_max, an inner
// node's overlapping children are pushed, the nearer on top) and reports
// whether the traversal has ended.  The answer is r_max.  Inner nodes are read
// from the child-in-parent records (children's boxes and cones from the SNCH
// record, their primitive references from the BVH record): one round of
// dependent fetches per popped node, the same nodes in the same order.
struct Rho3 {
    float r_max;
    int sp;
};
DWD DCone3 rho_cone3(float ax, float ay, float az, float ha, float radius) {
    DCone3 c{};
    c.ax = ax, c.ay = ay, c.az = az, c.half_angle = ha, c.radius = radius;
    return c;
}
DWD bool rho3_begin(Rho3& r, Ovf& o, const SubDev& S, F3 x, float r0) {
    r.r_max = r0;
    r.sp = 0;
    float dmin, dmax;
    box_dist(f3(S.root_mn[0], S.root_mn[1], S.root_mn[2]), f3(S.root_mx[0], S.root_mx[1], S.root_mx[2]), x, dmin, dmax);
    if (!(dmin <= r0)) return false;
    stk_push(r.sp, o, S.root_ref_bvh, dmin);
    return true;
}
DWD bool rho3_iter(Rho3& r, Ovf& o, const SubDev& S, F3 x, F3 e, float rho_max) {
    int ref;
    float cd;
    stk_pop(r.sp, o, ref, cd);
    if (cd <= r.r_max) {
        if (ref < 0) {
            int off, cnt;
            leaf_range(S, ref, off, cnt);
            for (int j = 0; j < cnt; ++j) {
                float rr = tri_star_radius2(gload<DTri>(S.prims, off + j), x, e, rho_max, r.r_max);
                if (rr < r.r_max) r.r_max = rr;
            }
        } else {
            const RSnch3 rc = gload<RSnch3>(S.rec_snch, ref);
            const int2 br = gload<int2>(S.rec_bvh, ref * 8 + 6);  // RBvh3 ref_l, ref_r
            float d0, d1;
            bool h0 = snch_visit3(f3(rc.h.lmnx, rc.h.lmny, rc.h.lmnz), f3(rc.h.lmxx, rc.h.lmxy, rc.h.lmxz),
                                  rho_cone3(rc.c.lax, rc.c.lay, rc.c.laz, rc.h.lha, rc.c.lrad), x, e, rho_max, r.r_max, d0);
            bool h1 = snch_visit3(f3(rc.h.rmnx, rc.h.rmny, rc.h.rmnz), f3(rc.h.rmxx, rc.h.rmxy, rc.h.rmxz),
                                  rho_cone3(rc.c.rax, rc.c.ray, rc.c.raz, rc.h.rha, rc.c.rrad), x, e, rho_max, r.r_max, d1);
            if (h0 && h1) {
                int closer = br.x, other = br.y;
                if (d1 < d0) {
                    closer = br.y, other = br.x;
                    float tmp = d0;
                    d0 = d1, d1 = tmp;
                }
                stk_push(r.sp, o, other, d1);
                stk_push(r.sp, o, closer, d0);
            } else if (h0) {
                stk_push(r.sp, o, br.x, d0);
            } else if (h1) {
                stk_push(r.sp, o, br.y, d1);
            }
        }
    }
    return r.sp == 0;
}
// Mesh.star_radius_2 (mesh.py:349-361): the star radius, then SNCH3D.star_radius_3
// (snch3d.py:742-811) over the triangles
DWD float star_radius3_rho(const SubDev& S, F3 x, F3 e, float rho_max) {
    Rho3 r;
    Ovf o;
    if (rho3_begin(r, o, S, x, star_radius3<true>(S, x, kInf)))
        while (!rho3_iter(r, o, S, x, e, rho_max)) {
        }
    return r.r_max;
}

DWD void grad_init(const SceneDev& sc, GradWalk<3>& g, F3 x, F3 e) {
    g.ip = x;
    g.in = f3(0.0f, 0.0f, 0.0f);
    g.e = e;
    g.ob = false;
    g.result = 0.0f;
    g.active = inside3(sc, x);
    g.i = 0;
    g.thr = 1.0f;
}
// One step of WoStGrad3D.dude (wost_grad_3d.py:93-206) on an all-Neumann
// scene, in the parts around its three queries (the star radius's silhouette
// search and rho-bounded traversal, then the ray), so a scheduler can run the
// queries one traversal iteration at a time.  GradStep3 carries a step's
// values from one part to the next.
struct GradStep3 {
    float R0;  // the silhouette search radius (star_bound)
    float R;   // the ball radius
    F3 dir, org, prev;
};
// Step start: a point on the boundary moves 1e-3 inside before the star radius.
DWD void grad3_start(const SceneDev& sc, GradWalk<3>& g) {
    if (sc.sub[2].present && g.ob) g.ip = g.ip - g.in * 1e-3f;
}
// With the star radius Rrho of the Neumann part: the ball radius, the boundary
// and silhouette samples (Mesh.sample_boundary / sample_silhouette with
// handle_boundary / handle_silhouette) and the direction; the ray is then
// (s.org, s.dir) up to s.R.
DWD void grad3_sample(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, GradWalk<3>& g, Rng& rng, float Rrho,
                      GradStep3& s) {
    const SubDev& Nm = sc.sub[2];
    const F3 ip = g.ip, in = g.in, e = g.e;
    const bool ob = g.ob;
    float result = g.result;
    bool inflate = false;
    float R = kInf;
    if (Nm.present) {
        R = fmn(R, Rrho);
        if (ob) {
            if (R < P.min_R) inflate = true;
            R = fmx(R, P.min_R);
        }
    }
    float alpha = ob ? 2.0f : 1.0f;
    if (Nm.present) {
        {  // Mesh.sample_boundary (mesh.py:474-490) + handle_boundary (wost_grad_3d.py:22-45)
            float rnd, pmf;
            int idx = cdf_sample(Nm, rng.next(), rnd, pmf);
            const float* tp = Nm.local_prims + 9 * idx;
            F3 a = f3(tp[0], tp[1], tp[2]), b = f3(tp[3], tp[4], tp[5]), c = f3(tp[6], tp[7], tp[8]);
            F3 cr = cross(b - a, c - a);
            float area = norm(cr) / 2.0f;
            float r1 = rng.next(), r2 = rng.next();
            float su = __builtin_sqrtf(r1) * (1.0f - r2), sv = __builtin_sqrtf(r1) * r2;
            F3 bp = (b * su + c * sv) + a * ((1.0f - su) - sv);
            F3 bn = normalize(cr);
            float bpdf = pmf / area;
            if (!inflate) {
                float Gv = clampf(greens3_G(ip, R, ip, bp), -100.0f, 100.0f);
                float res = 0.0f;
                if (norm(ip - bp) <= R) {
                    float e_n = dot(e, bn);
                    F3 e_t = e - bn * e_n;
                    F3 H[3];
                    bc_hess3(bc, bp, H);
                    F3 He = f3(dot(H[0], e_t), dot(H[1], e_t), dot(H[2], e_t));
                    float f_tn = dot(He, bn);
                    res += f_tn - e_n * src_f(bc, bp);
                    if (P.cv) res -= dot(f3(0.0f, 0.0f, 0.0f), bn);
                }
                result += alpha * (-Gv * res / bpdf);
            }
        }
        if (Nm.n_psil3 > 0) {  // Mesh.sample_silhouette (mesh.py:500-516) + handle_silhouette (:47-71)
            float rnd, pmf;
            int idx = psil3_sample(Nm, rng.next(), rnd, pmf);
            const float* q = Nm.psil3 + 12 * idx;
            F3 a = f3(q[0], q[1], q[2]), b = f3(q[3], q[4], q[5]), c = f3(q[6], q[7], q[8]), d = f3(q[9], q[10], q[11]);
            float length = norm(b - a);
            F3 n1 = normalize(cross(b - a, c - a));
            F3 n2 = normalize(cross(d - a, b - a));
            F3 t1 = normalize(cross(b - a, n1));
            F3 t2 = normalize(cross(a - b, n2));
            F3 sp = lerp(a, b, rng.next());
            float spdf = pmf / length;
            if (!inflate) {
                float Gv = clampf(greens3_G(ip, R, ip, sp), -100.0f, 100.0f);
                float res = 0.0f;
                if (norm(sp - ip) <= R) {
                    F3 gu = bc_grad(bc, sp);
                    float h1 = dot(gu, n1), h2 = dot(gu, n2);
                    float e1 = dot(e, t1), e2 = dot(e, t2);
                    if (P.cv) {
                        h1 = h1 - dot(f3(0.0f, 0.0f, 0.0f), n1);
                        h2 = h2 - dot(f3(0.0f, 0.0f, 0.0f), n2);
                    }
                    res -= h1 * e1;
                    res -= h2 * e2;
                }
                result += alpha * (-Gv * res / spdf);
            }
        }
    }
    F3 dir = uniform_sphere(rng);
    if (ob && (dot(dir, in) > 0.0f)) dir = -dir;
    s.prev = ip;
    s.org = ob ? ip - in * 1e-5f : ip;  // Mesh.intersect (mesh.py:384-415)
    s.dir = dir;
    s.R = R;
    g.result = result;
}
// With the ray's hit: the move, handle_source, the tangent reflection of e,
// Russian roulette.
DWD void grad3_finish(const BCDev& bc, const GradDev& G, GradWalk<3>& g, Rng& rng, const GradStep3& s, Hit its) {
    const F3 o = s.org, dir = s.dir, _p = s.prev;
    const float R = s.R;
    F3 ip, in, e = g.e;
    bool ob, active = g.active;
    float result = g.result;
    int i = g.i;
    if (!its.valid) {
        its.p = o + dir * R;
        its.n = dir;
        its.d = R;
        its.on_boundary = false;
    }
    ip = its.p;
    in = its.n;
    ob = its.on_boundary;
    result += dot(_p - ip, f3(0.0f, 0.0f, 0.0f));
    if (!G.ignore_source) {  // handle_source (wost_grad_3d.py:73-90)
        F3 spt;
        float sr_, spdf;
        greens3_sample(_p, R, dir, rng, spt, sr_, spdf);
        float res = 0.0f;
        if (norm(spt - _p) < its.d) res += (R * R / 6.0f) * dot(src_grad(bc), e);
        result += res;
    }
    if (ob) {
        float e_n = dot(e, in);
        F3 e_t = e - in * e_n;
        float dir_n = dot(dir, in);
        F3 dir_t = dir - in * dir_n;
        F3 ref_t = f3(0.0f, 0.0f, 0.0f) - in * dot(f3(0.0f, 0.0f, 0.0f), in);
        result += dot(dir_t, ref_t) * e_n / dir_n;
        result += e_n * dot(in, bc_grad(bc, ip));
        e = e_t - (dir_t * e_n) / dir_n;
    }
    if (norm(e) < 1.0f) {  // Russian roulette
        if (norm(e) < rng.next())
            active = false;
        else
            e = e / norm(e);
    }
    i += 1;
    if (bad(result) || __builtin_fabsf(result) > 1000.0f) result = 0.0f;
    g.ip = ip, g.in = in, g.e = e;
    g.ob = ob, g.active = active;
    g.result = result;
    g.i = i;
}
DWD void grad_step3(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, const GradDev& G, GradWalk<3>& g,
                    Rng& rng, int& steps) {
    const SubDev& Nm = sc.sub[2];
    grad3_start(sc, g);
    float Rrho = kInf;
    if (Nm.present) Rrho = star_radius3_rho(Nm, g.ip, g.e, G.rho_max);
    GradStep3 s;
    grad3_sample(sc, bc, P, g, rng, Rrho, s);
    grad3_finish(bc, G, g, rng, s, intersect3(sc.sub[0], s.org, s.dir, s.R));
}
DWD float grad_dude3(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, const GradDev& G, F3 x, F3 e,
                     Rng& rng, int& steps) {
    GradWalk<3> g;
    grad_init(sc, g, x, e);
    while (grad_running(g, P)) grad_step3(sc, bc, P, G, g, rng, steps);
    steps += g.i;
    return g.result;
}

}  // namespace dwd
