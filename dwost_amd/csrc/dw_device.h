// dw_device.h -- device-side geometry, sampling and traversal routines for
// the Walk-on-Stars kernels (gfx950).
//
// Arithmetic contract (shared with the CPU restatement used as the test
// oracle): fp32, no implicit contraction (-ffp-contract=off); Dr.Jit's fused
// forms are explicit fmaf (dot = fma chain, lerp = fma(b,t,fma(-a,t,a)),
// 3D cross = fma(u.y,v.z,-(u.z*v.y)) ...); sqrt and division are correctly
// rounded; log/exp/sin/cos/acos/asin/atan2/cbrt are evaluated in fp64 and
// rounded once (gfx950 runs fp64 at half the fp32 rate).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dw_layout.h"

namespace dwd {

#define DWD __device__ __forceinline__

constexpr float kInf = __builtin_inff();
constexpr float kPi = 3.14159265358979323846f;
constexpr float kPi2 = 1.57079632679489661923f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kFourPi = 12.5663706143591729538f;
constexpr float kDrEps = 5.9604644775390625e-08f;  // 2^-24
constexpr float kNegInv2Pi = -0.159154943091895335768f;

struct F2 { float x, y; };
struct F3 { float x, y, z; };

DWD F2 f2(float x, float y) { return F2{x, y}; }
DWD F3 f3(float x, float y, float z) { return F3{x, y, z}; }
DWD F2 operator+(F2 a, F2 b) { return f2(a.x + b.x, a.y + b.y); }
DWD F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
DWD F2 operator-(F2 a, F2 b) { return f2(a.x - b.x, a.y - b.y); }
DWD F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
DWD F2 operator*(F2 a, float s) { return f2(a.x * s, a.y * s); }
DWD F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
DWD F2 operator/(F2 a, float s) { return f2(a.x / s, a.y / s); }
DWD F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
DWD F2 operator-(F2 a) { return f2(-a.x, -a.y); }
DWD F3 operator-(F3 a) { return f3(-a.x, -a.y, -a.z); }
DWD float dot(F2 a, F2 b) { return __builtin_fmaf(a.y, b.y, a.x * b.x); }
DWD float dot(F3 a, F3 b) { return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)); }
DWD float norm(F2 a) { return __builtin_sqrtf(dot(a, a)); }
DWD float norm(F3 a) { return __builtin_sqrtf(dot(a, a)); }
DWD F2 normalize(F2 a) { return a * (1.0f / __builtin_sqrtf(dot(a, a))); }
DWD F3 normalize(F3 a) { return a * (1.0f / __builtin_sqrtf(dot(a, a))); }
DWD F3 cross(F3 u, F3 v) {
    return f3(__builtin_fmaf(u.y, v.z, -(u.z * v.y)), __builtin_fmaf(u.z, v.x, -(u.x * v.z)),
              __builtin_fmaf(u.x, v.y, -(u.y * v.x)));
}
DWD float cross2(F2 u, F2 v) { return u.x * v.y - u.y * v.x; }
DWD float lerpf(float a, float b, float t) { return __builtin_fmaf(b, t, __builtin_fmaf(-a, t, a)); }
DWD F2 lerp(F2 a, F2 b, float t) { return f2(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t)); }
DWD F3 lerp(F3 a, F3 b, float t) { return f3(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t)); }
DWD float fmn(float a, float b) { return __builtin_fminf(a, b); }
DWD float fmx(float a, float b) { return __builtin_fmaxf(a, b); }
DWD float clampf(float x, float lo, float hi) { return fmx(fmn(x, hi), lo); }
DWD F2 vmin(F2 a, F2 b) { return f2(fmn(a.x, b.x), fmn(a.y, b.y)); }
DWD F2 vmax(F2 a, F2 b) { return f2(fmx(a.x, b.x), fmx(a.y, b.y)); }
DWD F3 vmin(F3 a, F3 b) { return f3(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
DWD F3 vmax(F3 a, F3 b) { return f3(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
DWD F2 vabs(F2 a) { return f2(__builtin_fabsf(a.x), __builtin_fabsf(a.y)); }
DWD F3 vabs(F3 a) { return f3(__builtin_fabsf(a.x), __builtin_fabsf(a.y), __builtin_fabsf(a.z)); }
DWD bool bad(float v) { return __builtin_isnan(v) || __builtin_isinf(v); }

// fp64-evaluated, once-rounded transcendentals
#ifdef DW_FP32_TRANSCENDENTALS  // timing experiments only: breaks bit parity with the oracle
DWD float cr_log(float x) { return logf(x); }
DWD float cr_exp(float x) { return expf(x); }
DWD float cr_sin(float x) { return sinf(x); }
DWD float cr_cos(float x) { return cosf(x); }
DWD float cr_acos(float x) { return acosf(x); }
DWD float cr_asin(float x) { return asinf(x); }
DWD float cr_atan2(float y, float x) { return atan2f(y, x); }
DWD float cr_cbrt(float x) { return cbrtf(x); }
#else
DWD float cr_log(float x) { return (float)log((double)x); }
DWD float cr_exp(float x) { return (float)exp((double)x); }
DWD float cr_sin(float x) { return (float)sin((double)x); }
DWD float cr_cos(float x) { return (float)cos((double)x); }
DWD float cr_acos(float x) { return (float)acos((double)x); }
DWD float cr_asin(float x) { return (float)asin((double)x); }
DWD float cr_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
DWD float cr_cbrt(float x) { return (float)cbrt((double)x); }
#endif

// ---------------------------------------------------------------------------
// TEA seeding (diff_wost/core/math.py:137-157) and PCG32 (Dr.Jit PCG32)
// ---------------------------------------------------------------------------
DWD void tea4(uint32_t v0, uint32_t v1, uint32_t& o0, uint32_t& o1) {
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        s += 0x9E3779B9u;
        v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + s) ^ ((v1 >> 5) + 0xC8013EA4u);
        v1 += ((v0 << 4) + 0xAD90777Du) ^ (v0 + s) ^ ((v0 >> 5) + 0x7E95761Eu);
    }
    o0 = v0;
    o1 = v1;
}

struct Rng {
    uint64_t state, inc;
    DWD uint32_t next_u32() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    DWD float next() { return __uint_as_float((next_u32() >> 9) | 0x3f800000u) - 1.0f; }
    // Jump ahead by delta draws: the LCG's affine map composed by squaring
    // (O'Neill's pcg_advance_lcg_64, after Brown, "Random number generation
    // with arbitrary strides", 1994).
    DWD void advance(uint64_t delta) {
        uint64_t cur_mult = 0x5851f42d4c957f2dULL, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1u) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1u) * cur_plus;
            cur_mult *= cur_mult;
            delta >>= 1;
        }
        state = acc_mult * state + acc_plus;
    }
    // PCG32(size, initstate, initseq) lane `lane`: seed(initstate+lane, initseq+lane)
    DWD void seed_lane(uint32_t v0, uint32_t v1, uint64_t lane) {
        uint64_t initstate = (uint64_t)v0 + lane, initseq = (uint64_t)v1 + lane;
        state = 0;
        inc = (initseq << 1) | 1u;
        next_u32();
        state += initstate;
        next_u32();
    }
};

// ---------------------------------------------------------------------------
// Records
// ---------------------------------------------------------------------------
struct CP {  // closest point: 2D uses t, 3D uses (u, v)
    F3 p, n;
    float t, u, v, d;
    int prim;
};
struct Hit {
    bool valid, on_boundary;
    F3 p, n;
    float t, u, v, d;
    int prim;
};

// BoundingBox.distance (bvh.py:31-36)
DWD void box_dist(F2 mn, F2 mx, F2 p, float& dmin, float& dmax) {
    F2 u = mn - p, v = p - mx;
    dmin = norm(vmax(vmax(u, v), f2(0.0f, 0.0f)));
    dmax = norm(vmin(u, v));
}
DWD void box_dist(F3 mn, F3 mx, F3 p, float& dmin, float& dmax) {
    F3 u = mn - p, v = p - mx;
    dmin = norm(vmax(vmax(u, v), f3(0.0f, 0.0f, 0.0f)));
    dmax = norm(vmin(u, v));
}
DWD float box_dmin(F2 mn, F2 mx, F2 p) {
    F2 u = mn - p, v = p - mx;
    return norm(vmax(vmax(u, v), f2(0.0f, 0.0f)));
}
DWD float box_dmin(F3 mn, F3 mx, F3 p) {
    F3 u = mn - p, v = p - mx;
    return norm(vmax(vmax(u, v), f3(0.0f, 0.0f, 0.0f)));
}
// BoundingBox.intersect (bvh.py:42-59)
DWD bool box_ray(F2 mn, F2 mx, F2 x, F2 v, float rmax, float& tmin) {
    F2 t0 = f2((mn.x - x.x) / v.x, (mn.y - x.y) / v.y);
    F2 t1 = f2((mx.x - x.x) / v.x, (mx.y - x.y) / v.y);
    F2 tn = vmin(t0, t1), tf = vmax(t0, t1);
    float a = fmx(fmx(tn.x, tn.y), 0.0f);
    float b = fmn(fmn(tf.x, tf.y), rmax);
    tmin = a > b ? kInf : a;
    return !(a > b);
}
DWD bool box_ray(F3 mn, F3 mx, F3 x, F3 v, float rmax, float& tmin) {
    F3 t0 = f3((mn.x - x.x) / v.x, (mn.y - x.y) / v.y, (mn.z - x.z) / v.z);
    F3 t1 = f3((mx.x - x.x) / v.x, (mx.y - x.y) / v.y, (mx.z - x.z) / v.z);
    F3 tn = vmin(t0, t1), tf = vmax(t0, t1);
    float a = fmx(fmx(fmx(tn.x, tn.y), tn.z), 0.0f);
    float b = fmn(fmn(fmn(tf.x, tf.y), tf.z), rmax);
    tmin = a > b ? kInf : a;
    return !(a > b);
}

DWD F2 nmin(const DNode2& n) { return f2(n.mnx, n.mny); }
DWD F2 nmax(const DNode2& n) { return f2(n.mxx, n.mxy); }
DWD F3 nmin(const DNode3& n) { return f3(n.mnx, n.mny, n.mnz); }
DWD F3 nmax(const DNode3& n) { return f3(n.mxx, n.mxy, n.mxz); }

// ---------------------------------------------------------------------------
// Traversal stack: the bottom kLdsStack entries of each lane live in LDS
// (entry k of lane t at [k * blockDim.x + t], conflict-free), deeper entries in
// a private overflow array.  Traversals descend into the nearer child and push
// only the farther one, so the stack never holds more than the tree depth.
// Kernels that traverse must launch with kStackLdsBytes of dynamic LDS.
// ---------------------------------------------------------------------------
constexpr int kLdsStack = 16;
constexpr int kStackLdsBytesPerThread = kLdsStack * 8;
extern __shared__ int dw_lds[];

struct Stk {
    int* si;
    float* sd;
    int stride, sp;
    int oi[DW_MAX_STACK - kLdsStack];
    float od[DW_MAX_STACK - kLdsStack];
    DWD void push(int r, float d) {
        if (sp < kLdsStack) {
            si[sp * stride] = r;
            sd[sp * stride] = d;
        } else {
            oi[sp - kLdsStack] = r;
            od[sp - kLdsStack] = d;
        }
        ++sp;
    }
    DWD void pop(int& r, float& d) {
        --sp;
        if (sp < kLdsStack) {
            r = si[sp * stride];
            d = sd[sp * stride];
        } else {
            r = oi[sp - kLdsStack];
            d = od[sp - kLdsStack];
        }
    }
};
DWD void stack_init(Stk& st) {
    st.stride = (int)blockDim.x;
    st.si = dw_lds + threadIdx.x;
    st.sd = (float*)(dw_lds + kLdsStack * blockDim.x) + threadIdx.x;
    st.sp = 0;
}

// leaf reference -> [off, off + cnt)
DWD void leaf_range(const SubDev& S, int ref, int& off, int& cnt) {
    uint32_t u = (uint32_t)ref;
    int c = (int)((u >> 27) & 15u);
    if (c != DW_LEAF_TABLE) {
        cnt = c;
        off = (int)(u & 0x7FFFFFFu);
    } else {
        int k = (int)(u & 0x7FFFFFFu);
        off = S.leaf_table[2 * k];
        cnt = S.leaf_table[2 * k + 1];
    }
}

// ---------------------------------------------------------------------------
// 2D: BVH closest point (bvh.py:111-188 + line_segment.py:74-85).  Visit order
// and pruning are the reference's: both children tested in order with r_max
// tightened by each d_max, nearer child first, `cd > r_max` skips.
// ---------------------------------------------------------------------------
DWD CP closest2(const SubDev& S, F2 p) {
    const RBvh2* recs = (const RBvh2*)S.rec_bvh;
    const DSeg* prims = (const DSeg*)S.prims;
    Stk st;
    stack_init(st);
    CP its;
    its.p = f3(p.x, p.y, 0.0f);
    its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = kInf;
    its.u = its.v = 0.0f;
    its.d = kInf;
    its.prim = -1;
    int best = -1;
    float r_max = kInf;
    int ref = S.root_ref_bvh;
    float cd = kInf;
    while (true) {
        if (!(cd > r_max)) {
            if (ref < 0) {
                int off, cnt;
                leaf_range(S, ref, off, cnt);
                for (int j = 0; j < cnt; ++j) {
                    DSeg sg = prims[off + j];
                    F2 a = f2(sg.ax, sg.ay), b = f2(sg.bx, sg.by);
                    F2 pa = p - a, ba = b - a;
                    float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
                    F2 i = a + ba * h;
                    float d = norm(i - p);
                    if (d < its.d) {
                        its.d = d;
                        its.t = h;
                        its.p = f3(i.x, i.y, 0.0f);
                        best = off + j;
                    }
                }
                r_max = fmn(r_max, its.d);
            } else {
                RBvh2 rc = recs[ref];
                float dminL, dmaxL, dminR, dmaxR;
                box_dist(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), p, dminL, dmaxL);
                bool hitL = dminL <= r_max;
                r_max = fmn(r_max, dmaxL);
                box_dist(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), p, dminR, dmaxR);
                bool hitR = dminR <= r_max;
                r_max = fmn(r_max, dmaxR);
                if (hitL && hitR) {
                    bool sw = dminR < dminL;  // `d_max_right < d_max_right` (bvh.py:163) never swaps
                    st.push(sw ? rc.ref_l : rc.ref_r, sw ? dminL : dminR);
                    ref = sw ? rc.ref_r : rc.ref_l;
                    cd = sw ? dminR : dminL;
                    continue;
                } else if (hitL) {
                    ref = rc.ref_l;
                    cd = dminL;
                    continue;
                } else if (hitR) {
                    ref = rc.ref_r;
                    cd = dminR;
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        st.pop(ref, cd);
    }
    if (best >= 0) {  // normal of the winning segment (LineSegment.closest_point)
        DSeg sg = prims[best];
        F2 t = normalize(f2(sg.bx - sg.ax, sg.by - sg.ay));
        its.n = f3(t.y, -t.x, 0.0f);
        its.prim = S.prim_index[best];
    }
    return its;
}

// BVH.intersect (bvh.py:190-257): the first valid hit in traversal order wins
// (`r_max = its.d` reads the zero-initialised record, :216), so the traversal
// stops at the first hit.
DWD Hit intersect2(const SubDev& S, F2 x, F2 v, float r_max) {
    const RBvh2* recs = (const RBvh2*)S.rec_bvh;
    const DSeg* prims = (const DSeg*)S.prims;
    Hit its;
    its.valid = false;
    its.on_boundary = false;
    its.p = its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = its.u = its.v = its.d = 0.0f;
    its.prim = 0;
    float tmin;
    if (!box_ray(f2(S.root_mn[0], S.root_mn[1]), f2(S.root_mx[0], S.root_mx[1]), x, v, r_max, tmin)) return its;
    Stk st;
    stack_init(st);
    int ref = S.root_ref_bvh;
    float cd = tmin;
    while (true) {
        if (cd <= r_max) {
            if (ref < 0) {
                int off, cnt;
                leaf_range(S, ref, off, cnt);
                for (int j = 0; j < cnt; ++j) {  // LineSegment.ray_intersect (line_segment.py:47-72)
                    DSeg sg = prims[off + j];
                    F2 a = f2(sg.ax, sg.ay), b = f2(sg.bx, sg.by);
                    F2 u = a - x, w = b - a;
                    float dv = v.x * w.y - v.y * w.x;
                    if (!(__builtin_fabsf(dv) > kDrEps)) continue;
                    float ud = u.x * v.y - u.y * v.x;
                    float s_ = ud / dv;
                    if (!((s_ >= 0.0f) && (s_ <= 1.0f))) continue;
                    float uv = u.x * w.y - u.y * w.x;
                    float t = uv / dv;
                    if (!((t >= 0.0f) && (t <= r_max))) continue;
                    float d = __builtin_fabsf(t);
                    if (!(d < r_max)) continue;
                    F2 n = normalize(f2(w.y, -w.x));
                    F2 hp = x + v * t;
                    its.valid = true;
                    its.on_boundary = true;
                    its.p = f3(hp.x, hp.y, 0.0f);
                    its.n = f3(n.x, n.y, 0.0f);
                    its.t = s_;
                    its.d = d;
                    its.prim = S.prim_index[off + j];
                    return its;
                }
            } else {
                RBvh2 rc = recs[ref];
                float t0, t1;
                bool h0 = box_ray(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), x, v, r_max, t0);
                bool h1 = box_ray(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), x, v, r_max, t1);
                if (h0 && h1) {
                    bool sw = t1 < t0;
                    st.push(sw ? rc.ref_l : rc.ref_r, sw ? t0 : t1);
                    ref = sw ? rc.ref_r : rc.ref_l;
                    cd = sw ? t1 : t0;
                    continue;
                } else if (h0) {
                    ref = rc.ref_l;
                    cd = t0;
                    continue;
                } else if (h1) {
                    ref = rc.ref_r;
                    cd = t1;
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        st.pop(ref, cd);
    }
    return its;
}

// project_to_plane (bvh.py:16-20, snch3d.py:30-45)
DWD float project_to_plane(F2 n, F2 e) { return __builtin_fabsf(dot(e, vabs(f2(-n.y, n.x)))); }
DWD float project_to_plane(F3 n, F3 e) {
    float sign = __builtin_copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    F3 b1 = f3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    F3 b2 = f3(b, sign + n.y * n.y * a, -n.y);
    float r1 = dot(e, vabs(b1));
    float r2 = dot(e, vabs(b2));
    return __builtin_sqrtf(r1 * r1 + r2 * r2);
}
DWD bool in_range(float x, float a, float b) { return (x >= a) && (x <= b); }

// SNCHNode.overlaps (snch.py:46-97, snch3d.py:65-115)
template <typename V>
DWD bool snch_overlaps(V mn, V mx, V axis, float ha, float radius, V x, float r_max, float& d_min) {
    float d_max;
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    if ((ha >= 0.0f) && ov) {
        ov = (ha > kPi2) && (d_min < kDrEps);
        if (!ov) {
            V cen = (mn + mx) / 2.0f;
            V vca = cen - x;
            float len = norm(vca);
            vca = vca / len;
            float dax = cr_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
            ov = in_range(kPi2, dax - ha, dax + ha);
            if (!ov) {
                if (len > radius) {
                    float hs = ha + cr_asin(radius / len);
                    ov = (hs > kPi2) || in_range(kPi2, dax - hs, dax + hs);
                } else {
                    V e = mx - cen;
                    float d = dot(e, vabs(vca));
                    float s = len - d;
                    ov = s < 0.0f;
                    if (!ov) {
                        d = project_to_plane(vca, e);
                        float hs = ha + cr_atan2(d, s);
                        ov = (hs > kPi2) || in_range(kPi2, dax - hs, dax + hs);
                    }
                }
            }
        }
    }
    return ov;
}

// The same decision without transcendentals.  With c = |axis . vca| = |cos dax|
// and theta < pi/2:  |dax - pi/2| <= theta  <=>  c <= sin(theta);
// ha + vch > pi/2  <=>  cos(ha + vch) < 0.  sin/cos of the view-cone half angle
// come from radius/len (asin branch) or (d, s)/hypot(d, s) (atan2 branch).
// kConeMargin makes every comparison conservative: a node the literal test
// visits is always visited (its fp32 rounding error is below 1e-6 in cosine
// space), so results differ only where the literal test's own rounding
// prunes a node that does hold the nearest silhouette.
constexpr float kConeMargin = 4e-6f;
template <typename V>
DWD bool cone_fast(V mn, V mx, V axis, float ha, float sin_ha, float cos_ha, float radius, V x) {
    if (ha > kPi2 - kConeMargin) return true;
    V cen = (mn + mx) / 2.0f;
    V vca = cen - x;
    float len = norm(vca);
    vca = vca / len;
    float c = __builtin_fabsf(clampf(dot(axis, vca), -1.0f, 1.0f));
    if (c <= sin_ha + kConeMargin) return true;
    float sv, cv;
    if (len > radius) {
        sv = radius / len;
        cv = __builtin_sqrtf(fmx(0.0f, 1.0f - sv * sv));
    } else {
        V e = mx - cen;
        float d = dot(e, vabs(vca));
        float s = len - d;
        if (s < 0.0f) return true;
        d = project_to_plane(vca, e);
        float h = __builtin_sqrtf(d * d + s * s);
        sv = d / h;
        cv = s / h;
    }
    if (cos_ha * cv - sin_ha * sv < kConeMargin) return true;
    return c <= (sin_ha * cv + cos_ha * sv) + kConeMargin;
}
template <typename V>
DWD bool snch_overlaps_fast(V mn, V mx, V axis, float ha, float sin_ha, float cos_ha, float radius, V x, float r_max,
                            float& d_min) {
    float d_max;
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    if (!((ha >= 0.0f) && ov)) return ov;
    return cone_fast(mn, mx, axis, ha, sin_ha, cos_ha, radius, x);
}

// Face-plane bound of a 3D SNCH node (built on the host with the fp32 error
// of the silhouette test folded in): true when every face adjacent to the
// node's edges sees x on the same side, so the node holds no silhouette.
DWD bool face_pruned(F3 mn, F3 mx, F3 axis, float rho, float elo, float ehi, F3 x) {
    F3 v = x - (mn + mx) / 2.0f;
    float cv = dot(axis, v);
    float sl = rho * norm(v);
    return ((cv - sl) - ehi > 0.0f) || ((cv + sl) - elo < 0.0f);
}
// One child of a 3D SNCH record: box overlap, face-plane bound, cone test.
template <bool X>
DWD bool snch3_child(F3 mn, F3 mx, F3 axis, float ha, float sin_ha, float cos_ha, float radius, float rho, float elo,
                     float ehi, F3 x, float r_max, float& d_min) {
    if constexpr (X) {
        return snch_overlaps(mn, mx, axis, ha, radius, x, r_max, d_min) &&
               !face_pruned(mn, mx, axis, rho, elo, ehi, x);
    } else {
        float d_max;
        box_dist(mn, mx, x, d_min, d_max);
        if (!(d_min <= r_max)) return false;
        if (face_pruned(mn, mx, axis, rho, elo, ehi, x)) return false;
        if (!(ha >= 0.0f)) return true;
        return cone_fast(mn, mx, axis, ha, sin_ha, cos_ha, radius, x);
    }
}

// SNCH.closest_silhouette 2D (snch.py:312-378 + silhouette_vertex.py:80-93)
template <bool X = true>
DWD bool closest_sil2(const SubDev& S, F2 x, float r_max, float& d_out, int& ref_out, F2& p_out) {
    const RSnch2* recs = (const RSnch2*)S.rec_snch;
    const DSilV* sils = (const DSilV*)S.sils;
    bool found = false;
    d_out = 0.0f;
    ref_out = -1;
    float dmin = box_dmin(f2(S.root_mn[0], S.root_mn[1]), f2(S.root_mx[0], S.root_mx[1]), x);
    if (!(dmin <= r_max)) return false;
    Stk st;
    stack_init(st);
    int ref = S.root_ref_snch;
    float cd = dmin;
    while (true) {
        if (cd <= r_max) {
            if (ref < 0) {
                int off, cnt;
                leaf_range(S, ref, off, cnt);
                for (int j = 0; j < cnt; ++j) {
                    DSilV sv = sils[off + j];
                    F2 a = f2(sv.ax, sv.ay), b = f2(sv.bx, sv.by), c = f2(sv.cx, sv.cy);
                    float d = norm(x - b);
                    if (sv.both && (d < r_max)) {
                        if (cross2(b - a, x - a) * cross2(c - b, x - b) < 0.0f) {
                            r_max = fmn(r_max, d);
                            found = true;
                            d_out = d;
                            ref_out = off + j;
                            p_out = b;
                        }
                    }
                }
            } else {
                RSnch2 rc = recs[ref];
                float d0, d1;
                bool h0, h1;
                if constexpr (X) {
                    h0 = snch_overlaps(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), f2(rc.lax, rc.lay), rc.lha, rc.lrad,
                                       x, r_max, d0);
                    h1 = snch_overlaps(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), f2(rc.rax, rc.ray), rc.rha, rc.rrad,
                                       x, r_max, d1);
                } else {
                    h0 = snch_overlaps_fast(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), f2(rc.lax, rc.lay), rc.lha,
                                            rc.lsin, rc.lcos, rc.lrad, x, r_max, d0);
                    h1 = snch_overlaps_fast(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), f2(rc.rax, rc.ray), rc.rha,
                                            rc.rsin, rc.rcos, rc.rrad, x, r_max, d1);
                }
                if (h0 && h1) {
                    bool sw = d1 < d0;
                    st.push(sw ? rc.ref_l : rc.ref_r, sw ? d0 : d1);
                    ref = sw ? rc.ref_r : rc.ref_l;
                    cd = sw ? d1 : d0;
                    continue;
                } else if (h0) {
                    ref = rc.ref_l;
                    cd = d0;
                    continue;
                } else if (h1) {
                    ref = rc.ref_r;
                    cd = d1;
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        st.pop(ref, cd);
    }
    return found;
}

// SNCH3D.closest_silhouette (snch3d.py:594-663 + silhouette_edge.py:83-97)
template <bool X = true>
DWD bool closest_sil3(const SubDev& S, F3 x, float r_max, float& d_out, int& ref_out, F3& p_out) {
    const RSnch3* recs = (const RSnch3*)S.rec_snch;
    const DSilE* sils = (const DSilE*)S.sils;
    bool found = false;
    d_out = 0.0f;
    ref_out = -1;
    float dmin = box_dmin(f3(S.root_mn[0], S.root_mn[1], S.root_mn[2]), f3(S.root_mx[0], S.root_mx[1], S.root_mx[2]), x);
    if (!(dmin <= r_max)) return false;
    Stk st;
    stack_init(st);
    int ref = S.root_ref_snch;
    float cd = dmin;
    while (true) {
        if (cd <= r_max) {
            if (ref < 0) {
                int off, cnt;
                leaf_range(S, ref, off, cnt);
                for (int j = 0; j < cnt; ++j) {
                    DSilE se = sils[off + j];
                    F3 a = f3(se.ax, se.ay, se.az), b = f3(se.bx, se.by, se.bz);
                    F3 u = b - a, w = x - a;
                    float t = clampf(dot(w, u) / dot(u, u), 0.0f, 1.0f);
                    F3 pt = lerp(a, b, t);
                    float d = norm(pt - x);
                    if (d < r_max) {
                        bool is = se.boundary != 0;
                        if (!is) {
                            F3 c = f3(se.cx, se.cy, se.cz), dd = f3(se.dx, se.dy, se.dz);
                            F3 n0 = normalize(cross(b - a, c - a));
                            F3 n1 = normalize(cross(dd - a, b - a));
                            F3 view = pt - x;
                            is = dot(view, n0) * dot(view, n1) < 0.0f;
                        }
                        if (is) {
                            r_max = fmn(r_max, d);
                            found = true;
                            d_out = d;
                            ref_out = off + j;
                            p_out = pt;
                        }
                    }
                }
            } else {
                RSnch3 rc = recs[ref];
                F3 lmn = f3(rc.lmnx, rc.lmny, rc.lmnz), lmx = f3(rc.lmxx, rc.lmxy, rc.lmxz);
                F3 rmn = f3(rc.rmnx, rc.rmny, rc.rmnz), rmx = f3(rc.rmxx, rc.rmxy, rc.rmxz);
                float d0, d1;
                bool h0 = snch3_child<X>(lmn, lmx, f3(rc.lax, rc.lay, rc.laz), rc.lha, rc.lsin, rc.lcos, rc.lrad,
                                         rc.lrho, rc.lelo, rc.lehi, x, r_max, d0);
                bool h1 = snch3_child<X>(rmn, rmx, f3(rc.rax, rc.ray, rc.raz), rc.rha, rc.rsin, rc.rcos, rc.rrad,
                                         rc.rrho, rc.relo, rc.rehi, x, r_max, d1);
                if (h0 && h1) {
                    bool sw = d1 < d0;
                    st.push(sw ? rc.ref_l : rc.ref_r, sw ? d0 : d1);
                    ref = sw ? rc.ref_r : rc.ref_l;
                    cd = sw ? d1 : d0;
                    continue;
                } else if (h0) {
                    ref = rc.ref_l;
                    cd = d0;
                    continue;
                } else if (h1) {
                    ref = rc.ref_r;
                    cd = d1;
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        st.pop(ref, cd);
    }
    return found;
}

// Polyline.star_radius (polyline.py:304-314) / Mesh.star_radius (mesh.py:337-346)
template <bool X = true>
DWD float star_radius2(const SubDev& S, F2 x, float r_max) {
    F2 m = f2(fmx(__builtin_fabsf(x.x - S.bb_min[0]), __builtin_fabsf(x.x - S.bb_max[0])),
              fmx(__builtin_fabsf(x.y - S.bb_min[1]), __builtin_fabsf(x.y - S.bb_max[1])));
    float R = fmn(norm(m), r_max);
    if (!S.py_sils) return R;
    float d;
    int ref;
    F2 p;
    return closest_sil2<X>(S, x, R, d, ref, p) ? d : R;
}
template <bool X = true>
DWD float star_radius3(const SubDev& S, F3 x, float r_max) {
    F3 m = f3(fmx(__builtin_fabsf(x.x - S.bb_min[0]), __builtin_fabsf(x.x - S.bb_max[0])),
              fmx(__builtin_fabsf(x.y - S.bb_min[1]), __builtin_fabsf(x.y - S.bb_max[1])),
              fmx(__builtin_fabsf(x.z - S.bb_min[2]), __builtin_fabsf(x.z - S.bb_max[2])));
    float R = fmn(norm(m), r_max);
    if (!S.py_sils) return R;
    float d;
    int ref;
    F3 p;
    return closest_sil3<X>(S, x, R, d, ref, p) ? d : R;
}

// ---------------------------------------------------------------------------
// 3D: closest point on triangles (math.py:257-322) and SNCH3D.closest_point
// ---------------------------------------------------------------------------
DWD void tri_closest(const DTri& tr, F3 p, F3& pt, float& u_out, float& v_out, float& d) {
    F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
    F3 ab = b - a, ac = c - a;
    F3 ax = p - a;
    float d1 = dot(ab, ax), d2 = dot(ac, ax);
    F3 bx = p - b;
    float d3 = dot(ab, bx), d4 = dot(ac, bx);
    F3 cx = p - c;
    float d5 = dot(ab, cx), d6 = dot(ac, cx);
    float u0, u1;
    if ((d1 <= 0.0f) && (d2 <= 0.0f)) {
        pt = a, u0 = 1.0f, u1 = 0.0f;
    } else if ((d3 >= 0.0f) && (d4 <= d3)) {
        pt = b, u0 = 0.0f, u1 = 1.0f;
    } else if ((d6 >= 0.0f) && (d5 <= d6)) {
        pt = c, u0 = 0.0f, u1 = 0.0f;
    } else {
        float vc = d1 * d4 - d3 * d2;
        float vb = d5 * d2 - d1 * d6;
        float va = d3 * d6 - d5 * d4;
        if ((vc <= 0.0f) && (d1 >= 0.0f) && (d3 <= 0.0f)) {
            float v = d1 / (d1 - d3);
            pt = a + ab * v;
            u0 = 1.0f - v, u1 = v;
        } else if ((vb <= 0.0f) && (d2 >= 0.0f) && (d6 <= 0.0f)) {
            float w = d2 / (d2 - d6);
            pt = a + ac * w;
            u0 = 1.0f - w, u1 = 0.0f;
        } else if ((va <= 0.0f) && ((d4 - d3) >= 0.0f) && ((d5 - d6) >= 0.0f)) {
            float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            pt = b + (c - b) * w;
            u0 = 0.0f, u1 = 1.0f - w;
        } else {
            float denom = 1.0f / (va + vb + vc);
            float v = vb * denom, w = vc * denom;
            pt = (a + ab * v) + ac * w;
            u0 = 1.0f - v - w, u1 = v;
        }
    }
    d = norm(p - pt);
    u_out = u1;
    v_out = 1.0f - u0 - u1;
}

DWD CP closest3(const SubDev& S, F3 p) {
    const RBvh3* recs = (const RBvh3*)S.rec_bvh;
    const DTri* prims = (const DTri*)S.prims;
    Stk st;
    stack_init(st);
    CP its;
    its.p = its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = its.u = its.v = 0.0f;
    its.d = kInf;
    its.prim = -1;
    int best = -1;
    float r_max = kInf;
    int ref = S.root_ref_bvh;
    float cd = kInf;
    while (true) {
        if (cd <= r_max) {
            if (ref < 0) {
                int off, cnt;
                leaf_range(S, ref, off, cnt);
                for (int j = 0; j < cnt; ++j) {
                    DTri tr = prims[off + j];
                    F3 pt;
                    float u, v, d;
                    tri_closest(tr, p, pt, u, v, d);
                    if (d < its.d) {
                        its.d = d;
                        its.p = pt;
                        its.u = u;
                        its.v = v;
                        best = off + j;
                        r_max = fmn(r_max, d);
                    }
                }
            } else {
                RBvh3 rc = recs[ref];
                float dminL, dmaxL, dminR, dmaxR;
                box_dist(f3(rc.lmnx, rc.lmny, rc.lmnz), f3(rc.lmxx, rc.lmxy, rc.lmxz), p, dminL, dmaxL);
                bool hitL = dminL <= r_max;
                r_max = fmn(r_max, dmaxL);
                box_dist(f3(rc.rmnx, rc.rmny, rc.rmnz), f3(rc.rmxx, rc.rmxy, rc.rmxz), p, dminR, dmaxR);
                bool hitR = dminR <= r_max;
                r_max = fmn(r_max, dmaxR);
                if (hitL && hitR) {
                    // snch3d.py:383-388: equal zero d_min -> nearer d_max first; then d_min order
                    int closer = rc.ref_l, other = rc.ref_r;
                    if ((dminL == 0.0f) && (dminR == 0.0f) && (dmaxR < dmaxL)) closer = rc.ref_r, other = rc.ref_l;
                    float dc = dminL, dO = dminR;
                    if (dminR < dminL) {
                        int t = closer;
                        closer = other;
                        other = t;
                        dc = dminR;
                        dO = dminL;
                    }
                    st.push(other, dO);
                    ref = closer;
                    cd = dc;
                    continue;
                } else if (hitL) {
                    ref = rc.ref_l;
                    cd = dminL;
                    continue;
                } else if (hitR) {
                    ref = rc.ref_r;
                    cd = dminR;
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        st.pop(ref, cd);
    }
    if (best >= 0) {
        DTri tr = prims[best];
        F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
        its.n = normalize(cross(b - a, c - a));
        its.prim = tr.index;
    }
    return its;
}

// SNCH3D.intersect (snch3d.py:411-476) + Triangle.ray_intersect (triangle.py:59-86);
// prim = BVH-sorted index (triangle.py:82)
DWD Hit intersect3(const SubDev& S, F3 x, F3 v, float r_max) {
    const RBvh3* recs = (const RBvh3*)S.rec_bvh;
    const DTri* prims = (const DTri*)S.prims;
    Hit its;
    its.valid = false;
    its.on_boundary = false;
    its.p = its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = its.u = its.v = its.d = 0.0f;
    its.prim = 0;
    int best = -1;
    float bv = 0.0f, bw = 0.0f;
    float tmin;
    if (!box_ray(f3(S.root_mn[0], S.root_mn[1], S.root_mn[2]), f3(S.root_mx[0], S.root_mx[1], S.root_mx[2]), x, v,
                 r_max, tmin))
        return its;
    Stk st;
    stack_init(st);
    int ref = S.root_ref_bvh;
    float cd = tmin;
    while (true) {
        if (cd <= r_max) {
            if (ref < 0) {
                int off, cnt;
                leaf_range(S, ref, off, cnt);
                for (int j = 0; j < cnt; ++j) {
                    DTri tr = prims[off + j];
                    F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
                    F3 v1 = b - a, v2 = c - a;
                    F3 pp = cross(v, v2);
                    float det = dot(v1, pp);
                    if (!(__builtin_fabsf(det) > kDrEps)) continue;
                    float inv = 1.0f / det;
                    F3 s = x - a;
                    float bu = dot(s, pp) * inv;
                    if (!((bu >= 0.0f) && (bu <= 1.0f))) continue;
                    F3 q = cross(s, v1);
                    float bw_ = dot(v, q) * inv;
                    if (!((bw_ >= 0.0f) && (bu + bw_ <= 1.0f))) continue;
                    float t = dot(v2, q) * inv;
                    if (!((t >= 0.0f) && (t <= r_max))) continue;
                    if (!(t < r_max)) continue;
                    r_max = t;
                    best = off + j;
                    bv = bu;
                    bw = bw_;
                }
            } else {
                RBvh3 rc = recs[ref];
                float t0, t1;
                bool h0 = box_ray(f3(rc.lmnx, rc.lmny, rc.lmnz), f3(rc.lmxx, rc.lmxy, rc.lmxz), x, v, r_max, t0);
                bool h1 = box_ray(f3(rc.rmnx, rc.rmny, rc.rmnz), f3(rc.rmxx, rc.rmxy, rc.rmxz), x, v, r_max, t1);
                if (h0 && h1) {
                    bool sw = t1 < t0;
                    st.push(sw ? rc.ref_l : rc.ref_r, sw ? t0 : t1);
                    ref = sw ? rc.ref_r : rc.ref_l;
                    cd = sw ? t1 : t0;
                    continue;
                } else if (h0) {
                    ref = rc.ref_l;
                    cd = t0;
                    continue;
                } else if (h1) {
                    ref = rc.ref_r;
                    cd = t1;
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        st.pop(ref, cd);
    }
    if (best >= 0) {
        DTri tr = prims[best];
        F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
        F3 v1 = b - a, v2 = c - a;
        its.valid = true;
        its.on_boundary = true;
        its.p = (a + v1 * bv) + v2 * bw;
        its.n = normalize(cross(v1, v2));
        its.u = 1.0f - bv - bw;
        its.v = bv;
        its.d = r_max;
        its.prim = best;
    }
    return its;
}

}  // namespace dwd
