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
// Rarely taken paths kept out of line, so their registers do not count
// against the hot loops that call them.
#define DWD_COLD __device__ __attribute__((noinline))

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
// Bracketed decisions.  The SNCH cone tests use the once-rounded acos / asin /
// atan2 / cos above only through comparisons.  The fp32 library functions
// acosf, asinf, atanf and cosf stay within kTrSteps float steps of the
// once-rounded values on every float of the domains below (dw_math_check
// measures the largest distance over each whole domain), so one fp32
// evaluation gives an interval FI that holds the value the test computes.
// Correctly rounded +, -, *, / and fmin / fmax are monotone in each operand,
// so the test's own fp32 expressions evaluated at interval ends bound every
// later value exactly as the test rounds it.  A comparison whose interval lies
// on one side of its threshold has the test's outcome (kT / kF); otherwise the
// outcome is open (kU) and the caller runs the fp64 form.  NaN anywhere gives
// kU.
// ---------------------------------------------------------------------------
constexpr int kTrSteps = 4;                     // bound assumed for the four functions
constexpr float kCosLo = -4.0f, kCosHi = 8.0f;  // cosf domain the bound covers
enum { kF = 0, kT = 1, kU = 2 };
struct FI {
    float lo, hi;
};
DWD FI fi(float x) { return FI{x, x}; }
DWD FI fi_nan() { return FI{__builtin_nanf(""), __builtin_nanf("")}; }
DWD bool fi_bad(FI a) { return __builtin_isnan(a.lo) || __builtin_isnan(a.hi); }
// [f - r, f + r] evaluated in fp64, widened by more than the final rounding
DWD FI fi_around(double f, double r) {
    double s = (__builtin_fabs(f) + r) * 0x1p-22 + 0x1p-148;
    return FI{(float)(f - r - s), (float)(f + r + s)};
}
// once-rounded value of a function whose fp32 version f is within kTrSteps
// steps (plus `extra` steps of argument rounding)
DWD FI fi_steps(float f, int extra) {
    double a = __builtin_fabs((double)f);
    return fi_around((double)f, (double)(kTrSteps + 1 + extra) * (a * 0x1p-23 + 0x1p-149));
}
DWD FI fi_acos(float c) { return fi_steps(acosf(c), 0); }  // c in [-1, 1]
DWD FI fi_asin(float q) { return fi_steps(asinf(q), 0); }  // q in [-1, 1]
// atan2(y, x) for y >= 0, x > 0 from atanf(y / x): y / x rounds within half a
// step and atan(z(1 + d)) - atan(z) <= |d| atan(z) for z >= 0
DWD FI fi_atan2(float y, float x) {
    if (!((y >= 0.0f) && (x > 0.0f))) return fi_nan();
    return fi_steps(atanf(y / x), 2);
}
// cos over an interval: Lipschitz 1 around the lower end
DWD FI fi_cos(FI a) {
    if (!((a.lo >= kCosLo) && (a.hi <= kCosHi) && (a.lo <= a.hi))) return fi_nan();
    double f = (double)cosf(a.lo);
    double w = (double)a.hi - (double)a.lo;
    FI r = fi_around(f, w + (double)(kTrSteps + 2) * ((__builtin_fabs(f) + w) * 0x1p-23 + 0x1p-149));
    return FI{fmx(r.lo, -1.0f), fmn(r.hi, 1.0f)};
}
DWD FI fi_add(FI a, FI b) { return FI{a.lo + b.lo, a.hi + b.hi}; }
DWD FI fi_sub(FI a, FI b) { return FI{a.lo - b.hi, a.hi - b.lo}; }
DWD FI fi_hull(float p0, float p1, float p2, float p3) {
    if (__builtin_isnan(p0) || __builtin_isnan(p1) || __builtin_isnan(p2) || __builtin_isnan(p3)) return fi_nan();
    return FI{fmn(fmn(p0, p1), fmn(p2, p3)), fmx(fmx(p0, p1), fmx(p2, p3))};
}
DWD FI fi_mul(FI a, FI b) { return fi_hull(a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi); }
DWD FI fi_div(FI a, FI b) {
    if (!((b.lo > 0.0f) || (b.hi < 0.0f))) return fi_nan();
    return fi_hull(a.lo / b.lo, a.lo / b.hi, a.hi / b.lo, a.hi / b.hi);
}
DWD FI fi_sq(FI a) {  // a * a of one value
    if (fi_bad(a)) return fi_nan();
    float l = a.lo * a.lo, h = a.hi * a.hi;
    if (a.lo >= 0.0f) return FI{l, h};
    if (a.hi <= 0.0f) return FI{h, l};
    return FI{0.0f, fmx(l, h)};
}
DWD FI fi_min(FI a, FI b) {
    if (fi_bad(a) || fi_bad(b)) return fi_nan();
    return FI{fmn(a.lo, b.lo), fmn(a.hi, b.hi)};
}
DWD FI fi_max(FI a, FI b) {
    if (fi_bad(a) || fi_bad(b)) return fi_nan();
    return FI{fmx(a.lo, b.lo), fmx(a.hi, b.hi)};
}
DWD int t_gt(FI a, float t) { return a.lo > t ? kT : (a.hi <= t ? kF : kU); }
DWD int t_ge(FI a, float t) { return a.lo >= t ? kT : (a.hi < t ? kF : kU); }
DWD int t_lt(FI a, float t) { return a.hi < t ? kT : (a.lo >= t ? kF : kU); }
DWD int t_le(FI a, float t) { return a.hi <= t ? kT : (a.lo > t ? kF : kU); }
DWD int t_and(int a, int b) { return (a == kF || b == kF) ? kF : ((a == kT && b == kT) ? kT : kU); }
DWD int t_or(int a, int b) { return (a == kT || b == kT) ? kT : ((a == kF && b == kF) ? kF : kU); }
// t in [lo, hi] (in_range below)
DWD int t_in(float t, FI lo, FI hi) { return t_and(t_le(lo, t), t_ge(hi, t)); }

// ---------------------------------------------------------------------------
// TEA seeding (diff_wost/core/math.py:87-107) and PCG32 (Dr.Jit PCG32)
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
// (stk_slot; blocks are whole waves), deeper entries in a private overflow
// array.  Traversals descend into the nearer child and push
// only the farther one, so the stack never holds more than the tree depth.
// Kernels that traverse must launch with kStackLdsBytes of dynamic LDS.
// ---------------------------------------------------------------------------
constexpr int kLdsStack = 16;
constexpr int kStackLdsBytesPerThread = kLdsStack * 8;
extern __shared__ int dw_lds[];

// Stack entries below the LDS part (per lane, private memory).  Kept apart
// from the traversal state so that state stays in registers.
struct Ovf {
    int oi[DW_MAX_STACK - kLdsStack];
    float od[DW_MAX_STACK - kLdsStack];
};
// LDS slots of this lane: each wave owns 2 * kLdsStack rows of 64 dwords
// (node references, then bounds), entry k of lane l in row k, column l, so a
// wave's accesses are conflict-free.  The slot depends on threadIdx.x only (no
// block-size read) and is addressed from dw_lds at each access so the accesses
// stay LDS instructions.
DWD int stk_slot(int k) {
    int tid = (int)threadIdx.x;
    return (tid & ~63) * (2 * kLdsStack) + k * 64 + (tid & 63);
}
DWD void stk_push(int& sp, Ovf& o, int r, float d) {
    if (sp < kLdsStack) {
        dw_lds[stk_slot(sp)] = r;
        ((float*)dw_lds)[stk_slot(sp + kLdsStack)] = d;
    } else {
        o.oi[sp - kLdsStack] = r;
        o.od[sp - kLdsStack] = d;
    }
    ++sp;
}
DWD void stk_pop(int& sp, Ovf& o, int& r, float& d) {
    --sp;
    if (sp < kLdsStack) {
        r = dw_lds[stk_slot(sp)];
        d = ((float*)dw_lds)[stk_slot(sp + kLdsStack)];
    } else {
        r = o.oi[sp - kLdsStack];
        d = o.od[sp - kLdsStack];
    }
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

template <int D>
struct VecT;
template <>
struct VecT<2> {
    using T = F2;
};
template <>
struct VecT<3> {
    using T = F3;
};
// ---------------------------------------------------------------------------
// Resumable traversals.  A query starts with *_begin, advances by one loop
// iteration (an interior record's two child tests, or one leaf) per *_iter
// call, which returns true once the traversal has finished, and is read with
// *_end.  The synchronous queries run the iterations back to back; the
// persistent walk kernels interleave them with the other lanes' walk logic, so
// a lane whose traversal ends early moves on instead of idling until the
// wave's longest traversal ends.  During the traversal only the winning
// primitive's slot is kept: *_end recomputes its record with the same
// arithmetic (the primitive tests are pure functions of query and primitive).
// ---------------------------------------------------------------------------
enum : int { TQ_CLOSEST = 0, TQ_SIL = 1, TQ_ISECT = 2, TQ_CLOSEST_W = 3 };  // _W: two-level records
struct Trav {
    const SubDev* S;
    const void* rec;    // child-in-parent records of the tree walked (BVH or SNCH)
    const void* prims;  // its primitives (segments / triangles) or silhouettes
    F3 x;         // query point / ray origin (2D: z = 0)
    F3 v;         // ray direction (intersection)
    float r_max;  // pruning radius; silhouettes: the running nearest distance
    float d;      // closest point: best distance so far
    float cd;     // lower bound of the pending node
    int ref;      // pending node (>= 0) or leaf (< 0) reference
    int best;     // winning primitive / silhouette slot, -1: none
    int kind;     // TQ_*
    int sp;       // traversal stack depth
};

// Tree records and primitives are read through global-address-space pointers:
// global loads wait on the vector-memory counter only, not on LDS traffic.
#define DW_GLOBAL __attribute__((address_space(1)))
template <typename T>
DWD T gload(const void* base, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const DW_GLOBAL T*)base)[i];
#else  // host pass of the device code: never executed
    return ((const T*)base)[i];
#endif
}

// 2D: BVH closest point (bvh.py:111-188 + line_segment.py:74-85).  Visit order
// and pruning are the reference's: both children tested in order with r_max
// tightened by each d_max, nearer child first, `cd > r_max` skips.
// r0 < inf starts the pruning radius at a known bound on the answer (the
// reference query starts at inf and pushes the root at that distance).
// wide: read the two-level records (node width 4), which visit the same nodes in
// the same order with the same arithmetic.
DWD void closest2_begin(Trav& t, const SubDev& S, F2 p, float r0 = kInf, bool wide = false) {
    t.kind = wide ? TQ_CLOSEST_W : TQ_CLOSEST;
    t.S = &S;
    t.rec = wide ? S.rec_wide : S.rec_bvh;
    t.prims = S.prims;
    t.x = f3(p.x, p.y, 0.0f);
    t.d = kInf;
    t.best = -1;
    t.r_max = r0;
    t.ref = S.root_ref_bvh;
    t.cd = r0 < kInf ? 0.0f : kInf;
    t.sp = 0;
}
// A 2D closest-point leaf: the nearest of its segments (first on ties).
DWD void closest2_leaf(Trav& t, F2 p) {
    int off, cnt;
    leaf_range(*t.S, t.ref, off, cnt);
    for (int j = 0; j < cnt; ++j) {
        DSeg sg = gload<DSeg>(t.prims, off + j);
        F2 a = f2(sg.ax, sg.ay), b = f2(sg.bx, sg.by);
        F2 pa = p - a, ba = b - a;
        float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        F2 i = a + ba * h;
        float d = norm(i - p);
        if (d < t.d) {
            t.d = d;
            t.best = off + j;
        }
    }
    t.r_max = fmn(t.r_max, t.d);
}
// The two children of an interior node (boxes (min x, min y, max x, max y) and
// references): both tested in order with r_max tightened by each d_max, the
// nearer descended into (the left one on ties: `d_max_right < d_max_right`,
// bvh.py:163, never swaps), the other pushed.  False when neither is hit.
DWD bool closest2_pair(Trav& t, Ovf& o, F2 p, float4 bl, float4 br, int ref_l, int ref_r) {
    float dminL, dmaxL, dminR, dmaxR;
    box_dist(f2(bl.x, bl.y), f2(bl.z, bl.w), p, dminL, dmaxL);
    bool hitL = dminL <= t.r_max;
    t.r_max = fmn(t.r_max, dmaxL);
    box_dist(f2(br.x, br.y), f2(br.z, br.w), p, dminR, dmaxR);
    bool hitR = dminR <= t.r_max;
    t.r_max = fmn(t.r_max, dmaxR);
    if (hitL && hitR) {
        bool sw = dminR < dminL;
        stk_push(t.sp, o, sw ? ref_l : ref_r, sw ? dminL : dminR);
        t.ref = sw ? ref_r : ref_l;
        t.cd = sw ? dminR : dminL;
        return true;
    } else if (hitL) {
        t.ref = ref_l;
        t.cd = dminL;
        return true;
    } else if (hitR) {
        t.ref = ref_r;
        t.cd = dminR;
        return true;
    }
    return false;
}
DWD bool closest2_iter(Trav& t, Ovf& o) {
    F2 p = f2(t.x.x, t.x.y);
    if (!(t.cd > t.r_max)) {
        if (t.ref < 0) {
            closest2_leaf(t, p);
        } else {
            RBvh2 rc = gload<RBvh2>(t.rec, t.ref);
            if (closest2_pair(t, o, p, make_float4(rc.lmnx, rc.lmny, rc.lmxx, rc.lmxy),
                              make_float4(rc.rmnx, rc.rmny, rc.rmxx, rc.rmxy), rc.ref_l, rc.ref_r))
                return false;
        }
    }
    if (t.sp == 0) return true;
    stk_pop(t.sp, o, t.ref, t.cd);
    return false;
}
// The same query over the two-level records: a record's first pair is the
// node's children; when the descent continues into an interior child, that
// child's pair (in the same record) is tested in the same iteration, so two
// binary levels cost one dependent fetch.  A descended-into child always
// passes the pending-node test (its d_min is at most every d_max that
// tightened r_max), so the node sequence, the stack and r_max evolve exactly
// as in closest2_iter.
DWD bool closest2w_iter(Trav& t, Ovf& o) {
    F2 p = f2(t.x.x, t.x.y);
    if (!(t.cd > t.r_max)) {
        if (t.ref < 0) {
            closest2_leaf(t, p);
        } else {
            const char* rec = (const char*)t.rec + (size_t)t.ref * sizeof(RWide2);
            float4 b0 = gload<float4>(rec, 0), b1 = gload<float4>(rec, 1);
            int4 r03 = gload<int4>(rec, 6);
            if (closest2_pair(t, o, p, b0, b1, r03.x, r03.y)) {
                if (t.ref < 0) return false;  // a leaf child: next iteration
                // the chosen child's pair, from the line just fetched
                const bool right = t.ref == r03.y;
                float4 c0 = gload<float4>(rec, right ? 4 : 2), c1 = gload<float4>(rec, right ? 5 : 3);
                int2 cr = gload<int2>(rec + 96, right ? 2 : 1);
                if (closest2_pair(t, o, p, c0, c1, cr.x, cr.y)) return false;
            }
        }
    }
    if (t.sp == 0) return true;
    stk_pop(t.sp, o, t.ref, t.cd);
    return false;
}
DWD CP closest2_end(const Trav& t) {
    const SubDev& S = *t.S;
    F2 p = f2(t.x.x, t.x.y);
    CP its;
    its.p = f3(p.x, p.y, 0.0f);
    its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = kInf;
    its.u = its.v = 0.0f;
    its.d = kInf;
    its.prim = -1;
    if (t.best >= 0) {  // the winner's record and normal (LineSegment.closest_point)
        DSeg sg = ((const DSeg*)S.prims)[t.best];
        F2 a = f2(sg.ax, sg.ay), b = f2(sg.bx, sg.by);
        F2 pa = p - a, ba = b - a;
        float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        F2 i = a + ba * h;
        its.d = norm(i - p);
        its.t = h;
        its.p = f3(i.x, i.y, 0.0f);
        F2 tg = normalize(f2(sg.bx - sg.ax, sg.by - sg.ay));
        its.n = f3(tg.y, -tg.x, 0.0f);
        its.prim = S.prim_index[t.best];
    }
    return its;
}
DWD CP closest2(const SubDev& S, F2 p) {
    Trav t;
    Ovf o;
    closest2_begin(t, S, p);
    while (!closest2_iter(t, o)) {
    }
    return closest2_end(t);
}

// BVH.intersect (bvh.py:190-257): the first valid hit in traversal order wins
// (`r_max = its.d` reads the zero-initialised record, :216), so the traversal
// stops at the first hit.  LineSegment.ray_intersect (line_segment.py:47-72).
DWD bool seg_hit(const DSeg& sg, F2 x, F2 v, float r_max, float& s_, float& t_) {
    F2 a = f2(sg.ax, sg.ay), b = f2(sg.bx, sg.by);
    F2 u = a - x, w = b - a;
    float dv = v.x * w.y - v.y * w.x;
    if (!(__builtin_fabsf(dv) > kDrEps)) return false;
    float ud = u.x * v.y - u.y * v.x;
    s_ = ud / dv;
    if (!((s_ >= 0.0f) && (s_ <= 1.0f))) return false;
    float uv = u.x * w.y - u.y * w.x;
    t_ = uv / dv;
    if (!((t_ >= 0.0f) && (t_ <= r_max))) return false;
    return __builtin_fabsf(t_) < r_max;
}
DWD bool isect2_begin(Trav& t, const SubDev& S, F2 x, F2 v, float r_max) {
    t.kind = TQ_ISECT;
    t.S = &S;
    t.rec = S.rec_bvh;
    t.prims = S.prims;
    t.x = f3(x.x, x.y, 0.0f);
    t.v = f3(v.x, v.y, 0.0f);
    t.r_max = r_max;
    t.best = -1;
    float tmin;
    if (!box_ray(f2(S.root_mn[0], S.root_mn[1]), f2(S.root_mx[0], S.root_mx[1]), x, v, r_max, tmin)) return false;
    t.sp = 0;
    t.ref = S.root_ref_bvh;
    t.cd = tmin;
    return true;
}
DWD bool isect2_iter(Trav& t, Ovf& o) {
    const SubDev& S = *t.S;
    F2 x = f2(t.x.x, t.x.y), v = f2(t.v.x, t.v.y);
    if (t.cd <= t.r_max) {
        if (t.ref < 0) {
            int off, cnt;
            leaf_range(S, t.ref, off, cnt);
            for (int j = 0; j < cnt; ++j) {
                float s_, t_;
                if (seg_hit(gload<DSeg>(t.prims, off + j), x, v, t.r_max, s_, t_)) {
                    t.best = off + j;
                    return true;
                }
            }
        } else {
            RBvh2 rc = gload<RBvh2>(t.rec, t.ref);
            float t0, t1;
            bool h0 = box_ray(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), x, v, t.r_max, t0);
            bool h1 = box_ray(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), x, v, t.r_max, t1);
            if (h0 && h1) {
                bool sw = t1 < t0;
                stk_push(t.sp, o, sw ? rc.ref_l : rc.ref_r, sw ? t0 : t1);
                t.ref = sw ? rc.ref_r : rc.ref_l;
                t.cd = sw ? t1 : t0;
                return false;
            } else if (h0) {
                t.ref = rc.ref_l;
                t.cd = t0;
                return false;
            } else if (h1) {
                t.ref = rc.ref_r;
                t.cd = t1;
                return false;
            }
        }
    }
    if (t.sp == 0) return true;
    stk_pop(t.sp, o, t.ref, t.cd);
    return false;
}
DWD Hit isect2_end(const Trav& t) {
    const SubDev& S = *t.S;
    Hit its;
    its.valid = false;
    its.on_boundary = false;
    its.p = its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = its.u = its.v = its.d = 0.0f;
    its.prim = 0;
    if (t.best >= 0) {
        DSeg sg = ((const DSeg*)S.prims)[t.best];
        F2 x = f2(t.x.x, t.x.y), v = f2(t.v.x, t.v.y);
        float s_ = 0.0f, t_ = 0.0f;
        seg_hit(sg, x, v, t.r_max, s_, t_);
        F2 w = f2(sg.bx - sg.ax, sg.by - sg.ay);
        F2 n = normalize(f2(w.y, -w.x));
        F2 hp = x + v * t_;
        its.valid = true;
        its.on_boundary = true;
        its.p = f3(hp.x, hp.y, 0.0f);
        its.n = f3(n.x, n.y, 0.0f);
        its.t = s_;
        its.d = __builtin_fabsf(t_);
        its.prim = S.prim_index[t.best];
    }
    return its;
}
DWD Hit intersect2(const SubDev& S, F2 x, F2 v, float r_max) {
    Trav t;
    Ovf o;
    if (isect2_begin(t, S, x, v, r_max))
        while (!isect2_iter(t, o)) {
        }
    return isect2_end(t);
}

// Nearest-hit form of the 2D ray query: the answer of the reference's
// brute-force Polyline.intersect_baseline (polyline.py:468-496), found through
// the BVH the way SNCH3D.intersect does in 3D (snch3d.py:435): each hit below
// the running r_max becomes the new r_max, so farther nodes are pruned and the
// first of equally near hits in traversal order is kept.
DWD bool isect2n_iter(Trav& t, Ovf& o) {
    const SubDev& S = *t.S;
    F2 x = f2(t.x.x, t.x.y), v = f2(t.v.x, t.v.y);
    if (t.cd <= t.r_max) {
        if (t.ref < 0) {
            int off, cnt;
            leaf_range(S, t.ref, off, cnt);
            for (int j = 0; j < cnt; ++j) {
                float s_, t_;
                if (seg_hit(gload<DSeg>(t.prims, off + j), x, v, t.r_max, s_, t_)) {
                    t.r_max = t_;
                    t.best = off + j;
                }
            }
        } else {
            RBvh2 rc = gload<RBvh2>(t.rec, t.ref);
            float t0, t1;
            bool h0 = box_ray(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), x, v, t.r_max, t0);
            bool h1 = box_ray(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), x, v, t.r_max, t1);
            if (h0 && h1) {
                bool sw = t1 < t0;
                stk_push(t.sp, o, sw ? rc.ref_l : rc.ref_r, sw ? t0 : t1);
                t.ref = sw ? rc.ref_r : rc.ref_l;
                t.cd = sw ? t1 : t0;
                return false;
            } else if (h0) {
                t.ref = rc.ref_l;
                t.cd = t0;
                return false;
            } else if (h1) {
                t.ref = rc.ref_r;
                t.cd = t1;
                return false;
            }
        }
    }
    if (t.sp == 0) return true;
    stk_pop(t.sp, o, t.ref, t.cd);
    return false;
}
DWD Hit intersect2_nearest(const SubDev& S, F2 x, F2 v, float r_max) {
    Trav t;
    Ovf o;
    if (isect2_begin(t, S, x, v, r_max))
        while (!isect2n_iter(t, o)) {
        }
    t.r_max = r_max;  // the record re-tests the winner against the caller's r_max
    return isect2_end(t);
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

// SNCHNode.overlaps (snch.py:46-97, snch3d.py:65-115): the box test, then for a
// valid cone (ha >= 0) the literal cone test below.
template <typename V>
DWD bool snch_cone_literal(V mn, V mx, V axis, float ha, float radius, V x, float d_min);
template <typename V>
DWD bool snch_overlaps(V mn, V mx, V axis, float ha, float radius, V x, float r_max, float& d_min) {
    float d_max;
    box_dist(mn, mx, x, d_min, d_max);
    return (d_min <= r_max) && (!(ha >= 0.0f) || snch_cone_literal(mn, mx, axis, ha, radius, x, d_min));
}
// The literal test with its transcendentals bracketed: kT / kF is the literal
// test's outcome, kU leaves it open.
template <typename V>
DWD int cone_literal_bracketed(V mn, V mx, V axis, float ha, float radius, V x, float d_min) {
    if ((ha > kPi2) && (d_min < kDrEps)) return kT;
    V cen = (mn + mx) / 2.0f;
    V vca = cen - x;
    float len = norm(vca);
    vca = vca / len;
    FI dax = fi_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
    int t0 = t_in(kPi2, fi_sub(dax, fi(ha)), fi_add(dax, fi(ha)));
    if (t0 == kT) return kT;
    FI vh;
    if (len > radius) {
        vh = fi_asin(radius / len);
    } else {
        V e = mx - cen;
        float d = dot(e, vabs(vca));
        float s = len - d;
        if (s < 0.0f) return kT;
        d = project_to_plane(vca, e);
        vh = fi_atan2(d, s);
    }
    FI hs = fi_add(fi(ha), vh);
    return t_or(t0, t_or(t_gt(hs, kPi2), t_in(kPi2, fi_sub(dax, hs), fi_add(dax, hs))));
}
template <typename V>
DWD_COLD bool snch_cone_literal_fp64(V mn, V mx, V axis, float ha, float radius, V x, float d_min);
template <typename V>
DWD bool snch_cone_literal(V mn, V mx, V axis, float ha, float radius, V x, float d_min) {
    int t = cone_literal_bracketed(mn, mx, axis, ha, radius, x, d_min);
    if (t != kU) return t == kT;
    return snch_cone_literal_fp64(mn, mx, axis, ha, radius, x, d_min);
}
template <typename V>
DWD_COLD bool snch_cone_literal_fp64(V mn, V mx, V axis, float ha, float radius, V x, float d_min) {
    bool ov;
    {
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
// One child of a 3D SNCH record: box overlap (d_min from box_dist), face-plane
// bound, cone test.
template <bool X>
DWD bool snch3_child(F3 mn, F3 mx, F3 axis, float ha, float sin_ha, float cos_ha, float radius, float rho, float elo,
                     float ehi, F3 x, float r_max, float d_min) {
    if constexpr (X) {
        return (d_min <= r_max) && (!(ha >= 0.0f) || snch_cone_literal(mn, mx, axis, ha, radius, x, d_min)) &&
               !face_pruned(mn, mx, axis, rho, elo, ehi, x);
    } else {
        if (!(d_min <= r_max)) return false;
        if (face_pruned(mn, mx, axis, rho, elo, ehi, x)) return false;
        if (!(ha >= 0.0f)) return true;
        return cone_fast(mn, mx, axis, ha, sin_ha, cos_ha, radius, x);
    }
}

// SNCH.closest_silhouette 2D (snch.py:312-378 + silhouette_vertex.py:80-93) and
// SNCH3D.closest_silhouette (snch3d.py:594-663 + silhouette_edge.py:83-97).
// A silhouette is accepted only below the running r_max, which then drops to
// its distance, so the nearest distance found is r_max at the end.
template <int D>
DWD bool sil_begin(Trav& t, const SubDev& S, typename VecT<D>::T x, float r_max);
template <>
DWD bool sil_begin<2>(Trav& t, const SubDev& S, F2 x, float r_max) {
    t.kind = TQ_SIL;
    t.S = &S;
    t.rec = S.rec_snch;
    t.prims = S.sils;
    t.x = f3(x.x, x.y, 0.0f);
    t.r_max = r_max;
    t.best = -1;
    float dmin = box_dmin(f2(S.root_mn[0], S.root_mn[1]), f2(S.root_mx[0], S.root_mx[1]), x);
    if (!(dmin <= r_max)) return false;
    t.sp = 0;
    t.ref = S.root_ref_snch;
    t.cd = dmin;
    return true;
}
template <>
DWD bool sil_begin<3>(Trav& t, const SubDev& S, F3 x, float r_max) {
    t.kind = TQ_SIL;
    t.S = &S;
    t.rec = S.rec_snch;
    t.prims = S.sils;
    t.x = x;
    t.r_max = r_max;
    t.best = -1;
    float dmin = box_dmin(f3(S.root_mn[0], S.root_mn[1], S.root_mn[2]), f3(S.root_mx[0], S.root_mx[1], S.root_mx[2]), x);
    if (!(dmin <= r_max)) return false;
    t.sp = 0;
    t.ref = S.root_ref_snch;
    t.cd = dmin;
    return true;
}
template <bool X>
DWD bool sil2_iter(Trav& t, Ovf& o) {
    const SubDev& S = *t.S;
    F2 x = f2(t.x.x, t.x.y);
    if (t.cd <= t.r_max) {
        if (t.ref < 0) {
            int off, cnt;
            leaf_range(S, t.ref, off, cnt);
            for (int j = 0; j < cnt; ++j) {
                DSilV sv = gload<DSilV>(t.prims, off + j);
                F2 a = f2(sv.ax, sv.ay), b = f2(sv.bx, sv.by), c = f2(sv.cx, sv.cy);
                float d = norm(x - b);
                if (sv.both && (d < t.r_max)) {
                    if (cross2(b - a, x - a) * cross2(c - b, x - b) < 0.0f) {
                        t.r_max = fmn(t.r_max, d);
                        t.best = off + j;
                    }
                }
            }
        } else {
            RSnch2 rc = gload<RSnch2>(t.rec, t.ref);
            float d0, d1;
            bool h0, h1;
            if constexpr (X) {
                h0 = snch_overlaps(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), f2(rc.lax, rc.lay), rc.lha, rc.lrad, x,
                                   t.r_max, d0);
                h1 = snch_overlaps(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), f2(rc.rax, rc.ray), rc.rha, rc.rrad, x,
                                   t.r_max, d1);
            } else {
                h0 = snch_overlaps_fast(f2(rc.lmnx, rc.lmny), f2(rc.lmxx, rc.lmxy), f2(rc.lax, rc.lay), rc.lha, rc.lsin,
                                        rc.lcos, rc.lrad, x, t.r_max, d0);
                h1 = snch_overlaps_fast(f2(rc.rmnx, rc.rmny), f2(rc.rmxx, rc.rmxy), f2(rc.rax, rc.ray), rc.rha, rc.rsin,
                                        rc.rcos, rc.rrad, x, t.r_max, d1);
            }
            if (h0 && h1) {
                bool sw = d1 < d0;
                stk_push(t.sp, o, sw ? rc.ref_l : rc.ref_r, sw ? d0 : d1);
                t.ref = sw ? rc.ref_r : rc.ref_l;
                t.cd = sw ? d1 : d0;
                return false;
            } else if (h0) {
                t.ref = rc.ref_l;
                t.cd = d0;
                return false;
            } else if (h1) {
                t.ref = rc.ref_r;
                t.cd = d1;
                return false;
            }
        }
    }
    if (t.sp == 0) return true;
    stk_pop(t.sp, o, t.ref, t.cd);
    return false;
}
// silhouette_edge.py:83-97: distance to the edge, then the two-face test
DWD bool sile_test(const DSilE& se, F3 x, float r_max, float& d, F3& pt) {
    F3 a = f3(se.ax, se.ay, se.az), b = f3(se.bx, se.by, se.bz);
    F3 u = b - a, w = x - a;
    float tt = clampf(dot(w, u) / dot(u, u), 0.0f, 1.0f);
    pt = lerp(a, b, tt);
    d = norm(pt - x);
    if (!(d < r_max)) return false;
    if (se.boundary != 0) return true;
    F3 c = f3(se.cx, se.cy, se.cz), dd = f3(se.dx, se.dy, se.dz);
    F3 n0 = normalize(cross(b - a, c - a));
    F3 n1 = normalize(cross(dd - a, b - a));
    F3 view = pt - x;
    return dot(view, n0) * dot(view, n1) < 0.0f;
}
template <bool X = true>
DWD bool closest_sil2(const SubDev& S, F2 x, float r_max, float& d_out, int& ref_out, F2& p_out) {
    Trav t;
    Ovf o;
    if (sil_begin<2>(t, S, x, r_max))
        while (!sil2_iter<X>(t, o)) {
        }
    d_out = 0.0f;
    ref_out = t.best;
    if (t.best < 0) return false;
    const DSilV& sv = ((const DSilV*)S.sils)[t.best];
    d_out = t.r_max;
    p_out = f2(sv.bx, sv.by);
    return true;
}
template <bool X = true>
DWD bool closest_sil3(const SubDev& S, F3 x, float r_max, float& d_out, int& ref_out, F3& p_out) {
    Trav t;
    Ovf o;
    if (sil_begin<3>(t, S, x, r_max))
        while (!trav3_iter<X>(t, o)) {
        }
    d_out = 0.0f;
    ref_out = t.best;
    if (t.best < 0) return false;
    float d;
    sile_test(((const DSilE*)S.sils)[t.best], x, kInf, d, p_out);
    d_out = t.r_max;
    return true;
}

// Polyline.star_radius (polyline.py:304-314) / Mesh.star_radius (mesh.py:337-346):
// the silhouette search runs inside min(|x - farthest box corner|, r_max).
DWD float star_bound(const SubDev& S, F2 x, float r_max) {
    F2 m = f2(fmx(__builtin_fabsf(x.x - S.bb_min[0]), __builtin_fabsf(x.x - S.bb_max[0])),
              fmx(__builtin_fabsf(x.y - S.bb_min[1]), __builtin_fabsf(x.y - S.bb_max[1])));
    return fmn(norm(m), r_max);
}
DWD float star_bound(const SubDev& S, F3 x, float r_max) {
    F3 m = f3(fmx(__builtin_fabsf(x.x - S.bb_min[0]), __builtin_fabsf(x.x - S.bb_max[0])),
              fmx(__builtin_fabsf(x.y - S.bb_min[1]), __builtin_fabsf(x.y - S.bb_max[1])),
              fmx(__builtin_fabsf(x.z - S.bb_min[2]), __builtin_fabsf(x.z - S.bb_max[2])));
    return fmn(norm(m), r_max);
}
// Bounds [lo, hi] on every fp32 distance the queries compute from x to a point
// of the sub-scene: its vertex box widened by 1e-3 (1 + max |coordinate|), which
// holds every node box, closest point and silhouette point of the sub-scene;
// lo / hi take the per-axis nearest / farthest offsets through the same norm.
// fp32 subtraction, fma, sqrt, min and max round monotonically, so the box
// distances d_min / d_max of every node and the distance to every primitive
// or silhouette point lie in [lo, hi].  Only for |x_i| <= 2^59 / (1 + max
// |coordinate|): there no product of an edge with an offset from x, no square
// of such a product and no squared offset overflows (each stays below 2^126);
// otherwise returns false.
DWD bool dist_bounds(const SubDev& S, F3 x, float& lo, float& hi) {
    float big = fmx(fmx(fmx(__builtin_fabsf(S.bb_min[0]), __builtin_fabsf(S.bb_max[0])),
                        fmx(__builtin_fabsf(S.bb_min[1]), __builtin_fabsf(S.bb_max[1]))),
                    fmx(__builtin_fabsf(S.bb_min[2]), __builtin_fabsf(S.bb_max[2])));
    const float lim = 0x1p59f / (1.0f + big);
    if (!(__builtin_fabsf(x.x) <= lim && __builtin_fabsf(x.y) <= lim && __builtin_fabsf(x.z) <= lim)) return false;
    float m = 1e-3f * (1.0f + big);
    const float xs[3] = {x.x, x.y, x.z};
    float l[3], h[3];
    for (int k = 0; k < 3; ++k) {
        float a = S.bb_min[k] - m, b = S.bb_max[k] + m;
        l[k] = xs[k] < a ? a - xs[k] : (xs[k] > b ? xs[k] - b : 0.0f);
        h[k] = fmx(__builtin_fabsf(xs[k] - a), __builtin_fabsf(xs[k] - b));
    }
    lo = norm(f3(l[0], l[1], l[2]));
    hi = norm(f3(h[0], h[1], h[2]));
    return true;
}

template <bool X = true>
DWD float star_radius2(const SubDev& S, F2 x, float r_max) {
    float R = star_bound(S, x, r_max);
    if (!S.py_sils) return R;
    float d;
    int ref;
    F2 p;
    return closest_sil2<X>(S, x, R, d, ref, p) ? d : R;
}
template <bool X = true>
DWD float star_radius3(const SubDev& S, F3 x, float r_max) {
    float R = star_bound(S, x, r_max);
    if (!S.py_sils) return R;
    float d;
    int ref;
    F3 p;
    return closest_sil3<X>(S, x, R, d, ref, p) ? d : R;
}


// ---------------------------------------------------------------------------
// 3D: closest point on triangles (math.py:207-272) and SNCH3D.closest_point
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

// r0 < inf starts the pruning radius at a known bound on the answer (the
// reference query starts at inf).
DWD void closest3_begin(Trav& t, const SubDev& S, F3 p, bool wide = false, float r0 = kInf) {
    t.kind = wide ? TQ_CLOSEST_W : TQ_CLOSEST;
    t.S = &S;
    t.rec = wide ? S.rec_wide : S.rec_bvh;
    t.prims = S.prims;
    t.x = p;
    t.d = kInf;
    t.best = -1;
    t.r_max = r0;
    t.ref = S.root_ref_bvh;
    t.cd = r0 < kInf ? 0.0f : kInf;
    t.sp = 0;
}
DWD CP closest3_end(const Trav& t) {
    CP its;
    its.p = its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = its.u = its.v = 0.0f;
    its.d = kInf;
    its.prim = -1;
    if (t.best >= 0) {
        DTri tr = ((const DTri*)t.S->prims)[t.best];
        tri_closest(tr, t.x, its.p, its.u, its.v, its.d);
        F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
        its.n = normalize(cross(b - a, c - a));
        its.prim = tr.index;
    }
    return its;
}
// SNCH3D.intersect (snch3d.py:411-476) + Triangle.ray_intersect (triangle.py:59-86);
// the nearest hit wins (r_max drops to each accepted t); prim = BVH-sorted index
// (triangle.py:82)
DWD bool tri_hit(const DTri& tr, F3 x, F3 v, float r_max, float& bu, float& bw, float& t_) {
    F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
    F3 v1 = b - a, v2 = c - a;
    F3 pp = cross(v, v2);
    float det = dot(v1, pp);
    if (!(__builtin_fabsf(det) > kDrEps)) return false;
    float inv = 1.0f / det;
    F3 s = x - a;
    bu = dot(s, pp) * inv;
    if (!((bu >= 0.0f) && (bu <= 1.0f))) return false;
    F3 q = cross(s, v1);
    bw = dot(v, q) * inv;
    if (!((bw >= 0.0f) && (bu + bw <= 1.0f))) return false;
    t_ = dot(v2, q) * inv;
    if (!((t_ >= 0.0f) && (t_ <= r_max))) return false;
    return t_ < r_max;
}
DWD bool isect3_begin(Trav& t, const SubDev& S, F3 x, F3 v, float r_max) {
    t.kind = TQ_ISECT;
    t.S = &S;
    t.rec = S.rec_bvh;
    t.prims = S.prims;
    t.x = x;
    t.v = v;
    t.r_max = r_max;
    t.best = -1;
    float tmin;
    if (!box_ray(f3(S.root_mn[0], S.root_mn[1], S.root_mn[2]), f3(S.root_mx[0], S.root_mx[1], S.root_mx[2]), x, v,
                 r_max, tmin))
        return false;
    t.sp = 0;
    t.ref = S.root_ref_bvh;
    t.cd = tmin;
    return true;
}
// The two children of a 3D closest-point node (boxes and references): both
// tested with r_max tightened by each d_max in order; the nearer descended into
// (equal zero d_min: the nearer d_max first, snch3d.py:383-388), the other
// pushed.  False when neither is hit.
DWD bool closest3_pair(Trav& t, Ovf& o, F3 lmn, F3 lmx, F3 rmn, F3 rmx, int ref_l, int ref_r) {
    float d0, d1, dmaxL, dmaxR;
    box_dist(lmn, lmx, t.x, d0, dmaxL);
    box_dist(rmn, rmx, t.x, d1, dmaxR);
    bool h0 = d0 <= t.r_max;
    t.r_max = fmn(t.r_max, dmaxL);
    bool h1 = d1 <= t.r_max;
    t.r_max = fmn(t.r_max, dmaxR);
    bool sw = ((d0 == 0.0f) && (d1 == 0.0f) && (dmaxR < dmaxL)) || (d1 < d0);
    if (h0 && h1) {
        stk_push(t.sp, o, sw ? ref_l : ref_r, sw ? d0 : d1);
        t.ref = sw ? ref_r : ref_l;
        t.cd = sw ? d1 : d0;
        return true;
    } else if (h0) {
        t.ref = ref_l;
        t.cd = d0;
        return true;
    } else if (h1) {
        t.ref = ref_r;
        t.cd = d1;
        return true;
    }
    return false;
}
// One iteration of a 3D query: closest point (snch3d.py:349-409), closest
// silhouette (snch3d.py:594-663) or ray intersection (snch3d.py:411-476).  An
// interior step of any kind fetches the record head (children's boxes and
// references; SNCH records add their cones), so lanes of different kinds side
// by side in the interleaved walk loop share one fetch per iteration, and all
// kinds share the stack logic.
template <bool X>
DWD bool trav3_iter(Trav& t, Ovf& o) {
    const SubDev& S = *t.S;
    const int kind = t.kind;
    if (t.cd <= t.r_max) {
        if (t.ref < 0) {
            int off, cnt;
            leaf_range(S, t.ref, off, cnt);
            // each item's fetch is issued one item ahead (the last one repeats
            // its own), so the next record is in flight while this one is tested
            if (cnt <= 0) {
                // an SNCH leaf without silhouettes: nothing to fetch
            } else if (kind == TQ_SIL) {
                DSilE e = gload<DSilE>(t.prims, off);
                for (int j = 0; j < cnt; ++j) {
                    DSilE en = gload<DSilE>(t.prims, off + (j + 1 < cnt ? j + 1 : j));
                    float d;
                    F3 pt;
                    if (sile_test(e, t.x, t.r_max, d, pt)) {
                        t.r_max = fmn(t.r_max, d);
                        t.best = off + j;
                    }
                    e = en;
                }
            } else {
                DTri trn = gload<DTri>(t.prims, off);
                for (int j = 0; j < cnt; ++j) {
                    DTri tr = trn;
                    trn = gload<DTri>(t.prims, off + (j + 1 < cnt ? j + 1 : j));
                    if (kind != TQ_ISECT) {
                        F3 pt;
                        float u, v, d;
                        tri_closest(tr, t.x, pt, u, v, d);
                        if (d < t.d) {
                            t.d = d;
                            t.best = off + j;
                            t.r_max = fmn(t.r_max, d);
                        }
                    } else {
                        float bu, bw, t_;
                        if (tri_hit(tr, t.x, t.v, t.r_max, bu, bw, t_)) {
                            t.r_max = t_;
                            t.best = off + j;
                        }
                    }
                }
            }
        } else if (kind == TQ_CLOSEST_W) {
            // two-level record: the node's pair, then (in the same fetch) the
            // pair of the interior child descended into; node sequence, stack and
            // r_max as in the binary iteration (a descended-into child always
            // passes the pending-node test)
            const char* rec = (const char*)t.rec + (size_t)t.ref * sizeof(RWide3);
            // a pair of slots is 3 float4: (min.xyz, max.x | max.yz of the first,
            // min.xy of the second | min.z, max.xyz of the second)
            float4 q0 = gload<float4>(rec, 0), q1 = gload<float4>(rec, 1), q2 = gload<float4>(rec, 2);
            int2 r01 = gload<int2>(rec + 144, 0);
            if (closest3_pair(t, o, f3(q0.x, q0.y, q0.z), f3(q0.w, q1.x, q1.y), f3(q1.z, q1.w, q2.x),
                              f3(q2.y, q2.z, q2.w), r01.x, r01.y)) {
                if (t.ref < 0) return false;  // a leaf child: next iteration
                // the chosen child's pair (slots 2, 3 or 4, 5), from the lines just fetched
                const int k = t.ref == r01.y ? 1 : 0;
                float4 a0 = gload<float4>(rec, 3 + 3 * k), a1 = gload<float4>(rec, 4 + 3 * k),
                       a2 = gload<float4>(rec, 5 + 3 * k);
                int2 cr = gload<int2>(rec + 144, 1 + k);
                if (closest3_pair(t, o, f3(a0.x, a0.y, a0.z), f3(a0.w, a1.x, a1.y), f3(a1.z, a1.w, a2.x),
                                  f3(a2.y, a2.z, a2.w), cr.x, cr.y))
                    return false;
            }
        } else {
            const bool sn = kind == TQ_SIL;
            const char* rec = (const char*)t.rec + (size_t)t.ref * (sn ? sizeof(RSnch3) : sizeof(RHead3));
            RHead3 rc = gload<RHead3>(rec, 0);
            RSnchTail3 cn;
            if (sn) cn = gload<RSnchTail3>(rec + sizeof(RHead3), 0);
            F3 lmn = f3(rc.lmnx, rc.lmny, rc.lmnz), lmx = f3(rc.lmxx, rc.lmxy, rc.lmxz);
            F3 rmn = f3(rc.rmnx, rc.rmny, rc.rmnz), rmx = f3(rc.rmxx, rc.rmxy, rc.rmxz);
            float d0, d1, dmaxL, dmaxR;
            bool h0, h1, sw;  // sw: right child first
            // Box distances serve both the silhouette and the closest-point
            // kinds: one copy for the lanes of both.
            if (kind != TQ_ISECT) {
                box_dist(lmn, lmx, t.x, d0, dmaxL);
                box_dist(rmn, rmx, t.x, d1, dmaxR);
            }
            if (sn) {
                h0 = snch3_child<X>(lmn, lmx, f3(cn.lax, cn.lay, cn.laz), rc.lha, cn.lsin, cn.lcos, cn.lrad, cn.lrho,
                                    cn.lelo, cn.lehi, t.x, t.r_max, d0);
                h1 = snch3_child<X>(rmn, rmx, f3(cn.rax, cn.ray, cn.raz), rc.rha, cn.rsin, cn.rcos, cn.rrad, cn.rrho,
                                    cn.relo, cn.rehi, t.x, t.r_max, d1);
                sw = d1 < d0;
            } else if (kind == TQ_CLOSEST) {
                h0 = d0 <= t.r_max;
                t.r_max = fmn(t.r_max, dmaxL);
                h1 = d1 <= t.r_max;
                t.r_max = fmn(t.r_max, dmaxR);
                // snch3d.py:383-388: equal zero d_min -> nearer d_max first; then d_min order
                sw = ((d0 == 0.0f) && (d1 == 0.0f) && (dmaxR < dmaxL)) || (d1 < d0);
            } else {
                h0 = box_ray(lmn, lmx, t.x, t.v, t.r_max, d0);
                h1 = box_ray(rmn, rmx, t.x, t.v, t.r_max, d1);
                sw = d1 < d0;
            }
            if (h0 && h1) {
                stk_push(t.sp, o, sw ? rc.ref_l : rc.ref_r, sw ? d0 : d1);
                t.ref = sw ? rc.ref_r : rc.ref_l;
                t.cd = sw ? d1 : d0;
                return false;
            } else if (h0) {
                t.ref = rc.ref_l;
                t.cd = d0;
                return false;
            } else if (h1) {
                t.ref = rc.ref_r;
                t.cd = d1;
                return false;
            }
        }
    }
    if (t.sp == 0) return true;
    stk_pop(t.sp, o, t.ref, t.cd);
    return false;
}
DWD CP closest3(const SubDev& S, F3 p) {
    Trav t;
    Ovf o;
    closest3_begin(t, S, p);
    while (!trav3_iter<true>(t, o)) {
    }
    return closest3_end(t);
}

DWD Hit isect3_end(const Trav& t) {
    Hit its;
    its.valid = false;
    its.on_boundary = false;
    its.p = its.n = f3(0.0f, 0.0f, 0.0f);
    its.t = its.u = its.v = its.d = 0.0f;
    its.prim = 0;
    if (t.best >= 0) {
        DTri tr = ((const DTri*)t.S->prims)[t.best];
        float bv = 0.0f, bw = 0.0f, t_;
        tri_hit(tr, t.x, t.v, kInf, bv, bw, t_);
        F3 a = f3(tr.ax, tr.ay, tr.az), b = f3(tr.bx, tr.by, tr.bz), c = f3(tr.cx, tr.cy, tr.cz);
        F3 v1 = b - a, v2 = c - a;
        its.valid = true;
        its.on_boundary = true;
        its.p = (a + v1 * bv) + v2 * bw;
        its.n = normalize(cross(v1, v2));
        its.u = 1.0f - bv - bw;
        its.v = bv;
        its.d = t.r_max;
        its.prim = t.best;
    }
    return its;
}
DWD Hit intersect3(const SubDev& S, F3 x, F3 v, float r_max) {
    Trav t;
    Ovf o;
    if (isect3_begin(t, S, x, v, r_max))
        while (!trav3_iter<true>(t, o)) {
        }
    return isect3_end(t);
}

// One traversal iteration of whichever query t holds (the walk kernels' three
// query kinds); true when it has finished.
template <int D, bool X>
DWD bool trav_iter(Trav& t, Ovf& o) {
    if constexpr (D == 2) {
        if (t.kind == TQ_CLOSEST) return closest2_iter(t, o);
        if (t.kind == TQ_CLOSEST_W) return closest2w_iter(t, o);
        if (t.kind == TQ_SIL) return sil2_iter<X>(t, o);
        return isect2_iter(t, o);
    } else {
        return trav3_iter<X>(t, o);
    }
}
}  // namespace dwd
