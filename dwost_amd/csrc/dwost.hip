// dwost.hip -- kernels and the C ABI of libdwost (declared in include/dwost.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dwost.h"
#include "dw_layout.h"
#include "dw_walk.h"
#include "scene_build.h"

using namespace dwd;

// ===========================================================================
// Kernels
// ===========================================================================
namespace {

constexpr int kBlock = 256;
// Minimum resident waves per SIMD for the persistent walk kernels: the
// literal-cone kernels keep their fp64 acos/asin/atan2 without spilling at 2;
// the transcendental-free ones run best at 4 (a few VGPRs spill).
#ifndef DW_PERS_WAVES_FAST
#define DW_PERS_WAVES_FAST 4
#endif
#ifndef DW_PERS_WAVES_EXACT
#define DW_PERS_WAVES_EXACT 2
#endif

__global__ void __launch_bounds__(kBlock) k_closest(SceneDev sc, int sub, const float* __restrict__ pts, int64_t n,
                                                    int* prim, float* dist, float* point, float* normal, float* tuv) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const SubDev& S = sc.sub[sub];
    CP c;
    if (sc.dim == 2) {
        c = closest2(S, f2(pts[2 * q], pts[2 * q + 1]));
    } else {
        c = closest3(S, f3(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]));
    }
    int D = sc.dim;
    if (prim) prim[q] = c.prim;
    if (dist) dist[q] = c.d;
    if (point) {
        point[D * q] = c.p.x, point[D * q + 1] = c.p.y;
        if (D == 3) point[D * q + 2] = c.p.z;
    }
    if (normal) {
        normal[D * q] = c.n.x, normal[D * q + 1] = c.n.y;
        if (D == 3) normal[D * q + 2] = c.n.z;
    }
    if (tuv) {
        tuv[2 * q] = D == 2 ? c.t : c.u;
        tuv[2 * q + 1] = D == 2 ? 0.0f : c.v;
    }
}

__global__ void __launch_bounds__(kBlock) k_intersect(SceneDev sc, int sub, const float* __restrict__ org,
                                                      const float* __restrict__ dir, const float* __restrict__ rmax,
                                                      int64_t n, int* prim, float* dist, float* point, float* normal,
                                                      float* tuv) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const SubDev& S = sc.sub[sub];
    int D = sc.dim;
    Hit h;
    if (D == 2)
        h = intersect2(S, f2(org[2 * q], org[2 * q + 1]), f2(dir[2 * q], dir[2 * q + 1]), rmax[q]);
    else
        h = intersect3(S, f3(org[3 * q], org[3 * q + 1], org[3 * q + 2]),
                       f3(dir[3 * q], dir[3 * q + 1], dir[3 * q + 2]), rmax[q]);
    if (prim) prim[q] = h.valid ? h.prim : -1;
    if (dist) dist[q] = h.d;
    if (point) {
        point[D * q] = h.p.x, point[D * q + 1] = h.p.y;
        if (D == 3) point[D * q + 2] = h.p.z;
    }
    if (normal) {
        normal[D * q] = h.n.x, normal[D * q + 1] = h.n.y;
        if (D == 3) normal[D * q + 2] = h.n.z;
    }
    if (tuv) {
        tuv[2 * q] = D == 2 ? h.t : h.u;
        tuv[2 * q + 1] = D == 2 ? 0.0f : h.v;
    }
}

__global__ void __launch_bounds__(kBlock) k_closest_sil(SceneDev sc, int sub, const float* __restrict__ pts,
                                                        const float* __restrict__ rmax, int64_t n, int* sil,
                                                        float* dist, float* point) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const SubDev& S = sc.sub[sub];
    int D = sc.dim;
    float d;
    int ref;
    bool ok;
    if (D == 2) {
        F2 p;
        ok = closest_sil2(S, f2(pts[2 * q], pts[2 * q + 1]), rmax[q], d, ref, p);
        if (point) point[2 * q] = ok ? p.x : 0.0f, point[2 * q + 1] = ok ? p.y : 0.0f;
    } else {
        F3 p;
        ok = closest_sil3(S, f3(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]), rmax[q], d, ref, p);
        if (point)
            point[3 * q] = ok ? p.x : 0.0f, point[3 * q + 1] = ok ? p.y : 0.0f, point[3 * q + 2] = ok ? p.z : 0.0f;
    }
    if (sil) sil[q] = ok ? ref : -1;
    if (dist) dist[q] = ok ? d : 0.0f;
}

__global__ void __launch_bounds__(kBlock) k_star_radius(SceneDev sc, int sub, const float* __restrict__ pts,
                                                        const float* __restrict__ rmax, int64_t n, float* out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const SubDev& S = sc.sub[sub];
    out[q] = sc.dim == 2 ? star_radius2(S, f2(pts[2 * q], pts[2 * q + 1]), rmax[q])
                         : star_radius3(S, f3(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]), rmax[q]);
}

// Solver.solve (solver.py:36-71): lane = i*W + k owns the PCG32 stream
// PCG32(size=N*W, TEA(seed, i)) at offset `lane`; nsubwalks walks averaged.
__global__ void __launch_bounds__(kBlock) k_walk_u(SceneDev sc, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                                                   int64_t n, float* walk_u, int* walk_steps, int* walk_term) {
    int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= n * P.nwalks) return;
    int64_t i = lane / P.nwalks;
    uint32_t a, b;
    tea4(P.seed, (uint32_t)i, a, b);
    Rng rng;
    rng.seed_lane(a, b, (uint64_t)lane);
    float res = 0.0f;
    int steps = 0, term = -1;
    for (int s = 0; s < P.nsubwalks; ++s) {
        int st, tm;
        float v;
        if (sc.dim == 2)
            v = walk_u2(sc, bc, P, f2(pts[2 * i], pts[2 * i + 1]), rng, st, tm);
        else
            v = walk_u3(sc, bc, P, f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), rng, st, tm);
        res += v;
        steps += st;
        term = tm;
    }
    walk_u[lane] = res / (float)P.nsubwalks;
    if (walk_steps) walk_steps[lane] = steps;
    if (walk_term) walk_term[lane] = term;
}

// Per-point mean over W per-walk values: one wavefront per point, fp32 Kahan
// per lane, then a fixed-order Kahan combine across lanes (deterministic).
// Optional `comps` extra components per walk (gradient) use the same order.
__global__ void __launch_bounds__(kBlock) k_reduce_mean(const float* __restrict__ vals, int64_t n, int W, int comps,
                                                        float scale, float* out0, float* out_rest,
                                                        const int* __restrict__ steps,
                                                        unsigned long long* step_total) {
    int wave = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int lane = threadIdx.x & 63;
    if (wave >= n) return;
    int64_t base = (int64_t)wave * W;
    for (int c = 0; c < comps; ++c) {
        float s = 0.0f, e = 0.0f;
        for (int k = lane; k < W; k += 64) {
            float y = vals[(base + k) * comps + c] - e;
            float t = s + y;
            e = (t - s) - y;
            s = t;
        }
        for (int off = 32; off > 0; off >>= 1) {
            float so = __shfl_down(s, off, 64), eo = __shfl_down(e, off, 64);
            // TwoSum of (s - e) + (so - eo)
            float a = s - e, b = so - eo;
            float t = a + b;
            float bp = t - a;
            float err = (a - (t - bp)) + (b - bp);
            s = t;
            e = -err;
        }
        if (lane == 0) {
            float m = (s - e) * scale;
            if (c == 0)
                out0[wave] = m;
            else
                out_rest[(int64_t)wave * (comps - 1) + (c - 1)] = m;
        }
    }
    if (steps && step_total) {
        unsigned long long acc = 0;
        for (int k = lane; k < W; k += 64) acc += (unsigned long long)steps[base + k];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) atomicAdd(step_total, acc);
    }
}

// ---------------------------------------------------------------------------
// Persistent walk kernels with lane refill.  Each lane owns one job (a walk,
// or an antithetic pair of walks) as a step machine; whenever lanes fall idle
// the wave takes that many new job indices with one atomic and hands them out
// by lane rank, so a wave is never held by its longest walk.  A job's result
// depends only on its index (its PCG32 stream), not on the lane that runs it.
// ---------------------------------------------------------------------------
DWD uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Hands new job ids to the idle lanes of the wave; returns the id or -1 (no work left).
DWD int64_t refill(bool idle, unsigned long long* counter, int64_t total) {
    uint64_t m = __ballot(idle);
    if (m == 0) return -1;
    int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (!idle) return -1;
    unsigned long long mine = base + lane_rank(m);
    return mine < (unsigned long long)total ? (int64_t)mine : -2;
}

template <int D>
DWD typename VecT<D>::T load_pt(const float* pts, int64_t i);
template <>
DWD F2 load_pt<2>(const float* pts, int64_t i) {
    return f2(pts[2 * i], pts[2 * i + 1]);
}
template <>
DWD F3 load_pt<3>(const float* pts, int64_t i) {
    return f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

// Solver.solve jobs: job = i*W + k, nsubwalks walks on the job's stream.
template <int D, bool X>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST) k_walk_u_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P,
                                                        const float* __restrict__ pts, int64_t n, float* walk_u,
                                                        int* walk_steps, int* walk_term,
                                                        unsigned long long* counter) {
    const SceneDev& sc = *scp;
    using V = typename VecT<D>::T;
    const int64_t total = n * (int64_t)P.nwalks;
    int64_t job = -1;
    bool done = false;
    Walk<D> w;
    Rng rng;
    V x;
    float res = 0.0f;
    int steps = 0, sub = 0;
    while (true) {
        int64_t got = refill(!done && job < 0, counter, total);
        if (got == -2) done = true;
        if (got >= 0) {
            job = got;
            int64_t i = job / P.nwalks;
            uint32_t a, b;
            tea4(P.seed, (uint32_t)i, a, b);
            rng.seed_lane(a, b, (uint64_t)job);
            x = load_pt<D>(pts, i);
            walk_init<D>(w, x);
            res = 0.0f;
            steps = 0;
            sub = 0;
        }
        if (__ballot(!done) == 0) break;
        if (job >= 0 && !walk_step<X>(sc, bc, P, w, rng)) {
            res += walk_value(w);
            steps += w.i;
            if (++sub < P.nsubwalks) {
                walk_init<D>(w, x);
            } else {
                walk_u[job] = res / (float)P.nsubwalks;
                if (walk_steps) walk_steps[job] = steps;
                if (walk_term) walk_term[job] = w.term;
                job = -1;
            }
        }
    }
}

// Fused jobs: job = i*npairs + k, the antithetic pair of WoSt.dude's first ball.
template <int D, bool X>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST) k_fused_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P,
                                                       const float* __restrict__ pts, const float* __restrict__ Rp,
                                                       const int* __restrict__ insidep, int64_t n, int npairs,
                                                       float* pair_vals, int* pair_steps,
                                                       unsigned long long* counter) {
    const SceneDev& sc = *scp;
    using V = typename VecT<D>::T;
    const int64_t total = n * (int64_t)npairs;
    const float fsrc = src_f(bc);
    int64_t job = -1;
    bool done = false;
    Walk<D> w;
    Rng rng;
    V p, dir, gs;
    float R = 0.0f, up = 0.0f;
    uint64_t saved = 0;
    int phase = 0, steps = 0;
    while (true) {
        int64_t got = refill(!done && job < 0, counter, total);
        if (got == -2) done = true;
        if (got >= 0) {
            int64_t i = got / npairs;
            if (!insidep[i]) {
                float* o = pair_vals + got * (D + 1);
                for (int q = 0; q <= D; ++q) o[q] = 0.0f;
                pair_steps[got] = 0;
            } else {
                job = got;
                R = Rp[i];
                p = load_pt<D>(pts, i);
                uint32_t a, b;
                tea4(P.seed, (uint32_t)i, a, b);
                rng.seed_lane(a, b, (uint64_t)job);
                V p_src;
                float r_src, pdf_src;
                if constexpr (D == 2) {
                    dir = uniform_dir2(rng);
                    F2 dir_src = uniform_dir2(rng);
                    greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
                    gs = greens2_gradient(p, R, r_src, p_src) / (pdf_src * (R * R / 4.0f));
                } else {
                    dir = uniform_sphere(rng);
                    greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
                    gs = greens3_gradient(p, R, r_src, p_src) / (pdf_src * (R * R / 6.0f));
                }
                saved = rng.state;
                walk_init<D>(w, p + dir * R);
                phase = 0;
                steps = 0;
            }
        }
        if (__ballot(!done) == 0) break;
        if (job >= 0 && !walk_step<X>(sc, bc, P, w, rng)) {
            float v = walk_value(w);
            steps += w.i;
            if (phase == 0) {
                up = v;
                rng.state = saved;
                walk_init<D>(w, p - dir * R);
                phase = 1;
            } else {
                float um = v;
                float* o = pair_vals + job * (D + 1);
                if constexpr (D == 2) {
                    float s = (R * R / 4.0f) * fsrc;
                    float k = 2.0f / R * (up - um);
                    float g0 = k * dir.x + (s - s) * gs.x, g1 = k * dir.y + (s - s) * gs.y;
                    o[0] = (up + um) + (s + s);
                    o[1] = bad(g0) ? 0.0f : g0;
                    o[2] = bad(g1) ? 0.0f : g1;
                } else {
                    float u1 = (bad(up) || __builtin_fabsf(up) > 10000.0f) ? 0.0f : up;
                    float u2 = (bad(um) || __builtin_fabsf(um) > 10000.0f) ? 0.0f : um;
                    float s = (R * R / 6.0f) * fsrc;
                    float k = 3.0f / R * (u1 - u2);
                    float g0 = k * dir.x + (s - s) * gs.x, g1 = k * dir.y + (s - s) * gs.y,
                          g2_ = k * dir.z + (s - s) * gs.z;
                    o[0] = (u1 + u2) + (s + s);
                    o[1] = bad(g0) ? 0.0f : g0;
                    o[2] = bad(g1) ? 0.0f : g1;
                    o[3] = bad(g2_) ? 0.0f : g2_;
                }
                pair_steps[job] = steps;
                job = -1;
            }
        }
    }
}

// Distance to the full boundary and the inside flag per point (used by the
// gradient estimators, wost.py:43-50).
__global__ void __launch_bounds__(kBlock) k_point_prep(SceneDev sc, const float* __restrict__ pts, int64_t n,
                                                       float* R, int* inside) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (sc.dim == 2) {
        F2 x = f2(pts[2 * i], pts[2 * i + 1]);
        R[i] = __builtin_fabsf(closest2(sc.sub[0], x).d);
        inside[i] = inside2(sc, x) ? 1 : 0;
    } else {
        F3 x = f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        R[i] = __builtin_fabsf(closest3(sc.sub[0], x).d);
        inside[i] = inside3(sc, x) ? 1 : 0;
    }
}

// Fused u + grad u: one antithetic pair per lane (pair k of point i uses the
// stream of lane i*npairs + k).  Writes (usum, g[dim]) per pair and steps.
__global__ void __launch_bounds__(kBlock) k_fused(SceneDev sc, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                                                  const float* __restrict__ Rp, const int* __restrict__ insidep,
                                                  int64_t n, int npairs, float* pair_vals, int* pair_steps) {
    int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= n * npairs) return;
    int64_t i = lane / npairs;
    int D = sc.dim;
    float* o = pair_vals + lane * (D + 1);
    if (!insidep[i]) {
        for (int q = 0; q <= D; ++q) o[q] = 0.0f;
        pair_steps[lane] = 0;
        return;
    }
    float R = Rp[i];
    uint32_t a, b;
    tea4(P.seed, (uint32_t)i, a, b);
    Rng rng;
    rng.seed_lane(a, b, (uint64_t)lane);
    float fsrc = src_f(bc);
    int st1, st2, tm;
    if (D == 2) {
        F2 p = f2(pts[2 * i], pts[2 * i + 1]);
        F2 dir = uniform_dir2(rng);
        F2 dir_src = uniform_dir2(rng);
        F2 p_src;
        float r_src, pdf_src;
        greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
        float gnorm = R * R / 4.0f;
        uint64_t state = rng.state;
        float up = walk_u2(sc, bc, P, p + dir * R, rng, st1, tm);
        rng.state = state;
        float um = walk_u2(sc, bc, P, p - dir * R, rng, st2, tm);
        float s_p = gnorm * fsrc, s_m = gnorm * fsrc;
        F2 gs = greens2_gradient(p, R, r_src, p_src) / (pdf_src * gnorm);
        float k = 2.0f / R * (up - um);
        float g0 = k * dir.x + (s_p - s_m) * gs.x, g1 = k * dir.y + (s_p - s_m) * gs.y;
        o[0] = (up + um) + (s_p + s_m);
        o[1] = bad(g0) ? 0.0f : g0;
        o[2] = bad(g1) ? 0.0f : g1;
    } else {
        F3 p = f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        F3 dir = uniform_sphere(rng);
        F3 p_src;
        float r_src, pdf_src;
        greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
        float gnorm = R * R / 6.0f;
        uint64_t state = rng.state;
        float up = walk_u3(sc, bc, P, p + dir * R, rng, st1, tm);
        rng.state = state;
        float um = walk_u3(sc, bc, P, p - dir * R, rng, st2, tm);
        if (bad(up) || __builtin_fabsf(up) > 10000.0f) up = 0.0f;
        if (bad(um) || __builtin_fabsf(um) > 10000.0f) um = 0.0f;
        float s_p = gnorm * fsrc, s_m = gnorm * fsrc;
        F3 gs = greens3_gradient(p, R, r_src, p_src) / (pdf_src * gnorm);
        float k = 3.0f / R * (up - um);
        float g0 = k * dir.x + (s_p - s_m) * gs.x, g1 = k * dir.y + (s_p - s_m) * gs.y,
              g2_ = k * dir.z + (s_p - s_m) * gs.z;
        o[0] = (up + um) + (s_p + s_m);
        o[1] = bad(g0) ? 0.0f : g0;
        o[2] = bad(g1) ? 0.0f : g1;
        o[3] = bad(g2_) ? 0.0f : g2_;
    }
    pair_steps[lane] = st1 + st2;
}

// WoSt.solve_dude / WoSt3D.solve_dude (wost.py:21-102, wost_3d.py:108-175):
// one PCG32 stream per point, walks run serially inside the lane.
__global__ void __launch_bounds__(kBlock) k_dude(SceneDev sc, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                                                 const float* __restrict__ e, const float* __restrict__ Rp,
                                                 const int* __restrict__ insidep, int64_t n, float* out,
                                                 unsigned long long* step_total) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a, b;
    tea4(P.seed, (uint32_t)i, a, b);
    Rng rng;
    rng.seed_lane(a, b, (uint64_t)i);
    float R = Rp[i];
    bool active = insidep[i] != 0;
    float result = 0.0f;
    int wi = 0;
    unsigned long long steps = 0;
    int st, tm;
    if (sc.dim == 2) {
        F2 p = f2(pts[2 * i], pts[2 * i + 1]), ev = f2(e[2 * i], e[2 * i + 1]);
        float u_sum = 0.0f, u_count = 0.0f, src_sum = 0.0f, src_count = 0.0f;
        float gnorm = R * R / 4.0f;
        while ((wi < P.nwalks) && active) {
            F2 dir = uniform_dir2(rng);
            F2 dir_src = uniform_dir2(rng);
            F2 p_src;
            float r_src, pdf_src;
            greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
            int niters = P.av ? 2 : 1;
            uint64_t state = rng.state;
            int it = 0;
            while ((it < niters) && (wi < P.nwalks)) {
                if (it == 1) {
                    dir_src = -dir_src;
                    p_src = p * 2.0f - p_src;
                }
                float src = gnorm * src_f(bc);
                float src_grad = dot(greens2_gradient(p, R, r_src, p_src), ev) / (pdf_src * gnorm);
                src_sum += src;
                src_count += 1.0f;
                float _src = P.cv ? src - (src_sum / src_count) : src;
                result += _src * src_grad;
                if (it == 1) dir = -dir;
                F2 _p = p + dir * R;
                rng.state = state;
                float u = walk_u2(sc, bc, P, _p, rng, st, tm);
                steps += st;
                u_sum += u;
                u_count += 1.0f;
                float _u = P.cv ? u - u_sum / u_count : u;
                float r2 = 2.0f / R * _u * dot(ev, dir);
                if ((__builtin_fabsf(r2) > 1e6f) || bad(r2)) r2 = 0.0f;
                result += r2;
                wi += 1;
                it += 1;
            }
        }
    } else {
        F3 p = f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), ev = f3(e[3 * i], e[3 * i + 1], e[3 * i + 2]);
        float u_mean = 0.0f, u_count = 0.0f, src_mean = 0.0f, src_count = 0.0f;
        float gnorm = R * R / 6.0f;
        while ((wi < P.nwalks) && active) {
            F3 dir = uniform_sphere(rng);
            F3 p_src;
            float r_src, pdf_src;
            greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
            int niters = P.av ? 2 : 1;
            uint64_t state = rng.state;
            int it = 0;
            while ((it < niters) && (wi < P.nwalks)) {
                if (it == 1) {
                    dir = -dir;
                    p_src = p * 2.0f - p_src;
                }
                float src = gnorm * src_f(bc);
                float src_grad = dot(greens3_gradient(p, R, r_src, p_src), ev) / (pdf_src * gnorm);
                src_mean = src_mean * (src_count / (src_count + 1.0f)) + src / (src_count + 1.0f);
                src_count += 1.0f;
                float _src = P.cv ? src - src_mean : src;
                result += _src * src_grad;
                if (it == 1) dir = -dir;  // second flip (wost_3d.py:152-153)
                F3 _p = p + dir * R;
                rng.state = state;
                float u = walk_u3(sc, bc, P, _p, rng, st, tm);
                steps += st;
                if (bad(u) || __builtin_fabsf(u) > 10000.0f) u = 0.0f;
                u_mean = u_mean * (u_count / (u_count + 1.0f)) + u / (u_count + 1.0f);
                u_count += 1.0f;
                float _u = P.cv ? u - u_mean : u;
                result += 3.0f / R * _u * dot(ev, dir);
                wi += 1;
                it += 1;
            }
        }
        if (bad(result) || __builtin_fabsf(result) > 10000.0f) result = 0.0f;
    }
    out[i] = result / (float)P.nwalks;
    if (step_total) atomicAdd(step_total, steps);
}

}  // namespace

// ===========================================================================
// Host side: scene upload and the C ABI
// ===========================================================================
struct dw_scene {
    dwb::SceneHost host;
    SceneDev dev;
    std::vector<void*> allocs;
    const SceneDev* dev_ptr = nullptr;  // device copy of `dev`
    int device = 0;
    bool host_only = false;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define DW_HIP(call)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return fail(DW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int upload(dw_scene* s, const std::vector<T>& v, const void** out) {
    *out = nullptr;
    if (v.empty()) return DW_OK;
    void* p = nullptr;
    DW_HIP(hipMalloc(&p, v.size() * sizeof(T)));
    s->allocs.push_back(p);
    DW_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = p;
    return DW_OK;
}

int upload_sub(dw_scene* s, const dwb::SubHost& H, SubDev& S) {
    std::memset(&S, 0, sizeof(S));
    S.present = H.present ? 1 : 0;
    if (!H.present) return DW_OK;
    const int D = H.dim;
    size_t nn = H.n_nodes(), np = H.n_prims();
    S.n_nodes = (int)nn;
    S.n_prims = (int)np;
    S.n_refs = (int)H.refs.size();
    S.py_sils = H.py_sils ? 1 : 0;
    for (int d = 0; d < D; ++d) S.bb_min[d] = H.bb_min[d], S.bb_max[d] = H.bb_max[d];
    S.pmf_sum = H.pmf_sum;
    S.pmf_norm = H.pmf_norm;
    S.n_local_verts = (int)(H.V.size() / D);
    int rc;
    const void* p;
    if (D == 2) {
        std::vector<DNode2> nodes(nn);
        std::vector<DCone2> cones(nn);
        for (size_t i = 0; i < nn; ++i) {
            const float* b = &H.node_box[i * 4];
            nodes[i] = DNode2{b[0], b[1], b[2], b[3], H.node_off[i], H.node_nref[i], H.node_sil_off[i], H.node_nsil[i]};
            float ha = H.node_half_angle[i];
            cones[i] = DCone2{H.node_axis[2 * i], H.node_axis[2 * i + 1], ha, H.node_radius[i],
                              (float)std::sin((double)ha), (float)std::cos((double)ha), 0.0f, 0.0f};
        }
        std::vector<DSeg> prims(np);
        for (size_t j = 0; j < np; ++j) {
            int l = H.prim_order[j];
            const float* a = &H.V[2 * H.I[2 * l]];
            const float* b = &H.V[2 * H.I[2 * l + 1]];
            prims[j] = DSeg{a[0], a[1], b[0], b[1]};
        }
        std::vector<DSilV> sils(H.refs.size());
        for (size_t r = 0; r < H.refs.size(); ++r) {
            int si = H.refs[r];
            const float* q = &H.sil_pts[6 * si];
            const int* ix = &H.sil_idx[3 * si];
            sils[r] = DSilV{q[0], q[1], q[2], q[3], q[4], q[5], (ix[0] != -1 && ix[2] != -1) ? 1 : 0, 0};
        }
        if ((rc = upload(s, nodes, &S.nodes))) return rc;
        if ((rc = upload(s, cones, &S.cones))) return rc;
        if ((rc = upload(s, prims, &S.prims))) return rc;
        if ((rc = upload(s, sils, &S.sils))) return rc;
        if ((rc = upload(s, H.vnormal, &p))) return rc;
        S.vnormal = (const float*)p;
    } else {
        std::vector<DNode3> nodes(nn);
        std::vector<DCone3> cones(nn);
        for (size_t i = 0; i < nn; ++i) {
            const float* b = &H.node_box[i * 6];
            nodes[i] = DNode3{b[0], b[1], b[2], b[3], b[4], b[5], H.node_off[i], H.node_nref[i]};
            float ha = H.node_half_angle[i];
            cones[i] = DCone3{H.node_axis[3 * i], H.node_axis[3 * i + 1], H.node_axis[3 * i + 2], ha,
                              H.node_radius[i], H.node_sil_off[i], H.node_nsil[i], (float)std::sin((double)ha),
                              (float)std::cos((double)ha), 0.0f, 0.0f, 0.0f};
        }
        std::vector<DTri> prims(np);
        for (size_t j = 0; j < np; ++j) {
            int l = H.prim_order[j];
            const float* a = &H.V[3 * H.I[3 * l]];
            const float* b = &H.V[3 * H.I[3 * l + 1]];
            const float* c = &H.V[3 * H.I[3 * l + 2]];
            prims[j] = DTri{a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], l, 0, 0};
        }
        std::vector<DSilE> sils(H.refs.size());
        for (size_t r = 0; r < H.refs.size(); ++r) {
            int si = H.refs[r];
            const float* q = &H.sil_pts[12 * si];
            const int* ix = &H.sil_idx[4 * si];
            DSilE e;
            std::memset(&e, 0, sizeof(e));
            e.ax = q[0], e.ay = q[1], e.az = q[2];
            e.boundary = (ix[0] == -1 || ix[3] == -1) ? 1 : 0;
            e.bx = q[3], e.by = q[4], e.bz = q[5];
            e.cx = q[6], e.cy = q[7], e.cz = q[8];
            e.dx = q[9], e.dy = q[10], e.dz = q[11];
            sils[r] = e;
        }
        if ((rc = upload(s, nodes, &S.nodes))) return rc;
        if ((rc = upload(s, cones, &S.cones))) return rc;
        if ((rc = upload(s, prims, &S.prims))) return rc;
        if ((rc = upload(s, sils, &S.sils))) return rc;
    }
    if ((rc = upload(s, H.prim_order, &p))) return rc;
    S.prim_index = (const int*)p;
    if ((rc = upload(s, H.cdf, &p))) return rc;
    S.cdf = (const float*)p;
    if ((rc = upload(s, H.pmf, &p))) return rc;
    S.pmf = (const float*)p;
    std::vector<float> lp(np * D * D);
    std::vector<int> lvg(np * D), lvi(np * D);
    for (size_t j = 0; j < np; ++j)
        for (int k = 0; k < D; ++k) {
            int v = H.I[j * D + k];
            for (int d = 0; d < D; ++d) lp[(j * D + k) * D + d] = H.V[(size_t)v * D + d];
            lvg[j * D + k] = H.vglobal[v];
            lvi[j * D + k] = v;
        }
    if ((rc = upload(s, lp, &p))) return rc;
    S.local_prims = (const float*)p;
    if ((rc = upload(s, lvg, &p))) return rc;
    S.local_vglobal = (const int*)p;
    if ((rc = upload(s, lvi, &p))) return rc;
    S.local_vidx = (const int*)p;
    return DW_OK;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int check_scene(const dw_scene* s, int sub) {
    if (!s) return fail(DW_ERR_ARG, "null scene");
    if (s->host_only) return fail(DW_ERR_ARG, "scene was built with DW_SCENE_HOST_ONLY");
    if (sub < 0 || sub > 2) return fail(DW_ERR_ARG, "sub must be 0 (full), 1 (Dirichlet) or 2 (Neumann)");
    if (!s->dev.sub[sub].present) return fail(DW_ERR_ARG, "requested sub-scene is empty");
    return DW_OK;
}

BCDev to_bc(const dw_boundary* b) {
    BCDev bc;
    std::memset(&bc, 0, sizeof(bc));
    if (!b) return bc;
    bc.code = b->bc_code;
    for (int k = 0; k < 4; ++k) bc.c[k] = b->coeffs[k];
    bc.gv = b->g_vertex;
    bc.hv = b->h_vertex;
    bc.src_code = b->src_code;
    bc.src_c = b->src_value;
    return bc;
}
int check_bc(const dw_boundary* b) {
    if (!b) return fail(DW_ERR_ARG, "null boundary");
    if (b->bc_code < 0 || b->bc_code > 3) return fail(DW_ERR_ARG, "unknown boundary code");
    if (b->bc_code == DW_BC_PER_VERTEX && (!b->g_vertex || !b->h_vertex))
        return fail(DW_ERR_ARG, "per-vertex boundary needs g_vertex and h_vertex");
    if (b->src_code != DW_SRC_ZERO && b->src_code != DW_SRC_CONST) return fail(DW_ERR_ARG, "unknown source code");
    return DW_OK;
}
ParamsDev to_params(const dw_solver_params* p) {
    ParamsDev P;
    P.nwalks = p->nwalks;
    P.nsteps = p->nsteps;
    P.nsubwalks = p->nsubwalks;
    P.cv = p->control_variates;
    P.av = p->antithetic_variates;
    P.seed = p->seed;
    P.eps = p->eps;
    P.min_R = p->min_R;
    P.flags = p->flags;
    return P;
}
int check_params(const dw_solver_params* p) {
    if (!p) return fail(DW_ERR_ARG, "null params");
    if (p->nwalks < 1) return fail(DW_ERR_ARG, "nwalks must be >= 1");
    if (p->nsteps < 0) return fail(DW_ERR_ARG, "nsteps must be >= 0");
    if (p->nsubwalks < 1) return fail(DW_ERR_ARG, "nsubwalks must be >= 1");
    return DW_OK;
}

// Grow-only device workspace for per-walk / per-pair partial results.
struct Workspace {
    std::mutex mu;
    void* ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
};
Workspace g_ws;

// One device counter per device for the persistent kernels' job queue.
unsigned long long* g_counter[64] = {nullptr};
std::mutex g_counter_mu;

int counter_slot(unsigned long long** out) {
    int dev = 0;
    DW_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(DW_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_counter_mu);
    if (!g_counter[dev]) DW_HIP(hipMalloc((void**)&g_counter[dev], 256));
    *out = g_counter[dev];
    return DW_OK;
}

int workspace(size_t bytes, void** out) {
    int dev = 0;
    DW_HIP(hipGetDevice(&dev));
    if (g_ws.ptr && (g_ws.bytes < bytes || g_ws.device != dev)) {
        DW_HIP(hipDeviceSynchronize());
        DW_HIP(hipFree(g_ws.ptr));
        g_ws.ptr = nullptr;
        g_ws.bytes = 0;
    }
    if (!g_ws.ptr) {
        DW_HIP(hipMalloc(&g_ws.ptr, bytes));
        g_ws.bytes = bytes;
        g_ws.device = dev;
    }
    *out = g_ws.ptr;
    return DW_OK;
}

// Grid of a persistent kernel: as many resident blocks as the occupancy allows
// (DWOST_BLOCKS_PER_CU overrides the per-CU count for tuning).
template <typename K>
int persistent_grid(K kernel, unsigned& blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    DW_HIP(hipGetDevice(&dev));
    DW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    DW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0));
    if (const char* e = getenv("DWOST_BLOCKS_PER_CU")) per_cu = atoi(e);
    if (per_cu < 1) per_cu = 1;
    blocks = (unsigned)(per_cu * cus);
    return DW_OK;
}

bool use_simple_kernels() {
    const char* e = getenv("DWOST_KERNEL");
    return e && strcmp(e, "simple") == 0;
}

}  // namespace

extern "C" {

const char* dw_last_error(void) { return g_err.c_str(); }

int dw_version(void) { return DW_ABI_VERSION; }

int dw_scene_create(const dw_scene_desc* desc, dw_scene** out) {
    if (!desc || !out) return fail(DW_ERR_ARG, "null argument");
    *out = nullptr;
    if (!desc->vertices || !desc->indices || !desc->types) return fail(DW_ERR_ARG, "null scene arrays");
    dw_scene* s = new dw_scene();
    std::string err = dwb::build_scene(desc->dim, desc->vertices, desc->n_vertices, desc->indices, desc->types,
                                       desc->n_prims, s->host);
    if (!err.empty()) {
        delete s;
        return fail(DW_ERR_ARG, err);
    }
    s->dev.dim = desc->dim;
    if (desc->flags & DW_SCENE_HOST_ONLY) {
        s->host_only = true;
        *out = s;
        return DW_OK;
    }
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return fail(DW_ERR_HIP, "hipGetDevice failed");
    }
    s->dev.dim = desc->dim;
    for (int k = 0; k < 3; ++k) {
        int rc = upload_sub(s, s->host.sub[k], s->dev.sub[k]);
        if (rc) {
            std::string m = g_err;
            dw_scene_destroy(s);
            return fail(rc, m);
        }
    }
    {
        std::vector<SceneDev> one(1, s->dev);
        const void* p = nullptr;
        int rc = upload(s, one, &p);
        if (rc) {
            std::string m = g_err;
            dw_scene_destroy(s);
            return fail(rc, m);
        }
        s->dev_ptr = (const SceneDev*)p;
    }
    *out = s;
    return DW_OK;
}

int dw_scene_destroy(dw_scene* s) {
    if (!s) return DW_OK;
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
    return DW_OK;
}

int dw_scene_info(const dw_scene* s, int32_t sub, int64_t* out8) {
    if (!s || !out8 || sub < 0 || sub > 2) return fail(DW_ERR_ARG, "bad argument");
    const dwb::SubHost& H = s->host.sub[sub];
    int D = s->host.dim;
    out8[0] = H.present;
    out8[1] = H.present ? (int64_t)(H.V.size() / D) : 0;
    out8[2] = H.present ? (int64_t)H.n_prims() : 0;
    out8[3] = H.present ? (int64_t)H.n_nodes() : 0;
    out8[4] = (int64_t)H.refs.size();
    out8[5] = H.max_depth;
    out8[6] = H.py_sils;
    out8[7] = H.present ? (int64_t)(H.sil_idx.size() / (D == 2 ? 3 : 4)) : 0;
    return DW_OK;
}

int dw_scene_export(const dw_scene* s, int32_t sub, float* box, int32_t* ints, float* cone, int32_t* prim_index,
                    int32_t* refs) {
    if (!s || sub < 0 || sub > 2) return fail(DW_ERR_ARG, "bad argument");
    const dwb::SubHost& H = s->host.sub[sub];
    if (!H.present) return DW_OK;
    int D = s->host.dim;
    size_t nn = H.n_nodes();
    for (size_t i = 0; i < nn; ++i) {
        if (box)
            for (int d = 0; d < 2 * D; ++d) box[i * 2 * D + d] = H.node_box[i * 2 * D + d];
        if (ints) {
            ints[i * 4] = H.node_off[i];
            ints[i * 4 + 1] = H.node_nref[i];
            ints[i * 4 + 2] = H.node_sil_off[i];
            ints[i * 4 + 3] = H.node_nsil[i];
        }
        if (cone) {
            for (int d = 0; d < D; ++d) cone[i * (D + 2) + d] = H.node_axis[i * D + d];
            cone[i * (D + 2) + D] = H.node_half_angle[i];
            cone[i * (D + 2) + D + 1] = H.node_radius[i];
        }
    }
    if (prim_index)
        for (size_t j = 0; j < H.prim_order.size(); ++j) prim_index[j] = H.prim_order[j];
    if (refs)
        for (size_t r = 0; r < H.refs.size(); ++r) refs[r] = H.refs[r];
    return DW_OK;
}

int dw_closest_point(const dw_scene* s, int32_t sub, const float* pts, int64_t n, int32_t* prim, float* dist,
                     float* point, float* normal, float* tuv, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !pts)) return fail(DW_ERR_ARG, "bad points");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_closest, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, s->dev, sub, pts, n,
                       prim, dist, point, normal, tuv);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_intersect(const dw_scene* s, int32_t sub, const float* origin, const float* dir, const float* rmax, int64_t n,
                 int32_t* prim, float* dist, float* point, float* normal, float* tuv, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!origin || !dir || !rmax))) return fail(DW_ERR_ARG, "bad ray arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_intersect, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, s->dev, sub, origin,
                       dir, rmax, n, prim, dist, point, normal, tuv);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_closest_silhouette(const dw_scene* s, int32_t sub, const float* pts, const float* rmax, int64_t n,
                          int32_t* sil, float* dist, float* point, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!pts || !rmax))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_closest_sil, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, s->dev, sub, pts,
                       rmax, n, sil, dist, point);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_star_radius(const dw_scene* s, int32_t sub, const float* pts, const float* rmax, int64_t n, float* radius,
                   void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!pts || !rmax || !radius))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_star_radius, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, s->dev, sub, pts,
                       rmax, n, radius);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_solve_u(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
               const dw_boundary* boundary, float* u, float* walk_u, int32_t* walk_steps, int32_t* walk_term,
               uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (n < 0 || (n > 0 && (!pts || !u))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    if (!s->dev.sub[1].present && params->nsteps > 4096)
        return fail(DW_ERR_ARG, "scene without Dirichlet boundary: nsteps must be <= 4096");
    int64_t lanes = n * (int64_t)params->nwalks;
    float* wv = walk_u;
    int32_t* ws = walk_steps;
    if (!wv || (total_steps && !ws)) {
        void* p;
        if ((rc = workspace((size_t)lanes * 8, &p))) return rc;
        wv = (float*)p;
        ws = walk_steps ? walk_steps : (int32_t*)((char*)p + (size_t)lanes * 4);
        if (walk_u) wv = walk_u;
    }
    hipStream_t st = (hipStream_t)stream;
    if (use_simple_kernels()) {
        hipLaunchKernelGGL(k_walk_u, dim3(blocks_for(lanes)), dim3(kBlock), 0, st, s->dev, to_bc(boundary),
                           to_params(params), pts, n, wv, ws, walk_term);
    } else {
        unsigned long long* counter = nullptr;
        const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
        if ((rc = counter_slot(&counter))) return rc;
        DW_HIP(hipMemsetAsync(counter, 0, sizeof(unsigned long long), st));
        unsigned blocks;
        if (s->dev.dim == 2) {
            if (exact) {
                if ((rc = persistent_grid(k_walk_u_pers<2, true>, blocks))) return rc;
            } else {
                if ((rc = persistent_grid(k_walk_u_pers<2, false>, blocks))) return rc;
            }
            hipLaunchKernelGGL((exact ? k_walk_u_pers<2, true> : k_walk_u_pers<2, false>), dim3(blocks), dim3(kBlock), 0, st, s->dev_ptr, to_bc(boundary),
                               to_params(params), pts, n, wv, ws, walk_term, counter);
        } else {
            if (exact) {
                if ((rc = persistent_grid(k_walk_u_pers<3, true>, blocks))) return rc;
            } else {
                if ((rc = persistent_grid(k_walk_u_pers<3, false>, blocks))) return rc;
            }
            hipLaunchKernelGGL((exact ? k_walk_u_pers<3, true> : k_walk_u_pers<3, false>), dim3(blocks), dim3(kBlock), 0, st, s->dev_ptr, to_bc(boundary),
                               to_params(params), pts, n, wv, ws, walk_term, counter);
        }
    }
    DW_HIP(hipGetLastError());
    if (total_steps) DW_HIP(hipMemsetAsync(total_steps, 0, sizeof(uint64_t), st));
    hipLaunchKernelGGL(k_reduce_mean, dim3(blocks_for(n * 64)), dim3(kBlock), 0, st, wv, n, params->nwalks, 1,
                       1.0f / (float)params->nwalks, u, (float*)nullptr, total_steps ? ws : nullptr,
                       (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_solve_dude(const dw_scene* s, const float* pts, const float* e, int64_t n, const dw_solver_params* params,
                  const dw_boundary* boundary, float* out, uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (n < 0 || (n > 0 && (!pts || !e || !out))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    void* p;
    if ((rc = workspace((size_t)n * 8, &p))) return rc;
    float* R = (float*)p;
    int* ins = (int*)((char*)p + (size_t)n * 4);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(n)), dim3(kBlock), 0, st, s->dev, pts, n, R, ins);
    DW_HIP(hipGetLastError());
    if (total_steps) DW_HIP(hipMemsetAsync(total_steps, 0, sizeof(uint64_t), st));
    hipLaunchKernelGGL(k_dude, dim3(blocks_for(n)), dim3(kBlock), 0, st, s->dev, to_bc(boundary), to_params(params),
                       pts, e, R, ins, n, out, (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_solve_fused(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                   const dw_boundary* boundary, float* u, float* grad, float* pair_vals, int32_t* pair_steps,
                   uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (n < 0 || (n > 0 && (!pts || !u || !grad))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    int D = s->dev.dim;
    int npairs = (params->nwalks + 1) / 2;
    int64_t lanes = n * (int64_t)npairs;
    size_t need = (size_t)n * 8 + (pair_vals ? 0 : (size_t)lanes * (D + 1) * 4) + (pair_steps ? 0 : (size_t)lanes * 4);
    void* p;
    if ((rc = workspace(need, &p))) return rc;
    char* cur = (char*)p;
    float* R = (float*)cur;
    cur += (size_t)n * 4;
    int* ins = (int*)cur;
    cur += (size_t)n * 4;
    float* pv = pair_vals;
    if (!pv) pv = (float*)cur, cur += (size_t)lanes * (D + 1) * 4;
    int32_t* ps = pair_steps;
    if (!ps) ps = (int32_t*)cur;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(n)), dim3(kBlock), 0, st, s->dev, pts, n, R, ins);
    DW_HIP(hipGetLastError());
    if (use_simple_kernels()) {
        hipLaunchKernelGGL(k_fused, dim3(blocks_for(lanes)), dim3(kBlock), 0, st, s->dev, to_bc(boundary),
                           to_params(params), pts, R, ins, n, npairs, pv, ps);
    } else {
        unsigned long long* counter = nullptr;
        const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
        if ((rc = counter_slot(&counter))) return rc;
        DW_HIP(hipMemsetAsync(counter, 0, sizeof(unsigned long long), st));
        unsigned blocks;
        if (D == 2) {
            if (exact) {
                if ((rc = persistent_grid(k_fused_pers<2, true>, blocks))) return rc;
            } else {
                if ((rc = persistent_grid(k_fused_pers<2, false>, blocks))) return rc;
            }
            hipLaunchKernelGGL((exact ? k_fused_pers<2, true> : k_fused_pers<2, false>), dim3(blocks), dim3(kBlock), 0, st, s->dev_ptr, to_bc(boundary),
                               to_params(params), pts, R, ins, n, npairs, pv, ps, counter);
        } else {
            if (exact) {
                if ((rc = persistent_grid(k_fused_pers<3, true>, blocks))) return rc;
            } else {
                if ((rc = persistent_grid(k_fused_pers<3, false>, blocks))) return rc;
            }
            hipLaunchKernelGGL((exact ? k_fused_pers<3, true> : k_fused_pers<3, false>), dim3(blocks), dim3(kBlock), 0, st, s->dev_ptr, to_bc(boundary),
                               to_params(params), pts, R, ins, n, npairs, pv, ps, counter);
        }
    }
    DW_HIP(hipGetLastError());
    if (total_steps) DW_HIP(hipMemsetAsync(total_steps, 0, sizeof(uint64_t), st));
    // u = sum(usum) / (2 npairs); grad = sum(g) / (2 npairs)
    float scale = 1.0f / (float)(2 * npairs);
    hipLaunchKernelGGL(k_reduce_mean, dim3(blocks_for(n * 64)), dim3(kBlock), 0, st, pv, n, npairs, D + 1, scale, u,
                       grad, total_steps ? ps : nullptr, (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

}  // extern "C"
