// dwost.hip -- kernels and the C ABI of libdwost (declared in include/dwost.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dwost.h"
#include "dw_layout.h"
#include "dw_grad.h"
#include "dw_walk.h"
#include "scene_build.h"

using namespace dwd;

// ===========================================================================
// Kernels
// ===========================================================================
namespace {

constexpr int kBlock = 256;
// Waiting lanes of a wave that trigger a walk-logic pass in the interleaved
// persistent loop (run_jobs); 16-32 measured within 2% on the icosphere.
#ifndef DW_GATE
#define DW_GATE 24
#endif
constexpr int kGate = DW_GATE;

constexpr size_t kStackLds = (size_t)kBlock * kStackLdsBytesPerThread;  // traversal stacks
static_assert(kBlock % 64 == 0, "the LDS stack layout assumes whole waves per block");
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
                                                      int64_t n, int mode, int* prim, float* dist, float* point,
                                                      float* normal, float* tuv) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const SubDev& S = sc.sub[sub];
    int D = sc.dim;
    Hit h;
    if (D == 2)
        h = mode == DW_RAY_NEAREST
                ? intersect2_nearest(S, f2(org[2 * q], org[2 * q + 1]), f2(dir[2 * q], dir[2 * q + 1]), rmax[q])
                : intersect2(S, f2(org[2 * q], org[2 * q + 1]), f2(dir[2 * q], dir[2 * q + 1]), rmax[q]);
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

// Polyline.star_radius_2 (polyline.py:317-330) / Mesh.star_radius_2
// (mesh.py:349-361): the star radius bounded by the Poisson-kernel ratio
// rho_max along direction dirs[q].
__global__ void __launch_bounds__(kBlock) k_star_radius_rho(SceneDev sc, int sub, const float* __restrict__ pts,
                                                            const float* __restrict__ dirs, float rho_max, int64_t n,
                                                            float* out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const SubDev& S = sc.sub[sub];
    out[q] = sc.dim == 2 ? star_radius2_rho(S, f2(pts[2 * q], pts[2 * q + 1]), f2(dirs[2 * q], dirs[2 * q + 1]), rho_max)
                         : star_radius3_rho(S, f3(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]),
                                            f3(dirs[3 * q], dirs[3 * q + 1], dirs[3 * q + 2]), rho_max);
}

// Float order keys: consecutive floats have consecutive keys (-0 and +0 share 0).
__host__ __device__ inline int32_t fkey(float f) {
    int32_t b;
    memcpy(&b, &f, 4);
    return b < 0 ? (int32_t)(0x80000000u - (uint32_t)b) : b;
}
__device__ inline float key_float(int32_t k) {
    return __int_as_float(k < 0 ? (int32_t)(0x80000000u - (uint32_t)k) : k);
}
// Largest distance in float steps between the fp32 library function W
// (0 acosf, 1 asinf, 2 atanf, 3 cosf) and its once-rounded fp64 form, over the
// `count` floats from key k0 up.
template <int W>
__global__ void __launch_bounds__(kBlock) k_math_check(int32_t k0, int64_t count, int32_t* out) {
    int worst = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        float x = key_float((int32_t)(k0 + i));
        float f, e;
        if constexpr (W == 0) {
            f = acosf(x), e = cr_acos(x);
        } else if constexpr (W == 1) {
            f = asinf(x), e = cr_asin(x);
        } else if constexpr (W == 2) {
            f = atanf(x), e = (float)atan((double)x);
        } else {
            f = cosf(x), e = cr_cos(x);
        }
        int64_t d = (int64_t)fkey(f) - (int64_t)fkey(e);
        d = d < 0 ? -d : d;
        if (__builtin_isnan(f) || __builtin_isnan(e)) d = (__builtin_isnan(f) && __builtin_isnan(e)) ? 0 : INT32_MAX;
        worst = max(worst, (int)min(d, (int64_t)INT32_MAX));
    }
    for (int o = 32; o > 0; o >>= 1) worst = max(worst, __shfl_xor(worst, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out + W, worst);
}
__global__ void k_math_check_init(int32_t* out) {
    if (threadIdx.x < 4) out[threadIdx.x] = 0;
    if (threadIdx.x == 4) out[4] = kTrSteps;
}

// Solver.solve (solver.py:36-71): lane = i*W + k owns the PCG32 stream
// PCG32(size=N*W, TEA(seed, i)) at offset `lane`; nsubwalks walks averaged.
__global__ void __launch_bounds__(kBlock) k_walk_u(SceneDev sc, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                                                   int64_t n, float* walk_u, int* walk_steps, int* walk_term) {
    int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= n * P.nwalks) return;
    int64_t i = lane / P.nwalks;
    Rng rng;
    seed_stream(rng, P, i, P.nwalks, lane - i * P.nwalks);
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
DWD void reduce_point(const float* __restrict__ vals, int64_t point, int W, int comps, float scale, float* out0,
                      float* out_rest, const int* __restrict__ steps, unsigned long long* step_total, int lane) {
    int64_t base = point * W;
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
                out0[point] = m;
            else
                out_rest[point * (comps - 1) + (c - 1)] = m;
        }
    }
    if (steps && step_total) {
        unsigned long long acc = 0;
        for (int k = lane; k < W; k += 64) acc += (unsigned long long)steps[base + k];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) atomicAdd(step_total, acc);
    }
}
__global__ void __launch_bounds__(kBlock) k_reduce_mean(const float* __restrict__ vals, int64_t n, int W, int comps,
                                                        float scale, float* out0, float* out_rest,
                                                        const int* __restrict__ steps,
                                                        unsigned long long* step_total) {
    int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= n) return;
    reduce_point(vals, wave, W, comps, scale, out0, out_rest, steps, step_total, threadIdx.x & 63);
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

// Job queue of the persistent kernels.  A job is one walk (or walk pair) of
// one point; its random stream is keyed by (point, walk) only, so the order in
// which jobs run never changes a result.  Points are split into 8 contiguous
// ranges of a spatial (Morton-cell) order, one per XCD: a wave draws jobs from
// the range of the XCD it runs on, so each XCD's L2 holds the part of the scene
// near its points, and moves on to the next range when its own is exhausted.
struct JobQ {
    unsigned long long* ctr;  // 8 counters, 128 B apart (ctr[16 * range])
    const int* order;         // spatial position -> point index (nullptr: identity)
    int64_t n;                // points
    int64_t per;              // jobs per point
    int gate;                 // waiting lanes that trigger a walk-logic pass (run_jobs)
};
struct QState {
    int part;  // current range (wave-uniform)
    int left;  // ranges not yet found exhausted
};
DWD uint32_t xcc_id() {
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(v));
    return v;
}
DWD void q_start(QState& q) {
    q.part = (int)(xcc_id() & 7u);
    q.left = 8;
}
// Hands job ids (point * per + walk) to the idle lanes of the wave: one atomic
// per wave.  Returns the id, -1 (none this time) or -2 (all ranges exhausted).
DWD int64_t q_next(bool idle, const JobQ& Q, QState& q) {
    uint64_t m = __ballot(idle);
    if (m == 0) return -1;
    int leader = __ffsll((long long)m) - 1;
    uint32_t rank = lane_rank(m);
    while (q.left > 0) {
        int64_t p0 = (Q.n * q.part) >> 3, p1 = (Q.n * (q.part + 1)) >> 3;
        unsigned long long size = (unsigned long long)((p1 - p0) * Q.per);
        unsigned long long base = 0;
        if ((int)__lane_id() == leader) base = atomicAdd(Q.ctr + 16 * q.part, (unsigned long long)__popcll(m));
        base = __shfl(base, leader, 64);
        if (base < size) {
            if (!idle) return -1;
            unsigned long long j = base + rank;
            if (j >= size) return -1;
            int64_t pos = p0 + (int64_t)(j / (unsigned long long)Q.per);
            int64_t k = (int64_t)(j % (unsigned long long)Q.per);
            int64_t i = Q.order ? (int64_t)Q.order[pos] : pos;
            return i * Q.per + k;
        }
        q.part = (q.part + 1) & 7;
        q.left -= 1;
    }
    return idle ? -2 : -1;
}

// Spatial order of the query points: counting sort by a 12-bit Morton cell of
// the scene box (16^3 cells in 3D, 64^2 in 2D).
constexpr int kCells = 4096;
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 4 bits -> every third bit
    v &= 15u;
    v = (v | (v << 4)) & 0x0C3u;
    v = (v | (v << 2)) & 0x249u;
    return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {  // 6 bits -> every second bit
    v &= 63u;
    v = (v | (v << 4)) & 0x30Fu;
    v = (v | (v << 2)) & 0x333u;
    v = (v | (v << 1)) & 0x555u;
    return v;
}
__global__ void k_cell_hist(const float* __restrict__ pts, int64_t n, int dim, float lx, float ly, float lz, float sx,
                            float sy, float sz, int* __restrict__ cell, int* __restrict__ hist) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nb = dim == 3 ? 16 : 64;
    auto q = [&](float p, float lo, float sc) {
        float t = (p - lo) * sc;
        int c = t == t ? (int)fminf(fmaxf(t, 0.0f), (float)(nb - 1)) : 0;
        return (uint32_t)c;
    };
    uint32_t c;
    if (dim == 3)
        c = spread3(q(pts[3 * i], lx, sx)) | (spread3(q(pts[3 * i + 1], ly, sy)) << 1) |
            (spread3(q(pts[3 * i + 2], lz, sz)) << 2);
    else
        c = spread2(q(pts[2 * i], lx, sx)) | (spread2(q(pts[2 * i + 1], ly, sy)) << 1);
    cell[i] = (int)c;
    atomicAdd(hist + c, 1);
}
// Exclusive scan of kCells counts by one block of 1024 threads (4 cells each).
__global__ void __launch_bounds__(1024) k_cell_scan(int* __restrict__ hist) {
    __shared__ int part[1024];
    int t = threadIdx.x;
    int v[4], s = 0;
    for (int k = 0; k < 4; ++k) v[k] = hist[4 * t + k], s += v[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int k = 0; k < 4; ++k) hist[4 * t + k] = run, run += v[k];
}
__global__ void k_cell_place(const int* __restrict__ cell, int64_t n, int* __restrict__ offs, int* __restrict__ order) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    order[atomicAdd(offs + cell[i], 1)] = (int)i;
}
__global__ void k_fill_u64(unsigned long long* p, int count, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}
__global__ void k_fill_i32(int* p, int count, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
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

// Polyline.sdf / Mesh.sdf (polyline.py:188-194, mesh.py:290-293) on the full
// scene: the closest-point distance signed by the side of the interpolated
// vertex normal (2D) or the face normal (3D); inside is sdf < 0.
__global__ void __launch_bounds__(kBlock) k_sdf(SceneDev sc, const float* __restrict__ pts, int64_t n, float* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sc.dim == 2 ? sdf2(sc, load_pt<2>(pts, i)) : sdf3(sc, load_pt<3>(pts, i));
}

// Polyline.sample_boundary (polyline.py:498-512) / Mesh.sample_boundary
// (mesh.py:474-490) on one sub-scene, from caller-supplied uniform draws in
// [0, 1): 2D one per sample (the edge CDF; its reused remainder is t), 3D three
// (face CDF, r1, r2).  A draw outside [0, 1) yields prim -1 and NaN values.
__global__ void __launch_bounds__(kBlock) k_sample_boundary(SceneDev sc, int sub, const float* __restrict__ u,
                                                            int64_t n, float* p, float* nrm, float* tuv, float* pdf,
                                                            int* prim) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SubDev& S = sc.sub[sub];
    const int D = sc.dim;
    const float nan = __builtin_nanf("");
    float q[3] = {nan, nan, nan}, nn[3] = {nan, nan, nan}, uv[2] = {nan, nan}, pd = nan;
    int idx = -1;
    const float u0 = u[D == 2 ? i : 3 * i];
    bool ok = u0 >= 0.0f && u0 < 1.0f;
    if (D == 3) {
        const float r1 = u[3 * i + 1], r2 = u[3 * i + 2];
        ok = ok && r1 >= 0.0f && r1 < 1.0f && r2 >= 0.0f && r2 < 1.0f;
    }
    if (ok) {
        float rnd, pmf;
        idx = cdf_sample(S, u0, rnd, pmf);
        if (D == 2) {
            const float* lp = S.local_prims + 4 * idx;
            F2 a = f2(lp[0], lp[1]), b = f2(lp[2], lp[3]);
            F2 tg = b - a;
            F2 nv = normalize(f2(tg.y, -tg.x));
            F2 x = lerp(a, b, rnd);
            q[0] = x.x, q[1] = x.y, nn[0] = nv.x, nn[1] = nv.y;
            uv[0] = rnd, uv[1] = 0.0f;
            pd = pmf / norm(tg);
        } else {
            const float* lp = S.local_prims + 9 * idx;
            F3 a = f3(lp[0], lp[1], lp[2]), b = f3(lp[3], lp[4], lp[5]), c = f3(lp[6], lp[7], lp[8]);
            F3 cr = cross(b - a, c - a);
            float area = norm(cr) / 2.0f;
            float r1 = u[3 * i + 1], r2 = u[3 * i + 2];
            float su = __builtin_sqrtf(r1) * (1.0f - r2);
            float sv = __builtin_sqrtf(r1) * r2;
            F3 x = (b * su + c * sv) + a * ((1.0f - su) - sv);  // interpolate (math.py:306-307)
            F3 nv = normalize(cr);
            q[0] = x.x, q[1] = x.y, q[2] = x.z, nn[0] = nv.x, nn[1] = nv.y, nn[2] = nv.z;
            uv[0] = su, uv[1] = sv;
            pd = pmf / area;
        }
    }
    for (int d = 0; d < D; ++d) {
        if (p) p[D * i + d] = q[d];
        if (nrm) nrm[D * i + d] = nn[d];
    }
    if (tuv) tuv[2 * i] = uv[0], tuv[2 * i + 1] = uv[1];
    if (pdf) pdf[i] = pd;
    if (prim) prim[i] = idx;
}

// The reference's brute-force queries, one thread per query looping over the
// whole sub-scene in the reference's order: closest_point_baseline
// (polyline.py:246-279, mesh.py:312-335), intersect_baseline (polyline.py:468-496,
// mesh.py:440-472) and closest_silhouette_baseline (polyline.py:289-302,
// mesh.py:558-570) over the topology silhouettes.  The first of equally near
// candidates in that order wins (strict <); silhouettes come in reference-ref
// order and tie on their topology id.  prim = -1 when nothing qualifies.
__global__ void __launch_bounds__(kBlock) k_baseline(SceneDev sc, int sub, int kind, const float* __restrict__ pts,
                                                     const float* __restrict__ dir, const float* __restrict__ rmax,
                                                     int64_t n, int* prim, float* dist) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SubDev& S = sc.sub[sub];
    const int D = sc.dim;
    const float r0 = rmax ? rmax[i] : kInf;
    float dmin = kind == DW_BASE_CLOSEST ? kInf : r0;
    int best = -1;
    if (D == 2) {
        F2 x = load_pt<2>(pts, i);
        if (kind == DW_BASE_CLOSEST || kind == DW_BASE_RAY) {
            F2 v = kind == DW_BASE_RAY ? load_pt<2>(dir, i) : f2(0.0f, 0.0f);
            for (int j = 0; j < S.n_prims; ++j) {
                const float* lp = S.local_prims + 4 * j;
                F2 a = f2(lp[0], lp[1]), b = f2(lp[2], lp[3]);
                float d;
                if (kind == DW_BASE_CLOSEST) {
                    F2 pa = x - a, ba = b - a;
                    float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
                    d = norm(pa - ba * h);
                } else {  // ray_intersection (math.py:163-171)
                    F2 uu = b - a, w = x - a;
                    float dd = cross2(v, uu);
                    float s_ = cross2(v, w) / dd;
                    float t_ = cross2(uu, w) / dd;
                    d = ((t_ > 0.0f) && (s_ >= 0.0f) && (s_ <= 1.0f)) ? t_ : kInf;
                }
                if (d < dmin) dmin = d, best = j;
            }
        } else {
            const DSilV* sv = (const DSilV*)S.sils;
            for (int r = 0; r < S.n_refs; ++r) {
                F2 a = f2(sv[r].ax, sv[r].ay), b = f2(sv[r].bx, sv[r].by), c = f2(sv[r].cx, sv[r].cy);
                float d = norm(x - b);
                if (!(sv[r].both && d < r0)) continue;
                if (!(cross2(b - a, x - a) * cross2(c - b, x - b) < 0.0f)) continue;
                int topo = S.ref_topo[r];
                if (d < dmin || (d == dmin && topo < best)) dmin = d, best = topo;
            }
        }
    } else {
        F3 x = load_pt<3>(pts, i);
        if (kind == DW_BASE_CLOSEST || kind == DW_BASE_RAY) {
            F3 v = kind == DW_BASE_RAY ? load_pt<3>(dir, i) : f3(0.0f, 0.0f, 0.0f);
            for (int j = 0; j < S.n_prims; ++j) {
                const float* lp = S.local_prims + 9 * j;
                DTri tr;
                std::memset(&tr, 0, sizeof(tr));
                tr.ax = lp[0], tr.ay = lp[1], tr.az = lp[2], tr.bx = lp[3], tr.by = lp[4], tr.bz = lp[5];
                tr.cx = lp[6], tr.cy = lp[7], tr.cz = lp[8];
                float d;
                if (kind == DW_BASE_CLOSEST) {
                    F3 pt;
                    float cu, cv;
                    tri_closest(tr, x, pt, cu, cv, d);
                } else {
                    float bu, bw, t_;
                    d = tri_hit(tr, x, v, kInf, bu, bw, t_) ? t_ : kInf;
                }
                if (d < dmin) dmin = d, best = j;
            }
        } else {
            const DSilE* se = (const DSilE*)S.sils;
            for (int r = 0; r < S.n_refs; ++r) {
                float d;
                F3 pt;
                if (!sile_test(se[r], x, r0, d, pt)) continue;
                int topo = S.ref_topo[r];
                if (d < dmin || (d == dmin && topo < best)) dmin = d, best = topo;
            }
        }
    }
    prim[i] = best;
    if (dist) dist[i] = best >= 0 ? dmin : kInf;
}

// WoStGrad.solve_dude through Solver.solve_dude (solver.py:113-144) and
// walk_dude (:146-166): lane i * nwalks + k runs nsubwalks dudes on the PCG32
// stream of lane (point_base + i) * nwalks + k and writes their mean.
template <int D>
__global__ void __launch_bounds__(kBlock, 2) k_grad_dude(SceneDev sc, BCDev bc, ParamsDev P, GradDev G,
                                                      const float* __restrict__ pts, const float* __restrict__ ev,
                                                      int64_t n, float* walk_val, int* walk_steps) {
    int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= n * P.nwalks) return;
    int64_t i = lane / P.nwalks;
    Rng rng;
    seed_stream(rng, P, i, P.nwalks, lane - i * P.nwalks);
    auto x = load_pt<D>(pts, i), e = load_pt<D>(ev, i);
    float v = 0.0f;
    int st = 0;
    for (int s = 0; s < P.nsubwalks; ++s) {
        if constexpr (D == 2)
            v += grad_dude2(sc, bc, P, G, x, e, rng, st);
        else
            v += grad_dude3(sc, bc, P, G, x, e, rng, st);
    }
    walk_val[lane] = v / (float)P.nsubwalks;
    if (walk_steps) walk_steps[lane] = st;
}

// The same walks as persistent jobs: a lane takes job i * nwalks + k from the
// queue, runs its nsubwalks walks one step per loop iteration and takes the
// next job as soon as they end, so a wave is not held by its longest walk.  In
// 2D a Dirichlet point's nested dudn walk (up to 200 WoSt steps) also runs one
// step per iteration, beside the other lanes' WoStGrad steps.
template <int D>
__global__ void __launch_bounds__(kBlock, D == 2 ? 4 : 2) k_grad_dude_pers(SceneDev sc, BCDev bc, ParamsDev P, GradDev G,
                                                              const float* __restrict__ pts,
                                                              const float* __restrict__ ev, int64_t n,
                                                              float* walk_val, int* walk_steps, JobQ Q) {
    GradWalk<D> g;
    Rng rng;
    int64_t job = -1;
    float v = 0.0f;
    int st = 0, sub = 0;
    bool busy = false, done = false;
    bool nested = false;  // 2D: a dudn walk in d is pending
    Dudn2 d;
    const ParamsDev W = dudn_params(P);
    QState qs;
    q_start(qs);
    while (true) {
        int64_t got = q_next(!done && !busy, Q, qs);
        if (got == -2) done = true;
        if (got >= 0) {
            job = got;
            int64_t i = job / P.nwalks;
            seed_stream(rng, P, i, P.nwalks, job - i * P.nwalks);
            grad_init(sc, g, load_pt<D>(pts, i), load_pt<D>(ev, i));
            v = 0.0f;
            st = 0;
            sub = 0;
            busy = true;
        }
        if (__ballot(!done) == 0) break;
        if (busy) {
            if constexpr (D == 2) {
                if (nested) {
                    if (!walk_step<true>(sc, bc, W, d.w, rng)) {
                        grad_step2_end(sc, bc, P, g, rng, d, st);
                        nested = false;
                    }
                } else if (grad_running(g, P)) {
                    nested = grad_step2_begin(sc, bc, P, G, g, rng, d);
                }
            } else if (grad_running(g, P)) {
                grad_step3(sc, bc, P, G, g, rng, st);
            }
            if (!nested && !grad_running(g, P)) {  // the walk has ended
                st += g.i;
                v += g.result;
                if (++sub < P.nsubwalks) {
                    int64_t i = job / P.nwalks;
                    grad_init(sc, g, load_pt<D>(pts, i), load_pt<D>(ev, i));
                } else {
                    walk_val[job] = v / (float)P.nsubwalks;
                    if (walk_steps) walk_steps[job] = st;
                    busy = false;
                }
            }
        }
    }
}

// StreamlineTracer.next_position (examples/example_2d/trace.py:19-68) with a
// WoStGrad solver: sample j of point i on the stream of lane i * tn + j,
// walk_dude = the mean of nsubwalks WoStGrad dudes; the 2D step rule.
__global__ void __launch_bounds__(kBlock, 2) k_grad_trace(SceneDev sc, BCDev bc, ParamsDev P, GradDev G,
                                                       const float* __restrict__ pts, int64_t n, int tn, float step,
                                                       float* samples, unsigned long long* step_total) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n * tn) return;
    int64_t i = j / tn;
    Rng rng;
    seed_stream(rng, P, i, tn, j - i * tn);
    int st = 0;
    F2 p = load_pt<2>(pts, i);
    auto walk_dude = [&](F2 e) -> float {
        float r = 0.0f;
        for (int k = 0; k < P.nsubwalks; ++k) r += grad_dude2(sc, bc, P, G, p, e, rng, st);
        return r / (float)P.nsubwalks;
    };
    F2 cur = p;
    if (!(sdf2(sc, p) > 0.0f)) {
        float gx = walk_dude(f2(1.0f, 0.0f));
        float gy = walk_dude(f2(0.0f, 1.0f));
        cur = p + f2(gx, gy) * step;
        F2 yt = cur;
        if (!(sdf2(sc, cur) > 0.0f)) {
            F2 fy = yt - p, fyt = cur - yt;
            cur = p + (fy + fyt) * 0.5f;
        }
    }
    samples[2 * j] = cur.x;
    samples[2 * j + 1] = cur.y;
    if (step_total) atomicAdd(step_total, (unsigned long long)st);
}

// The same samples as persistent jobs: a lane takes sample j = i * tn + k,
// runs its x and then its y derivative walks one WoStGrad step (or one step of
// a Dirichlet point's nested dudn walk) per loop iteration, and takes the next
// sample when both are done.  Same draws in the same order as k_grad_trace.
__global__ void __launch_bounds__(kBlock, 3) k_grad_trace_pers(SceneDev sc, BCDev bc, ParamsDev P, GradDev G,
                                                            const float* __restrict__ pts, int64_t n, int tn,
                                                            float step, float* samples,
                                                            unsigned long long* step_total, JobQ Q) {
    GradWalk<2> g;
    Rng rng;
    Dudn2 d;
    const ParamsDev W = dudn_params(P);
    int64_t job = -1;
    F2 p = f2(0.0f, 0.0f);
    float acc = 0.0f, gx = 0.0f;
    int comp = 0, sub = 0, st = 0;  // comp: 0 while the x walks run, 1 for y
    bool busy = false, done = false, nested = false;
    QState qs;
    q_start(qs);
    while (true) {
        int64_t got = q_next(!done && !busy, Q, qs);
        if (got == -2) done = true;
        if (got >= 0) {
            job = got;
            int64_t i = job / tn;
            seed_stream(rng, P, i, tn, job - i * tn);
            p = load_pt<2>(pts, i);
            if (!(sdf2(sc, p) > 0.0f)) {
                grad_init(sc, g, p, f2(1.0f, 0.0f));
                comp = 0, sub = 0, st = 0, acc = 0.0f;
                busy = true;
            } else {  // outside: the sample stays at p
                samples[2 * job] = p.x;
                samples[2 * job + 1] = p.y;
            }
        }
        if (__ballot(!done) == 0) break;
        if (busy) {
            if (nested) {
                if (!walk_step<true>(sc, bc, W, d.w, rng)) {
                    grad_step2_end(sc, bc, P, g, rng, d, st);
                    nested = false;
                }
            } else if (grad_running(g, P)) {
                nested = grad_step2_begin(sc, bc, P, G, g, rng, d);
            }
            if (!nested && !grad_running(g, P)) {  // a walk has ended
                st += g.i;
                acc += g.result;
                if (++sub < P.nsubwalks) {
                    grad_init(sc, g, p, comp == 0 ? f2(1.0f, 0.0f) : f2(0.0f, 1.0f));
                } else if (comp == 0) {
                    gx = acc / (float)P.nsubwalks;
                    comp = 1, sub = 0, acc = 0.0f;
                    grad_init(sc, g, p, f2(0.0f, 1.0f));
                } else {
                    float gy = acc / (float)P.nsubwalks;
                    F2 cur = p + f2(gx, gy) * step;
                    F2 yt = cur;
                    if (!(sdf2(sc, cur) > 0.0f)) {
                        F2 fy = yt - p, fyt = cur - yt;
                        cur = p + (fy + fyt) * 0.5f;
                    }
                    samples[2 * job] = cur.x;
                    samples[2 * job + 1] = cur.y;
                    if (step_total) atomicAdd(step_total, (unsigned long long)st);
                    busy = false;
                }
            }
        }
    }
}

// Lane scheduling of the persistent walk kernels.  Each lane owns one job (one
// or more walks, see the kernels) and runs its walk as the step machine of
// walk_advance.  Two loops:
//  * step-synchronous: each loop iteration advances every lane's walk by one
//    whole step, its traversals run back to back;
//  * interleaved: the loop body runs one traversal iteration for every lane
//    with a query in flight.  A lane whose query has ended waits; once Q.gate
//    lanes of the wave wait (or none is traversing) the waiting lanes run their
//    walk logic up to their next query, idle lanes take new jobs, and all
//    rejoin the traversal.  A wave is then held neither by its longest walk nor
//    by its longest traversal, while the walk logic between queries (sampling,
//    transcendentals) still runs for many lanes at once.
// The interleaved loop pays off where traversal lengths diverge most: on the
// 81,920-triangle icosphere (3D, 262,144 points) it measured 1.3x the
// step-synchronous loop, on the 2D holes scene (65,536 points) 0.54x.
// start(job, w, rng): set up the job's first walk (false: nothing to walk).
// next(w, rng): the walk in w has ended; set up the job's next one (false: the
// job is complete).  advance(w, rng, t): walk_advance with the job's scene.
template <int D, bool X, bool I, typename Start, typename Next, typename Adv>
DWD void run_jobs(const JobQ& Q, Start& start, Next& next, Adv& advance) {
    Walk<D> w;
    Rng rng;
    Trav t;
    Ovf o;
    bool busy = false, trav = false, done = false;
    QState qs;
    q_start(qs);
    if constexpr (!I) {
        // one whole step per lane and loop iteration, traversals back to back
        while (true) {
            int64_t got = q_next(!done && !busy, Q, qs);
            if (got == -2) done = true;
            if (got >= 0) busy = start(got, w, rng);
            if (__ballot(!done) == 0) break;
            if (busy) {
                int r;
                while ((r = advance(w, rng, t)) == ADV_QUERY) {
                    while (!trav_iter<D, X>(t, o)) {
                    }
                }
                if (r == ADV_END && !next(w, rng)) busy = false;
            }
        }
        return;
    }
    while (true) {
        if (__ballot(trav) == 0 || __popcll(__ballot(!trav && !done)) >= Q.gate) {
            int64_t got = q_next(!done && !busy, Q, qs);
            if (got == -2) done = true;
            if (got >= 0) busy = start(got, w, rng);
            if (busy && !trav) {
                while (true) {
                    int r = advance(w, rng, t);
                    if (r == ADV_QUERY) {
                        trav = true;
                        break;
                    }
                    if (r == ADV_END && !next(w, rng)) {
                        busy = false;
                        break;
                    }
                }
            }
            if (__ballot(!done) == 0) break;
        }
        if (trav && trav_iter<D, X>(t, o)) trav = false;
    }
}
#define DW_LAMBDA __attribute__((always_inline))

// Solver.solve jobs: job = i*W + k, nsubwalks walks on the job's stream.
template <int D, bool X, bool I>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST)
    k_walk_u_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P, const float* __restrict__ pts, int64_t n,
                  float* walk_u, int* walk_steps, int* walk_term, uint8_t* walk_flags, JobQ Q) {
    const SceneDev& sc = *scp;
    using V = typename VecT<D>::T;
    int64_t job = -1;
    V x;
    float res = 0.0f;
    int steps = 0, sub = 0;
    uint32_t kept = 0;
    auto start = [&](int64_t got, Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        job = got;
        kept = 0;
        int64_t i = job / P.nwalks;
        seed_stream(rng, P, i, P.nwalks, job - i * P.nwalks);
        x = load_pt<D>(pts, i);
        walk_init<D>(w, x);
        res = 0.0f;
        steps = 0;
        sub = 0;
        return true;
    };
    auto next = [&](Walk<D>& w, Rng&) DW_LAMBDA -> bool {
        float v = walk_value(w, P);
        if (v == w.result) kept |= 1u << (sub & 7);
        res += v;
        steps += w.i;
        if (++sub < P.nsubwalks) {
            walk_init<D>(w, x);
            return true;
        }
        walk_u[job] = res / (float)P.nsubwalks;
        if (walk_steps) walk_steps[job] = steps;
        if (walk_term) walk_term[job] = w.term;
        if (walk_flags) walk_flags[job] = (uint8_t)kept;
        return false;
    };
    auto advance = [&](Walk<D>& w, Rng& rng, Trav& t) DW_LAMBDA -> int {
        return walk_advance<X, false>(sc, bc, P, w, rng, t, Adj{0.0f, nullptr, nullptr});
    };
    run_jobs<D, X, I>(Q, start, next, advance);
}


// Distance to the full boundary and the inside flag per point (used by the
// gradient estimators, wost.py:43-50).
// Rd (optional): the distance to the Dirichlet part (inf without one), from
// which the fused walks bound their first Dirichlet query (dirichlet_bound).
DWD void prep_point(const SceneDev& sc, const float* __restrict__ pts, int64_t i, float* R, int* inside,
                    float* Rd = nullptr) {
    if (sc.dim == 2) {
        F2 x = f2(pts[2 * i], pts[2 * i + 1]);
        R[i] = __builtin_fabsf(closest2(sc.sub[0], x).d);
        inside[i] = inside2(sc, x) ? 1 : 0;
        if (Rd) Rd[i] = sc.sub[1].present ? __builtin_fabsf(closest2(sc.sub[1], x).d) : kInf;
    } else {
        F3 x = f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        R[i] = __builtin_fabsf(closest3(sc.sub[0], x).d);
        inside[i] = inside3(sc, x) ? 1 : 0;
        if (Rd) Rd[i] = sc.sub[1].present ? __builtin_fabsf(closest3(sc.sub[1], x).d) : kInf;
    }
}
__global__ void __launch_bounds__(kBlock) k_point_prep(SceneDev sc, const float* __restrict__ pts, int64_t n,
                                                       float* R, int* inside, float* Rd = nullptr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    prep_point(sc, pts, i, R, inside, Rd);
}

// One scene's share of a fused launch: geometry, boundary data, walk
// parameters, its points, their first-ball radii / inside flags and where the
// per-pair partial results go.  A plain solve is a batch of one.
struct FusedItem {
    const SceneDev* sc;
    BCDev bc;
    ParamsDev P;
    const float* pts;
    const float* R;
    const int* inside;
    const float* Rd;   // distance to the Dirichlet part per point (nullptr: not known)
    float* pair_vals;  // [n * npairs][dim + 1]
    int* pair_steps;   // [n * npairs]
    uint8_t* pair_flags;
    float* u;
    float* grad;
    int64_t n;
    int64_t job0;  // first job of this item in the batch's job space
    int64_t pt0;   // first point of this item in the batch's point space
    int npairs;
    int per;  // jobs per point: ceil(npairs / walks_per_lane)
};
int jobs_per_point(int npairs, int walks_per_lane) { return (npairs + walks_per_lane - 1) / walks_per_lane; }
// Index of the item holding global index g, by the item starts `first` (sorted).
template <typename F>
DWD int find_item(int n_items, int64_t g, F first) {
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (first(mid) <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Fused jobs: job = i*per + k runs the antithetic pairs k*m .. k*m + m - 1 of
// point i (m = walks_per_lane, per = ceil(npairs / m)) back to back, each pair
// on its own stream (point i, pair j), so m only changes the scheduling.  A
// pair is WoSt.dude's first ball: its walks start at p + dir R and p - dir R.
// B = false: one item (kernel argument), jobs from the spatial queue.
// B = true: many items (device array); the queue runs over the concatenated
// job space and each lane looks up the item of its job.
template <int D, bool B, bool X, bool I>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST)
    k_fused_pers(FusedItem one, const FusedItem* __restrict__ items, int n_items, JobQ Q) {
    using V = typename VecT<D>::T;
#define ITEM(f) (B ? items[it].f : one.f)
    int it = 0;
    int64_t job = -1, pt = 0;  // current pair (i * npairs + j) and its point i
    int left = 0;              // pairs of the job after the current one
    // Held in registers across the pair: its direction, the stream state
    // after the first draws and the counters.  The point, its first-ball
    // radius and the pair's source terms stay in memory until the pair's end
    // (point and radius in the launch's arrays; (s+ + s-, (s+ - s-) gs) in the
    // pair's output slot, the first walk's value in its step-count slot), so
    // they take no registers during the walks.
    V dir;
    uint64_t saved = 0;
    int phase = 0, steps = 0;
    bool kept_plus = false;
    // a walk of the pair starts at x, |x - p| = R: its first Dirichlet query is
    // bounded by the point's Dirichlet distance + R
    auto start_walk = [&](Walk<D>& w, V x, float R) DW_LAMBDA {
        walk_init<D>(w, x);
        if (ITEM(Rd)) {
            w.d_dir = ITEM(Rd)[pt];
            w.R = R;
        }
    };
    // first walk of pair `job` of point `pt`
    auto begin_pair = [&](Walk<D>& w, Rng& rng) DW_LAMBDA {
        const int npairs = ITEM(npairs);
        const float R = ITEM(R)[pt];
        const V p = load_pt<D>(ITEM(pts), pt);
        seed_stream(rng, ITEM(P), pt, npairs, job - pt * npairs);
        V p_src, srcg;
        float r_src, pdf_src, s_sum;
        // the source at the sampled point and at its mirror 2p - p_src (the
        // antithetic sample, wost.py:62-77)
        if constexpr (D == 2) {
            dir = uniform_dir2(rng);
            F2 dir_src = uniform_dir2(rng);
            greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
            V gs = greens2_gradient(p, R, r_src, p_src) / (pdf_src * (R * R / 4.0f));
            float sp = (R * R / 4.0f) * src_f(ITEM(bc), p_src), sm = (R * R / 4.0f) * src_f(ITEM(bc), p * 2.0f - p_src);
            srcg = gs * (sp - sm);
            s_sum = sp + sm;
        } else {
            dir = uniform_sphere(rng);
            greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
            V gs = greens3_gradient(p, R, r_src, p_src) / (pdf_src * (R * R / 6.0f));
            float sp = (R * R / 6.0f) * src_f(ITEM(bc), p_src), sm = (R * R / 6.0f) * src_f(ITEM(bc), p * 2.0f - p_src);
            srcg = gs * (sp - sm);
            s_sum = sp + sm;
        }
        float* o = ITEM(pair_vals) + job * (D + 1);
        o[0] = s_sum;
        o[1] = srcg.x;
        o[2] = srcg.y;
        if constexpr (D == 3) o[3] = srcg.z;
        saved = rng.state;
        start_walk(w, p + dir * R, R);
        phase = 0;
        steps = 0;
    };
    auto start = [&](int64_t got, Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        if constexpr (B) {
            it = find_item(n_items, got, [&](int k) { return items[k].job0; });
            got -= items[it].job0;
        }
        const int npairs = ITEM(npairs), per = ITEM(per);
        const int m = ITEM(P).walks_per_lane;
        int64_t i = got / per;
        int j0 = (int)(got - i * per) * m;
        int cnt = npairs - j0 < m ? npairs - j0 : m;
        if (!ITEM(inside)[i]) {
            for (int j = j0; j < j0 + cnt; ++j) {
                int64_t q = i * npairs + j;
                float* o = ITEM(pair_vals) + q * (D + 1);
                for (int c = 0; c <= D; ++c) o[c] = 0.0f;
                ITEM(pair_steps)[q] = 0;
                if (ITEM(pair_flags)) ITEM(pair_flags)[q] = 0;
            }
            return false;
        }
        pt = i;
        job = i * npairs + j0;
        left = cnt - 1;
        begin_pair(w, rng);
        return true;
    };
    auto next = [&](Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        float v = walk_value(w, ITEM(P));
        steps += w.i;
        // walks the guards zero contribute nothing to the adjoint
        bool kept = D == 2 ? v == w.result : !(bad(v) || __builtin_fabsf(v) > 10000.0f);
        const float R = ITEM(R)[pt];
        if (phase == 0) {
            kept_plus = kept;
            ITEM(pair_steps)[job] = __float_as_int(v);  // the first walk's value until the pair ends
            rng.state = saved;
            start_walk(w, load_pt<D>(ITEM(pts), pt) - dir * R, R);
            phase = 1;
            return true;
        }
        const float up = __int_as_float(ITEM(pair_steps)[job]), um = v;
        float* o = ITEM(pair_vals) + job * (D + 1);
        const float s_sum = o[0];
        if constexpr (D == 2) {
            float k = 2.0f / R * (up - um);
            float g0 = k * dir.x + o[1], g1 = k * dir.y + o[2];
            o[0] = (up + um) + s_sum;
            o[1] = bad(g0) ? 0.0f : g0;
            o[2] = bad(g1) ? 0.0f : g1;
        } else {
            float u1 = (bad(up) || __builtin_fabsf(up) > 10000.0f) ? 0.0f : up;
            float u2 = (bad(um) || __builtin_fabsf(um) > 10000.0f) ? 0.0f : um;
            float k = 3.0f / R * (u1 - u2);
            float g0 = k * dir.x + o[1], g1 = k * dir.y + o[2], g2_ = k * dir.z + o[3];
            o[0] = (u1 + u2) + s_sum;
            o[1] = bad(g0) ? 0.0f : g0;
            o[2] = bad(g1) ? 0.0f : g1;
            o[3] = bad(g2_) ? 0.0f : g2_;
        }
        ITEM(pair_steps)[job] = steps;
        if (ITEM(pair_flags)) ITEM(pair_flags)[job] = (uint8_t)((kept_plus ? 1 : 0) | (kept ? 2 : 0));
        if (left > 0) {  // the job's next pair, same point
            --left;
            ++job;
            begin_pair(w, rng);
            return true;
        }
        return false;
    };
    auto advance = [&](Walk<D>& w, Rng& rng, Trav& t) DW_LAMBDA -> int {
        return walk_advance<X, false>(*ITEM(sc), ITEM(bc), ITEM(P), w, rng, t, Adj{0.0f, nullptr, nullptr});
    };
    run_jobs<D, X, I>(Q, start, next, advance);
#undef ITEM
}


// Batched point preparation and reduction: thread (wave) per point of the
// concatenated point space.
__global__ void __launch_bounds__(kBlock) k_point_prep_batch(const FusedItem* __restrict__ items, int n_items,
                                                             int64_t total) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int it = find_item(n_items, g, [&](int k) { return items[k].pt0; });
    const FusedItem& I = items[it];
    prep_point(*I.sc, I.pts, g - I.pt0, (float*)I.R, (int*)I.inside, (float*)I.Rd);
}
__global__ void __launch_bounds__(kBlock) k_reduce_batch(const FusedItem* __restrict__ items, int n_items,
                                                         int64_t total, int dim, unsigned long long* step_total) {
    int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (g >= total) return;
    int it = find_item(n_items, g, [&](int k) { return items[k].pt0; });
    const FusedItem& I = items[it];
    reduce_point(I.pair_vals, g - I.pt0, I.npairs, dim + 1, 1.0f / (float)(2 * I.npairs), I.u, I.grad,
                 step_total ? I.pair_steps : nullptr, step_total, threadIdx.x & 63);
}

// Adjoint of the fused estimator: replays every pair from its seed with the
// upstream weight of each walk, dL/du+- = (dL/du[i] +- (dim/R) dL/dgrad[i].dir)
// / (2 npairs), and scatter-adds dL/dg, dL/dh per vertex.  The walks are linear
// in the per-vertex boundary data for fixed random numbers; walks zeroed by the
// outlier guards in the forward pass (pair_flags) are skipped.
template <int D, bool X, bool I>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST)
    k_adjoint_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                   const float* __restrict__ Rp, const int* __restrict__ insidep, int64_t n, int npairs,
                   const float* __restrict__ dl_du, const float* __restrict__ dl_dgrad,
                   const uint8_t* __restrict__ pair_flags, float* dl_dg, float* dl_dh, JobQ Q) {
    const SceneDev& sc = *scp;
    using V = typename VecT<D>::T;
    const float inv = 1.0f / (float)(2 * npairs);
    V p, dir;
    float R = 0.0f, wm = 0.0f;
    Adj adj{0.0f, dl_dg, dl_dh};
    uint64_t saved = 0;
    int phase = 0;
    auto start = [&](int64_t got, Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        int64_t i = got / npairs;
        uint8_t fl = pair_flags[got];
        if (!(insidep[i] && fl)) return false;
        R = Rp[i];
        p = load_pt<D>(pts, i);
        seed_stream(rng, P, i, npairs, got - i * npairs);
        V p_src;
        float r_src, pdf_src;
        float proj;
        if constexpr (D == 2) {
            dir = uniform_dir2(rng);
            F2 dir_src = uniform_dir2(rng);
            greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
            proj = dl_dgrad[2 * i] * dir.x + dl_dgrad[2 * i + 1] * dir.y;
        } else {
            dir = uniform_sphere(rng);
            greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
            proj = (dl_dgrad[3 * i] * dir.x + dl_dgrad[3 * i + 1] * dir.y) + dl_dgrad[3 * i + 2] * dir.z;
        }
        float k = (float)D / R * proj;
        adj.w = (fl & 1) ? (dl_du[i] + k) * inv : 0.0f;
        wm = (fl & 2) ? (dl_du[i] - k) * inv : 0.0f;
        saved = rng.state;
        walk_init<D>(w, p + dir * R);
        phase = 0;
        return true;
    };
    auto next = [&](Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        if (phase != 0) return false;
        rng.state = saved;
        walk_init<D>(w, p - dir * R);
        adj.w = wm;
        phase = 1;
        return true;
    };
    auto advance = [&](Walk<D>& w, Rng& rng, Trav& t) DW_LAMBDA -> int {
        return walk_advance<X, true>(sc, bc, P, w, rng, t, adj);
    };
    run_jobs<D, X, I>(Q, start, next, advance);
}

// Adjoint of Solver.solve: job i*W + k replays its nsubwalks walks with weight
// dL/du[i] / (W * nsubwalks) each (walks zeroed by the guard are skipped).
template <int D, bool X, bool I>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST)
    k_adjoint_u_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                     int64_t n, const float* __restrict__ dl_du, const uint8_t* __restrict__ walk_flags, float* dl_dg,
                     float* dl_dh, JobQ Q) {
    const SceneDev& sc = *scp;
    using V = typename VecT<D>::T;
    const float inv = 1.0f / ((float)P.nwalks * (float)P.nsubwalks);
    V x;
    Adj adj{0.0f, dl_dg, dl_dh};
    int sub = 0;
    uint32_t fl = 0;
    float wi = 0.0f;
    auto start = [&](int64_t got, Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        int64_t i = got / P.nwalks;
        fl = walk_flags[got];
        if (!fl) return false;
        seed_stream(rng, P, i, P.nwalks, got - i * P.nwalks);
        x = load_pt<D>(pts, i);
        walk_init<D>(w, x);
        sub = 0;
        wi = dl_du[i] * inv;
        adj.w = (fl & 1u) ? wi : 0.0f;
        return true;
    };
    auto next = [&](Walk<D>& w, Rng&) DW_LAMBDA -> bool {
        if (++sub >= P.nsubwalks) return false;
        walk_init<D>(w, x);
        adj.w = (fl & (1u << (sub & 7))) ? wi : 0.0f;
        return true;
    };
    auto advance = [&](Walk<D>& w, Rng& rng, Trav& t) DW_LAMBDA -> int {
        return walk_advance<X, true>(sc, bc, P, w, rng, t, adj);
    };
    run_jobs<D, X, I>(Q, start, next, advance);
}


// The walk sequence of WoSt.dude / WoSt3D.dude (wost.py:30-102,
// wost_3d.py:108-175) and, with DW_SOLVE_WOS, WoS.dude / WoS3D.dude
// (wos.py:38-105, wos_3d.py:36-102) at one point for one direction, as a
// resumable machine: nwalks walks in order on the caller's stream, each
// antithetic pair restarting from the state its first-ball draws leave and the
// next pair continuing where the second walk ended (the arithmetic of dude2 /
// dude3 and of k_dude's WoS branch).  Running sums in 2D and for WoS, running
// means in WoSt3D; WoS has no outlier guards.
template <int D>
struct DudeRun {
    using V = typename VecT<D>::T;
    V p, ev, dir, dir_src, p_src;
    float R, r_src, pdf_src, result, u_acc, u_count, src_acc, src_count;
    uint64_t state;
    int wi, it, niters;

    DWD static bool sums(const ParamsDev& P) { return D == 2 || (P.flags & DW_SOLVE_WOS); }

    // Starts the estimate; true when its first walk is set up in w.
    DWD bool begin(V p_, V e_, float R_, bool active, const ParamsDev& P, const BCDev& bc, Walk<D>& w, Rng& rng) {
        p = p_;
        ev = e_;
        R = R_;
        result = u_acc = u_count = src_acc = src_count = 0.0f;
        wi = 0;
        if (!active || !begin_pair(P, rng)) return false;
        begin_walk(P, bc, w, rng);
        return true;
    }
    // The walk in w has ended; true when the next walk is set up.
    DWD bool walk_ended(const ParamsDev& P, const BCDev& bc, Walk<D>& w, Rng& rng, unsigned long long& steps) {
        float u = walk_value(w, P);
        steps += w.i;
        if (sums(P)) {
            const bool wos = (P.flags & DW_SOLVE_WOS) != 0;
            u_acc += u;
            u_count += 1.0f;
            float _u = P.cv ? u - u_acc / u_count : u;
            float r2 = (float)D / R * _u * dot(ev, dir);
            if (!wos && ((__builtin_fabsf(r2) > 1e6f) || bad(r2))) r2 = 0.0f;
            result += r2;
        } else {
            if (bad(u) || __builtin_fabsf(u) > 10000.0f) u = 0.0f;
            u_acc = u_acc * (u_count / (u_count + 1.0f)) + u / (u_count + 1.0f);
            u_count += 1.0f;
            float _u = P.cv ? u - u_acc : u;
            result += 3.0f / R * _u * dot(ev, dir);
        }
        wi += 1;
        it += 1;
        if ((it < niters && wi < P.nwalks) || begin_pair(P, rng)) {
            begin_walk(P, bc, w, rng);
            return true;
        }
        return false;
    }
    // The estimate: sum / nwalks (WoSt3D: outlier guard first).
    DWD float value(const ParamsDev& P) const {
        float r = result;
        if (!sums(P)) {
            if (bad(r) || __builtin_fabsf(r) > 10000.0f) r = 0.0f;
        }
        return r / (float)P.nwalks;
    }

    DWD bool begin_pair(const ParamsDev& P, Rng& rng) {
        if (!(wi < P.nwalks)) return false;
        if constexpr (D == 2) {
            dir = uniform_dir2(rng);
            dir_src = uniform_dir2(rng);
            greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
        } else if (P.flags & DW_SOLVE_WOS) {  // its own source direction
            dir = uniform_sphere(rng);
            dir_src = uniform_sphere(rng);
            greens3_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
        } else {
            dir = uniform_sphere(rng);
            greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
        }
        niters = P.av ? 2 : 1;
        state = rng.state;
        it = 0;
        return true;
    }
    DWD void begin_walk(const ParamsDev& P, const BCDev& bc, Walk<D>& w, Rng& rng) {
        const float gnorm = D == 2 ? R * R / 4.0f : R * R / 6.0f;
        float src, src_grad;
        if (sums(P)) {
            if (it == 1) {
                dir_src = -dir_src;
                p_src = p * 2.0f - p_src;
            }
            src = gnorm * src_f(bc, p_src);
            if constexpr (D == 2)
                src_grad = dot(greens2_gradient(p, R, r_src, p_src), ev) / (pdf_src * gnorm);
            else
                src_grad = dot(greens3_gradient(p, R, r_src, p_src), ev) / (pdf_src * gnorm);
            src_acc += src;
            src_count += 1.0f;
            float _src = P.cv ? src - (src_acc / src_count) : src;
            result += _src * src_grad;
        } else {
            if (it == 1) {
                dir = -dir;
                p_src = p * 2.0f - p_src;
            }
            src = gnorm * src_f(bc, p_src);
            if constexpr (D == 3) {
                src_grad = dot(greens3_gradient(p, R, r_src, p_src), ev) / (pdf_src * gnorm);
                src_acc = src_acc * (src_count / (src_count + 1.0f)) + src / (src_count + 1.0f);
                src_count += 1.0f;
                float _src = P.cv ? src - src_acc : src;
                result += _src * src_grad;
            }
        }
        if (it == 1) dir = -dir;  // WoSt3D: the second flip (wost_3d.py:152-153)
        rng.state = state;
        walk_init<D>(w, p + dir * R);
    }
};

// WoSt.solve_dude / WoSt3D.solve_dude (wost.py:21-102, wost_3d.py:108-175), and
// the WoS dudes, as persistent jobs: job = point, its walks on the point's one
// stream (DudeRun);
// the lane scheduler refills lanes as points finish.
template <int D, bool X, bool I>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST)
    k_dude_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                const float* __restrict__ e, const float* __restrict__ Rp, const int* __restrict__ insidep, float* out,
                unsigned long long* step_total, JobQ Q) {
    const SceneDev& sc = *scp;
    int64_t job = -1;
    DudeRun<D> dude;
    unsigned long long steps = 0;
    auto start = [&](int64_t got, Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        job = got;
        seed_stream(rng, P, got, 1, 0);
        if (dude.begin(load_pt<D>(pts, got), load_pt<D>(e, got), Rp[got], insidep[got] != 0, P, bc, w, rng))
            return true;
        out[job] = dude.value(P);
        return false;
    };
    auto next = [&](Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        if (dude.walk_ended(P, bc, w, rng, steps)) return true;
        out[job] = dude.value(P);
        return false;
    };
    auto advance = [&](Walk<D>& w, Rng& rng, Trav& t) DW_LAMBDA -> int {
        return walk_advance<X, false>(sc, bc, P, w, rng, t, Adj{0.0f, nullptr, nullptr});
    };
    run_jobs<D, X, I>(Q, start, next, advance);
    if (step_total && steps) atomicAdd(step_total, steps);
}

// StreamlineTracer.next_position (examples/example_2d/trace.py:19-68,
// examples/example_3d/trace.py:21-60) as persistent jobs: job j = i * tn + k is
// sample k of point i (the stream of lane j, TEA(seed, i)); its gradient
// components are walk_dude (nsubwalks DudeRuns) along each axis in turn on that
// stream, then the step rule of k_trace.  samples[j] = the sample's position.
template <int D, bool X, bool I>
__global__ void __launch_bounds__(kBlock, X ? DW_PERS_WAVES_EXACT : DW_PERS_WAVES_FAST)
    k_trace_pers(const SceneDev* __restrict__ scp, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                 const float* __restrict__ Rp, const int* __restrict__ insidep, int tn, float step, float* samples,
                 unsigned long long* step_total, JobQ Q) {
    using V = typename VecT<D>::T;
    const SceneDev& sc = *scp;
    int64_t job = -1;
    V p, g;
    float R = 0.0f, comp = 0.0f;
    bool active = false;
    int axis = 0, sub = 0;
    DudeRun<D> dude;
    unsigned long long steps = 0;
    auto put = [&](V q) DW_LAMBDA {
        float* o = samples + job * D;
        o[0] = q.x;
        o[1] = q.y;
        if constexpr (D == 3) o[2] = q.z;
    };
    auto unit = [&](int a) DW_LAMBDA -> V {
        V u;
        set_zero(u);
        if (a == 0) u.x = 1.0f;
        if (a == 1) u.y = 1.0f;
        if constexpr (D == 3) {
            if (a == 2) u.z = 1.0f;
        }
        return u;
    };
    auto finish_sample = [&]() DW_LAMBDA {
        V cur = p + g * step;
        if constexpr (D == 2) {
            F2 yt = cur;
            if (!(sdf2(sc, cur) > 0.0f)) {
                F2 fy = yt - p, fyt = cur - yt;
                cur = p + (fy + fyt) * 0.5f;
            }
        }
        put(cur);
    };
    // Dude value done: fold it into the axis component; starts dudes until one
    // has a walk (true) or the sample is complete (false).
    auto next_dude = [&](Walk<D>& w, Rng& rng, bool have_value) DW_LAMBDA -> bool {
        while (true) {
            if (have_value) {
                comp += dude.value(P);
                if (++sub == P.nsubwalks) {
                    float c = comp / (float)P.nsubwalks;
                    if (axis == 0) g.x = c;
                    if (axis == 1) g.y = c;
                    if constexpr (D == 3) {
                        if (axis == 2) g.z = c;
                    }
                    ++axis;
                    sub = 0;
                    comp = 0.0f;
                }
            }
            if (axis == D) {
                finish_sample();
                return false;
            }
            if (dude.begin(p, unit(axis), R, active, P, bc, w, rng)) return true;
            have_value = true;
        }
    };
    auto start = [&](int64_t got, Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        job = got;
        int64_t i = got / tn;
        seed_stream(rng, P, i, tn, got - i * tn);
        p = load_pt<D>(pts, i);
        R = Rp[i];
        active = insidep[i] != 0;
        set_zero(g);
        if constexpr (D == 2) {
            if (sdf2(sc, p) > 0.0f) {  // should_increment: only from inside
                put(p);
                return false;
            }
        }
        axis = 0;
        sub = 0;
        comp = 0.0f;
        return next_dude(w, rng, false);
    };
    auto next = [&](Walk<D>& w, Rng& rng) DW_LAMBDA -> bool {
        if (dude.walk_ended(P, bc, w, rng, steps)) return true;
        return next_dude(w, rng, true);
    };
    auto advance = [&](Walk<D>& w, Rng& rng, Trav& t) DW_LAMBDA -> int {
        return walk_advance<X, false>(sc, bc, P, w, rng, t, Adj{0.0f, nullptr, nullptr});
    };
    run_jobs<D, X, I>(Q, start, next, advance);
    if (step_total && steps) atomicAdd(step_total, steps);
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
    Rng rng;
    seed_stream(rng, P, i, npairs, lane - i * npairs);
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
        float s_p = gnorm * src_f(bc, p_src), s_m = gnorm * src_f(bc, p * 2.0f - p_src);
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
        float s_p = gnorm * src_f(bc, p_src), s_m = gnorm * src_f(bc, p * 2.0f - p_src);
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

// WoSt.dude (wost.py:30-102) / WoSt3D.dude (wost_3d.py:108-175) at one point
// for one direction e: nwalks walks (antithetic pairs) from the first ball of
// radius R on the caller's stream; returns the sum, the caller divides by nwalks.
DWD float dude2(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, F2 p, F2 ev, float R, bool active, Rng& rng,
                unsigned long long& steps) {
    float result = 0.0f;
    int wi = 0, st, tm;
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
            float src = gnorm * src_f(bc, p_src);
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
    return result;
}
DWD float dude3(const SceneDev& sc, const BCDev& bc, const ParamsDev& P, F3 p, F3 ev, float R, bool active, Rng& rng,
                unsigned long long& steps) {
    float result = 0.0f;
    int wi = 0, st, tm;
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
            float src = gnorm * src_f(bc, p_src);
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
    return result;
}

// WoSt.solve_dude / WoSt3D.solve_dude (wost.py:21-102, wost_3d.py:108-175):
// one PCG32 stream per point, walks run serially inside the lane.
__global__ void __launch_bounds__(kBlock) k_dude(SceneDev sc, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                                                 const float* __restrict__ e, const float* __restrict__ Rp,
                                                 const int* __restrict__ insidep, int64_t n, float* out,
                                                 unsigned long long* step_total) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng rng;
    seed_stream(rng, P, i, 1, 0);
    float R = Rp[i];
    bool active = insidep[i] != 0;
    float result = 0.0f;
    int wi = 0;
    unsigned long long steps = 0;
    int st, tm;
    if (P.flags & DW_SOLVE_WOS) {
        // WoS.dude (wos.py:38-105) / WoS3D.dude (wos_3d.py:36-102): separate
        // source direction, running-sum control variates, no outlier guards
        const int D = sc.dim;
        F3 p = f3(pts[D * i], pts[D * i + 1], D == 3 ? pts[D * i + 2] : 0.0f);
        F3 ev = f3(e[D * i], e[D * i + 1], D == 3 ? e[D * i + 2] : 0.0f);
        float u_sum = 0.0f, u_count = 0.0f, src_sum = 0.0f, src_count = 0.0f;
        float gnorm = D == 2 ? R * R / 4.0f : R * R / 6.0f;
        while ((wi < P.nwalks) && active) {
            F3 dir, dir_src, p_src;
            float r_src, pdf_src;
            if (D == 2) {
                F2 d2 = uniform_dir2(rng), s2 = uniform_dir2(rng), q2;
                greens2_sample(f2(p.x, p.y), R, s2, rng, q2, r_src, pdf_src);
                dir = f3(d2.x, d2.y, 0.0f), dir_src = f3(s2.x, s2.y, 0.0f), p_src = f3(q2.x, q2.y, 0.0f);
            } else {
                dir = uniform_sphere(rng);
                dir_src = uniform_sphere(rng);
                greens3_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
            }
            int niters = P.av ? 2 : 1;
            uint64_t state = rng.state;
            int it = 0;
            while ((it < niters) && (wi < P.nwalks)) {
                if (it == 1) {
                    dir_src = -dir_src;
                    p_src = p * 2.0f - p_src;
                }
                float src = gnorm * src_f(bc, p_src);
                float src_grad;
                if (D == 2) {
                    F2 gr = greens2_gradient(f2(p.x, p.y), R, r_src, f2(p_src.x, p_src.y));
                    src_grad = dot(gr, f2(ev.x, ev.y)) / (pdf_src * gnorm);
                } else {
                    src_grad = dot(greens3_gradient(p, R, r_src, p_src), ev) / (pdf_src * gnorm);
                }
                src_sum += src;
                src_count += 1.0f;
                float _src = P.cv ? src - (src_sum / src_count) : src;
                result += _src * src_grad;
                if (it == 1) dir = -dir;
                F3 _p = p + dir * R;
                rng.state = state;
                float u = D == 2 ? walk_u2(sc, bc, P, f2(_p.x, _p.y), rng, st, tm) : walk_u3(sc, bc, P, _p, rng, st, tm);
                steps += st;
                u_sum += u;
                u_count += 1.0f;
                float _u = P.cv ? u - u_sum / u_count : u;
                float de = D == 2 ? dot(f2(ev.x, ev.y), f2(dir.x, dir.y)) : dot(ev, dir);
                result += (D == 2 ? 2.0f : 3.0f) / R * _u * de;
                wi += 1;
                it += 1;
            }
        }
    } else if (sc.dim == 2) {
        result = dude2(sc, bc, P, f2(pts[2 * i], pts[2 * i + 1]), f2(e[2 * i], e[2 * i + 1]), R, active, rng, steps);
    } else {
        result = dude3(sc, bc, P, f3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), f3(e[3 * i], e[3 * i + 1], e[3 * i + 2]),
                       R, active, rng, steps);
    }
    out[i] = result / (float)P.nwalks;
    if (step_total) atomicAdd(step_total, steps);
}

// StreamlineTracer.next_position (examples/example_2d/trace.py:19-68,
// examples/example_3d/trace.py:21-60) with a WoSt / WoSt3D solver: sample
// j = i * tn + k of point i runs on the PCG32 stream of lane j (TEA(seed, i)),
// and the gradient components are Solver.walk_dude (solver.py:148-166) along
// the axes, one after the other on that stream.  2D steps only from inside
// (sdf <= 0); when the stepped point is inside too it keeps half the step (the
// reference averages fy with fy_tilde = current_p - y_tilde, which is zero
// there).  3D takes the full step.  samples[j] = the sample's new position.
template <int D>
__global__ void __launch_bounds__(kBlock) k_trace(SceneDev sc, BCDev bc, ParamsDev P, const float* __restrict__ pts,
                                                  int64_t n, int tn, float step, float* samples,
                                                  unsigned long long* step_total) {
    using V = typename VecT<D>::T;
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n * tn) return;
    int64_t i = j / tn;
    Rng rng;
    seed_stream(rng, P, i, tn, j - i * tn);
    unsigned long long steps = 0;
    V p = load_pt<D>(pts, i);
    auto walk_dude = [&](V ev) -> float {  // nsubwalks WoSt.dude estimates, each / nwalks
        float r = 0.0f;
        if constexpr (D == 2) {
            float R = __builtin_fabsf(closest2(sc.sub[0], p).d);
            bool active = inside2(sc, p);
            for (int s = 0; s < P.nsubwalks; ++s)
                r += dude2(sc, bc, P, p, ev, R, active, rng, steps) / (float)P.nwalks;
        } else {
            float R = __builtin_fabsf(closest3(sc.sub[0], p).d);
            bool active = inside3(sc, p);
            for (int s = 0; s < P.nsubwalks; ++s)
                r += dude3(sc, bc, P, p, ev, R, active, rng, steps) / (float)P.nwalks;
        }
        return r / (float)P.nsubwalks;
    };
    V cur = p;
    if constexpr (D == 2) {
        if (!(sdf2(sc, p) > 0.0f)) {
            float gx = walk_dude(f2(1.0f, 0.0f));
            float gy = walk_dude(f2(0.0f, 1.0f));
            cur = p + f2(gx, gy) * step;
            F2 yt = cur;
            if (!(sdf2(sc, cur) > 0.0f)) {
                F2 fy = yt - p, fyt = cur - yt;
                cur = p + (fy + fyt) * 0.5f;
            }
        }
        samples[2 * j] = cur.x;
        samples[2 * j + 1] = cur.y;
    } else {
        float gx = walk_dude(f3(1.0f, 0.0f, 0.0f));
        float gy = walk_dude(f3(0.0f, 1.0f, 0.0f));
        float gz = walk_dude(f3(0.0f, 0.0f, 1.0f));
        cur = p + f3(gx, gy, gz) * step;
        samples[3 * j] = cur.x;
        samples[3 * j + 1] = cur.y;
        samples[3 * j + 2] = cur.z;
    }
    if (step_total) atomicAdd(step_total, steps);
}
// Mean of each point's samples, summed in sample order.
__global__ void __launch_bounds__(kBlock) k_trace_mean(const float* __restrict__ samples, int64_t n, int tn, int dim,
                                                       float* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < dim; ++d) {
        float s = 0.0f;
        for (int k = 0; k < tn; ++k) s += samples[((int64_t)i * tn + k) * dim + d];
        out[(int64_t)i * dim + d] = tn > 1 ? s / (float)tn : s;
    }
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

// Builds the child-in-parent records of one sub-scene (see dw_layout.h).
struct RefEnc {
    std::vector<int> table;  // flattened (offset, count) pairs
    int leaf(int off, int cnt) {
        if (cnt >= 0 && cnt < DW_LEAF_TABLE && off >= 0 && off < (1 << 27))
            return (int)(DW_LEAF_BIT | ((unsigned)cnt << 27) | (unsigned)off);
        int k = (int)(table.size() / 2);
        table.push_back(off);
        table.push_back(cnt);
        return (int)(DW_LEAF_BIT | ((unsigned)DW_LEAF_TABLE << 27) | (unsigned)k);
    }
};

int upload_records(dw_scene* s, const dwb::SubHost& H, SubDev& S) {
    const int D = H.dim;
    size_t nn = H.n_nodes();
    RefEnc enc;
    auto bvh_ref = [&](int i) { return H.node_nref[i] > 0 ? enc.leaf(H.node_off[i], H.node_nref[i]) : i; };
    auto snch_ref = [&](int i) { return H.node_nref[i] > 0 ? enc.leaf(H.node_sil_off[i], H.node_nsil[i]) : i; };
    auto box = [&](int i, int k) { return H.node_box[(size_t)i * 2 * D + k]; };
    auto cs = [&](int i, bool sine) {
        double ha = H.node_half_angle[i];
        return (float)(sine ? std::sin(ha) : std::cos(ha));
    };
    int rc;
    const void* p;
    S.max_depth = H.max_depth;
    S.root_ref_bvh = bvh_ref(0);
    S.root_ref_snch = snch_ref(0);
    for (int d = 0; d < D; ++d) S.root_mn[d] = box(0, d), S.root_mx[d] = box(0, D + d);
    if (D == 2) {
        std::vector<RBvh2> rb(nn);
        std::vector<RSnch2> rs(nn);
        for (size_t i = 0; i < nn; ++i) {
            std::memset(&rb[i], 0, sizeof(RBvh2));
            std::memset(&rs[i], 0, sizeof(RSnch2));
            if (H.node_nref[i] > 0) continue;
            int L = (int)i + 1, R = (int)i + H.node_off[i];
            RBvh2& b = rb[i];
            b.lmnx = box(L, 0), b.lmny = box(L, 1), b.lmxx = box(L, 2), b.lmxy = box(L, 3);
            b.rmnx = box(R, 0), b.rmny = box(R, 1), b.rmxx = box(R, 2), b.rmxy = box(R, 3);
            b.ref_l = bvh_ref(L), b.ref_r = bvh_ref(R);
            RSnch2& c = rs[i];
            c.lmnx = b.lmnx, c.lmny = b.lmny, c.lmxx = b.lmxx, c.lmxy = b.lmxy;
            c.rmnx = b.rmnx, c.rmny = b.rmny, c.rmxx = b.rmxx, c.rmxy = b.rmxy;
            c.lax = H.node_axis[2 * L], c.lay = H.node_axis[2 * L + 1];
            c.lha = H.node_half_angle[L], c.lrad = H.node_radius[L];
            c.rax = H.node_axis[2 * R], c.ray = H.node_axis[2 * R + 1];
            c.rha = H.node_half_angle[R], c.rrad = H.node_radius[R];
            c.lsin = cs(L, true), c.lcos = cs(L, false), c.rsin = cs(R, true), c.rcos = cs(R, false);
            c.ref_l = snch_ref(L), c.ref_r = snch_ref(R);
        }
        if ((rc = upload(s, rb, &p))) return rc;
        S.rec_bvh = p;
        if ((rc = upload(s, rs, &p))) return rc;
        S.rec_snch = p;
    } else {
        std::vector<RBvh3> rb(nn);
        std::vector<RSnch3> rs(nn);
        for (size_t i = 0; i < nn; ++i) {
            std::memset(&rb[i], 0, sizeof(RBvh3));
            std::memset(&rs[i], 0, sizeof(RSnch3));
            if (H.node_nref[i] > 0) continue;
            int L = (int)i + 1, R = (int)i + H.node_off[i];
            float v[12] = {box(L, 0), box(L, 1), box(L, 2), box(L, 3), box(L, 4), box(L, 5),
                           box(R, 0), box(R, 1), box(R, 2), box(R, 3), box(R, 4), box(R, 5)};
            RBvh3& b = rb[i];
            std::memcpy(&b, v, sizeof(v));
            b.ref_l = bvh_ref(L), b.ref_r = bvh_ref(R);
            RSnch3& sn = rs[i];
            std::memcpy(&sn.h, v, sizeof(v));
            sn.h.ref_l = snch_ref(L), sn.h.ref_r = snch_ref(R);
            sn.h.lha = H.node_half_angle[L], sn.h.rha = H.node_half_angle[R];
            RSnchTail3& c = sn.c;
            c.lax = H.node_axis[3 * L], c.lay = H.node_axis[3 * L + 1], c.laz = H.node_axis[3 * L + 2];
            c.rax = H.node_axis[3 * R], c.ray = H.node_axis[3 * R + 1], c.raz = H.node_axis[3 * R + 2];
            c.lrad = H.node_radius[L], c.rrad = H.node_radius[R];
            c.lsin = cs(L, true), c.lcos = cs(L, false), c.rsin = cs(R, true), c.rcos = cs(R, false);
            c.lrho = H.node_pb[3 * L], c.lelo = H.node_pb[3 * L + 1], c.lehi = H.node_pb[3 * L + 2];
            c.rrho = H.node_pb[3 * R], c.relo = H.node_pb[3 * R + 1], c.rehi = H.node_pb[3 * R + 2];
        }
        if ((rc = upload(s, rb, &p))) return rc;
        S.rec_bvh = p;
        if ((rc = upload(s, rs, &p))) return rc;
        S.rec_snch = p;
    }
    // two-level records: slot k's children are slots 2k + 2, 2k + 3
    auto wide = [&](auto& rw, auto put) {
        for (size_t i = 0; i < nn; ++i) {
            std::memset(&rw[i], 0, sizeof(rw[i]));
            if (H.node_nref[i] > 0) continue;
            int slot_node[6] = {(int)i + 1, (int)i + H.node_off[i], -1, -1, -1, -1};
            for (int k = 0; k < 2; ++k) {
                int c = slot_node[k];
                if (H.node_nref[c] > 0) continue;
                slot_node[2 * k + 2] = c + 1;
                slot_node[2 * k + 3] = c + H.node_off[c];
            }
            for (int k = 0; k < 6; ++k) {
                int c = slot_node[k];
                if (c < 0) continue;
                put(rw[i], k, c);
                rw[i].ref[k] = bvh_ref(c);
            }
        }
    };
    if (D == 2) {
        std::vector<RWide2> rw(nn);
        wide(rw, [&](RWide2& r, int k, int c) {
            for (int q = 0; q < 4; ++q) r.b[k][q] = box(c, q);
        });
        if ((rc = upload(s, rw, &p))) return rc;
    } else {
        std::vector<RWide3> rw(nn);
        wide(rw, [&](RWide3& r, int k, int c) {
            for (int q = 0; q < 6; ++q) r.b[k][q] = box(c, q);
        });
        if ((rc = upload(s, rw, &p))) return rc;
    }
    S.rec_wide = p;
    enc.table.push_back(0);  // never empty
    enc.table.push_back(0);
    if ((rc = upload(s, enc.table, &p))) return rc;
    S.leaf_table = (const int*)p;
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
    if ((rc = upload_records(s, H, S))) return rc;
    if ((rc = upload(s, H.prim_order, &p))) return rc;
    S.prim_index = (const int*)p;
    if ((rc = upload(s, H.refs, &p))) return rc;
    S.ref_topo = (const int*)p;
    if (D == 2) {
        S.n_psil = (int)(H.psil.size() / 6);
        if ((rc = upload(s, H.psil, &p))) return rc;
        S.psil = (const float*)p;
        std::vector<DNode2> pn(H.pnode_off.size());
        for (size_t i = 0; i < pn.size(); ++i) {
            const float* b = &H.pnode_box[4 * i];
            pn[i] = DNode2{b[0], b[1], b[2], b[3], H.pnode_off[i], H.pnode_nref[i], 0, 0};
        }
        if ((rc = upload(s, pn, &S.pnodes))) return rc;
        if ((rc = upload(s, H.porder, &p))) return rc;
        S.porder = (const int*)p;
    } else {
        S.n_psil3 = (int)(H.psil3.size() / 12);
        S.ps_sum = H.ps_sum, S.ps_norm = H.ps_norm, S.ps_shift = H.ps_shift;
        if ((rc = upload(s, H.psil3, &p))) return rc;
        S.psil3 = (const float*)p;
        if ((rc = upload(s, H.ps_pmf, &p))) return rc;
        S.ps_pmf = (const float*)p;
        if ((rc = upload(s, H.ps_cdf, &p))) return rc;
        S.ps_cdf = (const float*)p;
        if ((rc = upload(s, H.ps_guide, &p))) return rc;
        S.ps_guide = (const int*)p;
    }
    if ((rc = upload(s, H.cdf, &p))) return rc;
    S.cdf = (const float*)p;
    if ((rc = upload(s, H.pmf, &p))) return rc;
    S.pmf = (const float*)p;
    if ((rc = upload(s, H.cdf_guide, &p))) return rc;
    S.cdf_guide = (const int*)p;
    S.guide_shift = H.guide_shift;
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
    for (int k = 0; k < 4; ++k) bc.sc[k] = b->src_coeffs[k];
    return bc;
}
int check_bc(const dw_boundary* b) {
    if (!b) return fail(DW_ERR_ARG, "null boundary");
    if (b->bc_code < 0 || b->bc_code > 3) return fail(DW_ERR_ARG, "unknown boundary code");
    if (b->bc_code == DW_BC_PER_VERTEX && (!b->g_vertex || !b->h_vertex))
        return fail(DW_ERR_ARG, "per-vertex boundary needs g_vertex and h_vertex");
    if (b->src_code != DW_SRC_ZERO && b->src_code != DW_SRC_CONST && b->src_code != DW_SRC_LINEAR)
        return fail(DW_ERR_ARG, "unknown source code");
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
    P.point_base = p->point_base;
    P.block_size = p->block_size;
    P.blocks_per_cu = p->blocks_per_cu;
    P.schedule = p->schedule;
    P.gate = p->gate;
    P.node_width = p->node_width ? p->node_width : 2;
    P.walks_per_lane = p->walks_per_lane ? p->walks_per_lane : 1;
    return P;
}
int check_params(const dw_solver_params* p) {
    if (!p) return fail(DW_ERR_ARG, "null params");
    if (p->nwalks < 1) return fail(DW_ERR_ARG, "nwalks must be >= 1");
    if (p->nsteps < 0) return fail(DW_ERR_ARG, "nsteps must be >= 0");
    if (p->nsubwalks < 1) return fail(DW_ERR_ARG, "nsubwalks must be >= 1");
    if (p->point_base < 0) return fail(DW_ERR_ARG, "point_base must be >= 0");
    if (p->block_size != 0 && p->block_size != 64 && p->block_size != 128 && p->block_size != 256)
        return fail(DW_ERR_ARG, "block_size must be 0, 64, 128 or 256");
    if (p->blocks_per_cu < 0) return fail(DW_ERR_ARG, "blocks_per_cu must be >= 0");
    if (p->spatial_order < 0 || p->spatial_order > 2) return fail(DW_ERR_ARG, "spatial_order must be 0, 1 or 2");
    if (p->schedule < 0 || p->schedule > 2) return fail(DW_ERR_ARG, "schedule must be 0, 1 or 2");
    if (p->gate < 0 || p->gate > 64) return fail(DW_ERR_ARG, "gate must be in 0..64");
    if (p->node_width != 0 && p->node_width != 2 && p->node_width != 4)
        return fail(DW_ERR_ARG, "node_width must be 0, 2 or 4");
    if (p->walks_per_lane != 0 && p->walks_per_lane != 1 && p->walks_per_lane != 2 && p->walks_per_lane != 4 &&
        p->walks_per_lane != 8)
        return fail(DW_ERR_ARG, "walks_per_lane must be 0, 1, 2, 4 or 8");
    return DW_OK;
}

// Per (device, stream) state: the job-queue counters and a grow-only workspace
// for partial results and the spatial order.  Launches on different streams
// never share either; launches on one stream are ordered by the stream.
struct StreamCtx {
    unsigned long long* ctr = nullptr;  // 8 range counters, 128 B apart
    void* ws = nullptr;
    size_t ws_bytes = 0;
};
std::mutex g_ctx_mu;
std::map<std::pair<int, hipStream_t>, StreamCtx> g_ctx;
constexpr int kQueueCtrs = 8 * 16;  // u64 slots

int stream_ctx(hipStream_t st, StreamCtx** out) {
    int dev = 0;
    DW_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    StreamCtx& c = g_ctx[std::make_pair(dev, st)];
    if (!c.ctr) DW_HIP(hipMalloc((void**)&c.ctr, kQueueCtrs * sizeof(unsigned long long)));
    *out = &c;
    return DW_OK;
}

int workspace(StreamCtx& c, size_t bytes, hipStream_t st, void** out) {
    if (c.ws && c.ws_bytes < bytes) {
        DW_HIP(hipStreamSynchronize(st));  // earlier launches on this stream may still read it
        DW_HIP(hipFree(c.ws));
        c.ws = nullptr;
        c.ws_bytes = 0;
    }
    if (!c.ws) {
        bytes = bytes < 4096 ? 4096 : bytes;
        DW_HIP(hipMalloc(&c.ws, bytes));
        c.ws_bytes = bytes;
    }
    *out = c.ws;
    return DW_OK;
}

// Bytes of workspace the spatial order of n points needs (0: identity order).
size_t order_bytes(int64_t n, const dw_solver_params* p) {
    if (p->spatial_order == DW_ORDER_OFF || n >= (int64_t)1 << 31) return 0;
    if (p->spatial_order != DW_ORDER_ON && n < 2 * kCells) return 0;
    return (size_t)n * 8 + kCells * 4;
}

// Builds the job queue of a persistent launch: zeroes the range counters (a
// fill kernel, so the sequence can be captured in a graph) and, when
// order_bytes(n) > 0, sorts the points by Morton cell into `ows`.
int make_queue(const dw_scene* s, unsigned long long* ctr, const float* pts, int64_t n, int64_t per, void* ows,
               hipStream_t st, JobQ& Q) {
    Q.ctr = ctr;
    Q.order = nullptr;
    Q.n = n;
    Q.per = per;
    Q.gate = kGate;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kQueueCtrs), 0, st, ctr, kQueueCtrs, 0ull);
    DW_HIP(hipGetLastError());
    if (!ows) return DW_OK;
    const dwb::SubHost& F = s->host.sub[0];
    const int D = s->dev.dim;
    const float nb = D == 3 ? 16.0f : 64.0f;
    float lo[3] = {0, 0, 0}, sc[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        float ext = F.bb_max[d] - F.bb_min[d];
        lo[d] = F.bb_min[d];
        sc[d] = ext > 0.0f ? nb / ext : 0.0f;
    }
    int* order = (int*)ows;
    int* cell = order + n;
    int* hist = cell + n;
    hipLaunchKernelGGL(k_fill_i32, dim3(kCells / 256), dim3(256), 0, st, hist, kCells, 0);
    hipLaunchKernelGGL(k_cell_hist, dim3(blocks_for(n)), dim3(kBlock), 0, st, pts, n, D, lo[0], lo[1], lo[2], sc[0],
                       sc[1], sc[2], cell, hist);
    hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, hist);
    hipLaunchKernelGGL(k_cell_place, dim3(blocks_for(n)), dim3(kBlock), 0, st, cell, n, hist, order);
    DW_HIP(hipGetLastError());
    Q.order = order;
    return DW_OK;
}

// Launch shape of the persistent kernels (dw_solver_params.block_size /
// blocks_per_cu; 0 picks 256 threads and as many resident blocks per CU as
// the occupancy allows).  Each lane's traversal stack takes 128 B of LDS.
struct LaunchCfg {
    int block = kBlock;
    int per_cu = 0;
    bool interleave = false;  // lane scheduling (DW_SCHED_*)
    int gate = kGate;
};
// DW_SCHED_AUTO: interleaved where traversal lengths diverge most.  On the
// 81,920-triangle icosphere (3D, 262,144 points) it measured 1.48x the
// step-synchronous loop, on the 2D holes scene (65,536 points) 0.54x.
LaunchCfg launch_cfg(const ParamsDev& P, int dim) {
    LaunchCfg c;
    if (P.block_size) c.block = P.block_size;
    c.per_cu = P.blocks_per_cu;
    c.interleave = P.schedule == DW_SCHED_INTERLEAVED || (P.schedule == DW_SCHED_AUTO && dim == 3);
    if (P.gate) c.gate = P.gate;
    return c;
}
template <typename K>
int persistent_grid(K kernel, const LaunchCfg& cfg, unsigned& blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    DW_HIP(hipGetDevice(&dev));
    DW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    DW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, cfg.block,
                                                       (size_t)cfg.block * kStackLdsBytesPerThread));
    if (cfg.per_cu > 0 && cfg.per_cu < per_cu) per_cu = cfg.per_cu;
    if (per_cu < 1) per_cu = 1;
    blocks = (unsigned)(per_cu * cus);
    return DW_OK;
}

// The four instances of a persistent kernel: exact / fast cones x interleaved /
// step-synchronous scheduling.
template <typename K>
struct Pers {
    K exact_i, fast_i, exact_s, fast_s;
};
#define DW_PERS(k, ...) \
    Pers<decltype(&k<__VA_ARGS__, true, true>)> { \
        k<__VA_ARGS__, true, true>, k<__VA_ARGS__, false, true>, k<__VA_ARGS__, true, false>, k<__VA_ARGS__, false, false> \
    }
// Launches the instance the cone mode and the schedule select on its grid.
template <typename K, typename... A>
int launch_pers(bool exact, const LaunchCfg& cfg, const Pers<K>& ks, hipStream_t st, A... args) {
    unsigned blocks = 0;
    int rc;
    const size_t lds = (size_t)cfg.block * kStackLdsBytesPerThread;
    K k = cfg.interleave ? (exact ? ks.exact_i : ks.fast_i) : (exact ? ks.exact_s : ks.fast_s);
    if ((rc = persistent_grid(k, cfg, blocks))) return rc;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(cfg.block), lds, st, args...);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int zero_u64(unsigned long long* p, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, p, 1, 0ull);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

// DWOST_KERNEL=simple selects the one-job-per-thread kernels (a cross-check
// of the persistent ones).  They evaluate the literal cone test and write no
// guard flags, so they serve only exact-cone solves that ask for no flags;
// every other call keeps the persistent kernels.
bool use_simple_kernels(const dw_solver_params* p, const void* flags_out) {
    const char* e = getenv("DWOST_KERNEL");
    return e && strcmp(e, "simple") == 0 && (p->flags & DW_SOLVE_EXACT_CONES) && !flags_out;
}

// Argument checks shared by the fused entry points.
int check_fused(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                const dw_boundary* boundary, const float* u, const float* grad) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (n < 0 || (n > 0 && (!pts || !u || !grad))) return fail(DW_ERR_ARG, "bad arrays");
    if (params->flags & DW_SOLVE_WOS) return fail(DW_ERR_ARG, "the fused estimator runs WoSt walks only");
    return DW_OK;
}

// Device buffers of one fused solve: queue counters, first-ball radius and
// inside flag per point, per-pair partial results and step counts, and the
// spatial order (nullptr: identity order).
struct FusedBufs {
    unsigned long long* ctr;
    float* R;
    int* ins;
    float* Rd;
    float* pv;
    int32_t* ps;
    void* order;
};

// The launch sequence of a fused solve: point preparation, the walk kernel
// (persistent, or the one-pair-per-thread form), per-point reduction.
int fused_launch(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                 const dw_boundary* boundary, float* u, float* grad, uint8_t* pair_flags, uint64_t* total_steps,
                 bool simple, const FusedBufs& b, hipStream_t st) {
    int rc;
    const int D = s->dev.dim;
    const int npairs = (params->nwalks + 1) / 2;
    const int64_t lanes = n * (int64_t)npairs;
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(n)), dim3(kBlock), kStackLds, st, s->dev, pts, n, b.R, b.ins,
                       simple ? nullptr : b.Rd);
    DW_HIP(hipGetLastError());
    if (simple) {
        hipLaunchKernelGGL(k_fused, dim3(blocks_for(lanes)), dim3(kBlock), kStackLds, st, s->dev, to_bc(boundary),
                           to_params(params), pts, b.R, b.ins, n, npairs, b.pv, b.ps);
        DW_HIP(hipGetLastError());
    } else {
        JobQ Q;
        const ParamsDev P = to_params(params);
        const int per = jobs_per_point(npairs, P.walks_per_lane);
        if ((rc = make_queue(s, b.ctr, pts, n, per, b.order, st, Q))) return rc;
        const LaunchCfg cfg = launch_cfg(P, s->dev.dim);
        Q.gate = cfg.gate;
        const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
        FusedItem one;
        std::memset(&one, 0, sizeof(one));
        one.sc = s->dev_ptr;
        one.bc = to_bc(boundary);
        one.P = to_params(params);
        one.pts = pts;
        one.R = b.R;
        one.inside = b.ins;
        one.Rd = b.Rd;
        one.pair_vals = b.pv;
        one.pair_steps = b.ps;
        one.pair_flags = pair_flags;
        one.n = n;
        one.npairs = npairs;
        one.per = per;
        const FusedItem* none = nullptr;
        if (D == 2)
            rc = launch_pers(exact, cfg, DW_PERS(k_fused_pers, 2, false), st, one, none, 1, Q);
        else
            rc = launch_pers(exact, cfg, DW_PERS(k_fused_pers, 3, false), st, one, none, 1, Q);
        if (rc) return rc;
    }
    if (total_steps && (rc = zero_u64((unsigned long long*)total_steps, st))) return rc;
    // u = sum(usum) / (2 npairs); grad = sum(g) / (2 npairs)
    float scale = 1.0f / (float)(2 * npairs);
    hipLaunchKernelGGL(k_reduce_mean, dim3(blocks_for(n * 64)), dim3(kBlock), 0, st, b.pv, n, npairs, D + 1, scale, u,
                       grad, total_steps ? b.ps : nullptr, (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

}  // namespace

// Scratch of dw_solve_fused_noalloc, for callers that cannot allocate: one
// static device array per GPU, held until the process ends.  Host code never
// uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device with that device current (scratch_base).
static __device__ __align__(256) unsigned char scratch[DW_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kScratchBytes = (size_t)DW_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

int scratch_base(void** out) {
    int dev = 0;
    DW_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(DW_ERR_ARG, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return DW_OK;
}

}  // namespace

// A prepared multi-scene fused solve: device item descriptors, the partial
// result buffers of every item, the queue counters and (optionally) the
// captured graph of its launch sequence.
struct dw_batch {
    int device = 0;
    int dim = 2;
    bool exact = false;
    bool use_graph = false;
    LaunchCfg cfg;
    int n_items = 0;
    int64_t total_jobs = 0, total_pts = 0;
    FusedItem* dev_items = nullptr;
    void* ws = nullptr;
    unsigned long long* ctr = nullptr;
    unsigned long long* total_steps = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void batch_free(dw_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->dev_items) (void)hipFree(b->dev_items);
    if (b->ws) (void)hipFree(b->ws);
    if (b->ctr) (void)hipFree(b->ctr);
    delete b;
}

// The launch sequence of a batch: point preparation, queue reset, one
// persistent launch over the concatenated job space, per-point reduction.
// It touches only device memory owned by the batch or its items, so it can be
// captured into a graph and replayed.
int batch_launch(dw_batch* b, hipStream_t st) {
    int rc;
    hipLaunchKernelGGL(k_point_prep_batch, dim3(blocks_for(b->total_pts)), dim3(kBlock), kStackLds, st,
                       (const FusedItem*)b->dev_items, b->n_items, b->total_pts);
    DW_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kQueueCtrs), 0, st, b->ctr, kQueueCtrs, 0ull);
    DW_HIP(hipGetLastError());
    JobQ Q;
    Q.ctr = b->ctr;
    Q.order = nullptr;
    Q.n = b->total_jobs;
    Q.per = 1;
    Q.gate = b->cfg.gate;
    FusedItem none;
    std::memset(&none, 0, sizeof(none));
    const FusedItem* items = b->dev_items;
    if (b->dim == 2)
        rc = launch_pers(b->exact, b->cfg, DW_PERS(k_fused_pers, 2, true), st, none, items,
                         b->n_items, Q);
    else
        rc = launch_pers(b->exact, b->cfg, DW_PERS(k_fused_pers, 3, true), st, none, items,
                         b->n_items, Q);
    if (rc) return rc;
    if (b->total_steps && (rc = zero_u64(b->total_steps, st))) return rc;
    hipLaunchKernelGGL(k_reduce_batch, dim3(blocks_for(b->total_pts * 64)), dim3(kBlock), 0, st, items, b->n_items,
                       b->total_pts, b->dim, b->total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
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

int dw_scene_prim_map(const dw_scene* s, int32_t sub, int32_t* out) {
    if (!s) return fail(DW_ERR_ARG, "null scene");
    if (sub < 0 || sub > 2) return fail(DW_ERR_ARG, "sub must be 0 (full), 1 (Dirichlet) or 2 (Neumann)");
    const dwb::SubHost& H = s->host.sub[sub];
    if (!H.present) return DW_OK;
    if (!out) return fail(DW_ERR_ARG, "null output");
    std::copy(H.pglobal.begin(), H.pglobal.end(), out);
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
    hipLaunchKernelGGL(k_closest, dim3(blocks_for(n)), dim3(kBlock), kStackLds, (hipStream_t)stream, s->dev, sub, pts, n,
                       prim, dist, point, normal, tuv);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_sdf(const dw_scene* s, const float* pts, int64_t n, float* sdf, void* stream) {
    int rc = check_scene(s, 0);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!pts || !sdf))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_sdf, dim3(blocks_for(n)), dim3(kBlock), kStackLds, (hipStream_t)stream, s->dev, pts, n, sdf);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_sample_boundary(const dw_scene* s, int32_t sub, const float* u, int64_t n, float* point, float* normal,
                       float* tuv, float* pdf, int32_t* prim, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !u)) return fail(DW_ERR_ARG, "bad draws");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_sample_boundary, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, s->dev, sub, u,
                       n, point, normal, tuv, pdf, prim);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_query_baseline(const dw_scene* s, int32_t sub, int32_t kind, const float* pts, const float* dir,
                      const float* rmax, int64_t n, int32_t* prim, float* dist, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (kind != DW_BASE_CLOSEST && kind != DW_BASE_RAY && kind != DW_BASE_SILHOUETTE)
        return fail(DW_ERR_ARG, "kind must be DW_BASE_CLOSEST, DW_BASE_RAY or DW_BASE_SILHOUETTE");
    if (n < 0 || (n > 0 && (!pts || !prim || (kind == DW_BASE_RAY && !dir)))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_baseline, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, s->dev, sub, kind, pts,
                       dir, rmax, n, prim, dist);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_intersect(const dw_scene* s, int32_t sub, const float* origin, const float* dir, const float* rmax, int64_t n,
                 int32_t mode, int32_t* prim, float* dist, float* point, float* normal, float* tuv, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (mode != DW_RAY_REFERENCE && mode != DW_RAY_NEAREST) return fail(DW_ERR_ARG, "ray mode must be 0 or 1");
    if (n < 0 || (n > 0 && (!origin || !dir || !rmax))) return fail(DW_ERR_ARG, "bad ray arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_intersect, dim3(blocks_for(n)), dim3(kBlock), kStackLds, (hipStream_t)stream, s->dev, sub, origin,
                       dir, rmax, n, mode, prim, dist, point, normal, tuv);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_closest_silhouette(const dw_scene* s, int32_t sub, const float* pts, const float* rmax, int64_t n,
                          int32_t* sil, float* dist, float* point, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!pts || !rmax))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_closest_sil, dim3(blocks_for(n)), dim3(kBlock), kStackLds, (hipStream_t)stream, s->dev, sub, pts,
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
    hipLaunchKernelGGL(k_star_radius, dim3(blocks_for(n)), dim3(kBlock), kStackLds, (hipStream_t)stream, s->dev, sub, pts,
                       rmax, n, radius);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_star_radius_rho(const dw_scene* s, int32_t sub, const float* pts, const float* dirs, float rho_max,
                       int64_t n, float* radius, void* stream) {
    int rc = check_scene(s, sub);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!pts || !dirs || !radius))) return fail(DW_ERR_ARG, "bad arrays");
    if (!(rho_max > 0.0f)) return fail(DW_ERR_ARG, "rho_max must be > 0");
    if (n == 0) return DW_OK;
    hipLaunchKernelGGL(k_star_radius_rho, dim3(blocks_for(n)), dim3(kBlock), kStackLds, (hipStream_t)stream, s->dev,
                       sub, pts, dirs, rho_max, n, radius);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_math_check(int32_t* out, void* stream) {
    if (!out) return fail(DW_ERR_ARG, "bad arrays");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_math_check_init, dim3(1), dim3(64), 0, st, out);
    const dim3 grid(4096), block(kBlock);
    const int32_t a = fkey(-1.0f), b = fkey(1.0f), z = fkey(0.0f), inf = fkey(kInf);
    const int32_t c0 = fkey(kCosLo), c1 = fkey(kCosHi);
    hipLaunchKernelGGL(k_math_check<0>, grid, block, 0, st, a, (int64_t)b - a + 1, out);
    hipLaunchKernelGGL(k_math_check<1>, grid, block, 0, st, a, (int64_t)b - a + 1, out);
    hipLaunchKernelGGL(k_math_check<2>, grid, block, 0, st, z, (int64_t)inf - z + 1, out);
    hipLaunchKernelGGL(k_math_check<3>, grid, block, 0, st, c0, (int64_t)c1 - c0 + 1, out);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_solve_u(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
               const dw_boundary* boundary, float* u, float* walk_u, int32_t* walk_steps, int32_t* walk_term,
               uint8_t* walk_flags, uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (n < 0 || (n > 0 && (!pts || !u))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    if (!s->dev.sub[1].present && params->nsteps > 4096)
        return fail(DW_ERR_ARG, "scene without Dirichlet boundary: nsteps must be <= 4096");
    if (walk_flags && params->nsubwalks > 8) return fail(DW_ERR_ARG, "walk_flags support at most 8 subwalks");
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    const bool simple = use_simple_kernels(params, walk_flags);
    int64_t lanes = n * (int64_t)params->nwalks;
    size_t obytes = simple ? 0 : order_bytes(n, params);
    size_t wbytes = (walk_u ? 0 : (size_t)lanes * 4) + ((total_steps && !walk_steps) ? (size_t)lanes * 4 : 0);
    void* p;
    if ((rc = workspace(*c, wbytes + obytes, st, &p))) return rc;
    char* cur = (char*)p;
    float* wv = walk_u;
    if (!wv) wv = (float*)cur, cur += (size_t)lanes * 4;
    int32_t* ws = walk_steps;
    if (!ws && total_steps) ws = (int32_t*)cur, cur += (size_t)lanes * 4;
    if (simple) {
        hipLaunchKernelGGL(k_walk_u, dim3(blocks_for(lanes)), dim3(kBlock), kStackLds, st, s->dev, to_bc(boundary),
                           to_params(params), pts, n, wv, ws, walk_term);
        DW_HIP(hipGetLastError());
    } else {
        JobQ Q;
        if ((rc = make_queue(s, c->ctr, pts, n, params->nwalks, obytes ? cur : nullptr, st, Q))) return rc;
        const LaunchCfg cfg = launch_cfg(to_params(params), s->dev.dim);
        Q.gate = cfg.gate;
        const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
        if (s->dev.dim == 2)
            rc = launch_pers(exact, cfg, DW_PERS(k_walk_u_pers, 2), st, s->dev_ptr, to_bc(boundary),
                             to_params(params), pts, n, wv, ws, walk_term, walk_flags, Q);
        else
            rc = launch_pers(exact, cfg, DW_PERS(k_walk_u_pers, 3), st, s->dev_ptr, to_bc(boundary),
                             to_params(params), pts, n, wv, ws, walk_term, walk_flags, Q);
        if (rc) return rc;
    }
    if (total_steps && (rc = zero_u64((unsigned long long*)total_steps, st))) return rc;
    hipLaunchKernelGGL(k_reduce_mean, dim3(blocks_for(n * 64)), dim3(kBlock), 0, st, wv, n, params->nwalks, 1,
                       1.0f / (float)params->nwalks, u, (float*)nullptr, total_steps ? ws : nullptr,
                       (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_grad_solve_dude(const dw_scene* s, const float* pts, const float* e, int64_t n, const dw_solver_params* params,
                       const dw_grad_params* grad, const dw_boundary* boundary, float* out, float* walk_out,
                       int32_t* walk_steps, uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (s->dev.dim == 3 && s->dev.sub[1].present)
        return fail(DW_ERR_ARG, "dw_grad_solve_dude: WoStGrad3D needs an all-Neumann scene (the reference's 3D walk "
                                "has no normal-derivative estimate at Dirichlet points)");
    if (!grad) return fail(DW_ERR_ARG, "null grad params");
    if (boundary->bc_code == DW_BC_PER_VERTEX)
        return fail(DW_ERR_ARG, "dw_grad_solve_dude needs an analytic boundary field (tangential derivatives)");
    if (params->flags & DW_SOLVE_WOS) return fail(DW_ERR_ARG, "DW_SOLVE_WOS does not apply to WoStGrad");
    if (n < 0 || (n > 0 && (!pts || !e || !out))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    const int64_t lanes = n * (int64_t)params->nwalks;
    float* wv = walk_out;
    int32_t* ws = walk_steps;
    size_t need = (walk_out ? 0 : (size_t)lanes * 4) + ((total_steps && !walk_steps) ? (size_t)lanes * 4 : 0);
    if (need) {
        void* p;
        if ((rc = workspace(*c, need, st, &p))) return rc;
        char* cur = (char*)p;
        if (!wv) wv = (float*)cur, cur += (size_t)lanes * 4;
        if (!ws && total_steps) ws = (int32_t*)cur;
    }
    GradDev G{grad->rho_max, grad->ignore_boundary, grad->ignore_silhouette, grad->ignore_source};
    const char* kenv = getenv("DWOST_KERNEL");
    if (kenv && strcmp(kenv, "simple") == 0) {  // one walk per thread (a cross-check of the persistent form)
        if (s->dev.dim == 2)
            hipLaunchKernelGGL(k_grad_dude<2>, dim3(blocks_for(lanes)), dim3(kBlock), kStackLds, st, s->dev,
                               to_bc(boundary), to_params(params), G, pts, e, n, wv, ws);
        else
            hipLaunchKernelGGL(k_grad_dude<3>, dim3(blocks_for(lanes)), dim3(kBlock), kStackLds, st, s->dev,
                               to_bc(boundary), to_params(params), G, pts, e, n, wv, ws);
    } else {
        JobQ Q;
        if ((rc = make_queue(s, c->ctr, pts, n, params->nwalks, nullptr, st, Q))) return rc;
        const ParamsDev P = to_params(params);
        const LaunchCfg cfg = launch_cfg(P, s->dev.dim);
        unsigned blocks = 0;
        const size_t lds = (size_t)cfg.block * kStackLdsBytesPerThread;
        if (s->dev.dim == 2) {
            if ((rc = persistent_grid(k_grad_dude_pers<2>, cfg, blocks))) return rc;
            hipLaunchKernelGGL(k_grad_dude_pers<2>, dim3(blocks), dim3(cfg.block), lds, st, s->dev, to_bc(boundary),
                               P, G, pts, e, n, wv, ws, Q);
        } else {
            if ((rc = persistent_grid(k_grad_dude_pers<3>, cfg, blocks))) return rc;
            hipLaunchKernelGGL(k_grad_dude_pers<3>, dim3(blocks), dim3(cfg.block), lds, st, s->dev, to_bc(boundary),
                               P, G, pts, e, n, wv, ws, Q);
        }
    }
    DW_HIP(hipGetLastError());
    if (total_steps && (rc = zero_u64((unsigned long long*)total_steps, st))) return rc;
    hipLaunchKernelGGL(k_reduce_mean, dim3(blocks_for(n * 64)), dim3(kBlock), 0, st, wv, n, params->nwalks, 1,
                       1.0f / (float)params->nwalks, out, (float*)nullptr, total_steps ? ws : nullptr,
                       (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_grad_trace_step(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                       const dw_grad_params* grad, const dw_boundary* boundary, int32_t samples_per_point,
                       float step_size, float* out, uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (s->dev.dim != 2) return fail(DW_ERR_ARG, "dw_grad_trace_step: WoStGrad runs on 2D scenes");
    if (!grad) return fail(DW_ERR_ARG, "null grad params");
    if (boundary->bc_code == DW_BC_PER_VERTEX)
        return fail(DW_ERR_ARG, "dw_grad_trace_step needs an analytic boundary field (tangential derivatives)");
    if (n < 0 || (n > 0 && (!pts || !out))) return fail(DW_ERR_ARG, "bad arrays");
    if (samples_per_point < 1) return fail(DW_ERR_ARG, "samples_per_point must be >= 1");
    if (n == 0) return DW_OK;
    const int64_t ns = n * (int64_t)samples_per_point;
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    void* p;
    if ((rc = workspace(*c, (size_t)ns * 8, st, &p))) return rc;
    float* samples = (float*)p;
    if (total_steps && (rc = zero_u64((unsigned long long*)total_steps, st))) return rc;
    GradDev G{grad->rho_max, grad->ignore_boundary, grad->ignore_silhouette, grad->ignore_source};
    const char* kenv = getenv("DWOST_KERNEL");
    if (kenv && strcmp(kenv, "simple") == 0) {  // one sample per thread (a cross-check of the persistent form)
        hipLaunchKernelGGL(k_grad_trace, dim3(blocks_for(ns)), dim3(kBlock), kStackLds, st, s->dev, to_bc(boundary),
                           to_params(params), G, pts, n, (int)samples_per_point, step_size, samples,
                           (unsigned long long*)total_steps);
    } else {
        JobQ Q;
        if ((rc = make_queue(s, c->ctr, pts, n, samples_per_point, nullptr, st, Q))) return rc;
        const ParamsDev P = to_params(params);
        const LaunchCfg cfg = launch_cfg(P, 2);
        unsigned blocks = 0;
        if ((rc = persistent_grid(k_grad_trace_pers, cfg, blocks))) return rc;
        hipLaunchKernelGGL(k_grad_trace_pers, dim3(blocks), dim3(cfg.block), (size_t)cfg.block * kStackLdsBytesPerThread,
                           st, s->dev, to_bc(boundary), P, G, pts, n, (int)samples_per_point, step_size, samples,
                           (unsigned long long*)total_steps, Q);
    }
    DW_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_trace_mean, dim3(blocks_for(n)), dim3(kBlock), 0, st, samples, n, (int)samples_per_point, 2,
                       out);
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
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    // dudes run as persistent point jobs (DWOST_KERNEL=simple: one point per thread)
    const bool persistent = !use_simple_kernels(params, nullptr);
    const size_t obytes = persistent ? order_bytes(n, params) : 0;
    void* p;
    if ((rc = workspace(*c, (size_t)n * 8 + obytes, st, &p))) return rc;
    float* R = (float*)p;
    int* ins = (int*)((char*)p + (size_t)n * 4);
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(n)), dim3(kBlock), kStackLds, st, s->dev, pts, n, R, ins);
    DW_HIP(hipGetLastError());
    if (total_steps && (rc = zero_u64((unsigned long long*)total_steps, st))) return rc;
    if (persistent) {
        JobQ Q;
        if ((rc = make_queue(s, c->ctr, pts, n, 1, obytes ? (char*)p + (size_t)n * 8 : nullptr, st, Q))) return rc;
        const LaunchCfg cfg = launch_cfg(to_params(params), s->dev.dim);
        Q.gate = cfg.gate;
        const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
        if (s->dev.dim == 2)
            return launch_pers(exact, cfg, DW_PERS(k_dude_pers, 2), st, s->dev_ptr, to_bc(boundary),
                               to_params(params), pts, e, (const float*)R, (const int*)ins, out,
                               (unsigned long long*)total_steps, Q);
        return launch_pers(exact, cfg, DW_PERS(k_dude_pers, 3), st, s->dev_ptr, to_bc(boundary), to_params(params),
                           pts, e, (const float*)R, (const int*)ins, out, (unsigned long long*)total_steps, Q);
    }
    hipLaunchKernelGGL(k_dude, dim3(blocks_for(n)), dim3(kBlock), kStackLds, st, s->dev, to_bc(boundary), to_params(params),
                       pts, e, R, ins, n, out, (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_trace_step(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                  const dw_boundary* boundary, int32_t samples_per_point, float step_size, float* out,
                  uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (n < 0 || (n > 0 && (!pts || !out))) return fail(DW_ERR_ARG, "bad arrays");
    if (samples_per_point < 1) return fail(DW_ERR_ARG, "samples_per_point must be >= 1");
    if (params->flags & DW_SOLVE_WOS) return fail(DW_ERR_ARG, "the tracer runs WoSt walks only");
    if (n == 0) return DW_OK;
    const int D = s->dev.dim;
    const int64_t ns = n * (int64_t)samples_per_point;
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    const bool persistent = !use_simple_kernels(params, nullptr);
    const size_t obytes = persistent ? order_bytes(n, params) : 0;
    void* p;
    if ((rc = workspace(*c, (size_t)ns * D * 4 + (size_t)n * 8 + obytes, st, &p))) return rc;
    float* samples = (float*)p;
    float* R = (float*)((char*)p + (size_t)ns * D * 4);
    int* ins = (int*)(R + n);
    if (total_steps && (rc = zero_u64((unsigned long long*)total_steps, st))) return rc;
    if (persistent) {  // samples as persistent jobs, grouped by point in the queue
        hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(n)), dim3(kBlock), kStackLds, st, s->dev, pts, n, R, ins);
        DW_HIP(hipGetLastError());
        JobQ Q;
        if ((rc = make_queue(s, c->ctr, pts, n, samples_per_point, obytes ? (char*)(ins + n) : nullptr, st, Q)))
            return rc;
        const LaunchCfg cfg = launch_cfg(to_params(params), D);
        Q.gate = cfg.gate;
        const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
        if (D == 2)
            rc = launch_pers(exact, cfg, DW_PERS(k_trace_pers, 2), st, s->dev_ptr, to_bc(boundary), to_params(params),
                             pts, (const float*)R, (const int*)ins, (int)samples_per_point, step_size, samples,
                             (unsigned long long*)total_steps, Q);
        else
            rc = launch_pers(exact, cfg, DW_PERS(k_trace_pers, 3), st, s->dev_ptr, to_bc(boundary), to_params(params),
                             pts, (const float*)R, (const int*)ins, (int)samples_per_point, step_size, samples,
                             (unsigned long long*)total_steps, Q);
        if (rc) return rc;
    } else if (D == 2)
        hipLaunchKernelGGL(k_trace<2>, dim3(blocks_for(ns)), dim3(kBlock), kStackLds, st, s->dev, to_bc(boundary),
                           to_params(params), pts, n, (int)samples_per_point, step_size, samples,
                           (unsigned long long*)total_steps);
    else
        hipLaunchKernelGGL(k_trace<3>, dim3(blocks_for(ns)), dim3(kBlock), kStackLds, st, s->dev, to_bc(boundary),
                           to_params(params), pts, n, (int)samples_per_point, step_size, samples,
                           (unsigned long long*)total_steps);
    DW_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_trace_mean, dim3(blocks_for(n)), dim3(kBlock), 0, st, samples, n, (int)samples_per_point, D,
                       out);
    DW_HIP(hipGetLastError());
    return DW_OK;
}

int dw_solve_fused(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                   const dw_boundary* boundary, float* u, float* grad, float* pair_vals, int32_t* pair_steps,
                   uint8_t* pair_flags, uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_fused(s, pts, n, params, boundary, u, grad))) return rc;
    if (n == 0) return DW_OK;
    const int D = s->dev.dim;
    const int64_t lanes = n * (int64_t)((params->nwalks + 1) / 2);
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    const bool simple = use_simple_kernels(params, pair_flags);
    FusedBufs b;
    b.ctr = c->ctr;
    size_t obytes = simple ? 0 : order_bytes(n, params);
    size_t need = (size_t)n * 12 + (pair_vals ? 0 : (size_t)lanes * (D + 1) * 4) +
                  (pair_steps ? 0 : (size_t)lanes * 4) + obytes;
    void* p;
    if ((rc = workspace(*c, need, st, &p))) return rc;
    char* cur = (char*)p;
    b.R = (float*)cur;
    cur += (size_t)n * 4;
    b.ins = (int*)cur;
    cur += (size_t)n * 4;
    b.Rd = (float*)cur;
    cur += (size_t)n * 4;
    b.pv = pair_vals;
    if (!b.pv) b.pv = (float*)cur, cur += (size_t)lanes * (D + 1) * 4;
    b.ps = pair_steps;
    if (!b.ps) b.ps = (int32_t*)cur, cur += (size_t)lanes * 4;
    b.order = obytes ? cur : nullptr;
    return fused_launch(s, pts, n, params, boundary, u, grad, pair_flags, total_steps, simple, b, st);
}

int dw_solve_fused_noalloc(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                           const dw_boundary* boundary, float* u, float* grad, uint64_t* total_steps, void* stream) {
    int rc;
    if ((rc = check_fused(s, pts, n, params, boundary, u, grad))) return rc;
    const int D = s->dev.dim;
    const int64_t lanes = n * (int64_t)((params->nwalks + 1) / 2);
    // scratch layout (each part 256-byte aligned): queue counters, first-ball
    // radii, inside flags, per-pair partial results and step counts, order,
    // Dirichlet distances
    const size_t obytes = order_bytes(n, params);
    const size_t parts[7] = {(size_t)kQueueCtrs * 8, (size_t)n * 4, (size_t)n * 4, (size_t)lanes * (D + 1) * 4,
                             (size_t)lanes * 4, obytes, (size_t)n * 4};
    size_t off[7], need = 0;
    for (int k = 0; k < 7; ++k) off[k] = need, need += (parts[k] + 255) & ~(size_t)255;
    if (need > kScratchBytes)
        return fail(DW_ERR_ARG, "dw_solve_fused_noalloc: n * ceil(nwalks / 2) needs " + std::to_string(need) +
                                    " bytes of scratch, above its fixed " + std::to_string(kScratchBytes) +
                                    " (DW_NOALLOC_SCRATCH_BYTES); use dw_solve_fused or fewer points per call");
    if (n == 0) return DW_OK;
    char* base;
    if ((rc = scratch_base((void**)&base))) return rc;
    FusedBufs b;
    b.ctr = (unsigned long long*)(base + off[0]);
    b.R = (float*)(base + off[1]);
    b.ins = (int*)(base + off[2]);
    b.pv = (float*)(base + off[3]);
    b.ps = (int32_t*)(base + off[4]);
    b.order = obytes ? base + off[5] : nullptr;
    b.Rd = (float*)(base + off[6]);
    return fused_launch(s, pts, n, params, boundary, u, grad, nullptr, total_steps, false, b, (hipStream_t)stream);
}

int dw_adjoint_fused(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                     const dw_boundary* boundary, const float* dl_du, const float* dl_dgrad, const uint8_t* pair_flags,
                     float* dl_dg, float* dl_dh, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (boundary->bc_code != DW_BC_PER_VERTEX)
        return fail(DW_ERR_ARG, "the adjoint needs per-vertex boundary data (DW_BC_PER_VERTEX)");
    if (params->flags & DW_SOLVE_WOS) return fail(DW_ERR_ARG, "the fused estimator runs WoSt walks only");
    if (n < 0 || (n > 0 && (!pts || !dl_du || !dl_dgrad || !pair_flags || !dl_dg || !dl_dh)))
        return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    const int D = s->dev.dim;
    const int npairs = (params->nwalks + 1) / 2;
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    size_t obytes = order_bytes(n, params);
    void* p;
    if ((rc = workspace(*c, (size_t)n * 8 + obytes, st, &p))) return rc;
    float* R = (float*)p;
    int* ins = (int*)((char*)p + (size_t)n * 4);
    hipLaunchKernelGGL(k_point_prep, dim3(blocks_for(n)), dim3(kBlock), kStackLds, st, s->dev, pts, n, R, ins);
    DW_HIP(hipGetLastError());
    JobQ Q;
    if ((rc = make_queue(s, c->ctr, pts, n, npairs, obytes ? (char*)p + (size_t)n * 8 : nullptr, st, Q))) return rc;
    const LaunchCfg cfg = launch_cfg(to_params(params), s->dev.dim);
    Q.gate = cfg.gate;
    const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
    if (D == 2)
        return launch_pers(exact, cfg, DW_PERS(k_adjoint_pers, 2), st, s->dev_ptr, to_bc(boundary),
                           to_params(params), pts, R, ins, n, npairs, dl_du, dl_dgrad, pair_flags, dl_dg, dl_dh, Q);
    return launch_pers(exact, cfg, DW_PERS(k_adjoint_pers, 3), st, s->dev_ptr, to_bc(boundary),
                       to_params(params), pts, R, ins, n, npairs, dl_du, dl_dgrad, pair_flags, dl_dg, dl_dh, Q);
}

int dw_adjoint_u(const dw_scene* s, const float* pts, int64_t n, const dw_solver_params* params,
                 const dw_boundary* boundary, const float* dl_du, const uint8_t* walk_flags, float* dl_dg,
                 float* dl_dh, void* stream) {
    int rc;
    if ((rc = check_scene(s, 0))) return rc;
    if ((rc = check_params(params)) || (rc = check_bc(boundary))) return rc;
    if (boundary->bc_code != DW_BC_PER_VERTEX)
        return fail(DW_ERR_ARG, "the adjoint needs per-vertex boundary data (DW_BC_PER_VERTEX)");
    if (params->nsubwalks > 8) return fail(DW_ERR_ARG, "the adjoint supports at most 8 subwalks");
    if (n < 0 || (n > 0 && (!pts || !dl_du || !walk_flags || !dl_dg || !dl_dh))) return fail(DW_ERR_ARG, "bad arrays");
    if (n == 0) return DW_OK;
    hipStream_t st = (hipStream_t)stream;
    StreamCtx* c;
    if ((rc = stream_ctx(st, &c))) return rc;
    size_t obytes = order_bytes(n, params);
    void* p = nullptr;
    if (obytes && (rc = workspace(*c, obytes, st, &p))) return rc;
    JobQ Q;
    if ((rc = make_queue(s, c->ctr, pts, n, params->nwalks, p, st, Q))) return rc;
    const LaunchCfg cfg = launch_cfg(to_params(params), s->dev.dim);
    Q.gate = cfg.gate;
    const bool exact = (params->flags & DW_SOLVE_EXACT_CONES) != 0;
    if (s->dev.dim == 2)
        return launch_pers(exact, cfg, DW_PERS(k_adjoint_u_pers, 2), st, s->dev_ptr,
                           to_bc(boundary), to_params(params), pts, n, dl_du, walk_flags, dl_dg, dl_dh, Q);
    return launch_pers(exact, cfg, DW_PERS(k_adjoint_u_pers, 3), st, s->dev_ptr, to_bc(boundary),
                       to_params(params), pts, n, dl_du, walk_flags, dl_dg, dl_dh, Q);
}

int dw_batch_create(const dw_batch_item* items, int32_t n_items, uint64_t* total_steps, int32_t flags,
                    dw_batch** out) {
    if (!out) return fail(DW_ERR_ARG, "null output");
    *out = nullptr;
    if (n_items < 1 || !items) return fail(DW_ERR_ARG, "a batch needs at least one item");
    int rc;
    int dev = 0;
    DW_HIP(hipGetDevice(&dev));
    const int dim = items[0].scene ? items[0].scene->dev.dim : 0;
    const int exact = items[0].params.flags & DW_SOLVE_EXACT_CONES;
    for (int k = 0; k < n_items; ++k) {
        const dw_batch_item& it = items[k];
        if ((rc = check_scene(it.scene, 0))) return rc;
        if (it.scene->device != dev) return fail(DW_ERR_ARG, "batch item " + std::to_string(k) + ": scene lives on another device");
        if (it.scene->dev.dim != dim) return fail(DW_ERR_ARG, "batch item " + std::to_string(k) + ": all scenes need the same dimension");
        if ((it.params.flags & DW_SOLVE_EXACT_CONES) != exact)
            return fail(DW_ERR_ARG, "batch item " + std::to_string(k) + ": all items need the same cone mode");
        if ((rc = check_params(&it.params)) || (rc = check_bc(&it.boundary))) return rc;
        if (it.params.flags & DW_SOLVE_WOS)
            return fail(DW_ERR_ARG, "batch item " + std::to_string(k) + ": the fused estimator runs WoSt walks only");
        if (it.n < 0 || (it.n > 0 && (!it.pts || !it.u || !it.grad)))
            return fail(DW_ERR_ARG, "batch item " + std::to_string(k) + ": bad arrays");
    }
    // Larger items first in the job space: their walks start early, so the
    // launch does not wait on a big scene's tail.
    std::vector<int> order;
    for (int k = 0; k < n_items; ++k)
        if (items[k].n > 0) order.push_back(k);
    auto jobs_of = [&](int k) { return items[k].n * (int64_t)((items[k].params.nwalks + 1) / 2); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return jobs_of(a) > jobs_of(c); });
    dw_batch* b = new dw_batch();
    b->device = dev;
    b->dim = dim;
    b->exact = exact != 0;
    b->use_graph = (flags & DW_BATCH_GRAPH) != 0;
    b->cfg = launch_cfg(to_params(&items[0].params), b->dim);
    b->n_items = (int)order.size();
    b->total_steps = (unsigned long long*)total_steps;
    size_t bytes = 0;
    for (int k : order) {
        int64_t j = jobs_of(k);
        bytes += (size_t)items[k].n * 12 + (size_t)j * (dim + 1) * 4 + (size_t)j * 4;
        bytes = (bytes + 255) & ~(size_t)255;
    }
    std::vector<FusedItem> host(order.size());
    if (bytes && hipMalloc(&b->ws, bytes) != hipSuccess) {
        batch_free(b);
        return fail(DW_ERR_HIP, "batch workspace allocation failed (" + std::to_string(bytes) + " bytes)");
    }
    char* cur = (char*)b->ws;
    for (size_t q = 0; q < order.size(); ++q) {
        const dw_batch_item& it = items[order[q]];
        FusedItem& f = host[q];
        std::memset(&f, 0, sizeof(f));
        f.sc = it.scene->dev_ptr;
        f.bc = to_bc(&it.boundary);
        f.P = to_params(&it.params);
        f.pts = it.pts;
        f.u = it.u;
        f.grad = it.grad;
        f.n = it.n;
        f.npairs = (it.params.nwalks + 1) / 2;
        f.per = jobs_per_point(f.npairs, f.P.walks_per_lane);
        f.job0 = b->total_jobs;
        f.pt0 = b->total_pts;
        int64_t j = it.n * (int64_t)f.npairs;
        f.R = (const float*)cur, cur += (size_t)it.n * 4;
        f.inside = (const int*)cur, cur += (size_t)it.n * 4;
        f.Rd = (const float*)cur, cur += (size_t)it.n * 4;
        f.pair_vals = (float*)cur, cur += (size_t)j * (dim + 1) * 4;
        f.pair_steps = (int*)cur, cur += (size_t)j * 4;
        cur = (char*)b->ws + (((size_t)(cur - (char*)b->ws) + 255) & ~(size_t)255);
        b->total_jobs += it.n * (int64_t)f.per;
        b->total_pts += it.n;
    }
    if (hipMalloc((void**)&b->dev_items, sizeof(FusedItem) * (host.empty() ? 1 : host.size())) != hipSuccess ||
        hipMalloc((void**)&b->ctr, kQueueCtrs * sizeof(unsigned long long)) != hipSuccess) {
        batch_free(b);
        return fail(DW_ERR_HIP, "batch descriptor allocation failed");
    }
    if (!host.empty() &&
        hipMemcpy(b->dev_items, host.data(), sizeof(FusedItem) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        batch_free(b);
        return fail(DW_ERR_HIP, "batch descriptor upload failed");
    }
    *out = b;
    return DW_OK;
}

int dw_batch_solve(dw_batch* b, void* stream) {
    if (!b) return fail(DW_ERR_ARG, "null batch");
    int dev = 0;
    DW_HIP(hipGetDevice(&dev));
    if (dev != b->device) return fail(DW_ERR_ARG, "batch was created on device " + std::to_string(b->device) + ", current is " + std::to_string(dev));
    hipStream_t st = (hipStream_t)stream;
    if (b->total_pts == 0) {
        if (b->total_steps) return zero_u64(b->total_steps, st);
        return DW_OK;
    }
    if (!b->use_graph) return batch_launch(b, st);
    if (!b->exec) {
        if (!b->cap) DW_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        DW_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
        int rc = batch_launch(b, b->cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(DW_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
        b->graph = g;
        DW_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    DW_HIP(hipGraphLaunch(b->exec, st));
    return DW_OK;
}

void dw_batch_destroy(dw_batch* b) { batch_free(b); }

}  // extern "C"
