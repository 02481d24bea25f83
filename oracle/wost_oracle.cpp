// wost_oracle.cpp -- CPU restatement of the diff_wost Walk-on-Stars hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
// dwost_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  The product never links, loads or calls anything here.
//
// Every function restates a reference function literally (citations are to
// files of zihay/dwost).  Floating-point conventions, shared with the product
// so that the two agree bit for bit:
//   * fp32 everywhere, compiled with -ffp-contract=off (no implicit FMA).
//   * Dr.Jit's fused forms are written out as explicit fmaf():
//       dot(a,b)   = fma(a_n,b_n, ... fma(a_1,b_1, a_0*b_0))
//       lerp(a,b,t)= fma(b,t, fma(-a,t,a))
//       cross(u,v) = (fma(u.y,v.z,-(u.z*v.y)), ...)   [3D only]
//     Python-level expressions (u[0]*v[1]-u[1]*v[0], a+b*t, ...) stay plain.
//   * sqrt and division are IEEE correctly rounded.
//   * log/exp/sin/cos/acos/asin/atan2/cbrt are evaluated in fp64 and rounded
//     once to fp32 (correctly rounded except in ~2^-28 of inputs).
//   * Dr.Jit's own bit-level choices (its transcendental polynomials, its
//     fmadd usage inside dot, PCG32 lane offsets) are not observable in this
//     container; the choices above are this restatement's own, shared with the
//     HIP kernels.
//
// The C++ builder restates include/bvh.h + include/snch.h (Eigen reductions
// written out in Eigen's unrolled order: x0 op (x1 op x2) for 3-vectors).

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static const float kInf = std::numeric_limits<float>::infinity();
static const float kPi = (float)3.14159265358979323846;          // f32(pi)
static const float kPi2 = (float)(3.14159265358979323846 / 2.0);  // f32(pi/2)
static const float kTwoPi = (float)(2.0 * 3.14159265358979323846);
static const float kFourPi = (float)(4.0 * 3.14159265358979323846);
static const float kDrEps = 5.9604644775390625e-08f;  // dr.epsilon(Float32) = 2^-24

// ----------------------------------------------------------------------------
// fp32 helpers
// ----------------------------------------------------------------------------
struct V2 { float x, y; };
struct V3 { float x, y, z; };

static inline V2 v2(float x, float y) { return V2{x, y}; }
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V2 add(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V2 sub(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V2 mul(V2 a, float s) { return v2(a.x * s, a.y * s); }
static inline V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V2 divs(V2 a, float s) { return v2(a.x / s, a.y / s); }
static inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V2 neg(V2 a) { return v2(-a.x, -a.y); }
static inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot(V2 a, V2 b) { return std::fma(a.y, b.y, a.x * b.x); }
static inline float dot(V3 a, V3 b) { return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x)); }
static inline float norm(V2 a) { return std::sqrt(dot(a, a)); }
static inline float norm(V3 a) { return std::sqrt(dot(a, a)); }
static inline V2 normalize(V2 a) { return mul(a, 1.0f / std::sqrt(dot(a, a))); }
static inline V3 normalize(V3 a) { return mul(a, 1.0f / std::sqrt(dot(a, a))); }
static inline V3 cross(V3 u, V3 v) {
    return v3(std::fma(u.y, v.z, -(u.z * v.y)), std::fma(u.z, v.x, -(u.x * v.z)),
              std::fma(u.x, v.y, -(u.y * v.x)));
}
static inline float lerpf(float a, float b, float t) { return std::fma(b, t, std::fma(-a, t, a)); }
static inline V2 lerp(V2 a, V2 b, float t) { return v2(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t)); }
static inline V3 lerp(V3 a, V3 b, float t) {
    return v3(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t));
}
static inline float fminf_(float a, float b) { return std::fmin(a, b); }
static inline float fmaxf_(float a, float b) { return std::fmax(a, b); }
static inline float clampf(float x, float lo, float hi) { return fmaxf_(fminf_(x, hi), lo); }
static inline V2 vmin(V2 a, V2 b) { return v2(fminf_(a.x, b.x), fminf_(a.y, b.y)); }
static inline V2 vmax(V2 a, V2 b) { return v2(fmaxf_(a.x, b.x), fmaxf_(a.y, b.y)); }
static inline V3 vmin(V3 a, V3 b) { return v3(fminf_(a.x, b.x), fminf_(a.y, b.y), fminf_(a.z, b.z)); }
static inline V3 vmax(V3 a, V3 b) { return v3(fmaxf_(a.x, b.x), fmaxf_(a.y, b.y), fmaxf_(a.z, b.z)); }
static inline V2 vabs(V2 a) { return v2(std::fabs(a.x), std::fabs(a.y)); }
static inline V3 vabs(V3 a) { return v3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
// 2D cross as written in diff_wost/core/math.py:35-45 (plain ops)
static inline float cross2(V2 u, V2 v) { return u.x * v.y - u.y * v.x; }

// correctly rounded (via fp64) transcendentals
static inline float cr_log(float x) { return (float)std::log((double)x); }
static inline float cr_exp(float x) { return (float)std::exp((double)x); }
static inline float cr_sin(float x) { return (float)std::sin((double)x); }
static inline float cr_cos(float x) { return (float)std::cos((double)x); }
static inline float cr_acos(float x) { return (float)std::acos((double)x); }
static inline float cr_asin(float x) { return (float)std::asin((double)x); }
static inline float cr_atan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
static inline float cr_cbrt(float x) { return (float)std::cbrt((double)x); }

static inline bool bad(float v) { return std::isnan(v) || std::isinf(v); }

// ----------------------------------------------------------------------------
// TEA (diff_wost/core/math.py:87-107) and PCG32 (Dr.Jit PCG32; O'Neill's
// pcg32_srandom_r / pcg32_random_r; next_float32 = bits(u>>9 | 1.0f) - 1)
// ----------------------------------------------------------------------------
static inline void tea(uint32_t v0, uint32_t v1, int rounds, const uint32_t key[4], uint32_t* o0,
                       uint32_t* o1) {
    uint32_t s = 0;
    for (int r = 0; r < rounds; ++r) {
        s += 0x9E3779B9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s) ^ ((v0 >> 5) + key[3]);
    }
    *o0 = v0;
    *o1 = v1;
}
static const uint32_t kTeaKey[4] = {0xA341316Cu, 0xC8013EA4u, 0xAD90777Du, 0x7E95761Eu};

struct Pcg {
    uint64_t state, inc;
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    float next_float32() {
        uint32_t bits = (next_u32() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0;
        inc = (initseq << 1) | 1u;
        next_u32();
        state += initstate;
        next_u32();
    }
};
// PCG32(size, initstate=v0, initseq=v1) adds arange(size) to both inputs
// (Dr.Jit constructor semantics; solver.py:61-63).
static inline Pcg pcg_for_lane(uint32_t v0, uint32_t v1, uint64_t lane) {
    Pcg r;
    r.seed((uint64_t)v0 + lane, (uint64_t)v1 + lane);
    return r;
}
static inline Pcg pcg_for(uint32_t seed, uint32_t point, uint64_t lane) {
    uint32_t a, b;
    tea(seed, point, 4, kTeaKey, &a, &b);
    return pcg_for_lane(a, b, lane);
}

// ----------------------------------------------------------------------------
// Builder: include/bounding_volume.h, include/primitive.h, include/bvh.h,
// include/snch.h, include/silhouette_{vertex,edge}.h
// ----------------------------------------------------------------------------
static const float kBoxEps = (float)1e-9;  // fwd.h:14 `Epsilon` (Float=float)

static inline float emin(float a, float b) { return (b < a) ? b : a; }  // Eigen cwiseMin
static inline float emax(float a, float b) { return (a < b) ? b : a; }  // Eigen cwiseMax

template <int D>
struct Box {
    float mn[D], mx[D];
    Box() {
        for (int d = 0; d < D; ++d) {
            mn[d] = FLT_MAX;
            mx[d] = -FLT_MAX;
        }
    }
    static Box point(const float* p) {  // bounding_volume.h:112-116
        Box b;
        for (int d = 0; d < D; ++d) {
            b.mn[d] = p[d] - kBoxEps;
            b.mx[d] = p[d] + kBoxEps;
        }
        return b;
    }
    void expand(const float* p) {  // :119-123
        for (int d = 0; d < D; ++d) {
            mn[d] = emin(mn[d], p[d] - kBoxEps);
            mx[d] = emax(mx[d], p[d] + kBoxEps);
        }
    }
    void expand(const Box& b) {  // :126-129
        for (int d = 0; d < D; ++d) {
            mn[d] = emin(mn[d], b.mn[d]);
            mx[d] = emax(mx[d], b.mx[d]);
        }
    }
    float extent(int d) const { return mx[d] - mn[d]; }
    float surface_area() const {  // :191-194
        float e[D];
        for (int d = 0; d < D; ++d) e[d] = emax(extent(d), 1e-5f);
        if (D == 2) {
            float prod = e[0] * e[1];
            return 2.0f * (prod / e[0] + prod / e[1]);
        } else {
            float prod = e[0] * (e[1] * e[2]);
            return 2.0f * (prod / e[0] + (prod / e[1] + prod / e[2]));
        }
    }
    int max_dimension() const {  // :178-183 (first index of the max)
        int idx = 0;
        float best = mx[0] - mn[0];
        for (int d = 1; d < D; ++d) {
            float v = mx[d] - mn[d];
            if (v > best) {
                best = v;
                idx = d;
            }
        }
        return idx;
    }
    void centroid(float* c) const {
        for (int d = 0; d < D; ++d) c[d] = (mn[d] + mx[d]) * 0.5f;
    }
};

// Eigen-order reductions (size 2: x0+x1; size 3: x0+(x1+x2))
template <int D>
static inline float e_sqnorm(const float* v) {
    if (D == 2) return v[0] * v[0] + v[1] * v[1];
    return v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
}
template <int D>
static inline float e_dot(const float* a, const float* b) {
    if (D == 2) return a[0] * b[0] + a[1] * b[1];
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
template <int D>
static inline void e_normalized(float* v) {  // Eigen normalized(): v/sqrt(|v|^2) if |v|^2>0
    float z = e_sqnorm<D>(v);
    if (z > 0.0f) {
        float s = std::sqrt(z);
        for (int d = 0; d < D; ++d) v[d] = v[d] / s;
    }
}
static inline void e_cross(const float* u, const float* v, float* o) {  // Eigen cross (plain ops)
    o[0] = u[1] * v[2] - u[2] * v[1];
    o[1] = u[2] * v[0] - u[0] * v[2];
    o[2] = u[0] * v[1] - u[1] * v[0];
}

template <int D>
struct Prim {  // LineSegment (D=2) or Triangle (D=3), primitive.h:12-114
    float v[D][D];  // vertices a, b[, c]
    int pIndex;
    Box<D> bbox() const {
        Box<D> b = Box<D>::point(v[0]);
        for (int k = 1; k < D; ++k) b.expand(v[k]);
        return b;
    }
    void centroid(float* c) const {
        if (D == 2) {
            for (int d = 0; d < 2; ++d) c[d] = (v[0][d] + v[1][d]) * 0.5f;
        } else {
            for (int d = 0; d < 3; ++d) c[d] = ((v[0][d] + v[1][d]) + v[2][d]) / 3.0f;
        }
    }
};

template <int D>
struct Node {
    Box<D> box;
    int off;   // referenceOffset (leaf) / secondChildOffset (internal)
    int nref;  // > 0 for leaves
    // SNCH extension
    float axis[D];
    float half_angle;
    float radius;
    int sil_off, nsil;
};

template <int D>
struct Tree {
    std::vector<Node<D>> nodes;
    std::vector<Prim<D>> prims;  // sorted
    int max_depth = 0;
};

// Point primitive of the point BVH over silhouette vertices (primitive.h:116-152:
// the epsilon box of the point widened by radius 1e-4; centroid = the point)
template <int D>
struct PPoint {
    float v[1][D];
    int pIndex;
    Box<D> bbox() const {
        Box<D> b = Box<D>::point(v[0]);
        for (int d = 0; d < D; ++d) b.mn[d] = b.mn[d] - 1e-4f, b.mx[d] = b.mx[d] + 1e-4f;
        return b;
    }
    void centroid(float* c) const {
        for (int d = 0; d < D; ++d) c[d] = v[0][d];
    }
};

template <int D, typename P = Prim<D>>
struct Builder {  // bvh.h:359-652 (SurfaceArea, no leaf packing, leaf 4, 8 buckets)
    std::vector<P>& prims;
    std::vector<Box<D>> boxes;
    std::vector<std::array<float, D>> cents;
    std::vector<Node<D>> nodes;
    int max_depth = 0;
    static const int kLeaf = 4, kBuckets = 8;
    static const int UNTOUCHED = (int)0xffffffff, TOUCHED_TWICE = (int)0xfffffffd;
    explicit Builder(std::vector<P>& p) : prims(p) {}

    void build() {
        if (prims.empty()) return;
        size_t n = prims.size();
        boxes.resize(n);
        cents.resize(n);
        for (size_t i = 0; i < n; ++i) {
            boxes[i] = prims[i].bbox();
            prims[i].centroid(cents[i].data());
        }
        nodes.reserve(2 * n);
        recurse(-1, 0, (int)n, 0);
    }

    void recurse(int parent, int start, int end, int depth) {
        max_depth = std::max(max_depth, depth);
        Node<D> node;
        for (int d = 0; d < D; ++d) node.axis[d] = 0.0f;
        node.half_angle = 0.0f;
        node.radius = 0.0f;
        node.sil_off = node.nsil = 0;
        int current = (int)nodes.size();
        int nrefs = end - start;
        Box<D> bb, bc;
        for (int p = start; p < end; ++p) {
            bb.expand(boxes[p]);
            bc.expand(cents[p].data());
        }
        node.box = bb;
        if (nrefs <= kLeaf || depth == 64 - 2) {
            node.off = start;
            node.nref = nrefs;
        } else {
            node.off = UNTOUCHED;
            node.nref = 0;
        }
        nodes.push_back(node);
        if (parent >= 0) {
            nodes[parent].off--;
            if (nodes[parent].off == TOUCHED_TWICE) nodes[parent].off = (int)nodes.size() - 1 - parent;
        }
        if (node.nref > 0) return;
        int split_dim;
        float split_coord;
        object_split(bb, bc, start, end, split_dim, split_coord);
        int mid = start;
        for (int i = start; i < end; ++i) {  // performObjectSplit bvh.h:622-652
            if (cents[i][split_dim] < split_coord) {
                std::swap(prims[i], prims[mid]);
                std::swap(boxes[i], boxes[mid]);
                std::swap(cents[i], cents[mid]);
                mid++;
            }
        }
        if (mid == start || mid == end) mid = start + (end - start) / 2;
        recurse(current, start, mid, depth + 1);
        recurse(current, mid, end, depth + 1);
    }

    void object_split(const Box<D>& nb, const Box<D>& cb, int start, int end, int& split_dim,
                      float& split_coord) {  // computeObjectSplit bvh.h:537-620
        float split_cost = FLT_MAX;
        split_dim = -1;
        split_coord = 0.0f;
        for (int dim = 0; dim < D; ++dim) {
            float extent = nb.mx[dim] - nb.mn[dim];
            if (extent < 1e-6f) continue;
            float width = extent / (float)kBuckets;
            Box<D> bk[kBuckets];
            int cnt[kBuckets] = {0};
            for (int p = start; p < end; ++p) {
                int bi = (int)((cents[p][dim] - nb.mn[dim]) / width);
                bi = std::clamp(bi, 0, kBuckets - 1);
                bk[bi].expand(boxes[p]);
                cnt[bi] += 1;
            }
            Box<D> rbox[kBuckets];
            int rcnt[kBuckets] = {0};
            Box<D> acc;
            for (int b = kBuckets - 1; b > 0; --b) {
                acc.expand(bk[b]);
                rbox[b] = acc;
                rcnt[b] = cnt[b];
                if (b != kBuckets - 1) rcnt[b] += rcnt[b + 1];
            }
            Box<D> lbox;
            int nleft = 0;
            for (int b = 1; b < kBuckets; ++b) {
                lbox.expand(bk[b - 1]);
                nleft += cnt[b - 1];
                if (nleft > 0 && rcnt[b] > 0) {
                    float cost = (float)nleft * lbox.surface_area() + (float)rcnt[b] * rbox[b].surface_area();
                    if (cost < split_cost) {
                        split_cost = cost;
                        split_dim = dim;
                        split_coord = nb.mn[dim] + (float)b * width;
                    }
                }
            }
        }
        if (split_dim == -1) {
            split_dim = cb.max_dimension();
            split_coord = (cb.mn[split_dim] + cb.mx[split_dim]) * 0.5f;
        }
    }
};

// SNCH silhouettes (silhouette_vertex.h / silhouette_edge.h)
struct Sil2 {
    float a[2], b[2], c[2];
    int idx[3];  // prev, center, next
    int pIndex;
    bool has_face(int f) const { return f == 0 ? idx[2] != -1 : idx[0] != -1; }
    void face_normal(int f, bool nrm, float* n) const {  // silhouette_vertex.h:58-66
        const float* pa = f == 0 ? b : a;
        const float* pb = f == 0 ? c : b;
        float s0 = pb[0] - pa[0], s1 = pb[1] - pa[1];
        n[0] = s1;
        n[1] = -s0;
        if (nrm) e_normalized<2>(n);
    }
    void normal(float* n) const {
        n[0] = n[1] = 0.0f;
        float t[2];
        if (has_face(0)) {
            face_normal(0, false, t);
            n[0] += t[0];
            n[1] += t[1];
        }
        if (has_face(1)) {
            face_normal(1, false, t);
            n[0] += t[0];
            n[1] += t[1];
        }
        e_normalized<2>(n);
    }
    void centroid(float* o) const { o[0] = b[0], o[1] = b[1]; }
};
struct Sil3 {
    float a[3], b[3], c[3], d[3];
    int idx[4];  // c, a, b, d
    int pIndex;
    bool has_face(int f) const { return f == 0 ? idx[0] != -1 : idx[3] != -1; }
    void face_normal(int f, bool nrm, float* n) const {  // silhouette_edge.h:70-78
        float u[3], v[3];
        if (f == 0) {
            for (int k = 0; k < 3; ++k) u[k] = b[k] - a[k], v[k] = c[k] - a[k];
        } else {
            for (int k = 0; k < 3; ++k) u[k] = d[k] - a[k], v[k] = b[k] - a[k];
        }
        e_cross(u, v, n);
        if (nrm) e_normalized<3>(n);
    }
    void normal(float* n) const {
        n[0] = n[1] = n[2] = 0.0f;
        float t[3];
        if (has_face(0)) {
            face_normal(0, false, t);
            for (int k = 0; k < 3; ++k) n[k] += t[k];
        }
        if (has_face(1)) {
            face_normal(1, false, t);
            for (int k = 0; k < 3; ++k) n[k] += t[k];
        }
        e_normalized<3>(n);
    }
    void centroid(float* o) const {
        for (int k = 0; k < 3; ++k) o[k] = (a[k] + b[k]) * 0.5f;
    }
};

template <int D, typename S>
struct SnchData {
    std::vector<S> sils;       // topology silhouettes (one per vertex / edge)
    std::vector<int> refs;     // silhouette refs in node order (indices into sils)
    std::vector<S> ref_sils;   // copies in ref order (what traversal reads)
};

// buildTopology 2D (snch.h:407-443)
static void topology2(const std::vector<float>& V, const std::vector<int>& I, std::vector<Sil2>& sils) {
    size_t nv = V.size() / 2, ne = I.size() / 2;
    std::vector<std::vector<int>> adj(nv);
    for (size_t i = 0; i < ne; ++i) {
        int u = I[2 * i], w = I[2 * i + 1];
        adj[u].push_back(w);
        adj[w].push_back(u);
    }
    sils.resize(nv);
    for (size_t i = 0; i < nv; ++i) {
        int prev = adj[i].size() >= 1 ? adj[i][0] : -1;
        int next = adj[i].size() >= 2 ? adj[i][1] : -1;
        Sil2& s = sils[i];
        s.a[0] = prev != -1 ? V[2 * prev] : 0.0f;
        s.a[1] = prev != -1 ? V[2 * prev + 1] : 0.0f;
        s.b[0] = V[2 * i];
        s.b[1] = V[2 * i + 1];
        s.c[0] = next != -1 ? V[2 * next] : 0.0f;
        s.c[1] = next != -1 ? V[2 * next + 1] : 0.0f;
        s.idx[0] = prev;
        s.idx[1] = (int)i;
        s.idx[2] = next;
        s.pIndex = (int)i;
    }
}

// buildTopology 3D (snch.h:446-489); conn[f][k] = edge id
static void topology3(const std::vector<float>& V, const std::vector<int>& I, std::vector<Sil3>& sils,
                      std::vector<int>& conn) {
    size_t nf = I.size() / 3;
    std::map<std::pair<int, int>, int> emap;
    conn.assign(nf * 3, -1);
    for (size_t i = 0; i < nf; ++i) {
        for (int k = 0; k < 3; ++k) {
            int u = I[3 * i + k], v = I[3 * i + (k + 1) % 3], w = I[3 * i + (k + 2) % 3];
            std::pair<int, int> key = u < v ? std::make_pair(u, v) : std::make_pair(v, u);
            auto it = emap.find(key);
            if (it == emap.end()) {
                int e = (int)sils.size();
                emap[key] = e;
                Sil3 s;
                std::memset(&s, 0, sizeof(s));
                for (int q = 0; q < 3; ++q) {
                    s.a[q] = V[3 * u + q];
                    s.b[q] = V[3 * v + q];
                    s.c[q] = V[3 * w + q];
                    s.d[q] = 0.0f;
                }
                s.idx[0] = w;
                s.idx[1] = u;
                s.idx[2] = v;
                s.idx[3] = -1;
                s.pIndex = e;
                sils.push_back(s);
                conn[3 * i + k] = e;
            } else {
                int e = it->second;
                Sil3& s = sils[e];
                for (int q = 0; q < 3; ++q) s.d[q] = V[3 * w + q];
                s.idx[3] = w;
                conn[3 * i + k] = e;
            }
        }
    }
}

// assignGeometricDataToNodes (snch.h:174-306) + computeBoundingConesRecursive (:308-385)
template <int D, typename S>
static void snch_assign(Tree<D>& T, const std::vector<S>& sils, const std::vector<int>& conn,
                        SnchData<D, S>& out) {
    std::vector<float> refN, refFN;  // D per ref, 2*D per ref
    for (size_t i = 0; i < T.nodes.size(); ++i) {
        Node<D>& node = T.nodes[i];
        std::unordered_map<int, bool> seen;
        int start = (int)out.refs.size();
        for (int j = 0; j < node.nref; ++j) {
            const Prim<D>& prim = T.prims[node.off + j];
            int pi = prim.pIndex;
            for (int k = 0; k < D; ++k) {
                int si = conn[D * pi + k];
                if (seen.find(si) != seen.end()) continue;
                seen[si] = true;
                const S& s = sils[si];
                float n0[D], n1[D], n[D];
                for (int q = 0; q < D; ++q) n0[q] = n1[q] = 0.0f;
                if (s.has_face(0)) s.face_normal(0, true, n0);
                if (s.has_face(1)) s.face_normal(1, true, n1);
                s.normal(n);
                out.refs.push_back(si);
                for (int q = 0; q < D; ++q) refN.push_back(n[q]);
                for (int q = 0; q < D; ++q) refFN.push_back(n0[q]);
                for (int q = 0; q < D; ++q) refFN.push_back(n1[q]);
            }
        }
        node.sil_off = start;
        node.nsil = (int)out.refs.size() - start;
    }
    for (auto& nd : T.nodes) {  // default cone: axis 0, halfAngle pi, radius 0
        for (int q = 0; q < D; ++q) nd.axis[q] = 0.0f;
        nd.half_angle = kPi;
        nd.radius = 0.0f;
    }
    struct Rec {
        static void go(Tree<D>& T, const std::vector<S>& sils, const std::vector<int>& refs,
                       const std::vector<float>& refN, const std::vector<float>& refFN, int start, int end) {
            Node<D>& node = T.nodes[start];
            float cen[D];
            node.box.centroid(cen);
            float axis[D];
            for (int q = 0; q < D; ++q) axis[q] = 0.0f;
            float radius = 0.0f;
            bool any = false, two = true;
            for (int i = start; i < end; ++i) {
                const Node<D>& ch = T.nodes[i];
                for (int j = 0; j < ch.nsil; ++j) {
                    int r = ch.sil_off + j;
                    const S& s = sils[refs[r]];
                    for (int q = 0; q < D; ++q) axis[q] += refN[D * r + q];
                    float sc[D], dv[D];
                    s.centroid(sc);
                    for (int q = 0; q < D; ++q) dv[q] = sc[q] - cen[q];
                    float len = std::sqrt(e_sqnorm<D>(dv));
                    radius = std::max(radius, len);
                    two = two && s.has_face(0) && s.has_face(1);
                    any = true;
                }
            }
            if (!any) {
                node.half_angle = -kPi;
            } else if (!two) {
                node.half_angle = kPi;
            } else {
                float an = std::sqrt(e_sqnorm<D>(axis));
                if (an > kBoxEps) {
                    for (int q = 0; q < D; ++q) axis[q] /= an;
                    float ha = 0.0f;
                    for (int i = start; i < end; ++i) {
                        const Node<D>& ch = T.nodes[i];
                        for (int j = 0; j < ch.nsil; ++j) {
                            int r = ch.sil_off + j;
                            for (int k = 0; k < 2; ++k) {
                                const float* n = &refFN[(2 * r + k) * D];
                                float ang = cr_acos(std::max(-1.0f, std::min(1.0f, e_dot<D>(axis, n))));
                                ha = std::max(ha, ang);
                            }
                        }
                    }
                    for (int q = 0; q < D; ++q) node.axis[q] = axis[q];
                    node.half_angle = ha;
                    node.radius = radius;
                }
            }
            if (node.nref == 0) {
                go(T, sils, refs, refN, refFN, start + 1, start + node.off);
                go(T, sils, refs, refN, refFN, start + node.off, end);
            }
        }
    };
    if (!T.nodes.empty()) Rec::go(T, sils, out.refs, refN, refFN, 0, (int)T.nodes.size());
    out.ref_sils.resize(out.refs.size());
    for (size_t r = 0; r < out.refs.size(); ++r) out.ref_sils[r] = sils[out.refs[r]];
}

// ----------------------------------------------------------------------------
// Scenes (diff_wost/shapes/polyline.py, mesh.py)
// ----------------------------------------------------------------------------
enum { BC_EXPCOS = 0, BC_QUAD = 1, BC_LINEAR = 2, BC_PER_VERTEX = 3 };
enum { SRC_ZERO = 0, SRC_CONST = 1, SRC_LINEAR = 2 };

template <int D>
struct Sub {  // one Polyline / Mesh
    bool present = false;
    std::vector<float> V;        // local vertices
    std::vector<int> I;          // local indices
    std::vector<int> vglobal;    // local vertex -> global vertex
    std::vector<int> pglobal;    // local prim -> global prim
    float bb_min[D], bb_max[D];  // python bbox (no epsilon)
    bool py_sils = false;        // configure_silhouette left silhouettes != None
    std::vector<float> pmf, cdf; // edge lengths / doubled face areas
    float pmf_sum = 0.0f, pmf_norm = 0.0f;
    std::vector<float> vnormal;  // 2D vertex normals (for sdf)
    Tree<D> tree;
    std::vector<Sil2> sil2, ref2;
    std::vector<Sil3> sil3, ref3;
    std::vector<int> refs;
    // 2D: the polyline's own silhouette vertices (configure_silhouette,
    // polyline.py:100-165: a, b, c per vertex with both neighbours, near-collinear
    // ones dropped) and the point BVH over their b (PointBVH, bvh.py:375-...)
    std::vector<float> psil;
    std::vector<Node<D>> pnodes;
    std::vector<PPoint<D>> pprims;
    // 3D: the mesh's own silhouette edges (Mesh.configure_silhouette,
    // mesh.py:140-243: a, b, c, d per edge shared by exactly two faces, flat
    // ones dropped) and their length-weighted pmf
    std::vector<float> psil3, ps_pmf, ps_cdf;
    float ps_sum = 0.0f, ps_norm = 0.0f;
};

struct Scene {
    int dim = 2;
    Sub<2> s2[3];  // 0 full, 1 dirichlet, 2 neumann
    Sub<3> s3[3];
    std::vector<float> V;  // global vertices
    std::vector<int> I;
    std::vector<int> types;
};

template <int D>
static void build_sub(Sub<D>& S) {
    S.present = true;
    size_t nv = S.V.size() / D, np = S.I.size() / D;
    for (int d = 0; d < D; ++d) {
        S.bb_min[d] = kInf;
        S.bb_max[d] = -kInf;
    }
    for (size_t v = 0; v < nv; ++v)
        for (int d = 0; d < D; ++d) {
            S.bb_min[d] = std::min(S.bb_min[d], S.V[D * v + d]);
            S.bb_max[d] = std::max(S.bb_max[d], S.V[D * v + d]);
        }
    // pmf: polyline.py:166-171 (edge length), mesh.py:279-283 (|cross|)
    S.pmf.resize(np);
    S.cdf.resize(np);
    float acc = 0.0f;
    for (size_t p = 0; p < np; ++p) {
        float w;
        if (D == 2) {
            V2 a = v2(S.V[2 * S.I[2 * p]], S.V[2 * S.I[2 * p] + 1]);
            V2 b = v2(S.V[2 * S.I[2 * p + 1]], S.V[2 * S.I[2 * p + 1] + 1]);
            w = norm(sub(b, a));
        } else {
            const float* A = &S.V[3 * S.I[3 * p]];
            const float* B = &S.V[3 * S.I[3 * p + 1]];
            const float* C = &S.V[3 * S.I[3 * p + 2]];
            V3 a = v3(A[0], A[1], A[2]), b = v3(B[0], B[1], B[2]), c = v3(C[0], C[1], C[2]);
            w = norm(cross(sub(b, a), sub(c, a)));
        }
        S.pmf[p] = w;
        acc += w;  // dr.cumsum, sequential fp32
        S.cdf[p] = acc;
    }
    S.pmf_sum = np ? S.cdf[np - 1] : 0.0f;
    S.pmf_norm = 1.0f / S.pmf_sum;
    if (D == 2) {
        // configure_normals (polyline.py:173-186)
        S.vnormal.assign(nv * 2, 0.0f);
        std::vector<V2> en(np);
        for (size_t p = 0; p < np; ++p) {
            V2 a = v2(S.V[2 * S.I[2 * p]], S.V[2 * S.I[2 * p] + 1]);
            V2 b = v2(S.V[2 * S.I[2 * p + 1]], S.V[2 * S.I[2 * p + 1] + 1]);
            V2 d = sub(b, a);
            en[p] = normalize(v2(d.y, -d.x));
        }
        for (int pass = 0; pass < 2; ++pass)
            for (size_t p = 0; p < np; ++p) {
                int v = S.I[2 * p + pass];
                S.vnormal[2 * v] += en[p].x;
                S.vnormal[2 * v + 1] += en[p].y;
            }
        for (size_t v = 0; v < nv; ++v) {
            V2 n = normalize(v2(S.vnormal[2 * v], S.vnormal[2 * v + 1]));
            S.vnormal[2 * v] = n.x;
            S.vnormal[2 * v + 1] = n.y;
        }
        // configure_silhouette (polyline.py:102-164): None unless a vertex has prev and next
        std::vector<int> prv(nv, -1), nxt(nv, -1);
        for (size_t p = 0; p < np; ++p) {
            prv[S.I[2 * p + 1]] = S.I[2 * p];
            nxt[S.I[2 * p]] = S.I[2 * p + 1];
        }
        S.py_sils = false;
        for (size_t v = 0; v < nv; ++v)
            if (prv[v] != -1 && nxt[v] != -1) S.py_sils = true;
    } else {
        // mesh.py:129-199: None unless an undirected edge is shared by exactly two faces
        std::map<std::pair<int, int>, int> cnt;
        for (size_t p = 0; p < np; ++p)
            for (int k = 0; k < 3; ++k) {
                int u = S.I[3 * p + k], v = S.I[3 * p + (k + 1) % 3];
                cnt[std::make_pair(std::min(u, v), std::max(u, v))]++;
            }
        S.py_sils = false;
        for (auto& kv : cnt)
            if (kv.second == 2) S.py_sils = true;
    }
    // BVH/SNCH (bind_bvh.cpp / bind_snch.cpp -> bvh.h / snch.h)
    S.tree.prims.resize(np);
    for (size_t p = 0; p < np; ++p) {
        Prim<D>& pr = S.tree.prims[p];
        for (int k = 0; k < D; ++k)
            for (int d = 0; d < D; ++d) pr.v[k][d] = S.V[D * S.I[D * p + k] + d];
        pr.pIndex = (int)p;
    }
    Builder<D> B(S.tree.prims);
    B.build();
    S.tree.nodes = B.nodes;
    S.tree.max_depth = B.max_depth;
}

template <int D>
static void build_snch(Sub<D>& S);
template <>
void build_snch<2>(Sub<2>& S) {
    topology2(S.V, S.I, S.sil2);
    SnchData<2, Sil2> sd;
    snch_assign<2, Sil2>(S.tree, S.sil2, S.I, sd);  // connectivity = indices
    S.refs = sd.refs;
    S.ref2 = sd.ref_sils;
}
template <>
void build_snch<3>(Sub<3>& S) {
    std::vector<int> conn;
    topology3(S.V, S.I, S.sil3, conn);
    SnchData<3, Sil3> sd;
    snch_assign<3, Sil3>(S.tree, S.sil3, conn, sd);
    S.refs = sd.refs;
    S.ref3 = sd.ref_sils;
}

// _build_subscene_from_mask (polyline.py:672-692, mesh.py:617-635)
// Polyline.configure_silhouette (polyline.py:100-165): prev[b] = a and
// next[a] = b over the edges (a, b) in order (later edges overwrite); every
// vertex with both, in vertex order, is a silhouette (a = prev, b, c = next)
// unless |normalize(b - a) . normalize(c - b)| >= 1 - 1e-5 (near-collinear);
// then PointBVH(silhouettes.b) (bvh.py:375-395; C++ BVH<2, Point2>).
static void build_psil2(Sub<2>& S) {
    int nv = (int)(S.V.size() / 2), ne = (int)(S.I.size() / 2);
    std::vector<int> prv(nv, -1), nxt(nv, -1);
    for (int e = 0; e < ne; ++e) prv[S.I[2 * e + 1]] = S.I[2 * e], nxt[S.I[2 * e]] = S.I[2 * e + 1];
    S.psil.clear();
    for (int v = 0; v < nv; ++v) {
        if (prv[v] == -1 || nxt[v] == -1) continue;
        V2 a = v2(S.V[2 * prv[v]], S.V[2 * prv[v] + 1]), b = v2(S.V[2 * v], S.V[2 * v + 1]);
        V2 c = v2(S.V[2 * nxt[v]], S.V[2 * nxt[v] + 1]);
        V2 t1 = normalize(sub(b, a)), t2 = normalize(sub(c, b));
        if (!(std::fabs(dot(t1, t2)) < (1.0f - 1e-5f))) continue;
        float q[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
        S.psil.insert(S.psil.end(), q, q + 6);
    }
    S.pprims.clear();
    for (size_t k = 0; k < S.psil.size() / 6; ++k) {
        PPoint<2> pt;
        pt.v[0][0] = S.psil[6 * k + 2], pt.v[0][1] = S.psil[6 * k + 3];
        pt.pIndex = (int)k;
        S.pprims.push_back(pt);
    }
    Builder<2, PPoint<2>> b(S.pprims);
    b.build();
    S.pnodes = b.nodes;
}

// Mesh.configure_silhouette (mesh.py:140-243): per face (a, b, c) the half
// edges (min, max) of (a, b), (b, c), (c, a) with the opposite vertex, all
// first edges, then all second, then all third; a stable sort by the key
// u * V + v; edges met exactly twice give (a = V[u], b = V[v], c = opposite of
// the first, d = opposite of the second), kept unless
// |n0 . n1| >= 1 - 1e-5 (n0 = N((b - a) x (c - a)), n1 = N((d - a) x (b - a)));
// pmf = |a - b| with a sequential fp32 CDF (distr.py:44-50).
static void build_psil3(Sub<3>& S) {
    int F = (int)(S.I.size() / 3), nv = (int)(S.V.size() / 3);
    std::vector<int> u, v, opp;
    for (int k = 0; k < 3; ++k)
        for (int f = 0; f < F; ++f) {
            int a = S.I[3 * f + k], b = S.I[3 * f + (k + 1) % 3], c = S.I[3 * f + (k + 2) % 3];
            u.push_back(std::min(a, b)), v.push_back(std::max(a, b)), opp.push_back(c);
        }
    std::vector<int64_t> key(u.size());
    for (size_t i = 0; i < u.size(); ++i) key[i] = (int64_t)u[i] * nv + v[i];
    std::vector<int> order(u.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
    S.psil3.clear();
    S.ps_pmf.clear();
    size_t i = 0;
    while (i < order.size()) {
        size_t j = i;
        while (j < order.size() && key[order[j]] == key[order[i]]) ++j;
        if (j - i == 2) {
            int e0 = order[i], e1 = order[i + 1];
            auto P = [&](int id) { return v3(S.V[3 * id], S.V[3 * id + 1], S.V[3 * id + 2]); };
            V3 a = P(u[e0]), b = P(v[e0]), c = P(opp[e0]), d = P(opp[e1]);
            V3 n0 = normalize(cross(sub(b, a), sub(c, a))), n1 = normalize(cross(sub(d, a), sub(b, a)));
            if (std::fabs(dot(n0, n1)) < (1.0f - 1e-5f)) {
                float q[12] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z};
                S.psil3.insert(S.psil3.end(), q, q + 12);
                S.ps_pmf.push_back(norm(sub(a, b)));
            }
        }
        i = j;
    }
    S.ps_cdf.resize(S.ps_pmf.size());
    float acc = 0.0f;
    for (size_t k = 0; k < S.ps_pmf.size(); ++k) acc += S.ps_pmf[k], S.ps_cdf[k] = acc;
    S.ps_sum = S.ps_cdf.empty() ? 0.0f : S.ps_cdf.back();
    S.ps_norm = 1.0f / S.ps_sum;
}

template <int D>
static void make_sub(const Scene& sc, int want_type, Sub<D>& S) {
    size_t np = sc.types.size();
    std::vector<int> sel;
    for (size_t p = 0; p < np; ++p)
        if (want_type < 0 || sc.types[p] == want_type) sel.push_back((int)p);
    if (sel.empty()) {
        S.present = false;
        return;
    }
    std::vector<int> uniq;
    for (int p : sel)
        for (int k = 0; k < D; ++k) uniq.push_back(sc.I[D * p + k]);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    std::unordered_map<int, int> remap;
    for (size_t i = 0; i < uniq.size(); ++i) remap[uniq[i]] = (int)i;
    S.V.resize(uniq.size() * D);
    for (size_t i = 0; i < uniq.size(); ++i)
        for (int d = 0; d < D; ++d) S.V[D * i + d] = sc.V[D * uniq[i] + d];
    S.I.resize(sel.size() * D);
    for (size_t j = 0; j < sel.size(); ++j)
        for (int k = 0; k < D; ++k) S.I[D * j + k] = remap[sc.I[D * sel[j] + k]];
    S.vglobal = uniq;
    S.pglobal = sel;
    build_sub<D>(S);
    build_snch<D>(S);
    if constexpr (D == 2) build_psil2(S);
    if constexpr (D == 3) build_psil3(S);
}

// ----------------------------------------------------------------------------
// Counters for algorithmic bytes (reference SoA record sizes)
// ----------------------------------------------------------------------------
struct Cnt {
    uint64_t bytes = 0, nodes = 0, prims = 0, sils = 0, flops = 0;
    uint64_t ff = 0;  // steps taken from a NaN position (no record fetched)
    // closest-point queries and the loop iterations / interior-record fetches
    // the GPU traversal takes for them at node width 2 and 4
    uint64_t cq = 0, it2 = 0, it4 = 0, rf2 = 0, rf4 = 0;
    void add(const Cnt& o) {
        bytes += o.bytes, nodes += o.nodes, prims += o.prims, sils += o.sils, flops += o.flops, ff += o.ff;
        cq += o.cq, it2 += o.it2, it4 += o.it4, rf2 += o.rf2, rf4 += o.rf4;
    }
};
// Iterations of the GPU closest-point loop, replayed from this traversal's
// pops.  The GPU descends into the chosen child without a stack round trip and
// takes one pop (a pending node, processed or pruned) per iteration, so its
// iterations are these pops.  With two-level records (node width 4) the pair
// of an interior child descended into from a record's first level is tested
// in the same iteration, so such a visit costs no iteration of its own.
struct IterCount {
    Cnt* c;
    int desc = -1;  // child the last interior visit descends into
    int lvl = 0;    // 1: the last visit read a record's first level
    explicit IterCount(Cnt* c_) : c(c_) {
        if (c) c->cq++;
    }
    // a popped node: pruned, a leaf, or an interior node (chosen: the child it
    // descends into, -1: none)
    void pruned() {
        if (c) c->it2++, c->it4++;
        desc = -1, lvl = 0;
    }
    void leaf() { pruned(); }
    void interior(int ni, int chosen) {
        if (c) {
            c->it2++, c->rf2++;
            if (ni == desc && lvl == 1) {
                lvl = 2;
            } else {
                c->it4++, c->rf4++;
                lvl = 1;
            }
        }
        desc = chosen;
    }
};
enum { NODE2 = 24, SNODE2 = 48, SNODE3 = 60, SEG = 20, TRI = 40, SILV = 40, SILE = 68 };
// Algorithmic fp32 operations per record test, counted from the routines below
// (add, sub, mul, div, min, max, compare, sqrt and each transcendental count 1;
// a dot product of n terms counts 2n-1):
//   FB2/FB3    box_dist (2 sub-vectors, 2 max, norm, min, norm) + the r_max update and test
//   FR2/FR3    box_ray (2 x (sub, div) per axis, min/max per axis, 2 folds of dim)
//   FC2/FC3    snch_overlaps past the box: centre, view vector, norm, acos of the
//              axis angle, range tests, one asin or atan2 branch
//   FSEG/FTRI  seg_closest / tri_closest (Ericson's region tests: 6 dots, 3 combos,
//              barycentric divide and lerp, norm)
//   FSEGR/FTRIR seg_ray / tri_ray (Moller-Trumbore: 2 edges, 2 crosses, 3 dots, 3 divides)
//   FSILV/FSILE silv_closest (norm, 2 crosses) / sile_closest (segment point, norm,
//              half of the time two face normals and two dots)
//   FSTEP2/FSTEP3 the scalar logic of one walk step (ball radius, direction sample,
//              Green's function and its rejection sampler, Neumann and control-variate
//              terms, ray set-up)
enum {
    FB2 = 20, FB3 = 29, FR2 = 17, FR3 = 25, FC2 = 30, FC3 = 42, FSEG = 29, FTRI = 70, FSEGR = 30, FTRIR = 50,
    FSILV = 21, FSILE = 70, FSTEP2 = 90, FSTEP3 = 80
};

// ----------------------------------------------------------------------------
// 2D queries (diff_wost/shapes/bvh.py, snch.py, line_segment.py, silhouette_vertex.py)
// ----------------------------------------------------------------------------
struct CP {  // ClosestPointRecord (2D uses t, 3D uses uv)
    V3 p, n;
    float t, u, v, d;
    int prim;
};
struct Its {  // Intersection
    bool valid, on_boundary;
    V3 p, n;
    float t, u, v, d;
    int prim;
};
struct SilRec {
    bool valid;
    V3 p;
    float d;
    int prim;
};

static inline V2 node_min2(const Node<2>& n) { return v2(n.box.mn[0], n.box.mn[1]); }
static inline V2 node_max2(const Node<2>& n) { return v2(n.box.mx[0], n.box.mx[1]); }
static inline V3 node_min3(const Node<3>& n) { return v3(n.box.mn[0], n.box.mn[1], n.box.mn[2]); }
static inline V3 node_max3(const Node<3>& n) { return v3(n.box.mx[0], n.box.mx[1], n.box.mx[2]); }

// BoundingBox.distance (bvh.py:31-36)
static inline void box_dist(V2 mn, V2 mx, V2 p, float& dmin, float& dmax) {
    V2 u = sub(mn, p), v = sub(p, mx);
    dmin = norm(vmax(vmax(u, v), v2(0.0f, 0.0f)));
    dmax = norm(vmin(u, v));
}
static inline void box_dist(V3 mn, V3 mx, V3 p, float& dmin, float& dmax) {
    V3 u = sub(mn, p), v = sub(p, mx);
    dmin = norm(vmax(vmax(u, v), v3(0.0f, 0.0f, 0.0f)));
    dmax = norm(vmin(u, v));
}
// BoundingBox.intersect (bvh.py:42-59)
static inline bool box_ray(V2 mn, V2 mx, V2 x, V2 v, float rmax, float& tmin) {
    V2 t0 = v2((mn.x - x.x) / v.x, (mn.y - x.y) / v.y);
    V2 t1 = v2((mx.x - x.x) / v.x, (mx.y - x.y) / v.y);
    V2 tn = vmin(t0, t1), tf = vmax(t0, t1);
    float tnm = fmaxf_(fmaxf_(tn.x, tn.y), 0.0f);
    float tfm = fminf_(fminf_(tf.x, tf.y), rmax);
    if (tnm > tfm) {
        tmin = kInf;
        return false;
    }
    tmin = tnm;
    return true;
}
static inline bool box_ray(V3 mn, V3 mx, V3 x, V3 v, float rmax, float& tmin) {
    V3 t0 = v3((mn.x - x.x) / v.x, (mn.y - x.y) / v.y, (mn.z - x.z) / v.z);
    V3 t1 = v3((mx.x - x.x) / v.x, (mx.y - x.y) / v.y, (mx.z - x.z) / v.z);
    V3 tn = vmin(t0, t1), tf = vmax(t0, t1);
    float tnm = fmaxf_(fmaxf_(fmaxf_(tn.x, tn.y), tn.z), 0.0f);
    float tfm = fminf_(fminf_(fminf_(tf.x, tf.y), tf.z), rmax);
    if (tnm > tfm) {
        tmin = kInf;
        return false;
    }
    tmin = tnm;
    return true;
}

// LineSegment.closest_point (line_segment.py:74-85)
static inline void seg_closest(const Prim<2>& s, V2 p, CP& r) {
    V2 a = v2(s.v[0][0], s.v[0][1]), b = v2(s.v[1][0], s.v[1][1]);
    V2 pa = sub(p, a), ba = sub(b, a);
    V2 t = normalize(ba);
    V2 n = v2(t.y, -t.x);
    float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
    V2 i = add(a, mul(ba, h));
    float d = norm(sub(i, p));
    r.p = v3(i.x, i.y, 0.0f);
    r.n = v3(n.x, n.y, 0.0f);
    r.t = h;
    r.u = r.v = 0.0f;
    r.d = d;
    r.prim = s.pIndex;
}

// BVH.closest_point (bvh.py:111-188)
// r0 < inf starts the pruning radius below infinity (a bound the caller knows
// on the answer); the reference query is r0 = inf.
static CP bvh2_closest(const Tree<2>& T, V2 p, Cnt* c, float r0 = kInf) {
    CP its;
    its.p = v3(p.x, p.y, 0.0f);
    its.n = v3(0.0f, 0.0f, 0.0f);
    its.t = kInf;
    its.u = its.v = 0.0f;
    its.d = kInf;
    its.prim = -1;
    float r_max = r0;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = r0 < kInf ? 0.0f : r_max;
    IterCount ic(c);
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (cd > r_max) {
            ic.pruned();
            continue;
        }
        const Node<2>& node = T.nodes[ni];
        if (c) c->bytes += NODE2, c->nodes++;
        if (node.nref > 0) {
            ic.leaf();
            for (int j = 0; j < node.nref; ++j) {
                CP r;
                seg_closest(T.prims[node.off + j], p, r);
                if (c) c->bytes += SEG, c->prims++, c->flops += FSEG + 1;
                if (r.d < its.d) its = r;
            }
            r_max = fminf_(r_max, its.d);
        } else {
            const int sp0 = sp;
            int L = ni + 1, R = ni + node.off;
            float dminL, dmaxL, dminR, dmaxR;
            box_dist(node_min2(T.nodes[L]), node_max2(T.nodes[L]), p, dminL, dmaxL);
            bool hitL = dminL <= r_max;
            r_max = fminf_(r_max, dmaxL);
            box_dist(node_min2(T.nodes[R]), node_max2(T.nodes[R]), p, dminR, dmaxR);
            bool hitR = dminR <= r_max;
            r_max = fminf_(r_max, dmaxR);
            if (c) c->bytes += 2 * NODE2, c->nodes += 2, c->flops += 2 * FB2 + 1;
            if (hitL && hitR) {
                int closer = L, other = R;
                // `d_max_right < d_max_right` at bvh.py:163 never swaps
                if (dminR < dminL) {
                    std::swap(closer, other);
                    std::swap(dminL, dminR);
                }
                st_i[++sp] = other;
                st_d[sp] = dminR;
                st_i[++sp] = closer;
                st_d[sp] = dminL;
            } else if (hitL) {
                st_i[++sp] = L;
                st_d[sp] = dminL;
            } else if (hitR) {
                st_i[++sp] = R;
                st_d[sp] = dminR;
            }
            ic.interior(ni, sp > sp0 ? st_i[sp] : -1);
        }
    }
    return its;
}

// LineSegment.ray_intersect (line_segment.py:47-72)
static inline bool seg_ray(const Prim<2>& s, V2 x, V2 d, float rmax, Its& r) {
    V2 a = v2(s.v[0][0], s.v[0][1]), b = v2(s.v[1][0], s.v[1][1]);
    V2 u = sub(a, x), v = sub(b, a);
    float dv = d.x * v.y - d.y * v.x;
    if (!(std::fabs(dv) > kDrEps)) return false;
    float ud = u.x * d.y - u.y * d.x;
    float s_ = ud / dv;
    if (!((s_ >= 0.0f) && (s_ <= 1.0f))) return false;
    float uv = u.x * v.y - u.y * v.x;
    float t = uv / dv;
    if (!((t >= 0.0f) && (t <= rmax))) return false;
    V2 n = normalize(v2(v.y, -v.x));
    V2 p = add(x, mul(d, t));
    r.valid = true;
    r.on_boundary = true;
    r.p = v3(p.x, p.y, 0.0f);
    r.n = v3(n.x, n.y, 0.0f);
    r.t = s_;
    r.u = r.v = 0.0f;
    r.d = std::fabs(t);
    r.prim = s.pIndex;
    return true;
}

// BVH.intersect (bvh.py:190-257).  `r_max = its.d` (:216) reads the previous
// record (d = 0 initially), so the first valid hit in traversal order wins.
static Its bvh2_intersect(const Tree<2>& T, V2 x, V2 v, float r_max, Cnt* c) {
    Its its;
    std::memset(&its, 0, sizeof(its));
    its.prim = 0;
    float tmin;
    if (c) c->bytes += NODE2, c->nodes++, c->flops += FR2;
    if (!box_ray(node_min2(T.nodes[0]), node_max2(T.nodes[0]), x, v, r_max, tmin)) return its;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = tmin;
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (!(cd <= r_max)) continue;
        const Node<2>& node = T.nodes[ni];
        if (c) c->bytes += NODE2, c->nodes++;
        if (node.nref > 0) {
            for (int j = 0; j < node.nref; ++j) {
                Its r;
                if (c) c->bytes += SEG, c->prims++, c->flops += FSEGR;
                if (seg_ray(T.prims[node.off + j], x, v, r_max, r) && r.d < r_max) {
                    r_max = its.d;
                    its = r;
                }
            }
        } else {
            int L = ni + 1, R = ni + node.off;
            float t0, t1;
            bool h0 = box_ray(node_min2(T.nodes[L]), node_max2(T.nodes[L]), x, v, r_max, t0);
            bool h1 = box_ray(node_min2(T.nodes[R]), node_max2(T.nodes[R]), x, v, r_max, t1);
            if (c) c->bytes += 2 * 16, c->nodes += 2, c->flops += 2 * FR2 + 1;
            if (h0 && h1) {
                int closer = L, other = R;
                if (t1 < t0) {
                    std::swap(closer, other);
                    std::swap(t0, t1);
                }
                st_i[++sp] = other;
                st_d[sp] = t1;
                st_i[++sp] = closer;
                st_d[sp] = t0;
            } else if (h0) {
                st_i[++sp] = L;
                st_d[sp] = t0;
            } else if (h1) {
                st_i[++sp] = R;
                st_d[sp] = t1;
            }
        }
    }
    return its;
}

// project_to_plane (bvh.py:16-20) and the 3D basis version (snch3d.py:30-45)
static inline float project_to_plane(V2 n, V2 e) {
    V2 b = v2(-n.y, n.x);
    return std::fabs(dot(e, vabs(b)));
}
static inline float project_to_plane(V3 n, V3 e) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    V3 b1 = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    V3 b2 = v3(b, sign + n.y * n.y * a, -n.y);
    float r1 = dot(e, vabs(b1));
    float r2 = dot(e, vabs(b2));
    return std::sqrt(r1 * r1 + r2 * r2);
}
static inline bool in_range(float x, float a, float b) { return (x >= a) && (x <= b); }

// SNCHNode.overlaps (snch.py:46-97, snch3d.py:65-115)
template <typename VT>
static inline bool snch_overlaps(VT mn, VT mx, const float* axis_, float ha, float radius, VT x, float r_max,
                                 float& d_min) {
    float d_max;
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    if ((ha >= 0.0f) && ov) {
        ov = (ha > kPi2) && (d_min < kDrEps);
        if (!ov) {
            VT cen = divs(add(mn, mx), 2.0f);
            VT vca = sub(cen, x);
            float len = norm(vca);
            vca = divs(vca, len);
            VT axis;
            std::memcpy(&axis, axis_, sizeof(VT));
            float dax = cr_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
            ov = in_range(kPi2, dax - ha, dax + ha);
            if (!ov) {
                if (len > radius) {
                    float vch = cr_asin(radius / len);
                    float hs = ha + vch;
                    ov = (hs > kPi2) || in_range(kPi2, dax - hs, dax + hs);
                } else {
                    VT e = sub(mx, cen);
                    float d = dot(e, vabs(vca));
                    float s = len - d;
                    ov = s < 0.0f;
                    if (!ov) {
                        d = project_to_plane(vca, e);
                        float vch = cr_atan2(d, s);
                        float hs = ha + vch;
                        ov = (hs > kPi2) || in_range(kPi2, dax - hs, dax + hs);
                    }
                }
            }
        }
    }
    return ov;
}

// is_silhouette (math.py:144-160) + SilhouetteVertex.closest_silhouette (silhouette_vertex.py:80-93)
static inline bool silv_closest(const Sil2& s, int ref, V2 x, float r_max, SilRec& r) {
    V2 a = v2(s.a[0], s.a[1]), b = v2(s.b[0], s.b[1]), c = v2(s.c[0], s.c[1]);
    V2 view = sub(x, b);
    float d = norm(view);
    if ((s.idx[0] != -1) && (s.idx[2] != -1)) {
        if (d < r_max) {
            if (cross2(sub(b, a), sub(x, a)) * cross2(sub(c, b), sub(x, b)) < 0.0f) {
                r.valid = true;
                r.p = v3(b.x, b.y, 0.0f);
                r.d = d;
                r.prim = ref;
                return true;
            }
        }
    }
    return false;
}

// closest_point_line_segment (math.py:124-141) in 3D + silhouette_edge.py:8-13,83-97
static inline bool sile_closest(const Sil3& s, int ref, V3 x, float r_max, SilRec& r) {
    V3 a = v3(s.a[0], s.a[1], s.a[2]), b = v3(s.b[0], s.b[1], s.b[2]);
    V3 u = sub(b, a), v = sub(x, a);
    float t = clampf(dot(v, u) / dot(u, u), 0.0f, 1.0f);
    V3 pt = lerp(a, b, t);
    float d = norm(sub(pt, x));
    if (d < r_max) {
        bool is = (s.idx[0] == -1) || (s.idx[3] == -1);
        if (!is) {
            V3 c = v3(s.c[0], s.c[1], s.c[2]), dd = v3(s.d[0], s.d[1], s.d[2]);
            V3 n0 = normalize(cross(sub(b, a), sub(c, a)));
            V3 n1 = normalize(cross(sub(dd, a), sub(b, a)));
            V3 view = sub(pt, x);
            is = dot(view, n0) * dot(view, n1) < 0.0f;
        }
        if (is) {
            r.valid = true;
            r.p = pt;
            r.d = d;
            r.prim = ref;
            return true;
        }
    }
    return false;
}

// SNCH.closest_silhouette (snch.py:312-378, snch3d.py:594-663)
template <int D>
static SilRec snch_closest_sil(const Sub<D>& S, const float* xq, float r_max, Cnt* c) {
    typedef typename std::conditional<D == 2, V2, V3>::type VT;
    SilRec rec;
    std::memset(&rec, 0, sizeof(rec));
    VT x;
    std::memcpy(&x, xq, sizeof(VT));
    const Tree<D>& T = S.tree;
    const int NB = D == 2 ? SNODE2 : SNODE3;
    float dmin, dmax;
    VT mn, mx;
    std::memcpy(&mn, T.nodes[0].box.mn, sizeof(VT));
    std::memcpy(&mx, T.nodes[0].box.mx, sizeof(VT));
    box_dist(mn, mx, x, dmin, dmax);
    if (c) c->bytes += NB, c->nodes++, c->flops += D == 2 ? FB2 : FB3;
    if (!(dmin <= r_max)) return rec;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = dmin;
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (!(cd <= r_max)) continue;
        const Node<D>& node = T.nodes[ni];
        if (c) c->bytes += NB, c->nodes++;
        if (node.nref > 0) {
            for (int j = 0; j < node.nsil; ++j) {
                int ref = node.sil_off + j;
                SilRec r;
                bool ok;
                if constexpr (D == 2) {
                    ok = silv_closest(S.ref2[ref], ref, v2(xq[0], xq[1]), r_max, r);
                    if (c) c->bytes += SILV, c->sils++, c->flops += FSILV;
                } else {
                    ok = sile_closest(S.ref3[ref], ref, v3(xq[0], xq[1], xq[2]), r_max, r);
                    if (c) c->bytes += SILE, c->sils++, c->flops += FSILE;
                }
                if (ok && r.d < r_max) {
                    r_max = fminf_(r_max, r.d);
                    rec = r;
                }
            }
        } else {
            int L = ni + 1, R = ni + node.off;
            const Node<D>& nl = T.nodes[L];
            const Node<D>& nr = T.nodes[R];
            VT mnl, mxl, mnr, mxr;
            std::memcpy(&mnl, nl.box.mn, sizeof(VT));
            std::memcpy(&mxl, nl.box.mx, sizeof(VT));
            std::memcpy(&mnr, nr.box.mn, sizeof(VT));
            std::memcpy(&mxr, nr.box.mx, sizeof(VT));
            float d0, d1;
            bool h0 = snch_overlaps(mnl, mxl, nl.axis, nl.half_angle, nl.radius, x, r_max, d0);
            bool h1 = snch_overlaps(mnr, mxr, nr.axis, nr.half_angle, nr.radius, x, r_max, d1);
            if (c) {
                c->bytes += 2 * NB, c->nodes += 2;
                // the cone part runs where the box overlaps and the cone is not trivially open
                const int fb = D == 2 ? FB2 : FB3, fc = D == 2 ? FC2 : FC3;
                const Node<D>* ch[2] = {&nl, &nr};
                const float dd[2] = {d0, d1};
                for (int k = 0; k < 2; ++k) {
                    float ha = ch[k]->half_angle;
                    bool cone = ha >= 0.0f && dd[k] <= r_max && !(ha > kPi2 && dd[k] < kDrEps);
                    c->flops += fb + (cone ? fc : 0);
                }
            }
            if (h0 && h1) {
                int closer = L, other = R;
                if (d1 < d0) {
                    std::swap(closer, other);
                    std::swap(d0, d1);
                }
                st_i[++sp] = other;
                st_d[sp] = d1;
                st_i[++sp] = closer;
                st_d[sp] = d0;
            } else if (h0) {
                st_i[++sp] = L;
                st_d[sp] = d0;
            } else if (h1) {
                st_i[++sp] = R;
                st_d[sp] = d1;
            }
        }
    }
    return rec;
}

// Polyline.star_radius (polyline.py:304-314) / Mesh.star_radius (mesh.py:337-346)
template <int D>
static float star_radius(const Sub<D>& S, const float* x, float r_max, Cnt* c) {
    float m[D];
    for (int d = 0; d < D; ++d) m[d] = fmaxf_(std::fabs(x[d] - S.bb_min[d]), std::fabs(x[d] - S.bb_max[d]));
    float R;
    if (D == 2)
        R = norm(v2(m[0], m[1]));
    else
        R = norm(v3(m[0], m[1], m[2]));
    R = fminf_(R, r_max);
    if (!S.py_sils) return R;
    SilRec s = snch_closest_sil<D>(S, x, R, c);
    return s.valid ? s.d : R;
}

// ----------------------------------------------------------------------------
// 3D queries (snch3d.py:334-476, triangle.py, math.py:207-272)
// ----------------------------------------------------------------------------
static inline void tri_closest(const Prim<3>& tr, V3 p, CP& r) {  // closest_point_triangle
    V3 a = v3(tr.v[0][0], tr.v[0][1], tr.v[0][2]);
    V3 b = v3(tr.v[1][0], tr.v[1][1], tr.v[1][2]);
    V3 c = v3(tr.v[2][0], tr.v[2][1], tr.v[2][2]);
    V3 ab = sub(b, a), ac = sub(c, a);
    V3 ax = sub(p, a);
    float d1 = dot(ab, ax), d2 = dot(ac, ax);
    V3 pt;
    float u0, u1;
    V3 bx = sub(p, b);
    float d3 = dot(ab, bx), d4 = dot(ac, bx);
    V3 cx = sub(p, c);
    float d5 = dot(ab, cx), d6 = dot(ac, cx);
    float vc = d1 * d4 - d3 * d2;
    float vb = d5 * d2 - d1 * d6;
    float va = d3 * d6 - d5 * d4;
    if ((d1 <= 0.0f) && (d2 <= 0.0f)) {
        pt = a, u0 = 1.0f, u1 = 0.0f;
    } else if ((d3 >= 0.0f) && (d4 <= d3)) {
        pt = b, u0 = 0.0f, u1 = 1.0f;
    } else if ((d6 >= 0.0f) && (d5 <= d6)) {
        pt = c, u0 = 0.0f, u1 = 0.0f;
    } else if ((vc <= 0.0f) && (d1 >= 0.0f) && (d3 <= 0.0f)) {
        float v = d1 / (d1 - d3);
        pt = add(a, mul(ab, v));
        u0 = 1.0f - v, u1 = v;
    } else if ((vb <= 0.0f) && (d2 >= 0.0f) && (d6 <= 0.0f)) {
        float w = d2 / (d2 - d6);
        pt = add(a, mul(ac, w));
        u0 = 1.0f - w, u1 = 0.0f;
    } else if ((va <= 0.0f) && ((d4 - d3) >= 0.0f) && ((d5 - d6) >= 0.0f)) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        pt = add(b, mul(sub(c, b), w));
        u0 = 0.0f, u1 = 1.0f - w;
    } else {
        float denom = 1.0f / (va + vb + vc);
        float v = vb * denom, w = vc * denom;
        pt = add(add(a, mul(ab, v)), mul(ac, w));
        u0 = 1.0f - v - w, u1 = v;
    }
    r.p = pt;
    r.d = norm(sub(p, pt));
    r.u = u1;
    r.v = 1.0f - u0 - u1;
    r.t = 0.0f;
    r.n = normalize(cross(ab, ac));
    r.prim = tr.pIndex;
}

// SNCH3D.closest_point (snch3d.py:334-409)
// r0 < inf starts the pruning radius below infinity (a bound the caller knows
// on the answer); the reference query is r0 = inf.
static CP snch3_closest(const Tree<3>& T, V3 p, Cnt* c, float r0 = kInf) {
    CP its;
    std::memset(&its, 0, sizeof(its));
    its.d = kInf;
    its.prim = -1;
    float r_max = r0;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = r0 < kInf ? 0.0f : r_max;
    IterCount ic(c);
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (!(cd <= r_max)) {
            ic.pruned();
            continue;
        }
        const Node<3>& node = T.nodes[ni];
        if (c) c->bytes += SNODE3, c->nodes++;
        if (node.nref > 0) {
            ic.leaf();
            for (int j = 0; j < node.nref; ++j) {
                CP r;
                tri_closest(T.prims[node.off + j], p, r);
                if (c) c->bytes += TRI, c->prims++, c->flops += FTRI + 1;
                if (r.d < its.d) {
                    its = r;
                    r_max = fminf_(r_max, its.d);
                }
            }
        } else {
            int L = ni + 1, R = ni + node.off;
            float dminL, dmaxL, dminR, dmaxR;
            const int sp0 = sp;
            box_dist(node_min3(T.nodes[L]), node_max3(T.nodes[L]), p, dminL, dmaxL);
            bool hitL = dminL <= r_max;
            r_max = fminf_(r_max, dmaxL);
            box_dist(node_min3(T.nodes[R]), node_max3(T.nodes[R]), p, dminR, dmaxR);
            bool hitR = dminR <= r_max;
            r_max = fminf_(r_max, dmaxR);
            if (c) c->bytes += 2 * SNODE3, c->nodes += 2, c->flops += 2 * FB3 + 1;
            if (hitL && hitR) {
                int closer = L, other = R;
                if ((dminL == 0.0f) && (dminR == 0.0f)) {
                    if (dmaxR < dmaxL) std::swap(closer, other);
                }
                if (dminR < dminL) {
                    std::swap(closer, other);
                    std::swap(dminL, dminR);
                }
                st_i[++sp] = other;
                st_d[sp] = dminR;
                st_i[++sp] = closer;
                st_d[sp] = dminL;
            } else if (hitL) {
                st_i[++sp] = L;
                st_d[sp] = dminL;
            } else if (hitR) {
                st_i[++sp] = R;
                st_d[sp] = dminR;
            }
            ic.interior(ni, sp > sp0 ? st_i[sp] : -1);
        }
    }
    return its;
}

// Triangle.ray_intersect (triangle.py:59-86); prim_id = sorted index
static inline bool tri_ray(const Prim<3>& tr, int sorted, V3 x, V3 d, float rmax, Its& r) {
    V3 a = v3(tr.v[0][0], tr.v[0][1], tr.v[0][2]);
    V3 b = v3(tr.v[1][0], tr.v[1][1], tr.v[1][2]);
    V3 c = v3(tr.v[2][0], tr.v[2][1], tr.v[2][2]);
    V3 v1 = sub(b, a), v2_ = sub(c, a);
    V3 p = cross(d, v2_);
    float det = dot(v1, p);
    if (!(std::fabs(det) > kDrEps)) return false;
    float inv = 1.0f / det;
    V3 s = sub(x, a);
    float v = dot(s, p) * inv;
    if (!((v >= 0.0f) && (v <= 1.0f))) return false;
    V3 q = cross(s, v1);
    float w = dot(d, q) * inv;
    if (!((w >= 0.0f) && (v + w <= 1.0f))) return false;
    float t = dot(v2_, q) * inv;
    if (!((t >= 0.0f) && (t <= rmax))) return false;
    r.valid = true;
    r.on_boundary = true;
    r.p = add(add(a, mul(v1, v)), mul(v2_, w));
    r.n = normalize(cross(v1, v2_));
    r.u = 1.0f - v - w;
    r.v = v;
    r.t = 0.0f;
    r.d = t;
    r.prim = sorted;
    return true;
}

// SNCH3D.intersect (snch3d.py:411-476)
static Its snch3_intersect(const Tree<3>& T, V3 x, V3 v, float r_max, Cnt* c) {
    Its its;
    std::memset(&its, 0, sizeof(its));
    float tmin;
    if (c) c->bytes += SNODE3, c->nodes++, c->flops += FR3;
    if (!box_ray(node_min3(T.nodes[0]), node_max3(T.nodes[0]), x, v, r_max, tmin)) return its;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = tmin;
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (!(cd <= r_max)) continue;
        const Node<3>& node = T.nodes[ni];
        if (c) c->bytes += SNODE3, c->nodes++;
        if (node.nref > 0) {
            for (int j = 0; j < node.nref; ++j) {
                Its r;
                if (c) c->bytes += TRI, c->prims++, c->flops += FTRIR;
                if (tri_ray(T.prims[node.off + j], node.off + j, x, v, r_max, r) && r.d < r_max) {
                    r_max = r.d;
                    its = r;
                }
            }
        } else {
            int L = ni + 1, R = ni + node.off;
            float t0, t1;
            bool h0 = box_ray(node_min3(T.nodes[L]), node_max3(T.nodes[L]), x, v, r_max, t0);
            bool h1 = box_ray(node_min3(T.nodes[R]), node_max3(T.nodes[R]), x, v, r_max, t1);
            if (c) c->bytes += 2 * 24, c->nodes += 2, c->flops += 2 * FR3 + 1;
            if (h0 && h1) {
                int closer = L, other = R;
                if (t1 < t0) {
                    std::swap(closer, other);
                    std::swap(t0, t1);
                }
                st_i[++sp] = other;
                st_d[sp] = t1;
                st_i[++sp] = closer;
                st_d[sp] = t0;
            } else if (h0) {
                st_i[++sp] = L;
                st_d[sp] = t0;
            } else if (h1) {
                st_i[++sp] = R;
                st_d[sp] = t1;
            }
        }
    }
    return its;
}

// ----------------------------------------------------------------------------
// Brute-force queries: the reference's *_baseline methods
// (polyline.py:246-302,468-496; mesh.py:312-335,440-472,558-570)
// ----------------------------------------------------------------------------
static CP brute_closest2(const Sub<2>& S, V2 p) {
    float dmin = kInf;
    int idx = -1;
    size_t np = S.I.size() / 2;
    for (size_t i = 0; i < np; ++i) {
        V2 a = v2(S.V[2 * S.I[2 * i]], S.V[2 * S.I[2 * i] + 1]);
        V2 b = v2(S.V[2 * S.I[2 * i + 1]], S.V[2 * S.I[2 * i + 1] + 1]);
        V2 pa = sub(p, a), ba = sub(b, a);
        float h = clampf(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
        float d = norm(sub(pa, mul(ba, h)));
        if (d < dmin) dmin = d, idx = (int)i;
    }
    CP r;
    std::memset(&r, 0, sizeof(r));
    r.d = dmin;
    r.prim = idx;
    return r;
}
static Its brute_ray2(const Sub<2>& S, V2 x, V2 v, float rmax) {
    float dmin = rmax;
    int idx = -1;
    size_t np = S.I.size() / 2;
    for (size_t i = 0; i < np; ++i) {
        V2 a = v2(S.V[2 * S.I[2 * i]], S.V[2 * S.I[2 * i] + 1]);
        V2 b = v2(S.V[2 * S.I[2 * i + 1]], S.V[2 * S.I[2 * i + 1] + 1]);
        V2 u = sub(b, a), w = sub(x, a);
        float d = cross2(v, u);
        float s = cross2(v, w) / d;
        float t = cross2(u, w) / d;
        bool valid = (t > 0.0f) && (s >= 0.0f) && (s <= 1.0f);
        float dd = valid ? t : kInf;
        if (dd < dmin) dmin = dd, idx = (int)i;
    }
    Its r;
    std::memset(&r, 0, sizeof(r));
    r.valid = idx >= 0;
    r.d = dmin;
    r.prim = idx;
    return r;
}
// the tree's primitives in the sub-scene's index order (the baselines' loop order)
static std::vector<int> index_order(const Sub<3>& S) {
    std::vector<int> pos(S.tree.prims.size());
    for (size_t j = 0; j < S.tree.prims.size(); ++j) pos[S.tree.prims[j].pIndex] = (int)j;
    return pos;
}
static CP brute_closest3(const Sub<3>& S, V3 p) {  // mesh.py:312-335: i = 0 .. n-1
    float dmin = kInf;
    int idx = -1;
    std::vector<int> pos = index_order(S);
    for (size_t i = 0; i < pos.size(); ++i) {
        CP r;
        tri_closest(S.tree.prims[pos[i]], p, r);
        if (r.d < dmin) dmin = r.d, idx = r.prim;
    }
    CP r;
    std::memset(&r, 0, sizeof(r));
    r.d = dmin;
    r.prim = idx;
    return r;
}
static Its brute_ray3(const Sub<3>& S, V3 x, V3 v, float rmax) {  // mesh.py:440-472: i = 0 .. n-1
    float dmin = rmax;
    int idx = -1;
    std::vector<int> pos = index_order(S);
    for (size_t i = 0; i < pos.size(); ++i) {
        Its r;
        const Prim<3>& tr = S.tree.prims[pos[i]];
        if (tri_ray(tr, pos[i], x, v, kInf, r) && r.d < dmin) dmin = r.d, idx = tr.pIndex;
    }
    Its r;
    std::memset(&r, 0, sizeof(r));
    r.valid = idx >= 0;
    r.d = dmin;
    r.prim = idx;
    return r;
}
template <int D>
static SilRec brute_sil(const Sub<D>& S, const float* x, float r_max) {
    SilRec rec;
    std::memset(&rec, 0, sizeof(rec));
    float dmin = r_max;
    size_t n = D == 2 ? S.sil2.size() : S.sil3.size();
    for (size_t i = 0; i < n; ++i) {
        SilRec r;
        bool ok;
        if constexpr (D == 2)
            ok = silv_closest(S.sil2[i], (int)i, v2(x[0], x[1]), r_max, r);
        else
            ok = sile_closest(S.sil3[i], (int)i, v3(x[0], x[1], x[2]), r_max, r);
        if (ok && r.d < dmin) dmin = r.d, rec = r;
    }
    return rec;
}

// ----------------------------------------------------------------------------
// Boundary data (g = u(p) on Dirichlet, h = n . grad u on Neumann)
// ----------------------------------------------------------------------------
struct BC {
    int code;
    float c[4];
    const float* gv;  // per-vertex g (global vertex ids) when code == BC_PER_VERTEX
    const float* hv;
    int src_code;
    float src_c;
    float sc[4];  // SRC_LINEAR: f = sc0 + sc1 x + sc2 y + sc3 z
};
static inline float bc_u(const BC& bc, V3 p) {
    switch (bc.code) {
        case BC_EXPCOS: return cr_exp(p.x) * cr_cos(p.y);
        case BC_QUAD: return p.x * p.x - p.y * p.y;
        default: return ((bc.c[0] + bc.c[1] * p.x) + bc.c[2] * p.y) + bc.c[3] * p.z;
    }
}
static inline V3 bc_grad(const BC& bc, V3 p) {
    switch (bc.code) {
        case BC_EXPCOS: {
            float ex = cr_exp(p.x);
            return v3(ex * cr_cos(p.y), -ex * cr_sin(p.y), 0.0f);
        }
        case BC_QUAD: return v3(2.0f * p.x, -2.0f * p.y, 0.0f);
        default: return v3(bc.c[1], bc.c[2], bc.c[3]);
    }
}
// the source f (the reference's scene.f): zero, a constant, or linear
static inline float src_f(const BC& bc, V3 p) {
    if (bc.src_code == SRC_CONST) return bc.src_c;
    if (bc.src_code == SRC_LINEAR) return ((bc.sc[0] + bc.sc[1] * p.x) + bc.sc[2] * p.y) + bc.sc[3] * p.z;
    return 0.0f;
}
static inline float src_f(const BC& bc, V2 p) { return src_f(bc, v3(p.x, p.y, 0.0f)); }
static inline V3 src_grad(const BC& bc) {  // the reference's scene.dfdx
    return bc.src_code == SRC_LINEAR ? v3(bc.sc[1], bc.sc[2], bc.sc[3]) : v3(0.0f, 0.0f, 0.0f);
}

// g on a Dirichlet closest-point record of sub-scene S
template <int D>
static inline float g_of(const BC& bc, const Sub<D>& S, const CP& c) {
    if (bc.code != BC_PER_VERTEX) return bc_u(bc, c.p);
    int gp = S.pglobal[c.prim];
    (void)gp;
    if (D == 2) {
        float g0 = bc.gv[S.vglobal[S.I[2 * c.prim]]], g1 = bc.gv[S.vglobal[S.I[2 * c.prim + 1]]];
        return lerpf(g0, g1, c.t);
    } else {
        float ga = bc.gv[S.vglobal[S.I[3 * c.prim]]], gb = bc.gv[S.vglobal[S.I[3 * c.prim + 1]]],
              gc = bc.gv[S.vglobal[S.I[3 * c.prim + 2]]];
        return (gb * c.u + gc * c.v) + ga * ((1.0f - c.u) - c.v);
    }
}
// h at a point on prim `prim` of sub-scene S with parameters (t) or (u,v), normal n
template <int D>
static inline float h_of(const BC& bc, const Sub<D>& S, int prim, float t, float u, float v, V3 p, V3 n) {
    if (bc.code != BC_PER_VERTEX) {
        V3 g = bc_grad(bc, p);
        return D == 2 ? dot(v2(n.x, n.y), v2(g.x, g.y)) : dot(n, g);
    }
    if (D == 2) {
        float h0 = bc.hv[S.vglobal[S.I[2 * prim]]], h1 = bc.hv[S.vglobal[S.I[2 * prim + 1]]];
        return lerpf(h0, h1, t);
    } else {
        float ha = bc.hv[S.vglobal[S.I[3 * prim]]], hb = bc.hv[S.vglobal[S.I[3 * prim + 1]]],
              hc = bc.hv[S.vglobal[S.I[3 * prim + 2]]];
        return (hb * u + hc * v) + ha * ((1.0f - u) - v);
    }
}

// ----------------------------------------------------------------------------
// Sampling (distr.py, greens_fn.py, greens_fn_3d.py, math.py:275-303)
// ----------------------------------------------------------------------------
// DiscreteDistribution.sample_reuse_pmf (distr.py:52-107) with Dr.Jit binary_search
template <int D>
static inline int cdf_sample(const Sub<D>& S, float value, float& reuse, float& pmf, Cnt* c) {
    float sample = value * S.pmf_sum;
    int n = (int)S.cdf.size();
    uint32_t start = 0, end = (uint32_t)(n - 1);
    int iters = 0;
    if (start < end) {
        uint32_t x = end - start;
        int l2 = 31 - __builtin_clz(x);
        iters = l2 + 1;
    }
    for (int it = 0; it < iters; ++it) {
        uint32_t mid = (start + end) >> 1;
        float cv = S.cdf[mid];
        bool cond = ((cv < sample) || (cv == 0.0f)) && (cv != S.pmf_sum);
        if (cond)
            start = std::min(mid + 1, end);
        else
            end = mid;
    }
    if (c) c->bytes += 4 * iters + 8, c->flops += 2 * iters + 4;
    int idx = (int)start;
    pmf = S.pmf[idx] * S.pmf_norm;
    float cdf = idx > 0 ? S.cdf[idx - 1] * S.pmf_norm : 0.0f;
    reuse = (value - cdf) / pmf;
    return idx;
}

static const float kInv2Pi_neg = (float)(-1.0 / (2.0 * 3.14159265358979323846));

// GreensBall (greens_fn.py:14-287); ball centred at c
static inline float greens2_G(V2 c, float R, V2 x, V2 y) {
    x = sub(x, c);
    y = sub(y, c);
    x = divs(x, R);
    y = divs(y, R);
    float v = norm(sub(x, y));
    return kInv2Pi_neg * cr_log((1.0f - dot(x, y)) / v);
}
// GreensBall._sample rejection loop (greens_fn.py:173-206)
static inline void greens2_sample(V2 c, float R, V2 dir, Pcg& rng, V2& p, float& r, float& pdf) {
    float bound = 1.5f / R;
    int iteration = 0;
    bool active = true;
    r = 0.0f;
    pdf = 0.0f;
    while ((iteration < 1000) && active) {
        float u = rng.next_float32();
        r = rng.next_float32() * R;
        pdf = (cr_log(R / r) / kTwoPi) / (R * R / 4.0f);
        float pdf_radius = pdf / (1.0f / (kTwoPi * r));
        iteration += 1;
        if (u < pdf_radius / bound) active = false;
    }
    r = fmaxf_(r, 1e-6f);
    if (r > R) r = R / 2.0f;
    p = add(c, mul(dir, r));
}
static inline V2 greens2_gradient(V2 c, float R, float r, V2 y) {
    float gn = (1.0f / (r * r) - 1.0f / (R * R)) / kTwoPi;
    return mul(sub(y, c), gn);
}
static inline float greens3_G(V3 c, float R, V3 x, V3 y) {  // greens_fn_3d.py:10-12
    float r = norm(sub(y, x));
    return -(1.0f / r - R / (R * R - dot(sub(x, c), sub(y, c)))) / kFourPi;
}
static inline void greens3_sample(V3 c, float R, V3 dir, Pcg& rng, V3& p, float& r, float& pdf) {
    float u1 = rng.next_float32();
    float u2 = rng.next_float32();
    float phi = kTwoPi * u2;
    r = (1.0f + std::sqrt(1.0f - cr_cbrt(u1 * u1)) * cr_cos(phi)) * R / 2.0f;
    r = fmaxf_(r, 1e-6f);
    if (r > R) r = R / 2.0f;
    pdf = ((1.0f / r - 1.0f / R) / kFourPi) / (R * R / 6.0f);
    p = add(c, mul(dir, r));
}
static inline V3 greens3_gradient(V3 c, float R, float r, V3 y) {
    float r3 = r * r * r;
    float gn = (1.0f / r3 - 1.0f / (R * R * R)) / kFourPi;
    return mul(sub(y, c), gn);
}
static inline V2 uniform_dir2(Pcg& rng) {  // uniform_angle (math.py:275-277)
    float theta = (rng.next_float32() * 2.0f) * kPi;
    return v2(cr_cos(theta), cr_sin(theta));
}
static inline V3 uniform_sphere(Pcg& rng) {  // math.py:298-303
    float u0 = rng.next_float32();
    float u1 = rng.next_float32();
    float z = 1.0f - 2.0f * u0;
    float r = std::sqrt(fmaxf_(0.0f, 1.0f - z * z));
    float theta = kTwoPi * u1;
    return v3(r * cr_cos(theta), r * cr_sin(theta), z);
}

// ----------------------------------------------------------------------------
// Walks
// ----------------------------------------------------------------------------
struct Params {
    int nwalks, nsteps, nsubwalks;
    float eps, min_R;
    int cv, av;
    uint32_t seed;
    int64_t base;  // global index of point 0: streams are keyed by (base + i, walk)
    int wos;       // plain Walk on Spheres (wos.py / wos_3d.py) instead of WoSt
};
struct WalkOut {
    float value;
    int steps;
    int term_prim;  // Dirichlet prim the walk terminated on (-1: none)
    Cnt cnt;
};

// sdf / inside (polyline.py:188-234, mesh.py:290-297) on the full scene
static float sdf2(const Scene& sc, V2 x) {
    const Sub<2>& F = sc.s2[0];
    CP its = bvh2_closest(F.tree, x, nullptr);
    int i0 = F.I[2 * its.prim], i1 = F.I[2 * its.prim + 1];
    V2 n1 = v2(F.vnormal[2 * i0], F.vnormal[2 * i0 + 1]), n2 = v2(F.vnormal[2 * i1], F.vnormal[2 * i1 + 1]);
    V2 n = lerp(n1, n2, its.t);
    float s = dot(n, sub(x, v2(its.p.x, its.p.y))) >= 0.0f ? 1.0f : -1.0f;  // dr.sign: +1 at 0
    return its.d * s;
}
static float sdf3(const Scene& sc, V3 x) {
    CP c = snch3_closest(sc.s3[0].tree, x, nullptr);
    float s = dot(c.n, sub(x, c.p)) >= 0.0f ? 1.0f : -1.0f;
    return c.d * s;
}
static bool inside2(const Scene& sc, V2 x) { return sdf2(sc, x) < 0.0f; }
static bool inside3(const Scene& sc, V3 x) { return sdf3(sc, x) < 0.0f; }

// WoSt.u (wost.py:145-221)
static WalkOut wost_u2(const Scene& sc, const BC& bc, const Params& P, V2 x, Pcg& rng, bool ob0, V2 n0,
                       bool count) {
    WalkOut out;
    out.term_prim = -1;
    Cnt* c = count ? &out.cnt : nullptr;
    const Sub<2>& F = sc.s2[0];
    const Sub<2>& Dn = sc.s2[1];
    const Sub<2>& Nm = sc.s2[2];
    V2 ip = x, in = n0;
    bool ob = ob0;
    float result = 0.0f;
    int i = 0;
    bool active = true;
    while ((i < P.nsteps) && active) {
        if (c) c->flops += FSTEP2;
        float R = kInf, d_dir = kInf;
        if (Dn.present) {
            CP cp = bvh2_closest(Dn.tree, ip, c);
            R = d_dir = std::fabs(cp.d);
        }
        if (Nm.present) {
            V2 p = ob ? sub(ip, mul(in, 1e-5f)) : ip;
            float sr = star_radius<2>(Nm, &p.x, d_dir, c) * 0.98f;
            sr = fmaxf_(sr, P.min_R);
            R = fminf_(R, sr);
        }
        if (d_dir < P.eps) {
            CP cp = bvh2_closest(Dn.tree, ip, c);
            result += g_of<2>(bc, Dn, cp);
            out.term_prim = cp.prim;
            active = false;
        } else {
            V2 ref = v2(0.0f, 0.0f);
            if (Nm.present) {
                float rnd, pmf;
                int idx = cdf_sample<2>(Nm, rng.next_float32(), rnd, pmf, c);
                V2 a = v2(Nm.V[2 * Nm.I[2 * idx]], Nm.V[2 * Nm.I[2 * idx] + 1]);
                V2 b = v2(Nm.V[2 * Nm.I[2 * idx + 1]], Nm.V[2 * Nm.I[2 * idx + 1] + 1]);
                V2 tg = sub(b, a);
                V2 bn = normalize(v2(tg.y, -tg.x));
                V2 bp = lerp(a, b, rnd);
                float bpdf = pmf / norm(tg);
                // boundary_control_variates (wost.py:123-126): full-scene closest point
                CP cr = bvh2_closest(F.tree, ip, c);
                float hc = h_of<2>(bc, F, cr.prim, cr.t, 0.0f, 0.0f, cr.p, cr.n);
                ref = v2(hc * cr.n.x, hc * cr.n.y);
                // handle_boundary (wost.py:104-121)
                float alpha = ob ? 2.0f : 1.0f;
                float hb = 0.0f;
                float G = greens2_G(ip, R, ip, bp);
                G = clampf(G, -10.0f, 10.0f);
                if (norm(sub(ip, bp)) <= R) {
                    float h = h_of<2>(bc, Nm, idx, rnd, 0.0f, 0.0f, v3(bp.x, bp.y, 0.0f), v3(bn.x, bn.y, 0.0f));
                    if (P.cv) h = h - dot(ref, bn);
                    hb = 0.0f - G * h / bpdf;
                }
                result += alpha * hb;
            }
            V2 dir = uniform_dir2(rng);
            if (ob && (dot(dir, in) > 0.0f)) dir = neg(dir);
            V2 _p = ip;
            // Polyline.intersect (polyline.py:435-466)
            V2 o = ob ? sub(ip, mul(in, 1e-5f)) : ip;
            Its its = bvh2_intersect(F.tree, o, dir, R, c);
            if (!its.valid) {
                V2 q = add(o, mul(dir, R));
                its.p = v3(q.x, q.y, 0.0f);
                its.n = v3(dir.x, dir.y, 0.0f);
                its.d = R;
                its.on_boundary = false;
            }
            ip = v2(its.p.x, its.p.y);
            in = v2(its.n.x, its.n.y);
            ob = its.on_boundary;
            if (P.cv) result += dot(sub(_p, ip), ref);
            // handle_source (wost.py:128-143)
            V2 sp;
            float sr_, spdf;
            greens2_sample(_p, R, dir, rng, sp, sr_, spdf);
            float src = 0.0f;
            if (norm(sub(sp, _p)) < its.d) src += (R * R / 4.0f) * src_f(bc, sp);
            result += src;
        }
        i += 1;
    }
    if (bad(result) || std::fabs(result) > 100.0f) result = 0.0f;
    out.value = result;
    out.steps = i;
    return out;
}

// WoSt3D.u (wost_3d.py:43-106)
static WalkOut wost_u3(const Scene& sc, const BC& bc, const Params& P, V3 x, Pcg& rng, bool ob0, V3 n0,
                       bool count) {
    WalkOut out;
    out.term_prim = -1;
    Cnt* c0 = count ? &out.cnt : nullptr;
    const Sub<3>& F = sc.s3[0];
    const Sub<3>& Dn = sc.s3[1];
    const Sub<3>& Nm = sc.s3[2];
    V3 ip = x, in = n0;
    bool ob = ob0;
    float result = 0.0f;
    int i = 0;
    bool active = true;
    while ((i < P.nsteps) && active) {
        // A step from a NaN position (a walk that overflowed to infinity) has a
        // fixed outcome and needs no record: it counts as a step with no bytes
        // and is tallied in `ff`.
        const bool ff = std::isnan(ip.x) || std::isnan(ip.y) || std::isnan(ip.z);
        Cnt* c = ff ? nullptr : c0;
        if (c0) {
            if (ff)
                c0->ff++;
            else
                c0->flops += FSTEP3;
        }
        float R = kInf, d_dir = kInf;
        if (Dn.present) {
            CP cp = snch3_closest(Dn.tree, ip, c);
            R = d_dir = std::fabs(cp.d);
        }
        V3 p = ob ? sub(ip, mul(in, 1e-3f)) : ip;
        float sr = star_radius<3>(F, &p.x, d_dir, c) * 0.98f;
        sr = fmaxf_(sr, P.min_R);
        R = fminf_(R, sr);
        if (d_dir < P.eps) {
            CP cp = snch3_closest(Dn.tree, ip, c);
            result += g_of<3>(bc, Dn, cp);
            out.term_prim = cp.prim;
            active = false;
        } else {
            if (Nm.present) {
                // Mesh.sample_boundary (mesh.py:474-490)
                float rnd, pmf;
                int idx = cdf_sample<3>(Nm, rng.next_float32(), rnd, pmf, c);
                const float* A = &Nm.V[3 * Nm.I[3 * idx]];
                const float* B = &Nm.V[3 * Nm.I[3 * idx + 1]];
                const float* C = &Nm.V[3 * Nm.I[3 * idx + 2]];
                V3 a = v3(A[0], A[1], A[2]), b = v3(B[0], B[1], B[2]), cc = v3(C[0], C[1], C[2]);
                V3 cr = cross(sub(b, a), sub(cc, a));
                float area = norm(cr) / 2.0f;
                float r1 = rng.next_float32();
                float r2 = rng.next_float32();
                float su = std::sqrt(r1) * (1.0f - r2);
                float sv = std::sqrt(r1) * r2;
                V3 bp = add(add(mul(b, su), mul(cc, sv)), mul(a, (1.0f - su) - sv));
                V3 bn = normalize(cr);
                float bpdf = pmf / area;
                float G = greens3_G(ip, R, ip, bp);
                G = clampf(G, -100.0f, 100.0f);
                float alpha = ob ? 2.0f : 1.0f;
                if (norm(sub(ip, bp)) <= R) {
                    float h = h_of<3>(bc, Nm, idx, 0.0f, su, sv, bp, bn);
                    result -= G * h * alpha / bpdf;
                }
            }
            V3 dir = uniform_sphere(rng);
            if (ob && (dot(dir, in) > 0.0f)) dir = neg(dir);
            V3 _p = ip;
            V3 o = ob ? sub(ip, mul(in, 1e-5f)) : ip;
            Its its = snch3_intersect(F.tree, o, dir, R, c);
            if (!its.valid) {
                its.p = add(o, mul(dir, R));
                its.n = dir;
                its.d = R;
                its.on_boundary = false;
            }
            ip = its.p;
            in = its.n;
            ob = its.on_boundary;
            V3 sp;
            float sr_, spdf;
            greens3_sample(_p, R, dir, rng, sp, sr_, spdf);
            float src = 0.0f;
            if (norm(sub(sp, _p)) < its.d) src += (R * R / 6.0f) * src_f(bc, sp);
            result += src;
        }
        i += 1;
    }
    out.value = result;
    out.steps = i;
    return out;
}

// WoS.u (wos.py:16-36) / WoS3D.u (wos_3d.py:15-34): every boundary primitive
// is Dirichlet; step to a uniform point on the largest empty ball.
template <int D>
static WalkOut wos_u(const Scene& sc, const BC& bc, const Params& P, const float* x, Pcg& rng) {
    WalkOut out;
    out.term_prim = -1;
    float result = 0.0f;
    int i = 0;
    bool active = true;
    V2 p2 = v2(x[0], x[1]);
    V3 p3 = D == 3 ? v3(x[0], x[1], x[2]) : v3(0, 0, 0);
    while ((i < P.nsteps) && active) {
        CP cp = D == 2 ? bvh2_closest(sc.s2[0].tree, p2, nullptr) : snch3_closest(sc.s3[0].tree, p3, nullptr);
        float d = std::fabs(cp.d);  // scene.distance
        if (d < P.eps) {
            // scene.closest_point: the same query
            result += D == 2 ? g_of<2>(bc, sc.s2[0], cp) : g_of<3>(bc, sc.s3[0], cp);
            out.term_prim = cp.prim;
            active = false;
        } else if (D == 2) {
            V2 dir = uniform_dir2(rng);
            p2 = add(p2, mul(dir, d));
        } else {
            V3 dir = uniform_sphere(rng);
            p3 = add(p3, mul(dir, d));
        }
        i += 1;
    }
    out.value = result;
    out.steps = i;
    return out;
}

// Solver.walk (solver.py:73-92): nsubwalks sequential walks, averaged
static WalkOut walk_u(const Scene& sc, const BC& bc, const Params& P, const float* x, Pcg& rng, bool count) {
    WalkOut acc;
    acc.value = 0.0f;
    acc.steps = 0;
    acc.term_prim = -1;
    float res = 0.0f;
    for (int s = 0; s < P.nsubwalks; ++s) {
        WalkOut w = P.wos ? (sc.dim == 2 ? wos_u<2>(sc, bc, P, x, rng) : wos_u<3>(sc, bc, P, x, rng))
                    : sc.dim == 2 ? wost_u2(sc, bc, P, v2(x[0], x[1]), rng, false, v2(0, 0), count)
                                  : wost_u3(sc, bc, P, v3(x[0], x[1], x[2]), rng, false, v3(0, 0, 0), count);
        res += w.value;
        acc.steps += w.steps;
        acc.term_prim = w.term_prim;
        acc.cnt.add(w.cnt);
    }
    acc.value = res / (float)P.nsubwalks;
    return acc;
}

// WoSt.dude (wost.py:30-102); returns result / nwalks
static float dude2(const Scene& sc, const BC& bc, const Params& P, V2 p, V2 e, Pcg& rng, uint64_t* steps) {
    CP cp = bvh2_closest(sc.s2[0].tree, p, nullptr);
    float R = std::fabs(cp.d);
    float u_sum = 0.0f, u_count = 0.0f, src_sum = 0.0f, src_count = 0.0f, result = 0.0f;
    int i = 0;
    bool active = inside2(sc, p);
    float gnorm = R * R / 4.0f;
    while ((i < P.nwalks) && active) {
        V2 dir = uniform_dir2(rng);
        V2 dir_src = uniform_dir2(rng);
        V2 p_src;
        float r_src, pdf_src;
        greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
        int niters = P.av ? 2 : 1;
        uint64_t state = rng.state;
        int it = 0;
        while ((it < niters) && (i < P.nwalks)) {
            if (it == 1) {
                dir_src = neg(dir_src);
                p_src = sub(mul(p, 2.0f), p_src);
            }
            float src = gnorm * src_f(bc, p_src);
            float src_grad = dot(greens2_gradient(p, R, r_src, p_src), e) / (pdf_src * gnorm);
            src_sum += src;
            src_count += 1.0f;
            float _src = src;
            if (P.cv) _src = src - (src_sum / src_count);
            result += _src * src_grad;
            if (it == 1) dir = neg(dir);
            V2 _p = add(p, mul(dir, R));
            rng.state = state;
            WalkOut w = wost_u2(sc, bc, P, _p, rng, false, v2(0, 0), false);
            if (steps) *steps += (uint64_t)w.steps;
            float u = w.value;
            u_sum += u;
            u_count += 1.0f;
            float _u = u;
            if (P.cv) _u = u - u_sum / u_count;
            float r2 = 2.0f / R * _u * dot(e, dir);
            if ((std::fabs(r2) > 1e6f) || bad(r2)) r2 = 0.0f;
            result += r2;
            i += 1;
            it += 1;
        }
    }
    return result / (float)P.nwalks;
}

// WoSt3D.dude (wost_3d.py:108-175)
static float dude3(const Scene& sc, const BC& bc, const Params& P, V3 p, V3 e, Pcg& rng, uint64_t* steps) {
    CP cp = snch3_closest(sc.s3[0].tree, p, nullptr);
    float R = std::fabs(cp.d);
    float u_mean = 0.0f, u_count = 0.0f, src_mean = 0.0f, src_count = 0.0f, result = 0.0f;
    int i = 0;
    bool active = inside3(sc, p);
    float gnorm = R * R / 6.0f;
    while ((i < P.nwalks) && active) {
        V3 dir = uniform_sphere(rng);
        V3 p_src;
        float r_src, pdf_src;
        greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
        int niters = P.av ? 2 : 1;
        uint64_t state = rng.state;
        int it = 0;
        while ((it < niters) && (i < P.nwalks)) {
            if (it == 1) {
                dir = neg(dir);
                p_src = sub(mul(p, 2.0f), p_src);
            }
            float src = gnorm * src_f(bc, p_src);
            float src_grad = dot(greens3_gradient(p, R, r_src, p_src), e) / (pdf_src * gnorm);
            src_mean = src_mean * (src_count / (src_count + 1.0f)) + src / (src_count + 1.0f);
            src_count += 1.0f;
            float _src = src;
            if (P.cv) _src = src - src_mean;
            result += _src * src_grad;
            if (it == 1) dir = neg(dir);  // second flip (wost_3d.py:152-153)
            V3 _p = add(p, mul(dir, R));
            rng.state = state;
            WalkOut w = wost_u3(sc, bc, P, _p, rng, false, v3(0, 0, 0), false);
            if (steps) *steps += (uint64_t)w.steps;
            float u = w.value;
            if (bad(u) || std::fabs(u) > 10000.0f) u = 0.0f;
            u_mean = u_mean * (u_count / (u_count + 1.0f)) + u / (u_count + 1.0f);
            u_count += 1.0f;
            float _u = u;
            if (P.cv) _u = u - u_mean;
            result += 3.0f / R * _u * dot(e, dir);
            i += 1;
            it += 1;
        }
    }
    if (bad(result) || std::fabs(result) > 10000.0f) result = 0.0f;
    return result / (float)P.nwalks;
}

// WoS.dude (wos.py:38-105) / WoS3D.dude (wos_3d.py:36-102): antithetic pairs
// from the largest empty ball with running-sum control variates; the source
// direction is a separate draw in both dimensions; no outlier guards.
template <int D>
static float wos_dude(const Scene& sc, const BC& bc, const Params& P, const float* x, const float* ev, Pcg& rng,
                      uint64_t* steps) {
    V3 p = v3(x[0], x[1], D == 3 ? x[2] : 0.0f), e = v3(ev[0], ev[1], D == 3 ? ev[2] : 0.0f);
    float R = D == 2 ? std::fabs(bvh2_closest(sc.s2[0].tree, v2(p.x, p.y), nullptr).d)
                     : std::fabs(snch3_closest(sc.s3[0].tree, p, nullptr).d);
    bool active = D == 2 ? inside2(sc, v2(p.x, p.y)) : inside3(sc, p);
    float u_sum = 0.0f, u_count = 0.0f, src_sum = 0.0f, src_count = 0.0f, result = 0.0f;
    float gnorm = D == 2 ? R * R / 4.0f : R * R / 6.0f;
    int i = 0;
    while ((i < P.nwalks) && active) {
        V3 dir, dir_src, p_src;
        float r_src, pdf_src;
        if (D == 2) {
            V2 d2 = uniform_dir2(rng), s2 = uniform_dir2(rng), q2;
            greens2_sample(v2(p.x, p.y), R, s2, rng, q2, r_src, pdf_src);
            dir = v3(d2.x, d2.y, 0.0f), dir_src = v3(s2.x, s2.y, 0.0f), p_src = v3(q2.x, q2.y, 0.0f);
        } else {
            dir = uniform_sphere(rng);
            dir_src = uniform_sphere(rng);
            greens3_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
        }
        int niters = P.av ? 2 : 1;
        uint64_t state = rng.state;
        int it = 0;
        while ((it < niters) && (i < P.nwalks)) {
            if (it == 1) {
                dir_src = neg(dir_src);
                p_src = sub(mul(p, 2.0f), p_src);
            }
            float src = gnorm * src_f(bc, p_src);
            float src_grad;
            if (D == 2) {
                V2 gr = greens2_gradient(v2(p.x, p.y), R, r_src, v2(p_src.x, p_src.y));
                src_grad = dot(gr, v2(e.x, e.y)) / (pdf_src * gnorm);
            } else {
                src_grad = dot(greens3_gradient(p, R, r_src, p_src), e) / (pdf_src * gnorm);
            }
            src_sum += src;
            src_count += 1.0f;
            float _src = P.cv ? src - (src_sum / src_count) : src;
            result += _src * src_grad;
            if (it == 1) dir = neg(dir);
            V3 _p = add(p, mul(dir, R));
            rng.state = state;
            float xp[3] = {_p.x, _p.y, _p.z};
            WalkOut w = wos_u<D>(sc, bc, P, xp, rng);
            if (steps) *steps += (uint64_t)w.steps;
            float u = w.value;
            u_sum += u;
            u_count += 1.0f;
            float _u = P.cv ? u - u_sum / u_count : u;
            float de = D == 2 ? dot(v2(e.x, e.y), v2(dir.x, dir.y)) : dot(e, dir);
            result += (D == 2 ? 2.0f : 3.0f) / R * _u * de;
            i += 1;
            it += 1;
        }
    }
    return result / (float)P.nwalks;
}

// Fused u + grad u estimator (the product's throughput form of u and the
// first-ball gradient; no single reference function): per antithetic pair k of point i, a PCG32 stream seeded like
// Solver.solve with lane = i*npairs + k; direction (and the source radius) are
// drawn as in WoSt.dude / WoSt3D.dude, the state is saved, and the two walks
// from p +- R*dir replay the same stream.
struct PairOut {
    float usum;      // u(+) + u(-) + src(+) + src(-)
    float g[3];      // sum over the two walks of the gradient terms
    int steps;
};
static PairOut fused_pair(const Scene& sc, const BC& bc, const Params& P, const float* x, float R, bool inside,
                          Pcg& rng, Cnt* cnt = nullptr) {
    PairOut o;
    std::memset(&o, 0, sizeof(o));
    if (!inside) return o;
    int D = sc.dim;
    if (D == 2) {
        V2 p = v2(x[0], x[1]);
        V2 dir = uniform_dir2(rng);
        V2 dir_src = uniform_dir2(rng);
        V2 p_src;
        float r_src, pdf_src;
        greens2_sample(p, R, dir_src, rng, p_src, r_src, pdf_src);
        float gnorm = R * R / 4.0f;
        uint64_t state = rng.state;
        WalkOut wp = wost_u2(sc, bc, P, add(p, mul(dir, R)), rng, false, v2(0, 0), cnt != nullptr);
        rng.state = state;
        WalkOut wm = wost_u2(sc, bc, P, sub(p, mul(dir, R)), rng, false, v2(0, 0), cnt != nullptr);
        if (cnt) cnt->add(wp.cnt), cnt->add(wm.cnt);
        float s_p = gnorm * src_f(bc, p_src), s_m = gnorm * src_f(bc, sub(mul(p, 2.0f), p_src));
        V2 gs = divs(greens2_gradient(p, R, r_src, p_src), pdf_src * gnorm);
        float du = wp.value - wm.value;
        float k = 2.0f / R * du;
        o.usum = (wp.value + wm.value) + (s_p + s_m);
        o.g[0] = k * dir.x + (s_p - s_m) * gs.x;
        o.g[1] = k * dir.y + (s_p - s_m) * gs.y;
        o.steps = wp.steps + wm.steps;
    } else {
        V3 p = v3(x[0], x[1], x[2]);
        V3 dir = uniform_sphere(rng);
        V3 p_src;
        float r_src, pdf_src;
        greens3_sample(p, R, dir, rng, p_src, r_src, pdf_src);
        float gnorm = R * R / 6.0f;
        uint64_t state = rng.state;
        WalkOut wp = wost_u3(sc, bc, P, add(p, mul(dir, R)), rng, false, v3(0, 0, 0), cnt != nullptr);
        rng.state = state;
        WalkOut wm = wost_u3(sc, bc, P, sub(p, mul(dir, R)), rng, false, v3(0, 0, 0), cnt != nullptr);
        if (cnt) cnt->add(wp.cnt), cnt->add(wm.cnt);
        float up = wp.value, um = wm.value;
        if (bad(up) || std::fabs(up) > 10000.0f) up = 0.0f;
        if (bad(um) || std::fabs(um) > 10000.0f) um = 0.0f;
        float s_p = gnorm * src_f(bc, p_src), s_m = gnorm * src_f(bc, sub(mul(p, 2.0f), p_src));
        V3 gs = divs(greens3_gradient(p, R, r_src, p_src), pdf_src * gnorm);
        float k = 3.0f / R * (up - um);
        o.usum = (up + um) + (s_p + s_m);
        o.g[0] = k * dir.x + (s_p - s_m) * gs.x;
        o.g[1] = k * dir.y + (s_p - s_m) * gs.y;
        o.g[2] = k * dir.z + (s_p - s_m) * gs.z;
        o.steps = wp.steps + wm.steps;
    }
    for (int q = 0; q < 3; ++q)
        if (bad(o.g[q])) o.g[q] = 0.0f;
    return o;
}

// ============================================================================
// WoStGrad (diff_wost/solvers/wost_grad.py): the robust-derivative walk, 2D
// ============================================================================
struct GradParams {
    float rho_max;
    int ignore_boundary, ignore_silhouette, ignore_source;
    int plant;  // tests only: 0 = the reference; 1/2/3 plant a mistake (K's sign, e kept at reflections, dh/dt's sign)
};

// Second derivatives of the analytic fields (for dh/dt): (u_xx, u_xy, u_yy)
static inline void bc_hess2(const BC& bc, V2 p, float& hxx, float& hxy, float& hyy) {
    if (bc.code == BC_EXPCOS) {
        float ex = cr_exp(p.x), c = cr_cos(p.y), sn = cr_sin(p.y);
        hxx = ex * c, hxy = -(ex * sn), hyy = -(ex * c);
    } else if (bc.code == BC_QUAD) {
        hxx = 2.0f, hxy = 0.0f, hyy = -2.0f;
    } else {
        hxx = hxy = hyy = 0.0f;
    }
}
// the scene callbacks of the analytic field u (examples/example_2d/scene.py:
// g = u, h = n . grad u, and their derivatives along the tangent (-n.y, n.x))
static inline float grad_h2(const BC& bc, V2 p, V2 n) {
    V3 g = bc_grad(bc, v3(p.x, p.y, 0.0f));
    return dot(n, v2(g.x, g.y));
}
static inline float grad_dgdt2(const BC& bc, V2 p, V2 t) {
    V3 g = bc_grad(bc, v3(p.x, p.y, 0.0f));
    return dot(v2(g.x, g.y), t);
}
static inline float grad_dhdt2(const BC& bc, V2 p, V2 n) {
    float hxx, hxy, hyy;
    bc_hess2(bc, p, hxx, hxy, hyy);
    V2 t = v2(-n.y, n.x);
    return dot(n, v2(hxx * t.x + hxy * t.y, hxy * t.x + hyy * t.y));
}

// LineSegment.star_radius_2 (line_segment.py:169-212)
static float seg_star_radius2(V2 a, V2 b, V2 x, V2 e, float rho_max, float r_max) {
    float R = r_max;
    V2 tangent = normalize(sub(b, a));
    V2 normal = v2(tangent.y, -tangent.x);
    V2 dir = normalize(sub(a, x));
    if (dot(dir, normal) > 0.0f) {
        float en = dot(e, normal), et = dot(e, tangent), d = dot(sub(a, x), normal);
        float t1 = -kInf, t2 = kInf;
        V2 pt = add(x, mul(normal, d));
        if (std::fabs(en) > 1e-4f) {
            t1 = d / en * (rho_max + et);
            t2 = -d / en * (rho_max - et);
        }
        float ta = dot(sub(a, x), tangent), tb = dot(sub(b, x), tangent);
        float tmin = fminf_(ta, tb), tmax = fmaxf_(ta, tb);
        if ((t1 > tmin) && (t1 < tmax)) R = fminf_(R, norm(sub(add(pt, mul(tangent, t1)), x)));
        if ((t2 > tmin) && (t2 < tmax)) R = fminf_(R, norm(sub(add(pt, mul(tangent, t2)), x)));
        if (((t1 > tmax) && (t2 > tmax)) || ((t1 < tmin) && (t2 < tmin))) {
            R = fminf_(R, norm(sub(a, x)));
            R = fminf_(R, norm(sub(b, x)));
        }
    } else {
        R = fminf_(R, norm(sub(a, x)));
        R = fminf_(R, norm(sub(b, x)));
    }
    return R;
}

// SNCHNode.overlaps_2 (snch.py:100-174)
static bool snch_overlaps2(const Node<2>& nd, V2 x, V2 e, float rho_bound, float r_max, float& d_min) {
    float d_max;
    V2 mn = node_min2(nd), mx = node_max2(nd);
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    float ha = nd.half_angle;
    if ((ha >= 0.0f) && ov) {
        ov = (ha > kPi2) && (d_min < kDrEps);
        if (!ov) {
            V2 c = divs(add(mn, mx), 2.0f);
            V2 vca = sub(c, x);
            float length = norm(vca);
            vca = divs(vca, length);
            ov = length < nd.radius;
            if (!ov) {
                V2 axis = v2(nd.axis[0], nd.axis[1]);
                float axis_angle = cr_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
                float vch = cr_asin(nd.radius / length);
                float hs = ha + vch;
                float amin = axis_angle - hs, amax = axis_angle + hs;
                ov = (hs > kPi2) || in_range(kPi2, amin, amax);
                if (!ov) {
                    float c0 = cr_cos(amin), c1 = cr_cos(amax);
                    float dn_min = fminf_(c0, c1), dn_max = fmaxf_(c0, c1);
                    float ed_axis = cr_acos(clampf(dot(e, vca), -1.0f, 1.0f));
                    float e0 = cr_cos(ed_axis - ha), e1 = cr_cos(ed_axis + ha);
                    float ed_min = fminf_(e0, e1), ed_max = fmaxf_(e0, e1);
                    float en_axis = cr_acos(clampf(dot(e, axis), -1.0f, 1.0f));
                    float n0 = cr_cos(en_axis - ha), n1 = cr_cos(en_axis + ha);
                    float en_min = fminf_(n0, n1), en_max = fmaxf_(n0, n1);
                    float en_dn_max = en_max / dn_min;
                    float en_dn_min = en_min < 0.0f ? en_min / dn_min : en_min / dn_max;
                    auto rho = [](float ed, float q) { return (1.0f - 2.0f * ed * q) + q * q; };
                    float rm = rho(ed_min, en_dn_min);
                    rm = fmaxf_(rm, rho(ed_min, en_dn_max));
                    rm = fmaxf_(rm, rho(ed_max, en_dn_min));
                    rm = fmaxf_(rm, rho(ed_max, en_dn_max));
                    ov = rm > rho_bound * rho_bound;
                }
            }
        }
    }
    return ov;
}

// Polyline.star_radius_2 (polyline.py:317-330) -> SNCH.star_radius_2 (snch.py:380-453)
static float star_radius2_rho(const Sub<2>& S, V2 x, V2 e, float rho_max) {
    float m0 = fmaxf_(std::fabs(x.x - S.bb_min[0]), std::fabs(x.x - S.bb_max[0]));
    float m1 = fmaxf_(std::fabs(x.y - S.bb_min[1]), std::fabs(x.y - S.bb_max[1]));
    float r_max = norm(v2(m0, m1));
    const Tree<2>& T = S.tree;
    float dmin, dmax;
    box_dist(node_min2(T.nodes[0]), node_max2(T.nodes[0]), x, dmin, dmax);
    if (!(dmin <= r_max)) return r_max;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = dmin;
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (!(cd <= r_max)) continue;
        const Node<2>& nd = T.nodes[ni];
        if (nd.nref > 0) {
            for (int j = 0; j < nd.nref; ++j) {
                const Prim<2>& pr = T.prims[nd.off + j];
                float r = seg_star_radius2(v2(pr.v[0][0], pr.v[0][1]), v2(pr.v[1][0], pr.v[1][1]), x, e, rho_max,
                                           r_max);
                if (r < r_max) r_max = r;
            }
        } else {
            int L = ni + 1, R = ni + nd.off;
            float d0, d1;
            bool h0 = snch_overlaps2(T.nodes[L], x, e, rho_max, r_max, d0);
            bool h1 = snch_overlaps2(T.nodes[R], x, e, rho_max, r_max, d1);
            if (h0 && h1) {
                int closer = L, other = R;
                if (d1 < d0) {
                    std::swap(closer, other);
                    std::swap(d0, d1);
                }
                st_i[++sp] = other;
                st_d[sp] = d1;
                st_i[++sp] = closer;
                st_d[sp] = d0;
            } else if (h0) {
                st_i[++sp] = L;
                st_d[sp] = d0;
            } else if (h1) {
                st_i[++sp] = R;
                st_d[sp] = d1;
            }
        }
    }
    return r_max;
}

// One root-to-leaf descent of BVH._sample_boundary (bvh.py:289-369; the point
// BVH's twin, bvh.py:512-...): each inner node picks a child overlapping the
// ball with probability proportional to |G(x, R, box centre)|; the leaf draws
// one of the primitives that lie in the ball (all of them when the chosen box
// lies inside it) by reservoir sampling weighted by `weight` (segment length,
// or 1 for points).  Returns the sorted index (-1: none) and the pmf.
template <typename Acc, typename Wt>
static int bvh2_descend(const std::vector<Node<2>>& nodes, V2 x, float R, Pcg& rng, bool& hit, float& pdf,
                        Acc accept, Wt weight) {
    hit = false;
    pdf = 1.0f;
    int sorted = -1;
    float r_max = kInf;
    int ni = 0;
    while (true) {
        const Node<2>& nd = nodes[ni];
        if (nd.nref > 0) {
            float total = 0.0f;
            for (int j = 0; j < nd.nref; ++j) {
                int k = nd.off + j;
                float sa = weight(k);
                if ((r_max <= R) || accept(k)) {
                    hit = true;
                    total += sa;
                    float sel = sa / total;
                    float u = rng.next_float32();
                    if (u < sel) sorted = k;
                }
            }
            if (total > 0.0f) pdf *= (sorted >= 0 ? weight(sorted) : 0.0f) / total;
            return sorted;
        }
        int L = ni + 1, Rn = ni + nd.off;
        float d0, m0, d1, m1;
        V2 mn0 = node_min2(nodes[L]), mx0 = node_max2(nodes[L]), mn1 = node_min2(nodes[Rn]), mx1 = node_max2(nodes[Rn]);
        box_dist(mn0, mx0, x, d0, m0);
        box_dist(mn1, mx1, x, d1, m1);
        float w0 = (d0 <= R) ? 1.0f : 0.0f, w1 = (d1 <= R) ? 1.0f : 0.0f;
        if (w0 > 0.0f) w0 *= std::fabs(greens2_G(x, R, x, divs(add(mn0, mx0), 2.0f)));
        if (w1 > 0.0f) w1 *= std::fabs(greens2_G(x, R, x, divs(add(mn1, mx1), 2.0f)));
        float tw = w0 + w1;
        if (!(tw > 0.0f)) return sorted;
        float p0 = w0 / tw, p1 = 1.0f - p0;
        float u = rng.next_float32();
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
    V2 p, n;
    float t, pdf;
    int prim;
};
// Polyline.sample_boundary_bvh (polyline.py:515-517) -> BVH.sample_boundary (bvh.py:266-287)
static BSample2 sample_boundary_bvh2(const Sub<2>& S, V2 x, float R, Pcg& rng) {
    BSample2 b;
    std::memset(&b, 0, sizeof(b));
    const Tree<2>& T = S.tree;
    float dmin, dmax;
    box_dist(node_min2(T.nodes[0]), node_max2(T.nodes[0]), x, dmin, dmax);
    if (!(dmin <= R)) return b;
    auto seg = [&](int k, V2& a, V2& c) {
        if (k < 0) {
            a = c = v2(0.0f, 0.0f);
            return;
        }
        a = v2(T.prims[k].v[0][0], T.prims[k].v[0][1]);
        c = v2(T.prims[k].v[1][0], T.prims[k].v[1][1]);
    };
    auto weight = [&](int k) {  // LineSegment.surface_area (line_segment.py:32-33)
        V2 a, c;
        seg(k, a, c);
        return norm(sub(c, a));
    };
    auto accept = [&](int k) {  // LineSegment.sphere_intersect (line_segment.py:43-45)
        V2 a, c;
        seg(k, a, c);
        V2 u = sub(c, a), w = sub(x, a);
        float t = clampf(dot(w, u) / dot(u, u), 0.0f, 1.0f);
        return norm(sub(lerp(a, c, t), x)) <= R;
    };
    bool hit;
    float pdf;
    int k = bvh2_descend(T.nodes, x, R, rng, hit, pdf, accept, weight);
    if (!hit) return b;
    V2 a, c;
    seg(k, a, c);
    float t = rng.next_float32();  // LineSegment.sample_point (line_segment.py:36-40)
    V2 tg = sub(c, a);
    b.hit = true;
    b.p = add(a, mul(tg, t));
    b.n = normalize(v2(tg.y, -tg.x));
    b.t = t;
    b.pdf = pdf * (1.0f / norm(tg));
    b.prim = k >= 0 ? T.prims[k].pIndex : 0;
    return b;
}

struct SSample2 {
    bool hit;
    V2 p, n1, n2, t1, t2;
    float pdf;
    int prim;
};
// Polyline.sample_silhouette_bvh (polyline.py:519-535) -> PointBVH.sample_boundary
static SSample2 sample_silhouette_bvh2(const Sub<2>& S, V2 x, float R, Pcg& rng) {
    SSample2 r;
    std::memset(&r, 0, sizeof(r));
    if (S.pnodes.empty()) return r;
    float dmin, dmax;
    box_dist(node_min2(S.pnodes[0]), node_max2(S.pnodes[0]), x, dmin, dmax);
    if (!(dmin <= R)) return r;
    auto weight = [](int) { return 1.0f; };
    auto accept = [&](int k) {
        return norm(sub(v2(S.pprims[k].v[0][0], S.pprims[k].v[0][1]), x)) <= R;
    };
    bool hit;
    float pdf;
    int k = bvh2_descend(S.pnodes, x, R, rng, hit, pdf, accept, weight);
    if (!hit) return r;
    int si = k >= 0 ? S.pprims[k].pIndex : 0;
    const float* q = &S.psil[6 * si];
    V2 a = v2(q[0], q[1]), b = v2(q[2], q[3]), c = v2(q[4], q[5]);
    r.hit = true;
    r.t1 = normalize(sub(b, a));
    r.t2 = normalize(sub(b, c));
    r.n1 = v2(r.t1.y, -r.t1.x);
    r.n2 = v2(-r.t2.y, r.t2.x);
    r.p = b;
    r.pdf = pdf * 1.0f;
    r.prim = si;
    return r;
}

// Polyline.off_centered_ball (polyline.py:583-600): bisection for the largest
// ball tangent at p that the boundary does not cut
static void off_centered_ball2(const Sub<2>& F, V2 p, V2 n, V2& c, float& radius) {
    float b = fminf_(F.bb_max[0] - F.bb_min[0], F.bb_max[1] - F.bb_min[1]) / 2.0f;
    float a = 1e-3f;
    for (int i = 0; i < 10; ++i) {
        float m = (a + b) / 2.0f;
        float d = std::fabs(bvh2_closest(F.tree, sub(p, mul(n, m)), nullptr).d);
        if (std::fabs(d - m) < 1e-3f)
            a = m;
        else
            b = m;
    }
    c = sub(p, mul(n, a));
    radius = a;
}

// WoStGrad.dudn (wost_grad.py:207-254): the normal derivative at a Dirichlet
// point from one WoSt walk (nsteps 200, eps 1e-4) started on the off-centred ball
static float grad_dudn2(const Scene& sc, const BC& bc, int cv, const CP& its, Pcg& rng, int& nested_steps) {
    const Sub<2>& F = sc.s2[0];
    const Sub<2>& Dn = sc.s2[1];
    const Sub<2>& Nm = sc.s2[2];
    V2 ip = v2(its.p.x, its.p.y), in = v2(its.n.x, its.n.y);
    V2 c;
    float R;
    off_centered_ball2(F, ip, in, c, R);
    float theta = (rng.next_float32() * 2.0f) * kPi;  // uniform_angle
    theta = clampf(theta, 3e-1f, kTwoPi - 3e-1f);
    V2 pd = v2(-in.y, in.x);
    V2 dir = add(mul(in, cr_cos(theta)), mul(pd, cr_sin(theta)));
    V2 p = add(c, mul(dir, R));
    float u_ref = bc_u(bc, its.p);
    Params W;
    std::memset(&W, 0, sizeof(W));
    W.nwalks = 1, W.nsubwalks = 1, W.nsteps = 200, W.eps = 1e-4f, W.min_R = 1e-3f, W.cv = 1, W.av = 1;
    WalkOut w = wost_u2(sc, bc, W, p, rng, false, v2(0.0f, 0.0f), false);
    nested_steps += w.steps;
    float u = w.value;
    if (Nm.present) {
        float dd = bvh2_closest(Dn.tree, p, nullptr).d;
        float dnm = bvh2_closest(Nm.tree, p, nullptr).d;
        if ((dd < 2e-2f) && (dnm < 2e-2f)) u = u_ref;
    }
    float P = 1.0f / (cr_cos(theta) - 1.0f);
    float dudn = 0.0f;
    if (cv)
        dudn += P * (u - u_ref) / R;
    else
        dudn += P * u / R;
    float rr = std::sqrt(rng.next_float32());  // uniform_in_disk (math.py:285-288)
    float th = kTwoPi * rng.next_float32();
    V2 p_src = add(c, mul(v2(rr * cr_cos(th), rr * cr_sin(th)), R));
    float f_ref = src_f(bc, ip), f = src_f(bc, p_src);
    float x2 = dot(sub(ip, c), sub(ip, c));
    float y2 = dot(sub(p_src, c), sub(p_src, c));
    float Pq = (x2 - y2) / fmaxf_(dot(sub(ip, p_src), sub(ip, p_src)), 1e-6f) * R / 2.0f;
    if (cv)
        dudn -= (f - f_ref) * Pq + f_ref * R / 2.0f;
    else
        dudn -= f * Pq;
    return dudn;
}

// WoStGrad.dude (wost_grad.py:91-205)
static float grad_dude2(const Scene& sc, const BC& bc, const Params& P, const GradParams& G, V2 x, V2 e, Pcg& rng,
                        int& steps) {
    const Sub<2>& F = sc.s2[0];
    const Sub<2>& Dn = sc.s2[1];
    const Sub<2>& Nm = sc.s2[2];
    V2 ip = x, in = v2(0.0f, 0.0f);
    bool ob = false;
    float result = 0.0f;
    int i = 0;
    bool active = ob || inside2(sc, ip);
    float thr = 1.0f;
    while ((i < P.nsteps) && active) {
        bool inflate = false;
        float d_dir = kInf;
        if (Dn.present) d_dir = std::fabs(bvh2_closest(Dn.tree, ip, nullptr).d);
        if (d_dir < P.eps) {
            CP cr = bvh2_closest(F.tree, ip, nullptr);
            V2 nrm = v2(cr.n.x, cr.n.y), tg = v2(-nrm.y, nrm.x);
            float dudt = grad_dgdt2(bc, v2(cr.p.x, cr.p.y), tg);
            float dudn = grad_dudn2(sc, bc, P.cv, cr, rng, steps);
            result += thr * dudt * dot(e, tg);
            result += thr * dudn * dot(e, nrm);
            active = false;
        } else {
            float R = d_dir;
            if (Nm.present) R = fminf_(R, star_radius2_rho(Nm, ip, e, G.rho_max));
            if (ob) {
                if (R < P.min_R) inflate = true;
                R = fmaxf_(R, P.min_R);
            }
            float alpha = ob ? 2.0f : 1.0f;
            if (Nm.present) {
                if (!G.ignore_boundary) {
                    BSample2 b = sample_boundary_bvh2(Nm, ip, R, rng);
                    if (b.hit) {  // handle_boundary (wost_grad.py:28-51)
                        float Gv = clampf(greens2_G(ip, R, ip, b.p), -100.0f, 100.0f);
                        float res = 0.0f;
                        if (norm(sub(ip, b.p)) <= R) {
                            V2 tg = v2(-b.n.y, b.n.x);
                            float e_t = dot(e, tg), e_n = dot(e, b.n);
                            float f_tn = grad_dhdt2(bc, b.p, b.n);
                            if (G.plant == 3) f_tn = -f_tn;
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
                        if (norm(sub(sr.p, ip)) <= R) {
                            float h1 = grad_h2(bc, sr.p, sr.n1), h2 = grad_h2(bc, sr.p, sr.n2);
                            res -= h1 * dot(e, sr.t1);
                            res -= h2 * dot(e, sr.t2);
                        }
                        result += thr * alpha * (-Gv * res / sr.pdf);
                    }
                }
            }
            V2 dir = uniform_dir2(rng);
            if (ob && (dot(dir, in) > 0.0f)) dir = neg(dir);
            V2 _p = ip;
            V2 o = ob ? sub(ip, mul(in, 1e-5f)) : ip;  // Polyline.intersect (polyline.py:435-466)
            Its its = bvh2_intersect(F.tree, o, dir, R, nullptr);
            if (!its.valid) {
                V2 q = add(o, mul(dir, R));
                its.p = v3(q.x, q.y, 0.0f);
                its.n = v3(dir.x, dir.y, 0.0f);
                its.d = R;
                its.on_boundary = false;
            }
            ip = v2(its.p.x, its.p.y);
            in = v2(its.n.x, its.n.y);
            ob = its.on_boundary;
            if (!G.ignore_source) {  // handle_source (wost_grad.py:74-89)
                V2 sp;
                float sr_, spdf;
                greens2_sample(_p, R, dir, rng, sp, sr_, spdf);
                float res = 0.0f;
                V3 df = src_grad(bc);
                if (norm(sub(sp, _p)) < std::fabs(its.d)) res += (R * R / 4.0f) * dot(v2(df.x, df.y), e);
                result += thr * res;
            }
            if (ob) {
                V2 tg = v2(-in.y, in.x);
                float e_t = dot(e, tg), e_n = dot(e, in);
                float K = dot(dir, tg) / dot(dir, in);
                if (G.plant == 1) K = -K;
                result += thr * e_n * grad_h2(bc, ip, in);
                thr *= e_t - K * e_n;
                if (G.plant != 2) e = tg;
            }
            if (std::fabs(thr) < 1.0f) {  // Russian roulette
                if (rng.next_float32() > std::fabs(thr))
                    active = false;
                else
                    thr /= std::fabs(thr);
            }
        }
        i += 1;
        if (bad(result) || std::fabs(result) > 1000.0f) result = 0.0f;
    }
    steps += i;
    return result;
}

// ============================================================================
// WoStGrad3D (diff_wost/solvers/wost_grad_3d.py): the robust derivative walk,
// 3D.  The reference calls a dudn it never defines once a walk reaches a
// Dirichlet point, so the estimator runs on all-Neumann scenes only.
// ============================================================================
// Hessian of the analytic field, rows (u_x., u_y., u_z.)
static inline void bc_hess3(const BC& bc, V3 p, V3 H[3]) {
    if (bc.code == BC_EXPCOS) {
        float ex = cr_exp(p.x), c = cr_cos(p.y), sn = cr_sin(p.y);
        H[0] = v3(ex * c, -(ex * sn), 0.0f), H[1] = v3(-(ex * sn), -(ex * c), 0.0f), H[2] = v3(0.0f, 0.0f, 0.0f);
    } else if (bc.code == BC_QUAD) {
        H[0] = v3(2.0f, 0.0f, 0.0f), H[1] = v3(0.0f, -2.0f, 0.0f), H[2] = v3(0.0f, 0.0f, 0.0f);
    } else {
        H[0] = H[1] = H[2] = v3(0.0f, 0.0f, 0.0f);
    }
}

// DiscreteDistribution.sample_reuse_pmf (distr.py:52-107) over (cdf, pmf)
static inline int dist_sample(const std::vector<float>& cdfv, const std::vector<float>& pmfv, float sum, float nrm,
                              float value, float& reuse, float& pmf) {
    float sample = value * sum;
    int n = (int)cdfv.size();
    uint32_t start = 0, end = (uint32_t)(n - 1);
    int iters = start < end ? (32 - __builtin_clz(end - start)) : 0;
    for (int it = 0; it < iters; ++it) {
        uint32_t mid = (start + end) >> 1;
        float cv = cdfv[mid];
        bool cond = ((cv < sample) || (cv == 0.0f)) && (cv != sum);
        if (cond)
            start = std::min(mid + 1, end);
        else
            end = mid;
    }
    int idx = (int)start;
    pmf = pmfv[idx] * nrm;
    float c = idx > 0 ? cdfv[idx - 1] * nrm : 0.0f;
    reuse = (value - c) / pmf;
    return idx;
}

// SNCHNode3D.visit (snch3d.py:196-271)
static bool snch_visit3(const Node<3>& nd, V3 x, V3 v, float rho_max, float r_max, float& d_min) {
    float d_max;
    V3 mn = node_min3(nd), mx = node_max3(nd);
    box_dist(mn, mx, x, d_min, d_max);
    bool ov = d_min <= r_max;
    float ha = nd.half_angle;
    if ((ha >= 0.0f) && ov) {
        ov = (ha > kPi2) && (d_min < kDrEps);
        if (!ov) {
            V3 c = divs(add(mn, mx), 2.0f);
            V3 vca = sub(c, x);
            float l = norm(vca);
            vca = divs(vca, l);
            ov = l < nd.radius;
            if (!ov) {
                V3 axis = v3(nd.axis[0], nd.axis[1], nd.axis[2]);
                float axis_angle = cr_acos(clampf(dot(axis, vca), -1.0f, 1.0f));
                float vch = cr_asin(nd.radius / l);
                float hs = ha + vch;
                float amin = axis_angle - hs, amax = axis_angle + hs;
                ov = (hs > kPi2) || in_range(kPi2, amin, amax);
                if (!ov) {
                    float c0 = cr_cos(amin), c1 = cr_cos(amax);
                    float cmin_a = fminf_(c0, c1), cmax_a = fmaxf_(c0, c1);
                    if ((amin <= 0.0f) && (amax >= 0.0f)) cmax_a = fmaxf_(cmax_a, 1.0f);
                    float beta = cr_acos(clampf(dot(axis, v), -1.0f, 1.0f));
                    float bmin = beta - ha, bmax = beta + ha;
                    float b0 = cr_cos(bmin), b1 = cr_cos(bmax);
                    float cmin_b = fminf_(b0, b1), cmax_b = fmaxf_(b0, b1);
                    if ((bmin <= 0.0f) && (bmax >= 0.0f)) cmax_b = fmaxf_(cmax_b, 1.0f);
                    float gamma = cr_acos(clampf(dot(v, vca), -1.0f, 1.0f));
                    float gmin = gamma - vch, gmax = gamma + vch;
                    float g0 = cr_cos(gmin), g1 = cr_cos(gmax);
                    float cmin_g = fminf_(g0, g1), cmax_g = fmaxf_(g0, g1);
                    if ((gmin <= 0.0f) && (gmax >= 0.0f)) cmax_g = fmaxf_(cmax_g, 1.0f);
                    float a_min = cmin_b / cmax_a, a_max = cmax_b / cmin_a;
                    float b_min = cmin_g, b_max = cmax_g;
                    float rho2 = (1.0f - 2.0f * fminf_(fminf_(a_min * b_min, a_min * b_max),
                                                       fminf_(a_max * b_min, a_max * b_max))) +
                                 fmaxf_(a_min * a_min, a_max * a_max);
                    ov = rho2 > rho_max * rho_max;
                }
            }
        }
    }
    return ov;
}

// inside_triangle (math.py:311-331)
static bool inside_triangle3(V3 p, V3 a, V3 b, V3 c) {
    V3 v0 = sub(c, a), v1 = sub(b, a), v2 = sub(p, a);
    float d00 = dot(v0, v0), d01 = dot(v0, v1), d02 = dot(v0, v2), d11 = dot(v1, v1), d12 = dot(v1, v2);
    float denom = d00 * d11 - d01 * d01;
    if (denom == 0.0f) return false;
    float uu = (d11 * d02 - d01 * d12) / denom;
    float vv = (d00 * d12 - d01 * d02) / denom;
    return (uu >= 0.0f) && (vv >= 0.0f) && (uu + vv <= 1.0f);
}

// Triangle.star_radius_2 (triangle.py:109-151) with Plane.bounding_point
// (plane.py:11-34) and line_sphere_intersection (math.py:373-386)
static float tri_star_radius2(const Prim<3>& tr, V3 x, V3 e, float rho_max, float r_max) {
    V3 a = v3(tr.v[0][0], tr.v[0][1], tr.v[0][2]), b = v3(tr.v[1][0], tr.v[1][1], tr.v[1][2]);
    V3 cc = v3(tr.v[2][0], tr.v[2][1], tr.v[2][2]);
    float R = r_max;
    V3 n = normalize(cross(sub(b, a), sub(cc, a)));
    V3 c = v3(kInf, kInf, kInf), p = v3(kInf, kInf, kInf);
    float r = kInf;
    if (std::fabs(dot(e, n)) > kDrEps) {
        float d = dot(sub(a, x), n) / dot(e, n);
        c = add(x, mul(e, d));
        V3 xp = add(x, mul(n, dot(sub(a, x), n)));
        r = rho_max * std::fabs(d);
        if (norm(sub(xp, c)) > kDrEps) {
            p = add(c, mul(normalize(sub(xp, c)), r));
        } else if (norm(sub(n, v3(1.0f, 0.0f, 0.0f))) < 1e-4f) {
            p = add(c, mul(cross(n, v3(0.0f, 1.0f, 0.0f)), r));
        } else {
            p = add(c, mul(cross(n, v3(0.0f, 0.0f, 1.0f)), r));
        }
    }
    CP q;
    tri_closest(tr, x, q);
    bool inside_circle = (norm(sub(a, c)) < r) && (norm(sub(b, c)) < r) && (norm(sub(cc, c)) < r);
    if (inside_circle) {
    } else if (norm(sub(q.p, c)) > r) {
        R = fminf_(R, norm(sub(q.p, x)));
    } else if (inside_triangle3(p, a, b, cc)) {
        R = fminf_(R, norm(sub(p, x)));
    } else {
        V3 vs[3] = {a, b, cc};
        for (int k = 0; k < 3; ++k) {
            V3 s0 = vs[k], s1 = vs[(k + 1) % 3];
            V3 dd = sub(s1, s0), L = sub(s0, c);
            float qa = dot(dd, dd), qb = 2.0f * dot(dd, L), qc = dot(L, L) - r * r;
            float disc = qb * qb - 4.0f * qa * qc;
            if (disc > 0.0f) {
                float t0 = (-qb - std::sqrt(disc)) / (2.0f * qa);
                float t1 = (-qb + std::sqrt(disc)) / (2.0f * qa);
                if ((t0 >= 0.0f) && (t0 <= 1.0f)) R = fminf_(R, norm(sub(add(s0, mul(dd, t0)), x)));
                if ((t1 >= 0.0f) && (t1 <= 1.0f)) R = fminf_(R, norm(sub(add(s0, mul(dd, t1)), x)));
            }
        }
    }
    return R;
}

// Mesh.star_radius_2 (mesh.py:349-361): the star radius, then SNCH3D.star_radius_3
// (snch3d.py:742-811) over the triangles
static float star_radius3_rho(const Sub<3>& S, V3 x, V3 e, float rho_max) {
    float r_max = star_radius<3>(S, &x.x, kInf, nullptr);
    const Tree<3>& T = S.tree;
    float dmin, dmax;
    box_dist(node_min3(T.nodes[0]), node_max3(T.nodes[0]), x, dmin, dmax);
    if (!(dmin <= r_max)) return r_max;
    int st_i[64];
    float st_d[64];
    int sp = 0;
    st_i[0] = 0;
    st_d[0] = dmin;
    while (sp >= 0) {
        int ni = st_i[sp];
        float cd = st_d[sp];
        sp--;
        if (!(cd <= r_max)) continue;
        const Node<3>& nd = T.nodes[ni];
        if (nd.nref > 0) {
            for (int j = 0; j < nd.nref; ++j) {
                float r = tri_star_radius2(T.prims[nd.off + j], x, e, rho_max, r_max);
                if (r < r_max) r_max = r;
            }
        } else {
            int L = ni + 1, R = ni + nd.off;
            float d0, d1;
            bool h0 = snch_visit3(T.nodes[L], x, e, rho_max, r_max, d0);
            bool h1 = snch_visit3(T.nodes[R], x, e, rho_max, r_max, d1);
            if (h0 && h1) {
                int closer = L, other = R;
                if (d1 < d0) {
                    std::swap(closer, other);
                    std::swap(d0, d1);
                }
                st_i[++sp] = other;
                st_d[sp] = d1;
                st_i[++sp] = closer;
                st_d[sp] = d0;
            } else if (h0) {
                st_i[++sp] = L;
                st_d[sp] = d0;
            } else if (h1) {
                st_i[++sp] = R;
                st_d[sp] = d1;
            }
        }
    }
    return r_max;
}

// WoStGrad3D.dude (wost_grad_3d.py:93-206) on an all-Neumann scene
static float grad_dude3(const Scene& sc, const BC& bc, const Params& P, const GradParams& G, V3 x, V3 e, Pcg& rng,
                        int& steps) {
    const Sub<3>& F = sc.s3[0];
    const Sub<3>& Nm = sc.s3[2];
    V3 ip = x, in = v3(0.0f, 0.0f, 0.0f);
    bool ob = false;
    float result = 0.0f;
    bool active = inside3(sc, x);
    int i = 0;
    while ((i < P.nsteps) && active) {
        bool inflate = false;
        float R = kInf;
        if (Nm.present) {
            if (ob) ip = sub(ip, mul(in, 1e-3f));
            if (G.plant == 4) {  // tests only: the loop over all triangles (star_radius_2_baseline)
                float r0 = star_radius<3>(Nm, &ip.x, kInf, nullptr), m[3];
                for (int d = 0; d < 3; ++d)
                    m[d] = fmaxf_(std::fabs((&ip.x)[d] - Nm.bb_min[d]), std::fabs((&ip.x)[d] - Nm.bb_max[d]));
                float Rb = fminf_(norm(v3(m[0], m[1], m[2])), r0);
                for (size_t j = 0; j < Nm.tree.prims.size(); ++j)
                    Rb = fminf_(Rb, tri_star_radius2(Nm.tree.prims[j], ip, e, G.rho_max, Rb));
                R = fminf_(R, Rb);
            } else {
                R = fminf_(R, star_radius3_rho(Nm, ip, e, G.rho_max));
            }
            if (ob) {
                if (R < P.min_R) inflate = true;
                R = fmaxf_(R, P.min_R);
            }
        }
        float alpha = ob ? 2.0f : 1.0f;
        if (Nm.present) {
            {  // Mesh.sample_boundary (mesh.py:474-490) + handle_boundary (wost_grad_3d.py:22-45)
                float rnd, pmf;
                int idx = cdf_sample<3>(Nm, rng.next_float32(), rnd, pmf, nullptr);
                const float* A = &Nm.V[3 * Nm.I[3 * idx]];
                const float* B = &Nm.V[3 * Nm.I[3 * idx + 1]];
                const float* C = &Nm.V[3 * Nm.I[3 * idx + 2]];
                V3 a = v3(A[0], A[1], A[2]), b = v3(B[0], B[1], B[2]), c = v3(C[0], C[1], C[2]);
                V3 cr = cross(sub(b, a), sub(c, a));
                float area = norm(cr) / 2.0f;
                float r1 = rng.next_float32(), r2 = rng.next_float32();
                float su = std::sqrt(r1) * (1.0f - r2), sv = std::sqrt(r1) * r2;
                V3 bp = add(add(mul(b, su), mul(c, sv)), mul(a, (1.0f - su) - sv));
                V3 bn = normalize(cr);
                float bpdf = pmf / area;
                if (!inflate) {
                    float Gv = clampf(greens3_G(ip, R, ip, bp), -100.0f, 100.0f);
                    float res = 0.0f;
                    if (norm(sub(ip, bp)) <= R) {
                        float e_n = dot(e, bn);
                        V3 e_t = sub(e, mul(bn, e_n));
                        V3 H[3];
                        bc_hess3(bc, bp, H);
                        V3 He = v3(dot(H[0], e_t), dot(H[1], e_t), dot(H[2], e_t));
                        float f_tn = dot(He, bn);
                        res += f_tn - e_n * src_f(bc, bp);
                        if (P.cv) res -= dot(v3(0.0f, 0.0f, 0.0f), bn);
                    }
                    result += alpha * (-Gv * res / bpdf);
                }
            }
            if (!Nm.psil3.empty()) {  // Mesh.sample_silhouette (mesh.py:500-516) + handle_silhouette (:47-71)
                float rnd, pmf;
                int idx = dist_sample(Nm.ps_cdf, Nm.ps_pmf, Nm.ps_sum, Nm.ps_norm, rng.next_float32(), rnd, pmf);
                const float* q = &Nm.psil3[12 * idx];
                V3 a = v3(q[0], q[1], q[2]), b = v3(q[3], q[4], q[5]), c = v3(q[6], q[7], q[8]), d = v3(q[9], q[10], q[11]);
                float length = norm(sub(b, a));
                V3 n1 = normalize(cross(sub(b, a), sub(c, a)));
                V3 n2 = normalize(cross(sub(d, a), sub(b, a)));
                V3 t1 = normalize(cross(sub(b, a), n1));
                V3 t2 = normalize(cross(sub(a, b), n2));
                V3 sp = lerp(a, b, rng.next_float32());
                float spdf = pmf / length;
                if (!inflate) {
                    float Gv = clampf(greens3_G(ip, R, ip, sp), -100.0f, 100.0f);
                    float res = 0.0f;
                    if (norm(sub(sp, ip)) <= R) {
                        V3 gu = bc_grad(bc, sp);
                        float h1 = dot(gu, n1), h2 = dot(gu, n2);
                        float e1 = dot(e, t1), e2 = dot(e, t2);
                        if (P.cv) {
                            h1 = h1 - dot(v3(0.0f, 0.0f, 0.0f), n1);
                            h2 = h2 - dot(v3(0.0f, 0.0f, 0.0f), n2);
                        }
                        res -= h1 * e1;
                        res -= h2 * e2;
                    }
                    result += alpha * (-Gv * res / spdf);
                }
            }
        }
        V3 dir = uniform_sphere(rng);
        if (ob && (dot(dir, in) > 0.0f)) dir = neg(dir);
        V3 _p = ip;
        V3 o = ob ? sub(ip, mul(in, 1e-5f)) : ip;  // Mesh.intersect (mesh.py:384-415)
        Its its = snch3_intersect(F.tree, o, dir, R, nullptr);
        if (!its.valid) {
            its.p = add(o, mul(dir, R));
            its.n = dir;
            its.d = R;
            its.on_boundary = false;
        }
        ip = its.p;
        in = its.n;
        ob = its.on_boundary;
        result += dot(sub(_p, ip), v3(0.0f, 0.0f, 0.0f));
        if (!G.ignore_source) {  // handle_source (wost_grad_3d.py:73-90)
            V3 spt;
            float sr_, spdf;
            greens3_sample(_p, R, dir, rng, spt, sr_, spdf);
            float res = 0.0f;
            if (norm(sub(spt, _p)) < its.d) res += (R * R / 6.0f) * dot(src_grad(bc), e);
            result += res;
        }
        if (ob) {
            float e_n = dot(e, in);
            V3 e_t = sub(e, mul(in, e_n));
            float dir_n = dot(dir, in);
            V3 dir_t = sub(dir, mul(in, dir_n));
            V3 ref_t = sub(v3(0.0f, 0.0f, 0.0f), mul(in, dot(v3(0.0f, 0.0f, 0.0f), in)));
            result += dot(dir_t, ref_t) * e_n / dir_n;
            result += e_n * dot(in, bc_grad(bc, ip));
            e = sub(e_t, divs(mul(dir_t, e_n), dir_n));
        }
        if (norm(e) < 1.0f) {  // Russian roulette
            if (norm(e) < rng.next_float32())
                active = false;
            else
                e = divs(e, norm(e));
        }
        i += 1;
        if (bad(result) || std::fabs(result) > 1000.0f) result = 0.0f;
    }
    steps += i;
    return result;
}

}  // namespace orc

// ============================================================================
// C ABI for ctypes (tests / bench cpu_baseline only)
// ============================================================================
using namespace orc;

extern "C" {

void* orc_scene_create(int dim, const float* V, int64_t nv, const int32_t* I, const int32_t* types, int64_t np) {
    Scene* sc = new Scene();
    sc->dim = dim;
    sc->V.assign(V, V + nv * dim);
    sc->I.assign(I, I + np * dim);
    sc->types.assign(types, types + np);
    if (dim == 2) {
        make_sub<2>(*sc, -1, sc->s2[0]);
        make_sub<2>(*sc, 1, sc->s2[1]);
        make_sub<2>(*sc, 0, sc->s2[2]);
    } else {
        make_sub<3>(*sc, -1, sc->s3[0]);
        make_sub<3>(*sc, 1, sc->s3[1]);
        make_sub<3>(*sc, 0, sc->s3[2]);
    }
    return sc;
}
void orc_scene_destroy(void* s) { delete (Scene*)s; }

// sizes: [present, nverts, nprims, nnodes, nrefs(silhouette refs), max_depth, py_sils, nsil_topology]
void orc_sub_info(void* s, int sub, int64_t* out) {
    Scene* sc = (Scene*)s;
    if (sc->dim == 2) {
        Sub<2>& S = sc->s2[sub];
        out[0] = S.present;
        out[1] = S.V.size() / 2;
        out[2] = S.I.size() / 2;
        out[3] = S.tree.nodes.size();
        out[4] = S.refs.size();
        out[5] = S.tree.max_depth;
        out[6] = S.py_sils;
        out[7] = S.sil2.size();
    } else {
        Sub<3>& S = sc->s3[sub];
        out[0] = S.present;
        out[1] = S.V.size() / 3;
        out[2] = S.I.size() / 3;
        out[3] = S.tree.nodes.size();
        out[4] = S.refs.size();
        out[5] = S.tree.max_depth;
        out[6] = S.py_sils;
        out[7] = S.sil3.size();
    }
}

// node export: box[nn*2*dim] (min then max per node), ints[nn*4] (off, nref, sil_off, nsil),
// cone[nn*(dim+2)] (axis, half_angle, radius); prims_index[np] (sorted pIndex);
// refs[nrefs]
void orc_sub_export(void* s, int sub, float* box, int32_t* ints, float* cone, int32_t* prim_index, int32_t* refs) {
    Scene* sc = (Scene*)s;
    auto go = [&](auto& S, int D) {
        for (size_t i = 0; i < S.tree.nodes.size(); ++i) {
            auto& n = S.tree.nodes[i];
            for (int d = 0; d < D; ++d) {
                box[i * 2 * D + d] = n.box.mn[d];
                box[i * 2 * D + D + d] = n.box.mx[d];
                cone[i * (D + 2) + d] = n.axis[d];
            }
            cone[i * (D + 2) + D] = n.half_angle;
            cone[i * (D + 2) + D + 1] = n.radius;
            ints[i * 4 + 0] = n.off;
            ints[i * 4 + 1] = n.nref;
            ints[i * 4 + 2] = n.sil_off;
            ints[i * 4 + 3] = n.nsil;
        }
        for (size_t p = 0; p < S.tree.prims.size(); ++p) prim_index[p] = S.tree.prims[p].pIndex;
        for (size_t r = 0; r < S.refs.size(); ++r) refs[r] = S.refs[r];
    };
    if (sc->dim == 2)
        go(sc->s2[sub], 2);
    else
        go(sc->s3[sub], 3);
}

// Polyline.sdf / Mesh.sdf (polyline.py:188-194, mesh.py:290-293): out[n]
void orc_sdf(void* s, const float* pts, int64_t n, float* out) {
    Scene* sc = (Scene*)s;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i)
        out[i] = sc->dim == 2 ? sdf2(*sc, v2(pts[2 * i], pts[2 * i + 1]))
                              : sdf3(*sc, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}

// Polyline.sample_boundary (polyline.py:498-512) / Mesh.sample_boundary
// (mesh.py:474-490) with the sampler's draws given: u[n] (2D) or u[n][3] (3D).
// out: p[n*dim], nrm[n*dim], tuv[n*2], pdf[n], prim[n]
void orc_sample_boundary(void* s, int which, const float* u, int64_t n, float* p, float* nrm, float* tuv, float* pdf,
                         int32_t* prim) {
    Scene* sc = (Scene*)s;
    for (int64_t i = 0; i < n; ++i) {
        float rnd, pmf;
        if (sc->dim == 2) {
            const Sub<2>& S = sc->s2[which];
            int idx = cdf_sample<2>(S, u[i], rnd, pmf, nullptr);
            V2 a = v2(S.V[2 * S.I[2 * idx]], S.V[2 * S.I[2 * idx] + 1]);
            V2 b = v2(S.V[2 * S.I[2 * idx + 1]], S.V[2 * S.I[2 * idx + 1] + 1]);
            V2 tg = sub(b, a);
            V2 nn = normalize(v2(tg.y, -tg.x));
            V2 q = lerp(a, b, rnd);
            p[2 * i] = q.x, p[2 * i + 1] = q.y, nrm[2 * i] = nn.x, nrm[2 * i + 1] = nn.y;
            tuv[2 * i] = rnd, tuv[2 * i + 1] = 0.0f;
            pdf[i] = pmf / norm(tg);
            prim[i] = idx;
        } else {
            const Sub<3>& S = sc->s3[which];
            int idx = cdf_sample<3>(S, u[3 * i], rnd, pmf, nullptr);
            const float* A = &S.V[3 * S.I[3 * idx]];
            const float* B = &S.V[3 * S.I[3 * idx + 1]];
            const float* Cc = &S.V[3 * S.I[3 * idx + 2]];
            V3 a = v3(A[0], A[1], A[2]), b = v3(B[0], B[1], B[2]), c = v3(Cc[0], Cc[1], Cc[2]);
            V3 cr = cross(sub(b, a), sub(c, a));
            float area = norm(cr) / 2.0f;
            float r1 = u[3 * i + 1], r2 = u[3 * i + 2];
            float su = std::sqrt(r1) * (1.0f - r2);
            float sv = std::sqrt(r1) * r2;
            V3 q = add(add(mul(b, su), mul(c, sv)), mul(a, (1.0f - su) - sv));  // interpolate (math.py:306-307)
            V3 nn = normalize(cr);
            for (int k = 0; k < 3; ++k) p[3 * i + k] = (&q.x)[k], nrm[3 * i + k] = (&nn.x)[k];
            tuv[2 * i] = su, tuv[2 * i + 1] = sv;
            pdf[i] = pmf / area;
            prim[i] = idx;
        }
    }
}

// Far-field check (3D).  The walk kernels skip two traversals from points far
// outside the scene on the strength of distance bounds: lo / hi = the norms of
// the per-axis nearest / farthest offsets from x to the sub-scene's vertex box
// widened by 1e-3 (1 + max |coordinate|), for |x_i| <= 2^59 / (1 + max
// |coordinate|).  Claim: every
// distance the queries compute lies in [lo, hi].  For each point this runs the
// reference queries and reports: out[4i+0] = 1 if the bounds apply (2: and
// the Dirichlet box's lo == hi); +1 = the
// Dirichlet closest-point distance lies in [lo, hi] of the Dirichlet box (and
// equals lo when lo == hi); +2 = with r_max = star_bound(full, x, that distance)
// <= lo of the full box, the closest-silhouette search finds nothing (1), or
// the shortcut does not apply (2); +3 = every silhouette distance found lies in
// the full box's [lo, hi] (1) or no silhouette was found (2).
static bool far_bounds(const Sub<3>& S, V3 x, float& lo, float& hi) {
    float big = 0.0f;
    for (int k = 0; k < 3; ++k) big = fmaxf_(big, fmaxf_(std::fabs(S.bb_min[k]), std::fabs(S.bb_max[k])));
    const float lim = 576460752303423488.0f / (1.0f + big);  // 2^59 / (1 + max |coordinate|)
    if (!(std::fabs(x.x) <= lim && std::fabs(x.y) <= lim && std::fabs(x.z) <= lim)) return false;
    float m = 1e-3f * (1.0f + big);
    float xs[3] = {x.x, x.y, x.z}, l[3], h[3];
    for (int k = 0; k < 3; ++k) {
        float a = S.bb_min[k] - m, b = S.bb_max[k] + m;
        l[k] = xs[k] < a ? a - xs[k] : (xs[k] > b ? xs[k] - b : 0.0f);
        h[k] = fmaxf_(std::fabs(xs[k] - a), std::fabs(xs[k] - b));
    }
    lo = norm(v3(l[0], l[1], l[2]));
    hi = norm(v3(h[0], h[1], h[2]));
    return true;
}

void orc_far_check(void* s, const float* pts, int64_t n, int32_t* out) {
    Scene* sc = (Scene*)s;
    const Sub<3>& F = sc->s3[0];
    const Sub<3>& Dn = sc->s3[1];
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < n; ++i) {
        V3 x = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        int32_t* o = out + 4 * i;
        float lo, hi, flo, fhi;
        o[0] = far_bounds(Dn, x, lo, hi) && far_bounds(F, x, flo, fhi);
        o[1] = o[2] = o[3] = 0;
        if (!o[0]) continue;
        if (lo == hi) o[0] = 2;  // the Dirichlet bounds coincide
        CP cp = snch3_closest(Dn.tree, x, nullptr);
        float d = std::fabs(cp.d);
        o[1] = (d >= lo && d <= hi && (lo != hi || d == lo)) ? 1 : 0;
        float m3[3];
        for (int k = 0; k < 3; ++k) m3[k] = fmaxf_(std::fabs((&x.x)[k] - F.bb_min[k]), std::fabs((&x.x)[k] - F.bb_max[k]));
        float r_max = fminf_(norm(v3(m3[0], m3[1], m3[2])), d);  // star_radius's bound (mesh.py:337-346)
        SilRec r = snch_closest_sil<3>(F, &x.x, r_max, nullptr);
        if (r_max <= flo)
            o[2] = r.valid ? 0 : 1;
        else
            o[2] = 2;
        o[3] = r.valid ? ((r.d >= flo && r.d <= fhi) ? 1 : 0) : 2;
    }
}

// Queries.  kind: 0 BVH closest point, 1 ray intersect (BVH), 2 closest silhouette,
// 3 star radius, 10/11/12 brute force closest point / ray / silhouette.
// in: pts[n*dim], dirs[n*dim] (ray), rmax[n].  out: prim[n], d[n], p[n*dim], nrm[n*dim], tuv[n*2]
// The 2D boundary control variate's full-scene closest point (wost.py:123-126)
// queried with its pruning radius started at the Dirichlet distance of the same
// point, widened by relative 2^-16 plus 1e-6 of the coordinate scale (the root
// box), against the unbounded reference query: same[i] = 1 when the two records
// agree bit for bit; nodes[0..1] = nodes visited (unbounded, bounded).
// A walk step moves x_prev to x_new with |x_new - x_prev| <= R + 1e-5: the
// Dirichlet closest point at x_new queried with its pruning radius started at
// (d(x_prev) + R + 1e-5)(1 + 2^-16) + 1e-6 (1 + coordinate scale of the full
// scene), against the unbounded query.  same[n]: bitwise agreement of the
// records; nodes[2]: nodes visited unbounded, bounded.
void orc_step_bound_check(void* s, const float* prev, const float* next, const float* R, int64_t n, int32_t* same,
                          uint64_t* nodes) {
    Scene* sc = (Scene*)s;
    const int D = sc->dim;
    float scale = 0.0f;
    for (int k = 0; k < D; ++k) {
        float lo = D == 2 ? sc->s2[0].tree.nodes[0].box.mn[k] : sc->s3[0].tree.nodes[0].box.mn[k];
        float hi = D == 2 ? sc->s2[0].tree.nodes[0].box.mx[k] : sc->s3[0].tree.nodes[0].box.mx[k];
        scale = fmaxf_(scale, fmaxf_(std::fabs(lo), std::fabs(hi)));
    }
    uint64_t n0 = 0, n1 = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : n0, n1)
    for (int64_t i = 0; i < n; ++i) {
        Cnt ca, cb;
        CP a, b;
        float dp;
        if (D == 2) {
            const Tree<2>& T = sc->s2[1].tree;
            dp = std::fabs(bvh2_closest(T, v2(prev[2 * i], prev[2 * i + 1]), nullptr).d);
            float r = dp + R[i] + 1e-5f;
            float r0 = r < kInf ? r * (1.0f + 0x1p-16f) + 1e-6f * (1.0f + scale) : kInf;
            V2 x = v2(next[2 * i], next[2 * i + 1]);
            a = bvh2_closest(T, x, &ca);
            b = bvh2_closest(T, x, &cb, r0);
        } else {
            const Tree<3>& T = sc->s3[1].tree;
            dp = std::fabs(snch3_closest(T, v3(prev[3 * i], prev[3 * i + 1], prev[3 * i + 2]), nullptr).d);
            float r = dp + R[i] + 1e-5f;
            float r0 = r < kInf ? r * (1.0f + 0x1p-16f) + 1e-6f * (1.0f + scale) : kInf;
            V3 x = v3(next[3 * i], next[3 * i + 1], next[3 * i + 2]);
            a = snch3_closest(T, x, &ca);
            b = snch3_closest(T, x, &cb, r0);
        }
        bool eq = a.prim == b.prim && std::memcmp(&a.d, &b.d, 4) == 0 && std::memcmp(&a.t, &b.t, 4) == 0 &&
                  std::memcmp(&a.u, &b.u, 4) == 0 && std::memcmp(&a.v, &b.v, 4) == 0 &&
                  std::memcmp(&a.p, &b.p, sizeof(a.p)) == 0 && std::memcmp(&a.n, &b.n, sizeof(a.n)) == 0;
        same[i] = eq ? 1 : 0;
        n0 += ca.nodes;
        n1 += cb.nodes;
    }
    nodes[0] = n0;
    nodes[1] = n1;
}

void orc_cv_bound_check(void* s, const float* pts, int64_t n, int32_t* same, uint64_t* nodes) {
    Scene* sc = (Scene*)s;
    const Tree<2>& F = sc->s2[0].tree;
    const Node<2>& root = F.nodes[0];
    V2 mn = node_min2(root), mx = node_max2(root);
    float scale = fmaxf_(fmaxf_(std::fabs(mn.x), std::fabs(mx.x)), fmaxf_(std::fabs(mn.y), std::fabs(mx.y)));
    uint64_t n0 = 0, n1 = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : n0, n1)
    for (int64_t i = 0; i < n; ++i) {
        V2 p = v2(pts[2 * i], pts[2 * i + 1]);
        float dd = kInf;
        if (sc->s2[1].present) dd = std::fabs(bvh2_closest(sc->s2[1].tree, p, nullptr).d);
        float r0 = dd < kInf ? dd * (1.0f + 0x1p-16f) + 1e-6f * (1.0f + scale) : kInf;
        Cnt ca, cb;
        CP a = bvh2_closest(F, p, &ca);
        CP b = bvh2_closest(F, p, &cb, r0);
        bool eq = a.prim == b.prim && std::memcmp(&a.d, &b.d, 4) == 0 && std::memcmp(&a.t, &b.t, 4) == 0 &&
                  std::memcmp(&a.p, &b.p, sizeof(a.p)) == 0 && std::memcmp(&a.n, &b.n, sizeof(a.n)) == 0;
        same[i] = eq ? 1 : 0;
        n0 += ca.nodes;
        n1 += cb.nodes;
    }
    nodes[0] = n0;
    nodes[1] = n1;
}

// Per closest-point query: the GPU loop's iterations and interior-record
// fetches at node width 2 and 4, out[n][4] = {iters2, iters4, fetch2, fetch4}.
void orc_closest_iters(void* s, int sub, const float* pts, int64_t n, int64_t* out) {
    Scene* sc = (Scene*)s;
    int D = sc->dim;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < n; ++q) {
        const float* x = pts + q * D;
        Cnt c;
        if (D == 2)
            bvh2_closest(sc->s2[sub].tree, v2(x[0], x[1]), &c);
        else
            snch3_closest(sc->s3[sub].tree, v3(x[0], x[1], x[2]), &c);
        out[q * 4] = (int64_t)c.it2, out[q * 4 + 1] = (int64_t)c.it4;
        out[q * 4 + 2] = (int64_t)c.rf2, out[q * 4 + 3] = (int64_t)c.rf4;
    }
}

void orc_query(void* s, int sub, int kind, const float* pts, const float* dirs, const float* rmax, int64_t n,
               int32_t* prim, float* d, float* p, float* nrm, float* tuv) {
    Scene* sc = (Scene*)s;
    int D = sc->dim;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < n; ++q) {
        const float* x = pts + q * D;
        int pr = -1;
        float dd = 0, pp[3] = {0, 0, 0}, nn[3] = {0, 0, 0}, tt[2] = {0, 0};
        if (D == 2) {
            const Sub<2>& S = sc->s2[sub];
            V2 xv = v2(x[0], x[1]);
            if (kind == 0) {
                CP c = bvh2_closest(S.tree, xv, nullptr);
                pr = c.prim, dd = c.d, pp[0] = c.p.x, pp[1] = c.p.y, nn[0] = c.n.x, nn[1] = c.n.y, tt[0] = c.t;
            } else if (kind == 1) {
                V2 v = v2(dirs[q * 2], dirs[q * 2 + 1]);
                Its it = bvh2_intersect(S.tree, xv, v, rmax[q], nullptr);
                pr = it.valid ? it.prim : -1;
                dd = it.d, pp[0] = it.p.x, pp[1] = it.p.y, nn[0] = it.n.x, nn[1] = it.n.y, tt[0] = it.t;
            } else if (kind == 2) {
                SilRec r = snch_closest_sil<2>(S, x, rmax[q], nullptr);
                pr = r.valid ? r.prim : -1, dd = r.d, pp[0] = r.p.x, pp[1] = r.p.y;
            } else if (kind == 3) {
                dd = star_radius<2>(S, x, rmax[q], nullptr);
            } else if (kind == 10) {
                CP c = brute_closest2(S, xv);
                pr = c.prim, dd = c.d;
            } else if (kind == 11) {
                V2 v = v2(dirs[q * 2], dirs[q * 2 + 1]);
                Its it = brute_ray2(S, xv, v, rmax[q]);
                pr = it.valid ? it.prim : -1, dd = it.d;
            } else if (kind == 12) {
                SilRec r = brute_sil<2>(S, x, rmax[q]);
                pr = r.valid ? S.refs.empty() ? r.prim : r.prim : -1, dd = r.d;
            }
        } else {
            const Sub<3>& S = sc->s3[sub];
            V3 xv = v3(x[0], x[1], x[2]);
            if (kind == 0) {
                CP c = snch3_closest(S.tree, xv, nullptr);
                pr = c.prim, dd = c.d;
                pp[0] = c.p.x, pp[1] = c.p.y, pp[2] = c.p.z, nn[0] = c.n.x, nn[1] = c.n.y, nn[2] = c.n.z;
                tt[0] = c.u, tt[1] = c.v;
            } else if (kind == 1) {
                V3 v = v3(dirs[q * 3], dirs[q * 3 + 1], dirs[q * 3 + 2]);
                Its it = snch3_intersect(S.tree, xv, v, rmax[q], nullptr);
                pr = it.valid ? it.prim : -1, dd = it.d;
                pp[0] = it.p.x, pp[1] = it.p.y, pp[2] = it.p.z, nn[0] = it.n.x, nn[1] = it.n.y, nn[2] = it.n.z;
                tt[0] = it.u, tt[1] = it.v;
            } else if (kind == 2) {
                SilRec r = snch_closest_sil<3>(S, x, rmax[q], nullptr);
                pr = r.valid ? r.prim : -1, dd = r.d, pp[0] = r.p.x, pp[1] = r.p.y, pp[2] = r.p.z;
            } else if (kind == 3) {
                dd = star_radius<3>(S, x, rmax[q], nullptr);
            } else if (kind == 10) {
                CP c = brute_closest3(S, xv);
                pr = c.prim, dd = c.d;
            } else if (kind == 11) {
                V3 v = v3(dirs[q * 3], dirs[q * 3 + 1], dirs[q * 3 + 2]);
                Its it = brute_ray3(S, xv, v, rmax[q]);
                pr = it.valid ? it.prim : -1, dd = it.d;
            } else if (kind == 12) {
                SilRec r = brute_sil<3>(S, x, rmax[q]);
                pr = r.valid ? r.prim : -1, dd = r.d;
            }
        }
        if (prim) prim[q] = pr;
        if (d) d[q] = dd;
        for (int k = 0; k < D; ++k) {
            if (p) p[q * D + k] = pp[k];
            if (nrm) nrm[q * D + k] = nn[k];
        }
        if (tuv) tuv[q * 2] = tt[0], tuv[q * 2 + 1] = tt[1];
    }
}

// topology silhouette -> index used by brute force; returns the topology id of a ref
int32_t orc_ref_topology(void* s, int sub, int32_t ref) {
    Scene* sc = (Scene*)s;
    return sc->dim == 2 ? sc->s2[sub].refs[ref] : sc->s3[sub].refs[ref];
}

// src: the source's four coefficients (SRC_CONST: src[0]; SRC_LINEAR: all four)
static BC make_bc(int code, const float* c, const float* gv, const float* hv, int src_code, const float* src) {
    BC bc;
    bc.code = code;
    for (int k = 0; k < 4; ++k) bc.c[k] = c ? c[k] : 0.0f;
    bc.gv = gv;
    bc.hv = hv;
    bc.src_code = src_code;
    bc.src_c = src ? src[0] : 0.0f;
    for (int k = 0; k < 4; ++k) bc.sc[k] = src ? src[k] : 0.0f;
    return bc;
}
static Params make_params(const int32_t* ip, const float* fp) {
    Params P;
    P.nwalks = ip[0];
    P.nsteps = ip[1];
    P.nsubwalks = ip[2];
    P.cv = ip[3];
    P.av = ip[4];
    P.seed = (uint32_t)ip[5];
    P.base = ip[6];
    P.wos = ip[7];
    P.eps = fp[0];
    P.min_R = fp[1];
    return P;
}

// Solver.solve: per-walk values/steps (n*W) and per-point means (Kahan, double)
// ip = [nwalks, nsteps, nsubwalks, cv, av, seed], fp = [eps, min_R]
void orc_solve_u(void* s, const float* pts, int64_t n, const int32_t* ip, const float* fp, int bc_code,
                 const float* bc_c, const float* gv, const float* hv, int src_code, const float* src_c, float* mean,
                 float* walk_val, int32_t* walk_steps, int32_t* walk_term, uint64_t* bytes_total) {
    Scene* sc = (Scene*)s;
    Params P = make_params(ip, fp);
    BC bc = make_bc(bc_code, bc_c, gv, hv, src_code, src_c);
    int D = sc->dim;
    int64_t W = P.nwalks;
    uint64_t bytes = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bytes)
    for (int64_t i = 0; i < n; ++i) {
        uint32_t a, b;
        tea(P.seed, (uint32_t)(P.base + i), 4, kTeaKey, &a, &b);
        double acc = 0.0;
        for (int64_t k = 0; k < W; ++k) {
            uint64_t lane = (uint64_t)((P.base + i) * W + k);
            Pcg rng = pcg_for_lane(a, b, lane);
            WalkOut w = walk_u(*sc, bc, P, pts + i * D, rng, bytes_total != nullptr);
            acc += (double)w.value;
            bytes += w.cnt.bytes;
            int64_t j = i * W + k;  // output slot (the stream lane is global)
            if (walk_val) walk_val[j] = w.value;
            if (walk_steps) walk_steps[j] = w.steps;
            if (walk_term) walk_term[j] = w.term_prim;
        }
        if (mean) mean[i] = (float)(acc / (double)W);
    }
    if (bytes_total) *bytes_total = bytes;
}

// WoSt.solve_dude (wost.py:21-28): one stream per point
void orc_solve_dude(void* s, const float* pts, const float* e, int64_t n, const int32_t* ip, const float* fp,
                    int bc_code, const float* bc_c, const float* gv, const float* hv, int src_code, const float* src_c,
                    float* out, uint64_t* steps_total) {
    Scene* sc = (Scene*)s;
    Params P = make_params(ip, fp);
    BC bc = make_bc(bc_code, bc_c, gv, hv, src_code, src_c);
    int D = sc->dim;
    uint64_t steps = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : steps)
    for (int64_t i = 0; i < n; ++i) {
        Pcg rng = pcg_for(P.seed, (uint32_t)(P.base + i), (uint64_t)(P.base + i));
        uint64_t st = 0;
        if (P.wos)
            out[i] = D == 2 ? wos_dude<2>(*sc, bc, P, pts + 2 * i, e + 2 * i, rng, &st)
                            : wos_dude<3>(*sc, bc, P, pts + 3 * i, e + 3 * i, rng, &st);
        else if (D == 2)
            out[i] = dude2(*sc, bc, P, v2(pts[2 * i], pts[2 * i + 1]), v2(e[2 * i], e[2 * i + 1]), rng, &st);
        else
            out[i] = dude3(*sc, bc, P, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]),
                           v3(e[3 * i], e[3 * i + 1], e[3 * i + 2]), rng, &st);
        steps += st;
    }
    if (steps_total) *steps_total = steps;
}

// StreamlineTracer.next_position (examples/example_2d/trace.py:19-68,
// examples/example_3d/trace.py:21-60) with WoSt / WoSt3D: sample j of point i
// runs on the lane-(base+i)*tn+j stream of TEA(seed, base+i); each gradient
// component is Solver.walk_dude (solver.py:148-166): the mean of nsubwalks
// dude estimates, the axes one after the other on the sample's stream.  2D:
// _next_position's two should_increment tests (sdf <= 0); the corrector term
// fy_tilde = current_p - y_tilde is zero as written there, so a step that stays
// inside keeps half its length.  3D: one Euler step.  The samples of a point
// are averaged in sample order (the reference's scatter_reduce leaves the order
// open).  samples[n*tn*dim] optional.
void orc_trace_step(void* s, const float* pts, int64_t n, const int32_t* ip, const float* fp, int bc_code,
                    const float* bc_c, const float* gv, const float* hv, int src_code, const float* src_c, int32_t tn,
                    float step, float* out, float* samples, uint64_t* steps_total) {
    Scene* sc = (Scene*)s;
    Params P = make_params(ip, fp);
    BC bc = make_bc(bc_code, bc_c, gv, hv, src_code, src_c);
    const int D = sc->dim;
    std::vector<float> smp((size_t)n * tn * D);
    uint64_t steps = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : steps)
    for (int64_t j = 0; j < n * (int64_t)tn; ++j) {
        int64_t i = j / tn;
        Pcg rng = pcg_for(P.seed, (uint32_t)(P.base + i), (uint64_t)((P.base + i) * tn + (j - i * tn)));
        uint64_t st = 0;
        float* o = &smp[(size_t)j * D];
        if (D == 2) {
            V2 p = v2(pts[2 * i], pts[2 * i + 1]);
            auto walk_dude = [&](V2 e) {
                float r = 0.0f;
                for (int k = 0; k < P.nsubwalks; ++k) r += dude2(*sc, bc, P, p, e, rng, &st);
                return r / (float)P.nsubwalks;
            };
            V2 y = p, cur = p;
            if (!(sdf2(*sc, cur) > 0.0f)) {
                float gx = walk_dude(v2(1.0f, 0.0f));
                float gy = walk_dude(v2(0.0f, 1.0f));
                cur = add(cur, mul(v2(gx, gy), step));
                V2 yt = cur;
                if (!(sdf2(*sc, cur) > 0.0f)) {
                    V2 fy = sub(yt, y), fyt = sub(cur, yt);
                    cur = add(y, mul(add(fy, fyt), 0.5f));
                }
            }
            o[0] = cur.x, o[1] = cur.y;
        } else {
            V3 p = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
            auto walk_dude = [&](V3 e) {
                float r = 0.0f;
                for (int k = 0; k < P.nsubwalks; ++k) r += dude3(*sc, bc, P, p, e, rng, &st);
                return r / (float)P.nsubwalks;
            };
            float gx = walk_dude(v3(1.0f, 0.0f, 0.0f));
            float gy = walk_dude(v3(0.0f, 1.0f, 0.0f));
            float gz = walk_dude(v3(0.0f, 0.0f, 1.0f));
            V3 cur = add(p, mul(v3(gx, gy, gz), step));
            o[0] = cur.x, o[1] = cur.y, o[2] = cur.z;
        }
        steps += st;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < D; ++d) {
            float acc = 0.0f;
            for (int k = 0; k < tn; ++k) acc += smp[((size_t)i * tn + k) * D + d];
            out[i * D + d] = tn > 1 ? acc / (float)tn : acc;
        }
    if (samples) std::memcpy(samples, smp.data(), smp.size() * sizeof(float));
    if (steps_total) *steps_total = steps;
}

// WoStGrad.solve_dude through Solver.solve_dude (solver.py:113-144) and
// walk_dude (:146-166): lane i * nwalks + k runs nsubwalks dudes on PCG32 lane
// (base + i) * nwalks + k seeded with TEA(seed, base + i).  gp = {rho_max};
// gi = {ignore_boundary, ignore_silhouette, ignore_source}.  out[n] = mean over
// the walks; walk_val[n*nwalks] and walk_steps (outer steps + nested WoSt
// steps) optional.  gi[3] plants a mistake (tests only; 0: the reference).
// 2D scenes only.
void orc_grad_dude(void* s, const float* pts, const float* ev, int64_t n, const int32_t* ip, const float* fp,
                   int bc_code, const float* bc_c, int src_code, const float* src_c, const float* gp, const int32_t* gi,
                   float* out, float* walk_val, int32_t* walk_steps, uint64_t* steps_total) {
    Scene* sc = (Scene*)s;
    Params P = make_params(ip, fp);
    BC bc = make_bc(bc_code, bc_c, nullptr, nullptr, src_code, src_c);
    GradParams G{gp[0], gi[0], gi[1], gi[2], gi[3]};
    int64_t W = P.nwalks;
    uint64_t steps = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : steps)
    for (int64_t i = 0; i < n; ++i) {
        uint32_t a, b;
        tea(P.seed, (uint32_t)(P.base + i), 4, kTeaKey, &a, &b);
        const int D = sc->dim;
        double acc = 0.0;
        for (int64_t k = 0; k < W; ++k) {
            Pcg rng = pcg_for_lane(a, b, (uint64_t)((P.base + i) * W + k));
            float v = 0.0f;
            int st = 0;
            for (int sw = 0; sw < P.nsubwalks; ++sw) {
                if (D == 2)
                    v += grad_dude2(*sc, bc, P, G, v2(pts[2 * i], pts[2 * i + 1]), v2(ev[2 * i], ev[2 * i + 1]), rng,
                                    st);
                else
                    v += grad_dude3(*sc, bc, P, G, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]),
                                    v3(ev[3 * i], ev[3 * i + 1], ev[3 * i + 2]), rng, st);
            }
            v = v / (float)P.nsubwalks;
            acc += (double)v;
            steps += (uint64_t)st;
            if (walk_val) walk_val[i * W + k] = v;
            if (walk_steps) walk_steps[i * W + k] = st;
        }
        if (out) out[i] = (float)(acc / (double)W);
    }
    if (steps_total) *steps_total = steps;
}

// StreamlineTracer.next_position (examples/example_2d/trace.py:19-68) with a
// WoStGrad solver, as in the reference example (run.py): sample j of point i
// runs on PCG32 lane (base + i) * tn + j; walk_dude = the mean of nsubwalks
// WoStGrad dudes.  out[n*2] = mean of the tn samples; samples[n*tn*2] optional.
void orc_grad_trace_step(void* s, const float* pts, int64_t n, const int32_t* ip, const float* fp, int bc_code,
                         const float* bc_c, int src_code, const float* src_c, const float* gp, const int32_t* gi, int32_t tn,
                         float step, float* out, float* samples, uint64_t* steps_total) {
    Scene* sc = (Scene*)s;
    Params P = make_params(ip, fp);
    BC bc = make_bc(bc_code, bc_c, nullptr, nullptr, src_code, src_c);
    GradParams G{gp[0], gi[0], gi[1], gi[2], gi[3]};
    std::vector<float> smp((size_t)n * tn * 2);
    uint64_t steps = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : steps)
    for (int64_t j = 0; j < n * (int64_t)tn; ++j) {
        int64_t i = j / tn;
        Pcg rng = pcg_for(P.seed, (uint32_t)(P.base + i), (uint64_t)((P.base + i) * tn + (j - i * tn)));
        int st = 0;
        V2 p = v2(pts[2 * i], pts[2 * i + 1]);
        auto walk_dude = [&](V2 e) {
            float r = 0.0f;
            for (int k = 0; k < P.nsubwalks; ++k) r += grad_dude2(*sc, bc, P, G, p, e, rng, st);
            return r / (float)P.nsubwalks;
        };
        V2 y = p, cur = p;
        if (!(sdf2(*sc, cur) > 0.0f)) {
            float gx = walk_dude(v2(1.0f, 0.0f));
            float gy = walk_dude(v2(0.0f, 1.0f));
            cur = add(cur, mul(v2(gx, gy), step));
            V2 yt = cur;
            if (!(sdf2(*sc, cur) > 0.0f)) {
                V2 fy = sub(yt, y), fyt = sub(cur, yt);
                cur = add(y, mul(add(fy, fyt), 0.5f));
            }
        }
        smp[(size_t)j * 2] = cur.x, smp[(size_t)j * 2 + 1] = cur.y;
        steps += (uint64_t)st;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            float acc = 0.0f;
            for (int k = 0; k < tn; ++k) acc += smp[((size_t)i * tn + k) * 2 + d];
            out[i * 2 + d] = tn > 1 ? acc / (float)tn : acc;
        }
    if (samples) std::memcpy(samples, smp.data(), smp.size() * sizeof(float));
    if (steps_total) *steps_total = steps;
}

// The rho-bounded star radius of a 2D sub-scene two ways: through the SNCH
// (Polyline.star_radius_2, polyline.py:317-330) and by the reference's loop over
// every segment (star_radius_2_3, polyline.py:332-375).  tree[n], brute[n].
void orc_star_radius2_rho(void* s, int sub, const float* pts, const float* ev, int64_t n, float rho_max, float* tree,
                          float* brute) {
    Scene* sc = (Scene*)s;
    if (sc->dim == 3) {  // Mesh.star_radius_2 (mesh.py:349-361) vs star_radius_2_baseline (:363-380)
        const Sub<3>& S = sc->s3[sub];
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; ++i) {
            V3 x = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), e = v3(ev[3 * i], ev[3 * i + 1], ev[3 * i + 2]);
            tree[i] = star_radius3_rho(S, x, e, rho_max);
            float r0 = star_radius<3>(S, &x.x, kInf, nullptr);
            float m[3];
            for (int d = 0; d < 3; ++d) m[d] = fmaxf_(std::fabs((&x.x)[d] - S.bb_min[d]), std::fabs((&x.x)[d] - S.bb_max[d]));
            float R = fminf_(norm(v3(m[0], m[1], m[2])), r0);
            for (size_t j = 0; j < S.tree.prims.size(); ++j) R = fminf_(R, tri_star_radius2(S.tree.prims[j], x, e, rho_max, R));
            brute[i] = R;
        }
        return;
    }
    const Sub<2>& S = sc->s2[sub];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        V2 x = v2(pts[2 * i], pts[2 * i + 1]), e = v2(ev[2 * i], ev[2 * i + 1]);
        tree[i] = star_radius2_rho(S, x, e, rho_max);
        float m0 = fmaxf_(std::fabs(x.x - S.bb_min[0]), std::fabs(x.x - S.bb_max[0]));
        float m1 = fmaxf_(std::fabs(x.y - S.bb_min[1]), std::fabs(x.y - S.bb_max[1]));
        float R = norm(v2(m0, m1));
        for (size_t j = 0; j < S.I.size() / 2; ++j) {
            V2 a = v2(S.V[2 * S.I[2 * j]], S.V[2 * S.I[2 * j] + 1]);
            V2 b = v2(S.V[2 * S.I[2 * j + 1]], S.V[2 * S.I[2 * j + 1] + 1]);
            R = seg_star_radius2(a, b, x, e, rho_max, R);
        }
        brute[i] = R;
    }
}

// Fused u + grad u.  pair_u[n*np], pair_g[n*np*dim], pair_steps[n*np] optional.
void orc_solve_fused(void* s, const float* pts, int64_t n, const int32_t* ip, const float* fp, int bc_code,
                     const float* bc_c, const float* gv, const float* hv, int src_code, const float* src_c, float* u,
                     float* grad, float* pair_u, float* pair_g, int32_t* pair_steps, uint64_t* steps_total,
                     uint64_t* counts) {
    Scene* sc = (Scene*)s;
    Params P = make_params(ip, fp);
    BC bc = make_bc(bc_code, bc_c, gv, hv, src_code, src_c);
    int D = sc->dim;
    int64_t NP = (P.nwalks + 1) / 2;
    uint64_t steps = 0, cb = 0, cf = 0, cn = 0, cp_ = 0, cs = 0, cff = 0, ccq = 0, ci2 = 0, ci4 = 0, cr2 = 0, cr4 = 0;
#pragma omp parallel for schedule(dynamic, 1) \
    reduction(+ : steps, cb, cf, cn, cp_, cs, cff, ccq, ci2, ci4, cr2, cr4)
    for (int64_t i = 0; i < n; ++i) {
        const float* x = pts + i * D;
        float R;
        bool ins;
        if (D == 2) {
            R = std::fabs(bvh2_closest(sc->s2[0].tree, v2(x[0], x[1]), nullptr).d);
            ins = inside2(*sc, v2(x[0], x[1]));
        } else {
            R = std::fabs(snch3_closest(sc->s3[0].tree, v3(x[0], x[1], x[2]), nullptr).d);
            ins = inside3(*sc, v3(x[0], x[1], x[2]));
        }
        uint32_t a, b;
        tea(P.seed, (uint32_t)(P.base + i), 4, kTeaKey, &a, &b);
        double su = 0.0, sg[3] = {0, 0, 0};
        for (int64_t k = 0; k < NP; ++k) {
            uint64_t lane = (uint64_t)((P.base + i) * NP + k);
            Pcg rng = pcg_for_lane(a, b, lane);
            Cnt k_;
            PairOut o = fused_pair(*sc, bc, P, x, R, ins, rng, counts ? &k_ : nullptr);
            cb += k_.bytes, cf += k_.flops, cn += k_.nodes, cp_ += k_.prims, cs += k_.sils, cff += k_.ff;
            ccq += k_.cq, ci2 += k_.it2, ci4 += k_.it4, cr2 += k_.rf2, cr4 += k_.rf4;
            su += o.usum;
            for (int q = 0; q < D; ++q) sg[q] += o.g[q];
            steps += (uint64_t)o.steps;
            int64_t j = i * NP + k;  // output slot (the stream lane is global)
            if (pair_u) pair_u[j] = o.usum;
            if (pair_g)
                for (int q = 0; q < D; ++q) pair_g[j * D + q] = o.g[q];
            if (pair_steps) pair_steps[j] = o.steps;
        }
        double nw = (double)(2 * NP);
        if (u) u[i] = (float)(su / nw);
        if (grad)
            for (int q = 0; q < D; ++q) grad[i * D + q] = (float)(sg[q] / nw);
    }
    if (steps_total) *steps_total = steps;
    // counts[11] = {bytes, flops, nodes, primitives, silhouettes, NaN-position steps,
    // closest-point queries, their GPU loop iterations at node width 2 and 4,
    // their interior-record fetches at node width 2 and 4}
    if (counts) {
        counts[0] = cb, counts[1] = cf, counts[2] = cn, counts[3] = cp_, counts[4] = cs, counts[5] = cff;
        counts[6] = ccq, counts[7] = ci2, counts[8] = ci4, counts[9] = cr2, counts[10] = cr4;
    }
}

// RNG / hashing exports for known-answer tests
void orc_tea(uint32_t v0, uint32_t v1, int rounds, const uint32_t* key, uint32_t* out) {
    tea(v0, v1, rounds, key ? key : kTeaKey, &out[0], &out[1]);
}
void orc_pcg_seq(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    Pcg r;
    r.seed(initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = r.next_u32();
}
void orc_pcg_lane_floats(uint32_t seed, uint32_t point, uint64_t lane, int n, float* out) {
    Pcg r = pcg_for(seed, point, lane);
    for (int i = 0; i < n; ++i) out[i] = r.next_float32();
}
// math exports for pinning the correctly rounded functions
void orc_math(int fn, const float* x, const float* y, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        float a = x[i];
        switch (fn) {
            case 0: out[i] = cr_log(a); break;
            case 1: out[i] = cr_exp(a); break;
            case 2: out[i] = cr_sin(a); break;
            case 3: out[i] = cr_cos(a); break;
            case 4: out[i] = cr_acos(a); break;
            case 5: out[i] = cr_asin(a); break;
            case 6: out[i] = cr_atan2(a, y[i]); break;
            case 7: out[i] = cr_cbrt(a); break;
        }
    }
}
int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
