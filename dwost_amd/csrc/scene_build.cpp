// scene_build.cpp -- see scene_build.h.  Pure host C++ (no HIP), compiled with
// -ffp-contract=off so that every fp32 operation rounds exactly as written.
#include "scene_build.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <thread>
#include <utility>

namespace dwb {
namespace {

const float kEps = (float)1e-9;  // include/fwd.h:14 (Float = float)
const float kPiF = (float)3.14159265358979323846;

inline float cmin(float a, float b) { return (b < a) ? b : a; }  // Eigen cwiseMin
inline float cmax(float a, float b) { return (a < b) ? b : a; }  // Eigen cwiseMax

template <int D>
struct Bx {
    float lo[D], hi[D];
    void reset() {
        for (int d = 0; d < D; ++d) lo[d] = FLT_MAX, hi[d] = -FLT_MAX;
    }
    void grow_pt(const float* p) {  // bounding_volume.h:119-123 (epsilon-inflated point)
        for (int d = 0; d < D; ++d) {
            lo[d] = cmin(lo[d], p[d] - kEps);
            hi[d] = cmax(hi[d], p[d] + kEps);
        }
    }
    void grow(const Bx& o) {
        for (int d = 0; d < D; ++d) {
            lo[d] = cmin(lo[d], o.lo[d]);
            hi[d] = cmax(hi[d], o.hi[d]);
        }
    }
    float area() const {  // bounding_volume.h:191-194, Eigen reduction order
        float e0 = cmax(hi[0] - lo[0], 1e-5f), e1 = cmax(hi[1] - lo[1], 1e-5f);
        if (D == 2) {
            float p = e0 * e1;
            return 2.0f * (p / e0 + p / e1);
        }
        float e2 = cmax(hi[D - 1] - lo[D - 1], 1e-5f);
        float p = e0 * (e1 * e2);
        return 2.0f * (p / e0 + (p / e1 + p / e2));
    }
};

template <int D>
struct Bvh {
    int n = 0;
    std::vector<Bx<D>> pbox;
    std::vector<float> pcen;  // D per primitive
    std::vector<int> order;   // current position -> local primitive id
    // output (preorder)
    std::vector<Bx<D>> nbox;
    std::vector<int> noff, nref;
    int max_depth = 0;

    void run(const std::vector<float>& V, const std::vector<int>& I) {
        n = (int)(I.size() / D);
        pbox.resize(n);
        pcen.resize((size_t)n * D);
        order.resize(n);
        for (int p = 0; p < n; ++p) {
            const float* a = &V[(size_t)D * I[(size_t)D * p]];
            const float* b = &V[(size_t)D * I[(size_t)D * p + 1]];
            pbox[p].reset();
            pbox[p].grow_pt(a);
            pbox[p].grow_pt(b);
            if (D == 3) {
                const float* c = &V[(size_t)D * I[(size_t)D * p + 2]];
                pbox[p].grow_pt(c);
                for (int d = 0; d < 3; ++d) pcen[(size_t)3 * p + d] = ((a[d] + b[d]) + c[d]) / 3.0f;  // primitive.h:99-101
            } else {
                for (int d = 0; d < 2; ++d) pcen[(size_t)2 * p + d] = (a[d] + b[d]) * 0.5f;  // primitive.h:29-31
            }
            order[p] = p;
        }
        if (n == 0) return;
        nbox.reserve(2 * n);
        noff.reserve(2 * n);
        nref.reserve(2 * n);
        node(-1, 0, n, 0);
    }

    // point primitives (primitive.h:116-152): the epsilon box of each point
    // widened by its radius 1e-4; centroid = the point
    void run_points(const std::vector<float>& P) {
        n = (int)(P.size() / D);
        pbox.resize(n);
        pcen.resize((size_t)n * D);
        order.resize(n);
        for (int p = 0; p < n; ++p) {
            pbox[p].reset();
            pbox[p].grow_pt(&P[(size_t)D * p]);
            for (int d = 0; d < D; ++d) {
                pbox[p].lo[d] = pbox[p].lo[d] - 1e-4f;
                pbox[p].hi[d] = pbox[p].hi[d] + 1e-4f;
                pcen[(size_t)D * p + d] = P[(size_t)D * p + d];
            }
            order[p] = p;
        }
        if (n == 0) return;
        nbox.reserve(2 * n);
        noff.reserve(2 * n);
        nref.reserve(2 * n);
        node(-1, 0, n, 0);
    }

    void swap_prim(int i, int j) {
        std::swap(order[i], order[j]);
        std::swap(pbox[i], pbox[j]);
        for (int d = 0; d < D; ++d) std::swap(pcen[(size_t)D * i + d], pcen[(size_t)D * j + d]);
    }

    // buildRecursive (bvh.h:402-477)
    void node(int parent, int start, int end, int depth) {
        if (depth > max_depth) max_depth = depth;
        int me = (int)noff.size();
        int cnt = end - start;
        Bx<D> bb, bc;
        bb.reset();
        bc.reset();
        for (int p = start; p < end; ++p) {
            bb.grow(pbox[p]);
            bc.grow_pt(&pcen[(size_t)D * p]);
        }
        bool leaf = cnt <= 4 || depth == 62;
        nbox.push_back(bb);
        noff.push_back(leaf ? start : -1);  // -1 == UNTOUCHED
        nref.push_back(leaf ? cnt : 0);
        if (parent >= 0) {
            // second touch of the parent: right-child offset (bvh.h:446-454)
            noff[parent] -= 1;
            if (noff[parent] == -3) noff[parent] = me - parent;
        }
        if (leaf) return;
        int dim;
        float coord;
        split(bb, bc, start, end, dim, coord);
        int mid = start;  // performObjectSplit (bvh.h:622-652)
        for (int i = start; i < end; ++i)
            if (pcen[(size_t)D * i + dim] < coord) swap_prim(i, mid++);
        if (mid == start || mid == end) mid = start + (end - start) / 2;
        node(me, start, mid, depth + 1);
        node(me, mid, end, depth + 1);
    }

    // computeObjectSplit (bvh.h:537-620): 8 buckets per axis, strict-min SAH
    void split(const Bx<D>& bb, const Bx<D>& bc, int start, int end, int& dim, float& coord) {
        float best = FLT_MAX;
        dim = -1;
        coord = 0.0f;
        for (int d = 0; d < D; ++d) {
            float ext = bb.hi[d] - bb.lo[d];
            if (ext < 1e-6f) continue;
            float w = ext / 8.0f;
            Bx<D> bk[8];
            int bn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int b = 0; b < 8; ++b) bk[b].reset();
            for (int p = start; p < end; ++p) {
                int bi = (int)((pcen[(size_t)D * p + d] - bb.lo[d]) / w);
                bi = bi < 0 ? 0 : (bi > 7 ? 7 : bi);
                bk[bi].grow(pbox[p]);
                bn[bi] += 1;
            }
            Bx<D> rb[8];
            int rn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            Bx<D> acc;
            acc.reset();
            for (int b = 7; b > 0; --b) {
                acc.grow(bk[b]);
                rb[b] = acc;
                rn[b] = bn[b] + (b != 7 ? rn[b + 1] : 0);
            }
            Bx<D> lb;
            lb.reset();
            int ln = 0;
            for (int b = 1; b < 8; ++b) {
                lb.grow(bk[b - 1]);
                ln += bn[b - 1];
                if (ln > 0 && rn[b] > 0) {
                    float cost = (float)ln * lb.area() + (float)rn[b] * rb[b].area();
                    if (cost < best) {
                        best = cost;
                        dim = d;
                        coord = bb.lo[d] + (float)b * w;
                    }
                }
            }
        }
        if (dim == -1) {  // longest centroid axis, first index of the max
            dim = 0;
            float e = bc.hi[0] - bc.lo[0];
            for (int d = 1; d < D; ++d)
                if (bc.hi[d] - bc.lo[d] > e) e = bc.hi[d] - bc.lo[d], dim = d;
            coord = (bc.lo[dim] + bc.hi[dim]) * 0.5f;
        }
    }
};

// Eigen-order helpers for the SNCH set-up
template <int D>
inline float sqn(const float* v) {
    return D == 2 ? v[0] * v[0] + v[1] * v[1] : v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
}
template <int D>
inline void nrmz(float* v) {  // Eigen normalized()
    float z = sqn<D>(v);
    if (z > 0.0f) {
        float s = std::sqrt(z);
        for (int d = 0; d < D; ++d) v[d] /= s;
    }
}
template <int D>
inline float edot(const float* a, const float* b) {
    return D == 2 ? a[0] * b[0] + a[1] * b[1] : a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

// Face normals of a topology silhouette (silhouette_vertex.h:53-76,
// silhouette_edge.h:63-88).  pts: 2D a,b,c / 3D a,b,c,d; idx as in the header.
template <int D>
inline bool has_face(const int* idx, int f) {
    if (D == 2) return f == 0 ? idx[2] != -1 : idx[0] != -1;
    return f == 0 ? idx[0] != -1 : idx[3] != -1;
}
template <int D>
inline void face_normal(const float* pts, int f, float* n) {
    if (D == 2) {
        const float* pa = f == 0 ? pts + 2 : pts;
        const float* pb = f == 0 ? pts + 4 : pts + 2;
        n[0] = pb[1] - pa[1];
        n[1] = -(pb[0] - pa[0]);
    } else {
        const float *a = pts, *b = pts + 3, *c = pts + 6, *d = pts + 9;
        float u[3], v[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = f == 0 ? b[k] - a[k] : d[k] - a[k];
            v[k] = f == 0 ? c[k] - a[k] : b[k] - a[k];
        }
        n[0] = u[1] * v[2] - u[2] * v[1];
        n[1] = u[2] * v[0] - u[0] * v[2];
        n[2] = u[0] * v[1] - u[1] * v[0];
    }
}

template <int D>
void build_snch(SubHost& S, const std::vector<int>& conn) {
    const int P = D == 2 ? 6 : 12, X = D == 2 ? 3 : 4;
    size_t nn = S.n_nodes();
    S.node_sil_off.assign(nn, 0);
    S.node_nsil.assign(nn, 0);
    std::vector<float> refN, refF;
    std::vector<int> seen_stamp(S.sil_idx.size() / X, -1);
    for (size_t i = 0; i < nn; ++i) {  // assignGeometricDataToNodes (snch.h:174-306)
        int start = (int)S.refs.size();
        for (int j = 0; j < S.node_nref[i]; ++j) {
            int prim = S.prim_order[S.node_off[i] + j];
            for (int k = 0; k < D; ++k) {
                int si = conn[(size_t)D * prim + k];
                if (seen_stamp[si] == (int)i) continue;
                seen_stamp[si] = (int)i;
                const float* pts = &S.sil_pts[(size_t)P * si];
                const int* idx = &S.sil_idx[(size_t)X * si];
                float n0[3] = {0, 0, 0}, n1[3] = {0, 0, 0}, n[3] = {0, 0, 0}, t[3];
                if (has_face<D>(idx, 0)) face_normal<D>(pts, 0, n0), nrmz<D>(n0);
                if (has_face<D>(idx, 1)) face_normal<D>(pts, 1, n1), nrmz<D>(n1);
                for (int f = 0; f < 2; ++f)
                    if (has_face<D>(idx, f)) {
                        face_normal<D>(pts, f, t);
                        for (int q = 0; q < D; ++q) n[q] += t[q];
                    }
                nrmz<D>(n);
                S.refs.push_back(si);
                for (int q = 0; q < D; ++q) refN.push_back(n[q]);
                for (int q = 0; q < D; ++q) refF.push_back(n0[q]);
                for (int q = 0; q < D; ++q) refF.push_back(n1[q]);
            }
        }
        S.node_sil_off[i] = start;
        S.node_nsil[i] = (int)S.refs.size() - start;
    }
    S.node_axis.assign(nn * D, 0.0f);
    S.node_half_angle.assign(nn, kPiF);
    S.node_radius.assign(nn, 0.0f);
    // computeBoundingConesRecursive (snch.h:308-385); the subtree of node s is
    // the contiguous preorder range [s, e).  Explicit stack, same visit order.
    std::vector<std::pair<int, int>> todo;
    if (nn) todo.emplace_back(0, (int)nn);
    while (!todo.empty()) {
        auto [s, e] = todo.back();
        todo.pop_back();
        float cen[3];
        for (int d = 0; d < D; ++d) cen[d] = (S.node_box[(size_t)2 * D * s + d] + S.node_box[(size_t)2 * D * s + D + d]) * 0.5f;
        float axis[3] = {0, 0, 0}, radius = 0.0f;
        bool any = false, two = true;
        int r0 = S.node_sil_off[s], r1 = r0;
        for (int i = s; i < e; ++i)
            if (S.node_nsil[i]) r1 = S.node_sil_off[i] + S.node_nsil[i];
        // refs of the subtree are contiguous: [first nonempty node's offset, last end)
        for (int i = s; i < e; ++i) {
            if (!S.node_nsil[i]) continue;
            r0 = std::min(r0, S.node_sil_off[i]);
        }
        for (int i = s; i < e; ++i)
            for (int j = 0; j < S.node_nsil[i]; ++j) {
                int r = S.node_sil_off[i] + j;
                int si = S.refs[r];
                const float* pts = &S.sil_pts[(size_t)P * si];
                const int* idx = &S.sil_idx[(size_t)X * si];
                for (int q = 0; q < D; ++q) axis[q] += refN[(size_t)D * r + q];
                float sc[3], dv[3];
                if (D == 2) {
                    sc[0] = pts[2], sc[1] = pts[3];
                } else {
                    for (int q = 0; q < 3; ++q) sc[q] = (pts[q] + pts[3 + q]) * 0.5f;
                }
                for (int q = 0; q < D; ++q) dv[q] = sc[q] - cen[q];
                radius = std::max(radius, std::sqrt(sqn<D>(dv)));
                two = two && has_face<D>(idx, 0) && has_face<D>(idx, 1);
                any = true;
            }
        (void)r1;
        if (!any) {
            S.node_half_angle[s] = -kPiF;
        } else if (!two) {
            S.node_half_angle[s] = kPiF;
        } else {
            float an = std::sqrt(sqn<D>(axis));
            if (an > kEps) {
                for (int q = 0; q < D; ++q) axis[q] /= an;
                float ha = 0.0f;
                for (int i = s; i < e; ++i)
                    for (int j = 0; j < S.node_nsil[i]; ++j) {
                        int r = S.node_sil_off[i] + j;
                        for (int k = 0; k < 2; ++k) {
                            const float* fn = &refF[((size_t)2 * r + k) * D];
                            float c = std::max(-1.0f, std::min(1.0f, edot<D>(axis, fn)));
                            ha = std::max(ha, (float)std::acos((double)c));
                        }
                    }
                for (int q = 0; q < D; ++q) S.node_axis[(size_t)D * s + q] = axis[q];
                S.node_half_angle[s] = ha;
                S.node_radius[s] = radius;
            }
        }
        if (S.node_nref[s] == 0) {
            int r = s + S.node_off[s];
            todo.emplace_back(r, e);      // processed second
            todo.emplace_back(s + 1, r);  // processed first
        }
    }
}

// Face-plane bound of every 3D SNCH node.  An edge (a, b) with faces (a, b, c)
// and (a, d, b) is a silhouette for x iff the two face planes see x on
// opposite sides (silhouette_edge.py:8-13: the view vector from the closest
// point on the edge is dotted with both face normals).  With v = x - o,
//   s_f(x) = n_f . (x - a_f) in [c.v - rho |v| - e_hi, c.v + rho |v| - e_lo],
// so when that interval excludes 0 every face sees x on the same side and the
// subtree holds no silhouette for x.  The test only removes nodes without
// silhouettes, so query results (distances and ids) are unchanged.
//
// Error terms (eps = 2^-24), folded into rho and e: the fp32 test evaluates
// dot(pt - x, normalize(cross(u, v))); the computed unit normal is within
// (16 / sin(theta) + 8) eps of the exact one (theta the angle between u and v),
// pt is within 4 eps |a| of the edge, and |pt - x| <= |v| + ext with ext the
// largest distance from o to a vertex of the node's edges.  The fp32
// evaluation of the bound itself adds a few eps (|v| + |e| + |o|).
void face_bounds3(SubHost& S) {
    const double eps = std::ldexp(1.0, -24);
    size_t nn = S.n_nodes();
    S.node_pb.assign(nn * 3, 0.0f);
    std::vector<int> end(nn);
    for (size_t k = nn; k-- > 0;)
        end[k] = S.node_nref[k] > 0 ? (int)k + 1 : end[k + S.node_off[k]];
    // per-edge face data in double: unit normals, sin(theta), vertices
    size_t ne = S.sil_idx.size() / 4;
    std::vector<double> fn(ne * 6), fsin(ne * 2);
    std::vector<char> open(ne);
    for (size_t e = 0; e < ne; ++e) {
        const float* p = &S.sil_pts[12 * e];
        const int* idx = &S.sil_idx[4 * e];
        open[e] = idx[0] < 0 || idx[3] < 0;
        if (open[e]) continue;
        for (int f = 0; f < 2; ++f) {
            double u[3], v[3], n[3];
            for (int k = 0; k < 3; ++k) {
                u[k] = f == 0 ? (double)p[3 + k] - p[k] : (double)p[9 + k] - p[k];
                v[k] = f == 0 ? (double)p[6 + k] - p[k] : (double)p[3 + k] - p[k];
            }
            n[0] = u[1] * v[2] - u[2] * v[1];
            n[1] = u[2] * v[0] - u[0] * v[2];
            n[2] = u[0] * v[1] - u[1] * v[0];
            double ln = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            double lv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            fsin[2 * e + f] = (lu > 0 && lv > 0) ? ln / (lu * lv) : 0.0;
            for (int k = 0; k < 3; ++k) fn[6 * e + 3 * f + k] = ln > 0 ? n[k] / ln : 0.0;
        }
    }
    for (size_t s = 0; s < nn; ++s) {
        float* out = &S.node_pb[3 * s];
        out[0] = INFINITY, out[1] = -INFINITY, out[2] = INFINITY;
        if (!(S.node_half_angle[s] >= 0.0f)) {  // no silhouettes below: always pruned
            out[0] = 0.0f, out[1] = INFINITY, out[2] = -INFINITY;
            continue;
        }
        float o[3];
        for (int k = 0; k < 3; ++k) o[k] = (S.node_box[6 * s + k] + S.node_box[6 * s + 3 + k]) / 2.0f;
        double c[3] = {S.node_axis[3 * s], S.node_axis[3 * s + 1], S.node_axis[3 * s + 2]};
        double rho = 0.0, elo = INFINITY, ehi = -INFINITY, ext = 0.0, amax = 0.0, dn = 0.0;
        bool ok = true;
        for (int i = (int)s; i < end[s] && ok; ++i)
            for (int j = 0; j < S.node_nsil[i] && ok; ++j) {
                int e = S.refs[S.node_sil_off[i] + j];
                if (open[e] || fsin[2 * e] <= 0.0 || fsin[2 * e + 1] <= 0.0) {
                    ok = false;
                    break;
                }
                const float* p = &S.sil_pts[12 * (size_t)e];
                for (int q = 0; q < 2; ++q) {  // the edge's endpoints a, b
                    double r2 = 0.0, a2 = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        double d = (double)p[3 * q + k] - o[k];
                        r2 += d * d;
                        a2 += (double)p[3 * q + k] * p[3 * q + k];
                    }
                    ext = std::max(ext, std::sqrt(r2));
                    amax = std::max(amax, std::sqrt(a2));
                }
                for (int f = 0; f < 2; ++f) {
                    const double* n = &fn[6 * e + 3 * f];
                    double dc[3] = {n[0] - c[0], n[1] - c[1], n[2] - c[2]};
                    rho = std::max(rho, std::sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]));
                    double ef = n[0] * ((double)p[0] - o[0]) + n[1] * ((double)p[1] - o[1]) + n[2] * ((double)p[2] - o[2]);
                    elo = std::min(elo, ef);
                    ehi = std::max(ehi, ef);
                    dn = std::max(dn, (16.0 / fsin[2 * e + f] + 8.0) * eps);
                }
            }
        if (!ok || !(ehi >= elo)) continue;
        double onorm = std::sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
        double A = dn + 16.0 * eps * (1.0 + rho);
        double B = ext * (dn + 16.0 * eps) + 16.0 * eps * (amax + onorm + std::max(std::fabs(elo), std::fabs(ehi)) + 1.0);
        // widen by 1e-6 relative so the float rounding below stays outward
        double rw = (rho + A) * (1.0 + 1e-6), lw = elo - B, hw = ehi + B;
        out[0] = std::nextafter((float)rw, INFINITY);
        out[1] = std::nextafter((float)lw, -INFINITY);
        out[2] = std::nextafter((float)hw, INFINITY);
    }
}

// The index DiscreteDistribution.sample draws for a value u (distr.py:52-70):
// the first i whose predicate ((cdf < u*sum) || cdf == 0) && cdf != sum fails,
// n - 1 when none does.  The predicate is monotone in i (cdf is non-decreasing
// and non-negative) and in u, so the index is non-decreasing in u; a table of
// the indices at u = k/K therefore brackets the index of every u in
// [k/K, (k+1)/K), and a search inside the bracket finds the same index.
int cdf_first_false(const SubHost& S, float sample) {
    int lo = 0, hi = (int)S.cdf.size() - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        float cv = S.cdf[mid];
        bool cond = ((cv < sample) || (cv == 0.0f)) && (cv != S.pmf_sum);
        if (cond)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
void cdf_guide(SubHost& S) {
    size_t n = S.cdf.size();
    int shift = 0;
    while (shift < 16 && ((size_t)1 << shift) * 4 < n) ++shift;
    S.guide_shift = shift;
    int K = 1 << shift;
    S.cdf_guide.resize((size_t)K + 1);
    for (int k = 0; k <= K; ++k) {
        float u = (float)k / (float)K;  // exact: K is a power of two
        S.cdf_guide[k] = n ? cdf_first_false(S, u * S.pmf_sum) : 0;
    }
}

template <int D>
void finish_sub(SubHost& S) {
    S.present = true;
    S.dim = D;
    size_t nv = S.V.size() / D, np = S.I.size() / D;
    for (int d = 0; d < D; ++d) S.bb_min[d] = INFINITY, S.bb_max[d] = -INFINITY;
    for (size_t v = 0; v < nv; ++v)
        for (int d = 0; d < D; ++d) {
            S.bb_min[d] = std::min(S.bb_min[d], S.V[D * v + d]);
            S.bb_max[d] = std::max(S.bb_max[d], S.V[D * v + d]);
        }
    // sampling pmf + sequential fp32 cdf (distr.py:44-50)
    S.pmf.resize(np);
    S.cdf.resize(np);
    float acc = 0.0f;
    for (size_t p = 0; p < np; ++p) {
        float w;
        const float* a = &S.V[(size_t)D * S.I[D * p]];
        const float* b = &S.V[(size_t)D * S.I[D * p + 1]];
        if (D == 2) {
            float dx = b[0] - a[0], dy = b[1] - a[1];
            w = std::sqrt(std::fmaf(dy, dy, dx * dx));
        } else {
            const float* c = &S.V[(size_t)3 * S.I[3 * p + 2]];
            float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            float x = std::fmaf(u[1], v[2], -(u[2] * v[1]));
            float y = std::fmaf(u[2], v[0], -(u[0] * v[2]));
            float z = std::fmaf(u[0], v[1], -(u[1] * v[0]));
            w = std::sqrt(std::fmaf(z, z, std::fmaf(y, y, x * x)));
        }
        S.pmf[p] = w;
        acc += w;
        S.cdf[p] = acc;
    }
    S.pmf_sum = np ? S.cdf[np - 1] : 0.0f;
    S.pmf_norm = 1.0f / S.pmf_sum;
    cdf_guide(S);
    if (D == 2) {
        // vertex normals (polyline.py:173-186): scatter edge normals to .x then .y
        S.vnormal.assign(nv * 2, 0.0f);
        std::vector<float> en(np * 2);
        for (size_t p = 0; p < np; ++p) {
            const float* a = &S.V[2 * S.I[2 * p]];
            const float* b = &S.V[2 * S.I[2 * p + 1]];
            float nx = b[1] - a[1], ny = -(b[0] - a[0]);
            float inv = 1.0f / std::sqrt(std::fmaf(ny, ny, nx * nx));
            en[2 * p] = nx * inv;
            en[2 * p + 1] = ny * inv;
        }
        for (int pass = 0; pass < 2; ++pass)
            for (size_t p = 0; p < np; ++p) {
                int v = S.I[2 * p + pass];
                S.vnormal[2 * v] += en[2 * p];
                S.vnormal[2 * v + 1] += en[2 * p + 1];
            }
        for (size_t v = 0; v < nv; ++v) {
            float x = S.vnormal[2 * v], y = S.vnormal[2 * v + 1];
            float inv = 1.0f / std::sqrt(std::fmaf(y, y, x * x));
            S.vnormal[2 * v] = x * inv;
            S.vnormal[2 * v + 1] = y * inv;
        }
        // python-side silhouettes exist iff some vertex has both prev and next (polyline.py:118-130)
        std::vector<char> hp(nv, 0), hn(nv, 0);
        for (size_t p = 0; p < np; ++p) hp[S.I[2 * p + 1]] = 1, hn[S.I[2 * p]] = 1;
        S.py_sils = false;
        for (size_t v = 0; v < nv; ++v) S.py_sils = S.py_sils || (hp[v] && hn[v]);
    } else {
        // python-side silhouettes exist iff an undirected edge has exactly two faces (mesh.py:177-199)
        std::unordered_map<uint64_t, int> cnt;
        cnt.reserve(np * 3);
        for (size_t p = 0; p < np; ++p)
            for (int k = 0; k < 3; ++k) {
                uint64_t u = (uint32_t)S.I[3 * p + k], v = (uint32_t)S.I[3 * p + (k + 1) % 3];
                cnt[std::min(u, v) << 32 | std::max(u, v)]++;
            }
        S.py_sils = false;
        for (auto& kv : cnt) S.py_sils = S.py_sils || kv.second == 2;
    }
    // BVH
    Bvh<D> B;
    B.run(S.V, S.I);
    S.max_depth = B.max_depth;
    size_t nn = B.noff.size();
    S.node_box.resize(nn * 2 * D);
    for (size_t i = 0; i < nn; ++i)
        for (int d = 0; d < D; ++d) {
            S.node_box[i * 2 * D + d] = B.nbox[i].lo[d];
            S.node_box[i * 2 * D + D + d] = B.nbox[i].hi[d];
        }
    S.node_off = B.noff;
    S.node_nref = B.nref;
    S.prim_order = B.order;
    // SNCH topology (snch.h:407-489)
    std::vector<int> conn;
    if (D == 2) {
        std::vector<std::vector<int>> adj(nv);
        for (size_t p = 0; p < np; ++p) {
            adj[S.I[2 * p]].push_back(S.I[2 * p + 1]);
            adj[S.I[2 * p + 1]].push_back(S.I[2 * p]);
        }
        S.sil_pts.assign(nv * 6, 0.0f);
        S.sil_idx.assign(nv * 3, -1);
        for (size_t v = 0; v < nv; ++v) {
            int pv = adj[v].size() >= 1 ? adj[v][0] : -1, nx = adj[v].size() >= 2 ? adj[v][1] : -1;
            float* pts = &S.sil_pts[6 * v];
            if (pv != -1) pts[0] = S.V[2 * pv], pts[1] = S.V[2 * pv + 1];
            pts[2] = S.V[2 * v], pts[3] = S.V[2 * v + 1];
            if (nx != -1) pts[4] = S.V[2 * nx], pts[5] = S.V[2 * nx + 1];
            S.sil_idx[3 * v] = pv;
            S.sil_idx[3 * v + 1] = (int)v;
            S.sil_idx[3 * v + 2] = nx;
        }
        conn = S.I;
    } else {
        std::unordered_map<uint64_t, int> emap;
        emap.reserve(np * 2);
        conn.assign(np * 3, -1);
        for (size_t f = 0; f < np; ++f)
            for (int k = 0; k < 3; ++k) {
                int u = S.I[3 * f + k], v = S.I[3 * f + (k + 1) % 3], w = S.I[3 * f + (k + 2) % 3];
                uint64_t key = (uint64_t)(uint32_t)std::min(u, v) << 32 | (uint32_t)std::max(u, v);
                auto it = emap.find(key);
                if (it == emap.end()) {
                    int e = (int)(S.sil_idx.size() / 4);
                    emap.emplace(key, e);
                    for (int q = 0; q < 3; ++q) S.sil_pts.push_back(S.V[3 * u + q]);
                    for (int q = 0; q < 3; ++q) S.sil_pts.push_back(S.V[3 * v + q]);
                    for (int q = 0; q < 3; ++q) S.sil_pts.push_back(S.V[3 * w + q]);
                    for (int q = 0; q < 3; ++q) S.sil_pts.push_back(0.0f);
                    S.sil_idx.push_back(w);
                    S.sil_idx.push_back(u);
                    S.sil_idx.push_back(v);
                    S.sil_idx.push_back(-1);
                    conn[3 * f + k] = e;
                } else {
                    int e = it->second;
                    for (int q = 0; q < 3; ++q) S.sil_pts[(size_t)12 * e + 9 + q] = S.V[3 * w + q];
                    S.sil_idx[(size_t)4 * e + 3] = w;
                    conn[3 * f + k] = e;
                }
            }
    }
    build_snch<D>(S, conn);
    if (D == 3) face_bounds3(S);
}

// The polyline's own silhouette vertices (Polyline.configure_silhouette,
// polyline.py:100-165) -- distinct from the SNCH's topology silhouettes: over
// the edges (a, b) in order prev[b] = a and next[a] = b (later edges win);
// each vertex with both, in vertex order, gives (a = prev, b, c = next) unless
// |normalize(b - a) . normalize(c - b)| >= 1 - 1e-5; then the point BVH over
// their b (PointBVH, bvh.py:375-395, built by BVH<2, Point2>).  WoStGrad samples
// silhouettes through it.
void build_psil2(SubHost& S) {
    int nv = (int)(S.V.size() / 2), ne = (int)(S.I.size() / 2);
    std::vector<int> prv(nv, -1), nxt(nv, -1);
    for (int e = 0; e < ne; ++e) prv[S.I[2 * e + 1]] = S.I[2 * e], nxt[S.I[2 * e]] = S.I[2 * e + 1];
    auto unit = [](float x, float y, float& ux, float& uy) {  // normalize: v * (1 / sqrt(dot(v, v)))
        float inv = 1.0f / std::sqrt(std::fmaf(y, y, x * x));
        ux = x * inv, uy = y * inv;
    };
    std::vector<float> pts;
    for (int v = 0; v < nv; ++v) {
        if (prv[v] == -1 || nxt[v] == -1) continue;
        const float *a = &S.V[2 * prv[v]], *b = &S.V[2 * v], *c = &S.V[2 * nxt[v]];
        float t1x, t1y, t2x, t2y;
        unit(b[0] - a[0], b[1] - a[1], t1x, t1y);
        unit(c[0] - b[0], c[1] - b[1], t2x, t2y);
        if (!(std::fabs(std::fmaf(t1y, t2y, t1x * t2x)) < (1.0f - 1e-5f))) continue;
        S.psil.insert(S.psil.end(), {a[0], a[1], b[0], b[1], c[0], c[1]});
        pts.push_back(b[0]);
        pts.push_back(b[1]);
    }
    Bvh<2> tree;
    tree.run_points(pts);
    S.pnode_box.clear();
    for (size_t i = 0; i < tree.nbox.size(); ++i)
        S.pnode_box.insert(S.pnode_box.end(), {tree.nbox[i].lo[0], tree.nbox[i].lo[1], tree.nbox[i].hi[0],
                                               tree.nbox[i].hi[1]});
    S.pnode_off = tree.noff;
    S.pnode_nref = tree.nref;
    S.porder = tree.order;
}

// The mesh's own silhouette edges (Mesh.configure_silhouette, mesh.py:140-243)
// for WoStGrad3D: per face (a, b, c) the half edges (min, max) of (a, b), (b, c),
// (c, a) with the opposite vertex -- all first edges, then all second, then all
// third -- stably sorted by the key u * V + v; an edge met exactly twice gives
// (a = V[u], b = V[v], c = opposite of the first, d = opposite of the second),
// kept unless |n0 . n1| >= 1 - 1e-5 (n0 = N((b - a) x (c - a)),
// n1 = N((d - a) x (b - a))); sampled by length (sequential fp32 CDF).
void build_psil3(SubHost& S) {
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
    auto sub3 = [](const float* p, const float* q, float* o) { o[0] = p[0] - q[0], o[1] = p[1] - q[1], o[2] = p[2] - q[2]; };
    auto cross3 = [](const float* x, const float* y, float* o) {  // fma(u.y, v.z, -(u.z * v.y)), ...
        o[0] = std::fmaf(x[1], y[2], -(x[2] * y[1]));
        o[1] = std::fmaf(x[2], y[0], -(x[0] * y[2]));
        o[2] = std::fmaf(x[0], y[1], -(x[1] * y[0]));
    };
    auto dot3 = [](const float* x, const float* y) { return std::fmaf(x[2], y[2], std::fmaf(x[1], y[1], x[0] * y[0])); };
    auto unit3 = [&](float* x) {
        float inv = 1.0f / std::sqrt(dot3(x, x));
        x[0] *= inv, x[1] *= inv, x[2] *= inv;
    };
    S.psil3.clear();
    S.ps_pmf.clear();
    size_t i = 0;
    while (i < order.size()) {
        size_t j = i;
        while (j < order.size() && key[order[j]] == key[order[i]]) ++j;
        if (j - i == 2) {
            int e0 = order[i], e1 = order[i + 1];
            const float *a = &S.V[3 * u[e0]], *b = &S.V[3 * v[e0]], *c = &S.V[3 * opp[e0]], *d = &S.V[3 * opp[e1]];
            float ba[3], ca[3], da[3], n0[3], n1[3];
            sub3(b, a, ba), sub3(c, a, ca), sub3(d, a, da);
            cross3(ba, ca, n0), cross3(da, ba, n1);
            unit3(n0), unit3(n1);
            if (std::fabs(dot3(n0, n1)) < (1.0f - 1e-5f)) {
                S.psil3.insert(S.psil3.end(), {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]});
                float ab[3];
                sub3(a, b, ab);
                S.ps_pmf.push_back(std::sqrt(dot3(ab, ab)));
            }
        }
        i = j;
    }
    SubHost tmp;  // the guide table of the length distribution (see cdf_guide)
    float acc = 0.0f;
    for (float w : S.ps_pmf) acc += w, tmp.cdf.push_back(acc);
    tmp.pmf_sum = tmp.cdf.empty() ? 0.0f : tmp.cdf.back();
    cdf_guide(tmp);
    S.ps_cdf = tmp.cdf;
    S.ps_sum = tmp.pmf_sum;
    S.ps_norm = 1.0f / tmp.pmf_sum;
    S.ps_guide = tmp.cdf_guide;
    S.ps_shift = tmp.guide_shift;
}

// _build_subscene_from_mask (polyline.py:672-692, mesh.py:617-635)
template <int D>
void make_sub(const SceneHost& sc, int want, SubHost& S) {
    size_t np = sc.types.size();
    for (size_t p = 0; p < np; ++p)
        if (want < 0 || sc.types[p] == want) S.pglobal.push_back((int)p);
    if (S.pglobal.empty()) {
        S.present = false;
        S.dim = D;
        return;
    }
    size_t nvg = sc.V.size() / D;
    std::vector<int> used(nvg, 0);
    for (int p : S.pglobal)
        for (int k = 0; k < D; ++k) used[sc.I[(size_t)D * p + k]] = 1;
    std::vector<int> remap(nvg, -1);
    for (size_t v = 0; v < nvg; ++v)
        if (used[v]) {
            remap[v] = (int)S.vglobal.size();
            S.vglobal.push_back((int)v);
        }
    S.V.resize(S.vglobal.size() * D);
    for (size_t i = 0; i < S.vglobal.size(); ++i)
        for (int d = 0; d < D; ++d) S.V[i * D + d] = sc.V[(size_t)D * S.vglobal[i] + d];
    S.I.resize(S.pglobal.size() * D);
    for (size_t j = 0; j < S.pglobal.size(); ++j)
        for (int k = 0; k < D; ++k) S.I[j * D + k] = remap[sc.I[(size_t)D * S.pglobal[j] + k]];
    finish_sub<D>(S);
    if constexpr (D == 2) build_psil2(S);
    if constexpr (D == 3) build_psil3(S);
}

}  // namespace

std::string build_scene(int dim, const float* V, int64_t nv, const int32_t* I, const int32_t* types, int64_t np,
                        SceneHost& out) {
    if (dim != 2 && dim != 3) return "dim must be 2 or 3";
    if (nv <= 0 || np <= 0) return "scene needs at least one vertex and one primitive";
    for (int64_t i = 0; i < np * dim; ++i)
        if (I[i] < 0 || I[i] >= nv) return "primitive index out of range";
    for (int64_t i = 0; i < np; ++i)
        if (types[i] != 0 && types[i] != 1) return "types must be 0 (Neumann) or 1 (Dirichlet)";
    out.dim = dim;
    out.V.assign(V, V + nv * dim);
    out.I.assign(I, I + np * dim);
    out.types.assign(types, types + np);
    // The three sub-scenes (full, Dirichlet, Neumann) are independent builds
    // over the same read-only input: one thread each.
    auto build_one = [&out, dim](int s) {
        int want = s == 0 ? -1 : (s == 1 ? 1 : 0);
        if (dim == 2)
            make_sub<2>(out, want, out.sub[s]);
        else
            make_sub<3>(out, want, out.sub[s]);
    };
    std::thread t1(build_one, 1), t2(build_one, 2);
    build_one(0);
    t1.join();
    t2.join();
    return "";
}

}  // namespace dwb
