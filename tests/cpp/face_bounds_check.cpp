// Checks the 3D SNCH face-plane bound built by scene_build.cpp: for every query
// point and every node the bound prunes, no edge referenced below that node
// may pass the fp32 silhouette test (silhouette_edge.py:8-13, evaluated with
// the same fused multiply-adds as the device code).  Open edges count as
// silhouettes, so a pruned node holding one is a failure too.
//
// usage: face_bounds_check <scene.bin> <n_queries> <seed>
// scene.bin: int64 nv, int64 np, float32 V[nv*3], int32 I[np*3], int32 T[np]
#include "scene_build.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

namespace {
struct F3 {
    float x, y, z;
};
F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dotf(F3 a, F3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
F3 cross(F3 u, F3 v) {
    return {fmaf(u.y, v.z, -(u.z * v.y)), fmaf(u.z, v.x, -(u.x * v.z)), fmaf(u.x, v.y, -(u.y * v.x))};
}
F3 unit(F3 a) {
    float s = 1.0f / sqrtf(dotf(a, a));
    return {a.x * s, a.y * s, a.z * s};
}
float lerpf(float a, float b, float t) { return fmaf(b, t, fmaf(-a, t, a)); }
F3 at(const float* p) { return {p[0], p[1], p[2]}; }
}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t nv = 0, np = 0;
    if (fread(&nv, 8, 1, f) != 1 || fread(&np, 8, 1, f) != 1) return 2;
    std::vector<float> V(nv * 3);
    std::vector<int> I(np * 3), T(np);
    if (fread(V.data(), 4, V.size(), f) != V.size() || fread(I.data(), 4, I.size(), f) != I.size() ||
        fread(T.data(), 4, T.size(), f) != T.size())
        return 2;
    fclose(f);
    dwb::SceneHost H;
    std::string err = dwb::build_scene(3, V.data(), nv, I.data(), T.data(), np, H);
    if (!err.empty()) {
        printf("build failed: %s\n", err.c_str());
        return 2;
    }
    const dwb::SubHost& S = H.sub[0];
    size_t nn = S.n_nodes();
    if (S.node_pb.size() != 3 * nn) {
        printf("missing face bounds\n");
        return 1;
    }
    std::vector<int> end(nn);
    for (size_t k = nn; k-- > 0;) end[k] = S.node_nref[k] > 0 ? (int)k + 1 : end[k + S.node_off[k]];

    std::mt19937 rng((unsigned)atoi(argv[3]));
    std::uniform_int_distribution<int> pick(0, (int)np - 1);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double offsets[] = {0.0, 1e-7, -1e-7, 1e-6, -1e-6, 1e-4, -1e-4, 1e-3, -1e-3, 3e-2};
    long long pruned = 0, checked = 0, bad = 0;
    int nq = atoi(argv[2]);
    for (int q = 0; q < nq; ++q) {
        F3 x;
        if (q % 11 == 10) {  // anywhere in [-1.5, 1.5]^3
            x = {(float)(3 * U(rng) - 1.5), (float)(3 * U(rng) - 1.5), (float)(3 * U(rng) - 1.5)};
        } else {  // on or beside a random face
            int t = pick(rng);
            double w0 = U(rng), w1 = U(rng);
            if (w0 + w1 > 1.0) w0 = 1.0 - w0, w1 = 1.0 - w1;
            const float *a = &V[3 * I[3 * t]], *b = &V[3 * I[3 * t + 1]], *c = &V[3 * I[3 * t + 2]];
            double p[3], u[3], v[3], n[3];
            for (int k = 0; k < 3; ++k) {
                u[k] = (double)b[k] - a[k];
                v[k] = (double)c[k] - a[k];
                p[k] = a[k] + w0 * u[k] + w1 * v[k];
            }
            n[0] = u[1] * v[2] - u[2] * v[1];
            n[1] = u[2] * v[0] - u[0] * v[2];
            n[2] = u[0] * v[1] - u[1] * v[0];
            double ln = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double off = offsets[q % 10] / ln;
            x = {(float)(p[0] + off * n[0]), (float)(p[1] + off * n[1]), (float)(p[2] + off * n[2])};
        }
        for (size_t s = 0; s < nn; ++s) {
            const float* pb = &S.node_pb[3 * s];
            const float* bx = &S.node_box[6 * s];
            F3 o = {(bx[0] + bx[3]) / 2.0f, (bx[1] + bx[4]) / 2.0f, (bx[2] + bx[5]) / 2.0f};
            F3 v = sub(x, o);
            float cv = dotf(at(&S.node_axis[3 * s]), v);
            float sl = pb[0] * sqrtf(dotf(v, v));
            if (!(((cv - sl) - pb[2] > 0.0f) || ((cv + sl) - pb[1] < 0.0f))) continue;
            ++pruned;
            for (int i = (int)s; i < end[s]; ++i)
                for (int j = 0; j < S.node_nsil[i]; ++j) {
                    int e = S.refs[S.node_sil_off[i] + j];
                    const float* P = &S.sil_pts[12 * (size_t)e];
                    const int* idx = &S.sil_idx[4 * (size_t)e];
                    ++checked;
                    bool sil;
                    if (idx[0] < 0 || idx[3] < 0) {
                        sil = true;
                    } else {
                        F3 A = at(P), B = at(P + 3), C = at(P + 6), D = at(P + 9);
                        F3 ab = sub(B, A), ax = sub(x, A);
                        float t = fminf(fmaxf(dotf(ax, ab) / dotf(ab, ab), 0.0f), 1.0f);
                        F3 pt = {lerpf(A.x, B.x, t), lerpf(A.y, B.y, t), lerpf(A.z, B.z, t)};
                        F3 n0 = unit(cross(sub(B, A), sub(C, A))), n1 = unit(cross(sub(D, A), sub(B, A)));
                        F3 view = sub(pt, x);
                        sil = dotf(view, n0) * dotf(view, n1) < 0.0f;
                    }
                    if (sil && ++bad <= 5) printf("pruned silhouette: query %d node %zu edge %d\n", q, s, e);
                }
        }
    }
    printf("queries %d pruned nodes %lld edges checked %lld silhouettes pruned %lld\n", nq, pruned, checked, bad);
    return bad ? 1 : 0;
}
