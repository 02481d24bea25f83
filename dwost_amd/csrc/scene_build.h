// scene_build.h -- host-side scene construction (built once per scene, then
// uploaded).  Restates, without Eigen:
//   * the binned-SAH BVH of include/bvh.h:359-652 (8 buckets, leaves <= 4,
//     depth cap 62, preorder with right-child offsets),
//   * the SNCH of include/snch.h:159-489 (per-leaf silhouette references,
//     recursive bounding cones),
//   * the Polyline / Mesh set-up of diff_wost/shapes/polyline.py:60-186 and
//     mesh.py:82-283 (bbox, sampling pmf/cdf, vertex normals, sub-scenes).
// Arithmetic follows Eigen's fixed-size evaluation order (3-vector reductions
// are x0 op (x1 op x2)) so the tree is identical to the reference builder's.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dwb {

struct SubHost {
    bool present = false;
    int dim = 2;
    std::vector<float> V;        // local vertices
    std::vector<int> I;          // local indices
    std::vector<int> vglobal;    // local vertex -> global vertex id
    std::vector<int> pglobal;    // local primitive -> global primitive id
    float bb_min[3] = {0, 0, 0}, bb_max[3] = {0, 0, 0};
    bool py_sils = false;
    std::vector<float> pmf, cdf;
    float pmf_sum = 0.0f, pmf_norm = 0.0f;
    // guide for the CDF search: a draw u in [k/K, (k+1)/K) selects an index in
    // [cdf_guide[k], cdf_guide[k+1]] (K = 1 << guide_shift)
    std::vector<int> cdf_guide;
    int guide_shift = 0;
    std::vector<float> vnormal;  // 2D only
    // BVH / SNCH (node-major)
    int max_depth = 0;
    std::vector<float> node_box;   // 2*dim per node (min..., max...)
    std::vector<int> node_off, node_nref, node_sil_off, node_nsil;
    std::vector<float> node_axis;  // dim per node
    std::vector<float> node_half_angle, node_radius;
    // 3D face-plane bound per node (rho, e_lo, e_hi): every face f adjacent to
    // an edge of the subtree has |n_f - axis| <= rho and n_f . (a_f - o) in
    // [e_lo, e_hi] with o the node's box centre, widened by the fp32 error of
    // the silhouette test.  rho = +inf disables the bound (open edges).
    std::vector<float> node_pb;
    std::vector<int> prim_order;   // BVH order -> local primitive index
    std::vector<int> refs;         // SNCH silhouette refs (topology ids)
    // topology silhouettes (2D: one per local vertex; 3D: one per edge)
    std::vector<float> sil_pts;    // 2D: a,b,c (6 floats); 3D: a,b,c,d (12 floats)
    std::vector<int> sil_idx;      // 2D: prev,center,next; 3D: c,a,b,d
    // 2D: the polyline's own silhouette vertices (a, b, c: 6 floats each) and the
    // point BVH over their b: node boxes (min x, min y, max x, max y), offsets,
    // counts, and BVH order -> silhouette index
    std::vector<float> psil, pnode_box;
    std::vector<int> pnode_off, pnode_nref, porder;
    // 3D: the mesh's own silhouette edges (a, b, c, d: 12 floats each) and their
    // length distribution (pmf, sequential fp32 cdf, guide table)
    std::vector<float> psil3, ps_pmf, ps_cdf;
    float ps_sum = 0.0f, ps_norm = 0.0f;
    std::vector<int> ps_guide;
    int ps_shift = 0;
    size_t n_nodes() const { return node_off.size(); }
    size_t n_prims() const { return I.size() / dim; }
};

struct SceneHost {
    int dim = 2;
    std::vector<float> V;
    std::vector<int> I, types;
    SubHost sub[3];  // full, Dirichlet (type 1), Neumann (type 0)
};

// Builds all three sub-scenes.  Returns an empty string on success.
std::string build_scene(int dim, const float* V, int64_t nv, const int32_t* I, const int32_t* types, int64_t np,
                        SceneHost& out);

}  // namespace dwb
