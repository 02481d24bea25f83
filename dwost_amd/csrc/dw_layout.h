// dw_layout.h -- device-resident scene layout shared by the host builder and
// the HIP kernels.
//
// One scene holds up to three sub-scenes, mirroring diff_wost's Scene2D /
// Scene3D (polyline.py:623-692, mesh.py:592-635): the full boundary, the
// Dirichlet part and the Neumann part.  Each sub-scene owns a binary BVH in
// preorder (bvh.h:402-477) whose nodes also carry the SNCH cone and silhouette
// range (snch.h:21-36).  Records are 16-byte aligned SoA streams so every
// node / primitive fetch is one or two dwordx4 loads.
#pragma once
#include <stdint.h>

#define DW_MAX_STACK 64

struct alignas(16) DNode2 {  // 32 B
    float mnx, mny, mxx, mxy;
    int off;   // first primitive (leaf) or right-child offset (internal)
    int nref;  // > 0 for leaves
    int sil_off, nsil;
};
struct alignas(16) DCone2 {  // 32 B
    float ax, ay, half_angle, radius;
    float sin_ha, cos_ha;  // of half_angle (transcendental-free cone test)
    float pad0, pad1;
};
struct alignas(16) DNode3 {  // 32 B
    float mnx, mny, mnz, mxx;
    float mxy, mxz;
    int off, nref;
};
struct alignas(16) DCone3 {  // 48 B
    float ax, ay, az, half_angle;
    float radius;
    int sil_off, nsil;
    float sin_ha;
    float cos_ha, pad0, pad1, pad2;
};
struct alignas(16) DSeg {  // 2D primitive in BVH order, 16 B (+ index stream)
    float ax, ay, bx, by;
};
struct alignas(16) DTri {  // 3D primitive in BVH order, 48 B
    float ax, ay, az, bx;
    float by, bz, cx, cy;
    float cz;
    int index;  // original (sub-scene local) primitive index
    int pad0, pad1;
};
struct alignas(16) DSilV {  // 2D silhouette vertex in reference order, 32 B
    float ax, ay, bx, by;
    float cx, cy;
    int both;  // indices[0] != -1 && indices[2] != -1
    int pad;
};
struct alignas(16) DSilE {  // 3D silhouette edge in reference order, 64 B
    float ax, ay, az;
    int boundary;  // indices[0] == -1 || indices[3] == -1
    float bx, by, bz, p0;
    float cx, cy, cz, p1;
    float dx, dy, dz, p2;
};

// Child-in-parent records, one per internal node (indexed by the node's
// preorder index): both children's boxes (and, for the SNCH, cones) plus an
// encoded reference per child, so descending one level costs one dependent
// fetch.  Child reference encoding (int32):
//   >= 0                  internal child: its preorder node index
//   bit31 | c<<27 | off   leaf with c (0..14) items starting at off (< 2^27)
//   bit31 | 15<<27 | k    leaf described by leaf table entry k (int2 off, count)
// BVH records reference primitive ranges; SNCH records silhouette ranges.
#define DW_LEAF_BIT 0x80000000u
#define DW_LEAF_TABLE 15
struct alignas(16) RBvh2 {  // 48 B
    float lmnx, lmny, lmxx, lmxy;
    float rmnx, rmny, rmxx, rmxy;
    int ref_l, ref_r, pad0, pad1;
};
// Two-level BVH records (node width 4): one 128 B record per internal node
// holds its two children and, for each internal child, that child's two
// children.  Slots 0, 1 are the node's children (left, right); the children of
// slot k (k = 0, 1) are slots 2k + 2 (left) and 2k + 3 (right).  Grandchild
// slots of a leaf child are unused (zero).  Same reference encoding as RBvh2;
// an internal reference indexes both the binary and the two-level records.
struct alignas(16) RWide2 {  // 128 B
    float b[6][4];  // per slot: min x, min y, max x, max y
    int ref[6];
    int pad0, pad1;
};
struct alignas(64) RWide3 {  // 192 B: the same slots with 3D boxes; every record spans two 128 B lines
    float b[6][6];  // per slot: min x, y, z, max x, y, z
    int ref[6];
    int pad[6];
};
static_assert(sizeof(RWide2) == 128 && sizeof(RWide3) == 192, "two-level record layout");
struct alignas(16) RSnch2 {  // 96 B
    float lmnx, lmny, lmxx, lmxy;
    float rmnx, rmny, rmxx, rmxy;
    float lax, lay, lha, lrad;
    float rax, ray, rha, rrad;
    float lsin, lcos, rsin, rcos;
    int ref_l, ref_r, pad0, pad1;
};
// 3D records share a 64 B head (both children's boxes and references), so one
// fetch serves every query kind; an SNCH record appends its cones and
// face-plane bounds.
struct alignas(16) RHead3 {  // 64 B
    float lmnx, lmny, lmnz, lmxx;
    float lmxy, lmxz, rmnx, rmny;
    float rmnz, rmxx, rmxy, rmxz;
    int ref_l, ref_r;
    float lha, rha;  // SNCH: cone half angles (BVH: 0)
};
typedef RHead3 RBvh3;
struct alignas(16) RSnchTail3 {  // 80 B
    float lax, lay, laz, lrad;
    float rax, ray, raz, rrad;
    float lsin, lcos, rsin, rcos;
    float lrho, lelo, lehi, rrho;  // face-plane bounds (rho, e_lo, e_hi) per child
    float relo, rehi, pad0, pad1;
};
struct alignas(16) RSnch3 {  // 144 B
    RHead3 h;
    RSnchTail3 c;
};
static_assert(sizeof(RHead3) == 64 && sizeof(RSnch3) == 144, "3D record layout");

// Everything one sub-scene exposes to the kernels (device pointers).
struct SubDev {
    int present;
    int n_nodes, n_prims, n_refs;
    int py_sils;        // star_radius skips the SNCH when false
    float bb_min[3], bb_max[3];
    const void* nodes;  // DNode2 / DNode3
    const void* cones;  // DCone2 / DCone3
    const void* prims;  // DSeg / DTri (BVH order)
    const int* prim_index;  // BVH order -> local primitive index (2D)
    const void* sils;   // DSilV / DSilE (reference ref order)
    const int* ref_topo;  // silhouette ref -> topology silhouette id (2D: local vertex, 3D: edge)
    // 2D: the polyline's own silhouette vertices (a, b, c) and the point BVH
    // over their b (DNode2 records, leaves index `porder`), for WoStGrad
    const float* psil;
    const void* pnodes;
    const int* porder;
    int n_psil;
    // 3D: the mesh's own silhouette edges (a, b, c, d) and their length distribution
    const float* psil3;
    const float* ps_pmf;
    const float* ps_cdf;
    const int* ps_guide;
    float ps_sum, ps_norm;
    int ps_shift, n_psil3;
    // sampling (local primitive order)
    const float* cdf;
    const float* pmf;
    float pmf_sum, pmf_norm;
    const int* cdf_guide;  // (1 << guide_shift) + 1 bracket starts for the CDF search
    int guide_shift;
    const float* local_prims;  // dim*dim floats per primitive (a, b[, c])
    const int* local_vglobal;  // dim global vertex ids per primitive
    const float* vnormal;      // 2D vertex normals (local vertex order), full scene only
    const int* local_vidx;     // dim local vertex ids per primitive
    int n_local_verts;
    // child-in-parent records and the root
    const void* rec_bvh;        // RBvh2 / RBvh3
    const void* rec_snch;       // RSnch2 / RSnch3
    const void* rec_wide;       // RWide2 / RWide3 (two-level BVH records)
    const int* leaf_table;      // int2 (offset, count) per large leaf
    int root_ref_bvh, root_ref_snch;
    float root_mn[3], root_mx[3];
    int max_depth;
};

struct SceneDev {
    int dim;
    SubDev sub[3];  // 0 full, 1 Dirichlet, 2 Neumann
};

// Boundary data: g = u(p) on Dirichlet, h = n . grad u on Neumann.
#include "../../include/dwost.h"  // DW_BC_* / DW_SRC_* codes
struct BCDev {
    int code;
    float c[4];
    const float* gv;
    const float* hv;
    int src_code;
    float src_c;
    float sc[4];  // DW_SRC_LINEAR: f = sc0 + sc1 x + sc2 y + sc3 z
};
struct ParamsDev {
    int nwalks, nsteps, nsubwalks, cv, av;
    uint32_t seed;
    float eps, min_R;
    int flags;  // DW_SOLVE_*
    int64_t point_base;  // global index of point 0 (stream keys)
    int block_size, blocks_per_cu;  // launch shape (0: defaults)
    int schedule, gate;             // lane scheduling of the persistent kernels
    int node_width;                 // closest-point records: 2 binary, 4 two-level
    int walks_per_lane;             // fused solve: walk pairs per job (>= 1)
};
