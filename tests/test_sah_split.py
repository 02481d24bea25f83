"""The builder's binned-SAH split, pinned node by node (include/bvh.h:402-652).

For every internal node of small 2D and 3D scenes this recomputes, in float32,
the split the reference's computeObjectSplit chooses: per dimension with
extent >= 1e-6 (in order 0..dim-1), 8 buckets of width extent/8 over the
NODE box, bucket index int((c - pmin) / width) clamped to [0, 7], suffix boxes
swept right to left, cost nL SA(L) + nR SA(R) for b = 1..7 with a strict <
across all dimensions, split coordinate pmin + b * width; with no valid split,
the centroid box's longest axis (first maximum) at its midpoint.  The builder's
left child must then hold exactly the node's primitives with centroid < split
(performObjectSplit), or half of them when that partition is degenerate.
Planted misreadings (bins over the centroid box, a <= tie rule) must disagree
with the builder somewhere, so the check has the power to see them."""

import numpy as np
import pytest

F = np.float32
EPS = F(1e-9)  # bounding_volume.h:24


def _prim_boxes(V, I):
    """Reference boxes (expandToInclude of each vertex, inflated by epsilon) and
    centroids (segment (a+b)*0.5f, triangle (a+b+c)/3.0f; primitive.h:29-31,99-101)."""
    P = V[I]  # (n, k, dim) float32
    mn = (P - EPS).min(axis=1)
    mx = (P + EPS).max(axis=1)
    if I.shape[1] == 2:
        c = (P[:, 0] + P[:, 1]) * F(0.5)
    else:
        c = ((P[:, 0] + P[:, 1]) + P[:, 2]) / F(3.0)
    return mn.astype(F), mx.astype(F), c.astype(F)


def _surface_area(mn, mx):
    """bounding_volume.h:190-194: e = max(extent, 1e-5); 2 * sum(prod(e) / e_i),
    Eigen's fixed-size reductions evaluated as x0 op (x1 op x2)."""
    e = np.maximum(mx - mn, F(1e-5)).astype(F)
    if len(e) == 2:
        prod = F(e[0] * e[1])
        q = [F(prod / e[0]), F(prod / e[1])]
        return F(F(2.0) * F(q[0] + q[1]))
    prod = F(e[0] * F(e[1] * e[2]))
    q = [F(prod / e[k]) for k in range(3)]
    return F(F(2.0) * F(q[0] + F(q[1] + q[2])))


def _split(bmn, bmx, cen, node_mn, node_mx, variant=None):
    """computeObjectSplit (bvh.h:537-620) for the references of one node."""
    D = len(node_mn)
    best_cost, best_dim, best_coord = F(np.finfo(np.float32).max), -1, F(0)
    lo_box = node_mn
    ext_box = (node_mx - node_mn).astype(F)
    if variant == "centroid_bins":
        lo_box = cen.min(axis=0)
        ext_box = (cen.max(axis=0) - lo_box).astype(F)
    extent = (node_mx - node_mn).astype(F)
    for d in range(D):
        if extent[d] < F(1e-6):
            continue
        width = F(ext_box[d] / F(8.0))
        if not width > 0:
            continue
        idx = ((cen[:, d] - lo_box[d]) / width).astype(F)
        idx = np.clip(np.floor(idx).astype(np.int64), 0, 7)  # int() truncation, non-negative here
        counts = np.bincount(idx, minlength=8)
        bmin = np.full((8, D), np.finfo(np.float32).max, F)
        bmax = np.full((8, D), -np.finfo(np.float32).max, F)
        for b in range(8):
            m = idx == b
            if m.any():
                bmin[b] = bmn[m].min(axis=0)
                bmax[b] = bmx[m].max(axis=0)
        rmin = np.full((8, D), np.finfo(np.float32).max, F)
        rmax = np.full((8, D), -np.finfo(np.float32).max, F)
        rcnt = np.zeros(8, np.int64)
        cur_min, cur_max, cnt = rmin[0].copy(), rmax[0].copy(), 0
        for b in range(7, 0, -1):
            cur_min = np.minimum(cur_min, bmin[b])
            cur_max = np.maximum(cur_max, bmax[b])
            cnt += counts[b]
            rmin[b], rmax[b], rcnt[b] = cur_min, cur_max, cnt
        lmin, lmax, nl = rmin[0].copy(), rmax[0].copy(), 0
        for b in range(1, 8):
            lmin = np.minimum(lmin, bmin[b - 1])
            lmax = np.maximum(lmax, bmax[b - 1])
            nl += counts[b - 1]
            if nl > 0 and rcnt[b] > 0:
                cost = F(F(F(nl) * _surface_area(lmin, lmax)) + F(F(rcnt[b]) * _surface_area(rmin[b], rmax[b])))
                better = cost <= best_cost if variant == "le" else cost < best_cost
                if better:
                    best_cost, best_dim = cost, d
                    best_coord = F(lo_box[d] + F(F(b) * width))
    if best_dim == -1:  # LongestAxisCenter fallback on the centroid box (points inflated by epsilon)
        cmn, cmx = (cen - EPS).min(axis=0), (cen + EPS).max(axis=0)
        best_dim = int(np.argmax(cmx - cmn))
        best_coord = F(F(cmn[best_dim] + cmx[best_dim]) * F(0.5))
    return best_dim, best_coord


def _ranges(ex):
    """[start, end) of every node's references in BVH order, from the preorder
    layout (left child i + 1, right child i + offset; leaves hold offset, count)."""
    ints = ex["ints"]
    n = len(ints)
    rng = [None] * n

    def visit(i):
        off, cnt = int(ints[i][0]), int(ints[i][1])
        if cnt > 0:
            rng[i] = (off, off + cnt)
        else:
            a, b = visit(i + 1), visit(i + off)
            assert a[1] == b[0], "children ranges must be adjacent"
            rng[i] = (a[0], b[1])
        return rng[i]

    visit(0)
    return rng


def _check(spec_v, spec_i, sub_scene, variant=None):
    """Returns the number of internal nodes whose builder split differs from the
    restated rule (0 expected for variant None)."""
    ex = sub_scene
    V = np.ascontiguousarray(spec_v, F)
    I = np.ascontiguousarray(spec_i, np.int64)
    pidx = ex["prim_index"]
    bmn, bmx, cen = _prim_boxes(V, I[pidx])
    rng = _ranges(ex)
    ints, box = ex["ints"], ex["box"]
    D = V.shape[1]
    bad = internal = 0
    for i in range(len(ints)):
        s, e = rng[i]
        node_mn, node_mx = bmn[s:e].min(axis=0), bmx[s:e].max(axis=0)
        if variant is None:
            assert np.array_equal(box[i, :D].view(np.uint32), node_mn.view(np.uint32)), i
            assert np.array_equal(box[i, D:].view(np.uint32), node_mx.view(np.uint32)), i
        if int(ints[i][1]) > 0:
            assert e - s <= 4 or variant is not None
            continue
        assert e - s > 4
        internal += 1
        d, coord = _split(bmn[s:e], bmx[s:e], cen[s:e], node_mn, node_mx, variant)
        left = set(pidx[rng[i + 1][0]:rng[i + 1][1]].tolist())
        mine = set(pidx[s:e][cen[s:e, d] < coord].tolist())
        if mine and len(mine) < e - s:
            ok = left == mine
        else:  # degenerate partition: the first half (bvh.h:636-637)
            ok = len(left) == (e - s) // 2
        bad += not ok
    assert internal > 3
    return bad


def _scenes():
    from dwost_amd import scenes

    out = [scenes.star_scene_2d(n_query=4), scenes.holes_scene_2d(n_outer=512, n_holes=2, n_query=64),
           scenes.icosphere_scene(level=2, n_query=4), scenes.icosphere_scene(level=3, n_query=4, scene_seed=11)]
    # a flat scene (every segment on y = 0.25: the y extent is below 1e-6) with
    # duplicated segments (identical centroids force the fallback and the
    # degenerate half split)
    x = np.linspace(-1, 1, 41, dtype=F)
    V = np.stack([x, np.full_like(x, 0.25)], 1)
    I = np.stack([np.arange(40), np.arange(1, 41)], 1)
    I = np.concatenate([I, np.tile(np.array([[5, 6]]), (9, 1))])
    out.append(scenes.SceneSpec(2, V, I.astype(np.int32), np.ones(len(I), np.int32), 0))
    return out


def test_builder_split_is_the_reference_rule():
    from dwost_amd.shapes import Scene2D, Scene3D

    for spec in _scenes():
        cls = Scene2D if spec.dim == 2 else Scene3D
        s = cls(spec.vertices, spec.indices, spec.types, host_only=True)
        for sub in range(3):
            if not s.info(sub)["present"]:
                continue
            ex = s.export(sub)
            types = spec.types
            I = spec.indices if sub == 0 else spec.indices[types == (1 if sub == 1 else 0)]
            assert _check(spec.vertices, I, ex) == 0, (spec.dim, sub)


@pytest.mark.parametrize("variant", ["centroid_bins", "le"])
def test_planted_misreadings_disagree(variant):
    from dwost_amd.shapes import Scene2D, Scene3D

    total = 0
    for spec in _scenes()[:4]:
        cls = Scene2D if spec.dim == 2 else Scene3D
        s = cls(spec.vertices, spec.indices, spec.types, host_only=True)
        total += _check(spec.vertices, spec.indices, s.export(0), variant)
    print(f"{variant}: {total} internal nodes split differently")
    assert total > 0
