"""BVH / SNCH builder: structural invariants of the restated reference builder
(include/bvh.h, include/snch.h) and bit-exact agreement between the product's
host builder (libdwost, host-only mode) and the oracle's.  CPU only."""

import numpy as np
import pytest

from dwost_amd import scenes


def _walk(ints, nn):
    """Preorder walk: returns parent of each node and depth."""
    parent = np.full(nn, -1)
    depth = np.zeros(nn, int)
    stack = [0]
    while stack:
        i = stack.pop()
        off, nref = ints[i, 0], ints[i, 1]
        if nref == 0:
            L, R = i + 1, i + off
            for c in (L, R):
                parent[c] = i
                depth[c] = depth[i] + 1
                stack.append(c)
    return parent, depth


def _check_tree(e, spec, sub_prims_boxes, D):
    nn = e["nnodes"]
    box, ints = e["box"], e["ints"]
    parent, depth = _walk(ints, nn)
    # every node reached exactly once (preorder offsets are consistent)
    assert (parent[1:] >= 0).all()
    leaves = ints[:, 1] > 0
    # leaves partition the sorted primitive array
    covered = np.zeros(e["nprims"], int)
    for i in np.nonzero(leaves)[0]:
        covered[ints[i, 0]:ints[i, 0] + ints[i, 1]] += 1
        assert ints[i, 1] <= 4 or depth[i] == 62
    assert (covered == 1).all()
    assert np.array_equal(np.sort(e["prim_index"]), np.arange(e["nprims"]))
    # child boxes inside parent boxes; leaf boxes contain their primitives
    for i in range(1, nn):
        p = parent[i]
        assert (box[i, :D] >= box[p, :D]).all() and (box[i, D:] <= box[p, D:]).all()
    for i in np.nonzero(leaves)[0]:
        ids = e["prim_index"][ints[i, 0]:ints[i, 0] + ints[i, 1]]
        lo, hi = sub_prims_boxes
        assert (box[i, :D] <= lo[ids].min(0)).all() and (box[i, D:] >= hi[ids].max(0)).all()


def _sub_geometry(spec, sub):
    T = spec.types
    sel = np.arange(len(T)) if sub == 0 else np.nonzero(T == (1 if sub == 1 else 0))[0]
    P = spec.vertices[spec.indices[sel]]  # (n, D, D)
    return P.min(1), P.max(1)


@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_tree_invariants(oracle, which, star_spec, holes_spec, ico_spec):
    spec = dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which]
    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    D = spec.dim
    for sub in range(3):
        e = sc.export(sub)
        if not e["present"]:
            continue
        _check_tree(e, spec, _sub_geometry(spec, sub), D)


@pytest.mark.parametrize("which", ["star", "ico"])
def test_snch_cones_contain_face_normals(oracle, which, star_spec, ico_spec):
    """Every valid cone (0 <= half angle < pi) bounds the face normals of the
    silhouettes referenced in its subtree (snch.h:308-385)."""
    spec = dict(star=star_spec, ico=ico_spec)[which]
    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    D = spec.dim
    e = sc.export(0)
    V, I = spec.vertices.astype(np.float64), spec.indices
    nn, ints, cone = e["nnodes"], e["ints"], e["cone"]
    parent, _ = _walk(ints, nn)
    # subtree ranges are contiguous in preorder: end = next node whose parent chain leaves
    end = np.full(nn, nn)
    for i in range(nn):
        j = i + 1
        while j < nn:
            k, inside = j, False
            while k >= 0:
                if k == i:
                    inside = True
                    break
                k = parent[k]
            if not inside:
                break
            j += 1
        end[i] = j
    # face normals of a topology silhouette
    if D == 2:
        adj = [[] for _ in range(len(V))]
        for a, b in I:
            adj[a].append(b)
            adj[b].append(a)

        def normals(si):
            out = []
            pv = adj[si][0] if len(adj[si]) >= 1 else -1
            nx = adj[si][1] if len(adj[si]) >= 2 else -1
            if nx != -1:
                s = V[nx] - V[si]
                n = np.array([s[1], -s[0]])
                out.append(n / np.linalg.norm(n))
            if pv != -1:
                s = V[si] - V[pv]
                n = np.array([s[1], -s[0]])
                out.append(n / np.linalg.norm(n))
            return out
    else:
        emap = {}
        faces_of = {}
        for f, tri in enumerate(I):
            for k in range(3):
                u, v = tri[k], tri[(k + 1) % 3]
                key = (min(u, v), max(u, v))
                if key not in emap:
                    emap[key] = len(emap)
                faces_of.setdefault(emap[key], []).append(f)
        fnorm = np.cross(V[I[:, 1]] - V[I[:, 0]], V[I[:, 2]] - V[I[:, 0]])
        fnorm /= np.linalg.norm(fnorm, axis=1, keepdims=True)

        def normals(si):
            return [fnorm[f] for f in faces_of[si]]
    checked = 0
    for i in range(nn):
        ha = cone[i, D]
        if not (0.0 <= ha < np.pi - 1e-6):
            continue
        axis = cone[i, :D].astype(np.float64)
        for j in range(i, end[i]):
            for r in range(ints[j, 2], ints[j, 2] + ints[j, 3]):
                for n in normals(e["refs"][r]):
                    ang = np.arccos(np.clip(np.dot(axis, n), -1, 1))
                    assert ang <= ha + 1e-4
                    checked += 1
    assert checked > 0


def _same_trees(oracle, spec):
    from dwost_amd.shapes import Scene2D, Scene3D

    sc = oracle.OracleScene(spec.vertices, spec.indices, spec.types)
    ps = (Scene2D if spec.dim == 2 else Scene3D)(spec.vertices, spec.indices, spec.types, host_only=True)
    for sub in range(3):
        a, b = sc.export(sub), ps.export(sub)
        for k in ("present", "nnodes", "nrefs", "max_depth", "py_sils", "nsil_topology"):
            assert a[k] == b[k], (sub, k)
        if a["present"]:
            for k in ("box", "cone"):
                assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (sub, k)
            for k in ("ints", "prim_index", "refs"):
                assert np.array_equal(a[k], b[k]), (sub, k)


@pytest.mark.parametrize("which", ["star", "holes", "ico"])
def test_product_builder_matches_oracle_bit_exact(oracle, which, star_spec, holes_spec, ico_spec):
    _same_trees(oracle, dict(star=star_spec, holes=holes_spec, ico=ico_spec)[which])


def test_product_builder_matches_oracle_on_benchmark_scenes(oracle):
    """The full-size icosphere of the 3D workload and a spread of the uneven
    batch's scenes (64 to 16,384 segments, 0 to 4 holes): bit-exact trees."""
    from dwost_amd import scenes

    _same_trees(oracle, scenes.icosphere_scene(level=6, n_query=16))
    batch = scenes.batch_scenes_2d(n_scenes=64, seed=5, max_queries=16)
    for k in sorted(range(64), key=lambda i: len(batch[i].indices))[::8]:
        _same_trees(oracle, batch[k])


def test_builder_rejects_bad_input():
    from dwost_amd import DwostError
    from dwost_amd.shapes import Scene2D

    V = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    with pytest.raises(DwostError):
        Scene2D(V, np.array([[0, 5]], np.int32), np.array([1], np.int32), host_only=True)
    with pytest.raises(DwostError):
        Scene2D(V, np.array([[0, 1]], np.int32), np.array([7], np.int32), host_only=True)
