"""OBJ loading for scene construction (diff_wost/utils/obj_loader.py:4-113).

``load_obj_2d`` reads ``v x y ...`` vertices and turns ``f`` records into closed
loops of edges and ``l`` records into open chains; ``load_obj_3d`` reads
``v x y z`` vertices and ``f`` index lists (``i/t/n`` forms keep the vertex
index).  Indices are returned zero-based; ``normalize`` centres the vertices on
their mean and scales by the largest absolute coordinate.
"""

from __future__ import annotations

import numpy as np


def _normalize(v: np.ndarray) -> np.ndarray:
    if len(v) == 0:
        return v
    v = v - v.mean(axis=0)
    s = np.abs(v).max()
    return v / s if s > 0 else v


def _index(tok: str) -> int:
    return int(tok.split("/")[0]) - 1


def load_obj_2d(path: str, normalize: bool = False):
    """(vertices (N, 2) float64, edges (M, 2) int64)."""
    verts, edges = [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                verts.append((float(tok[1]), float(tok[2])))
            elif tok[0] == "f":
                idx = [_index(t) for t in tok[1:]]
                edges.extend((idx[k], idx[(k + 1) % len(idx)]) for k in range(len(idx)))
            elif tok[0] == "l" and len(tok) >= 3:
                idx = [_index(t) for t in tok[1:]]
                edges.extend((idx[k], idx[k + 1]) for k in range(len(idx) - 1))
    v = np.array(verts, dtype=np.float64).reshape(-1, 2)
    e = np.array(edges, dtype=np.int64).reshape(-1, 2)
    return (_normalize(v) if normalize else v), e


def load_obj_3d(path: str, normalize: bool = False):
    """(vertices (N, 3) float64, faces (M, k) int64)."""
    verts, faces = [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v" and len(tok) >= 4:
                verts.append((float(tok[1]), float(tok[2]), float(tok[3])))
            elif tok[0] == "f":
                faces.append([_index(t) for t in tok[1:]])
    v = np.array(verts, dtype=np.float64).reshape(-1, 3)
    f = np.array(faces, dtype=np.int64) if faces else np.zeros((0, 3), np.int64)
    return (_normalize(v) if normalize else v), f
