"""OBJ loading (diff_wost/utils/obj_loader.py:4-113): face loops, open chains,
slash forms, normalisation, and scenes built from the loaded arrays."""

import os

import numpy as np
import pytest

from dwost_amd import io

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_load_2d_faces_and_chains():
    v, e = io.load_obj_2d(os.path.join(GOLD, "square_hole.obj"))
    assert v.shape == (7, 2)
    assert e.tolist() == [[0, 1], [1, 2], [2, 3], [3, 0], [4, 5], [5, 6]]
    vn, _ = io.load_obj_2d(os.path.join(GOLD, "square_hole.obj"), normalize=True)
    assert np.isclose(np.abs(vn).max(), 1.0) and np.allclose(vn.mean(0), 0.0)


def test_load_3d_slash_forms():
    v, f = io.load_obj_3d(os.path.join(GOLD, "tetra.obj"))
    assert v.shape == (4, 3)
    assert f.tolist() == [[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]]


def test_loaded_mesh_builds_a_scene():
    import dwost_amd

    v, f = io.load_obj_3d(os.path.join(GOLD, "tetra.obj"))
    types = np.array([1, 0, 1, 0], np.int32)
    sc = dwost_amd.Scene3D(v.astype(np.float32), f.astype(np.int32), types, host_only=True)
    assert sc.info(0)["nprims"] == 4 and sc.info(1)["nprims"] == 2


@pytest.mark.skipif(not os.path.isdir("/root/reference/assets"), reason="reference assets not mounted")
def test_reference_assets_load():
    v, e = io.load_obj_2d("/root/reference/assets/bunny2d.obj")
    assert v.shape[1] == 2 and e.shape[1] == 2 and e.max() < len(v)
    v3, f3 = io.load_obj_3d("/root/reference/assets/cube.obj")
    assert v3.shape == (8, 3) and f3.shape == (12, 3)
