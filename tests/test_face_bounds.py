"""The 3D SNCH face-plane bound (built on the host) never prunes a node that
holds an edge the fp32 silhouette test accepts: checked exhaustively over all
nodes for points on, just beside and away from the surface, on a closed noisy
icosphere and on an open cap (whose boundary edges are always silhouettes)."""

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("fb") / "face_bounds_check"
    csrc = os.path.join(ROOT, "dwost_amd", "csrc")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", csrc, "-o", str(out),
           os.path.join(ROOT, "tests", "cpp", "face_bounds_check.cpp"), os.path.join(csrc, "scene_build.cpp")]
    subprocess.run(cmd, check=True)
    return str(out)


def _write(path, V, I, T):
    with open(path, "wb") as fh:
        np.array([len(V), len(I)], np.int64).tofile(fh)
        np.ascontiguousarray(V, np.float32).tofile(fh)
        np.ascontiguousarray(I, np.int32).tofile(fh)
        np.ascontiguousarray(T, np.int32).tofile(fh)


def _run(checker, path, nq, seed):
    r = subprocess.run([checker, path, str(nq), str(seed)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    pruned = int(r.stdout.split("pruned nodes")[1].split()[0])
    return pruned


def test_closed_noisy_icosphere(checker, tmp_path):
    from dwost_amd import scenes

    spec = scenes.icosphere_scene(level=3, n_query=8)
    p = str(tmp_path / "ico.bin")
    _write(p, spec.vertices, spec.indices, spec.types)
    assert _run(checker, p, 1500, 1) > 0  # the bound does prune


def test_open_cap(checker, tmp_path):
    from dwost_amd import scenes

    spec = scenes.icosphere_scene(level=3, n_query=8)
    V, F = spec.vertices, spec.indices
    keep = V[F].mean(1)[:, 2] < 0.4  # cut the top off: an open mesh
    p = str(tmp_path / "cap.bin")
    _write(p, V, F[keep], spec.types[keep])
    _run(checker, p, 800, 2)
