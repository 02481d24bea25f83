"""The C ABI: header is plain C, and libdwost.so exports every declared symbol.
CPU only (no compute calls)."""

import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dwost.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:const char\*|int)\s+(dw_\w+)\s*\(", src, re.M)))


def test_header_is_plain_c(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "dwost.h"\nint main(void){ dw_scene_desc d; (void)d; return 0; }\n')
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), "-c", str(c), "-o",
                           str(tmp_path / "t.o")])


def test_library_exports_declared_symbols():
    from dwost_amd import _lib

    names = declared()
    assert len(names) >= 11
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
    assert L.dw_version() == 1


def test_host_only_scene_rejects_queries():
    import numpy as np

    from dwost_amd import DwostError
    from dwost_amd.shapes import Scene2D

    V = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    I = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    s = Scene2D(V, I, np.array([1, 1, 0, 0], np.int32), host_only=True)
    assert s.info(1)["nprims"] == 2 and s.info(2)["nprims"] == 2
    L = s.handle
    from dwost_amd import _lib

    rc = _lib.lib().dw_closest_point(L, 0, None, 0, None, None, None, None, None, None)
    assert rc == -1 and b"HOST_ONLY" in _lib.lib().dw_last_error()


def test_ctypes_structs_and_constants_match_header(tmp_path):
    """The ctypes mirror in dwost_amd/_lib.py has the C layout of every struct
    (field offsets and sizes as gcc lays them out) and the header's constants."""
    import ctypes as C

    from dwost_amd import _lib, shapes

    structs = {"dw_scene_desc": _lib.SceneDesc, "dw_boundary": _lib.Boundary, "dw_solver_params": _lib.SolverParams,
               "dw_batch_item": _lib.BatchItem, "dw_grad_params": _lib.GradParams}
    consts = {"DW_OK": _lib.DW_OK, "DW_SCENE_HOST_ONLY": _lib.SCENE_HOST_ONLY,
              "DW_SOLVE_EXACT_CONES": _lib.SOLVE_EXACT_CONES, "DW_SOLVE_WOS": _lib.SOLVE_WOS,
              "DW_BATCH_GRAPH": _lib.BATCH_GRAPH, "DW_SUB_FULL": _lib.SUB_FULL, "DW_SUB_DIRICHLET": _lib.SUB_DIRICHLET,
              "DW_SUB_NEUMANN": _lib.SUB_NEUMANN, "DW_BC_EXPCOS": shapes.BC_EXPCOS, "DW_BC_QUAD": shapes.BC_QUAD,
              "DW_BC_LINEAR": shapes.BC_LINEAR, "DW_BC_PER_VERTEX": shapes.BC_PER_VERTEX,
              "DW_SRC_ZERO": shapes.SRC_ZERO, "DW_SRC_CONST": shapes.SRC_CONST, "DW_SRC_LINEAR": shapes.SRC_LINEAR,
              "DW_RAY_REFERENCE": _lib.RAY_REFERENCE, "DW_RAY_NEAREST": _lib.RAY_NEAREST,
              "DW_NOALLOC_SCRATCH_BYTES": _lib.NOALLOC_SCRATCH_BYTES}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "dwost.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    for k in consts:
        lines.append(f'printf("const {k} %d\\n", (int)({k}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), str(src), "-o",
                           str(exe)])
    got = {}
    for line in subprocess.check_output([str(exe)], text=True).splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    for k, v in consts.items():
        assert got[("const", k)] == v, k


def test_replay_solver_must_be_same_walk_family():
    """CustomSolver reuses the forward's guard flags only when the backward
    solver draws the same walks: same fields and the same walk family."""
    import dwost_amd as dw
    from dwost_amd.solvers import _same_walks

    assert _same_walks(dw.WoSt(), dw.WoSt())
    assert not _same_walks(dw.WoSt(), dw.WoS())  # equal fields, other walk
    assert not _same_walks(dw.WoSt3D(), dw.WoS3D())
    assert not _same_walks(dw.WoSt(), dw.WoSt(nwalks=7))
