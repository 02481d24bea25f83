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
