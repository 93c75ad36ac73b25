"""The ctypes mirrors of the C-ABI structs (cffc_amd/abi.py) match the
header's layout as a C compiler sees it: every field's offset and every
struct's size, from a small program gcc builds against
include/cffc_euler2d.h (CPU only)."""
import os
import shutil
import subprocess
import tempfile

import ctypes
import pytest

from cffc_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {"cffc_mesh_desc": abi.MeshDesc, "cffc_scheme_desc": abi.SchemeDesc, "cffc_mg_desc": abi.MGDesc}


def _c_layout():
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "cffc_euler2d.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True, capture_output=True, text=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lay = {}
    for line in out.splitlines():
        s, f, v = line.split()
        lay[(s, f)] = int(v)
    return lay


def test_ctypes_structs_match_the_header():
    lay = _c_layout()
    for cname, py in STRUCTS.items():
        assert lay[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert lay[(cname, f)] == getattr(py, f).offset, (cname, f)
