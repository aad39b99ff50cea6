"""CPU: the ctypes mirrors of the C-ABI structs (rvs_amd.py, serving.py) have
the layout the C headers declare -- every field at the same offset, every
struct the same size -- checked by compiling a small C program against
include/rvs_hip.h and include/rvs_host.h with gcc."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import rvs_amd as R
import serving as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = [
    ("rvs_cam", R.RvsCam), ("rvs_xform", R.RvsXform), ("rvs_source", R.RvsSource),
    ("rvs_target", R.RvsTarget), ("rvs_view_out", R.RvsViewOut),
    ("rvs_serve_opts", S.RvsServeOpts), ("rvs_request", S.RvsRequest), ("rvs_request_stats", S.RvsRequestStats),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not on PATH")
def test_ctypes_structs_match_headers(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rvs_host.h"', "int main(void) {"]
    for cname, py in STRUCTS:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for cname, py in STRUCTS:
        assert int(got[cname]) == C.sizeof(py), "%s: C %s bytes, ctypes %d" % (cname, got[cname], C.sizeof(py))
        for f in py._fields_:
            off = getattr(py, f[0]).offset
            assert int(got["%s.%s" % (cname, f[0])]) == off, "%s.%s at %s in C, %d in ctypes" % (
                cname, f[0], got["%s.%s" % (cname, f[0])], off)
