"""The ctypes mirrors in cuckoo_amd.py against the C structs of include/cuckoo_amd.h: every field's
offset and every struct's size, read from a probe compiled with the host compiler (no GPU)."""
import ctypes
import os
import shutil
import subprocess

import pytest

import cuckoo_amd as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [("cf_stats", cf.Stats), ("cf_instance", cf.Instance), ("cf_exact", cf.Exact), ("cf_tune", cf.Tune),
         ("cf_serve_stats", cf.ServeStats)]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no host C compiler")
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cuckoo_amd.h"', "int main(void) {"]
    for cname, py in PAIRS:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'    printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["    return 0;", "}"]
    src = d / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "probe"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.parametrize("cname,py", PAIRS, ids=[p[0] for p in PAIRS])
def test_struct_layout_matches_header(c_layout, cname, py):
    assert c_layout[(cname, "size")] == ctypes.sizeof(py)
    for f in py._fields_:
        assert c_layout[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
