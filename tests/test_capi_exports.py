"""The C-ABI library loads and exports every symbol its headers declare (CPU only)."""
import ctypes
import os
import re

import feelpp_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(feel_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(fa.LIBPATH)
    names = declared("feelpp_hip.h") + declared("feelpp_host.h")
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    names = set(declared("feelpp_hip.h") + declared("feelpp_host.h"))
    assert names == set(fa._SIGS), names ^ set(fa._SIGS)


def test_last_error_is_empty_string_initially():
    assert fa.lib().feel_hip_last_error() is not None


def test_ctypes_structs_match_the_header(tmp_path):
    # the Python mirrors of the C-ABI structs must have the header's size and
    # field offsets (a field added on one side only shifts every later one)
    import shutil
    import subprocess
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"feel_hip_bilinear_desc": fa.BilinearDesc, "feel_hip_linear_desc": fa.LinearDesc,
               "feel_hip_solve_opts": fa.SolveOpts, "feel_hip_solve_data": fa.SolveData,
               "feel_hip_function": fa.Function, "feel_hip_batch_opts": fa.BatchOpts}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "feelpp_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
