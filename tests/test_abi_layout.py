"""The ctypes structures of phoenix_amd/capi.py against the C structures of include/phx.h: the same
size and, field by field, the same offsets, as g++ lays the header out.  A field that the header
lacks fails to compile."""
import ctypes
import json
import os
import subprocess

import pytest

from phoenix_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [(capi.FreqOut, "phx_freqout"), (capi.Info, "phx_info"), (capi.Model, "phx_model"),
         (capi.Config, "phx_config"), (capi.Tune, "phx_tune"), (capi.ServerOpts, "phx_server_opts"),
         (capi.ServeStats, "phx_serve_stats")]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <cstddef>', '#include <cstdio>', '#include "phx.h"', "int main() {", '  std::printf("{");']
    for i, (cls, cname) in enumerate(PAIRS):
        sep = "," if i else ""
        lines.append(f'  std::printf("{sep}\\"{cname}\\": {{\\"size\\": %zu", sizeof({cname}));')
        for name, _ in cls._fields_:
            lines.append(f'  std::printf(", \\"{name}\\": %zu", offsetof({cname}, {name}));')
        lines.append('  std::printf("}");')
    lines += ['  std::printf("}\\n");', "  return 0;", "}"]
    src, exe = d / "layout.cpp", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


@pytest.mark.parametrize("cls,cname", PAIRS, ids=[c for _, c in PAIRS])
def test_ctypes_layout_matches_header(c_layout, cls, cname):
    c = c_layout[cname]
    assert ctypes.sizeof(cls) == c["size"], f"{cname}: ctypes {ctypes.sizeof(cls)} bytes, C {c['size']}"
    for name, _ in cls._fields_:
        assert getattr(cls, name).offset == c[name], f"{cname}.{name}"
