"""The ctypes mirrors in eegraph_amd/_lib.py have the layout of the C structs in
include/eeg_capi.h: a C program built with gcc prints sizeof and offsetof of every field, and
they must equal the ctypes sizes and offsets (a field added on one side only shifts the rest)."""
import ctypes as C
import json
import os
import shutil
import subprocess

import pytest

from eegraph_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {"eeg_opts": L.Opts, "eeg_stats": L.Stats, "eeg_graph_info": L.GraphInfo,
           "eeg_edgelist": L.EdgeList, "eeg_server_cfg": L.ServerCfg, "eeg_query_result": L.QueryResult}


def _c_layout(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "eeg_capi.h"', "int main(void) {",
             '    printf("{");']
    first = True
    for cname, py in STRUCTS.items():
        fields = ", ".join(f'\\"{f}\\": [%zu, %zu]' for f, _ in py._fields_)
        args = ", ".join(f"offsetof({cname}, {f}), sizeof((({cname}*)0)->{f})" for f, _ in py._fields_)
        sep = "" if first else ", "
        first = False
        lines.append(f'    printf("{sep}\\"{cname}\\": {{\\"size\\": %zu, {fields}}}", sizeof({cname}), {args});')
    lines += ['    printf("}\\n");', "    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_the_c_header(tmp_path):
    c = _c_layout(tmp_path)
    for cname, py in STRUCTS.items():
        assert c[cname]["size"] == C.sizeof(py), cname
        for f, _ in py._fields_:
            off, size = c[cname][f]
            assert off == getattr(py, f).offset, (cname, f)
            assert size == getattr(py, f).size, (cname, f)
