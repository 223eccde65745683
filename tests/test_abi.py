"""The C ABI boundary without a GPU: libpgdg.so loads, exports every function
include/pgdg.h declares, and the ctypes mirrors of the header's structs have
the header's layout (sizes and field offsets from a gcc-compiled probe)."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

import pgdg_p2_amd as P
from pgdg_p2_amd import mls, state

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pgdg.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", " ", HEADER.read_text(), flags=re.S)
    text = re.sub(r"#[^\n]*", " ", text)
    names = re.findall(r"\b(pgdg_[a-z0-9_]+)\s*\([^;{)]*\)\s*;", text, flags=re.S)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("pgdg_create", "pgdg_step", "pgdg_run_simulation", "pgdg_mgfm_batch",
                 "pgdg_mls_ls_fit", "pgdg_split_particle_fast", "pgdg_merge_particle_fast"):
        assert must in names
    assert len(names) >= 38


def test_library_exports_every_declared_symbol():
    L = P.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.pgdg_version().decode().startswith("pgdg-mi355x")


STRUCTS = {
    "pgdg_case": state.Case,
    "pgdg_particles": state.Particles,
    "pgdg_walls": state.Walls,
    "pgdg_cells": state.Cells,
    "pgdg_scalars": state.Scalars,
    "pgdg_eos_pair": P.EosPair,
    "pgdg_mgfm_opts": P.MgfmOpts,
    "pgdg_mls_problem": mls.MlsProblem,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc for the layout probe")
def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
