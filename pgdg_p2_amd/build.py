"""Build recipe for the HIP library (libpgdg.so) and the CPU oracle.

Everything is compiled in-tree so the built shared objects travel with the
source tree.  The HIP sources are cross-compiled for gfx950 with hipcc; no
GPU is needed to build.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libpgdg.so"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, so FP64 results are bit-identical to
#   the oracle's separately rounded operations.
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    # hipcc's host pass warns about the device-only target feature; keep the
    # rest of its output visible.
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        sys.stderr.write(out + "\n")


def build_product(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    if not srcs:
        raise RuntimeError("no sources under csrc/")
    cmds = []
    objs = []
    for s in srcs:
        o = OBJ / (s.name + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.suffix == ".hip" else []
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", str(s), "-o", str(o)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


HOST_TEST_SRC = ROOT / "tests" / "cpp" / "test_mpm_host.cpp"
HOST_TEST_BIN = ROOT / "tests" / "cpp" / "mpm_host_test"


def build_host_test() -> Path:
    """The C++ program that drives the step through include/pgdg_mpm.hpp
    (host code only: g++ against libpgdg.so)."""
    _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(HOST_TEST_SRC),
          f"-L{PKG}", "-lpgdg", "-Wl,-rpath,$ORIGIN/../../pgdg_p2_amd", "-o", str(HOST_TEST_BIN)])
    return HOST_TEST_BIN


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE), "-j8"])
    return ORACLE / "liboracle.so"


def build_all() -> None:
    build_oracle()
    build_product()
    build_host_test()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
