"""The C++ host class pgdg::MPM (include/pgdg_mpm.hpp): the reference's
member names and call sequence (S3:mpm.h:160-194, S3:solver.cpp:30-64)
forwarding to the C ABI.  tests/cpp/test_mpm_host.cpp is built by build()
with g++ against libpgdg.so; on a GPU it steps the 2-D triple point through
the class and compares every array with pgdg_step / pgdg_run_simulation."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "mpm_host_test"
HDR = Path(__file__).resolve().parent.parent / "include" / "pgdg_mpm.hpp"


def test_host_class_header_compiles_alone(tmp_path):
    # the header needs only pgdg.h and the C++ standard library
    src = tmp_path / "t.cpp"
    src.write_text('#include "pgdg_mpm.hpp"\nint main() { pgdg::communicator w; return w.rank(); }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{HDR.parent}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_host_class_drives_the_step(cuda):
    assert BIN.exists(), "tests/cpp/mpm_host_test not built (run __graft_entry__.build())"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mpm host ok" in r.stdout
