"""The reference's unit tests for tv_admm / vsharp / the pipeline dispatch,
restated in C++ over the header-only host mirror qsm.rs_amd/host/qsm.hpp
(tests/native/host_api_test.cpp, one function per Rust test, each citing it).
The binary is built by __graft_entry__.build() and runs the GPU library."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "native", "bin", "host_api_test")


@pytest.mark.gpu
def test_cpp_host_mirror_reference_unit_tests():
    assert os.path.exists(EXE), "tests/native/bin/host_api_test missing: run __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
