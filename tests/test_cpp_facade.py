"""The C++ facade (include/graphzeppelin_amd.hpp) with the reference's class names: the
reference-shaped test program builds on the CPU and passes on the GPU."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "cpp", "driver_test")


def test_cpp_facade_test_program_is_built():
    assert os.path.exists(BIN), "tests/cpp/driver_test missing: run __graft_entry__.build()"


@pytest.mark.gpu
def test_cpp_reference_shaped_driver_suite():
    print("running tests/cpp/driver_test", flush=True)
    res = subprocess.run([BIN, os.path.join(REPO, "tests", "golden")], capture_output=True, text=True, timeout=110)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all passed" in res.stdout
