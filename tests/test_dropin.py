"""The C++ drop-in CSSD_Matrix (include/mfs_cssd.hpp) used the way the
reference's drivers use Library/CSSD_Matrix.h."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
EXE = os.path.join(HERE, "test_cssd_dropin")


def _build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def test_dropin_host_methods():
    _build()
    out = subprocess.run([EXE, "cpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_dropin_solvers_on_gpu():
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build does)"
    out = subprocess.run([EXE, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "OK" in out.stdout, out.stdout + out.stderr
