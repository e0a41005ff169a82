"""The reference's test sequence written in C++ against include/sem.hpp
(namespace schro_hip), run as a separate program on the GPU."""
import os
import subprocess

import pytest

from conftest import ROOT, SMALL_MESH

pytestmark = pytest.mark.gpu


def test_cpp_drop_in_suite():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_drop_in")
    assert os.path.exists(exe), "build first: python -m schrodingersem_amd.build"
    r = subprocess.run([exe, SMALL_MESH], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "14 of 14 reference tests run" in r.stdout and "all tests passed! :)" in r.stdout
