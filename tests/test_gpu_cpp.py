"""The reference-shaped C++ classes (include/dynamap/dynamap.hpp) driven by a
C++ program, as the reference's own test programs drive its classes."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_builder_program():
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "test_builder")
    assert os.path.exists(exe), "build it with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
