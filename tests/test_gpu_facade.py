"""The C++ drop-in facade (include/libfly/gpu/facade.hpp): reference-style code compiled against it runs
the reference's EAM known-answer test, an L-BFGS relaxation and a dimer vacancy hop on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def test_facade_program():
    exe = os.path.join(ROOT, "examples", "facade_test")
    assert os.path.exists(exe), "examples/facade_test is built by __graft_entry__.build()"
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "wen.eam.fs")], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "facade ok" in r.stdout
