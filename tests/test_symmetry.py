"""Environment self-symmetries (env::site_symmetries, Catalogue::calc_self_syms catalogue.cpp:33-64 with the
GREEDY_PERM search of geometry.hpp:186-281) and Master::append_syms (find.cpp:1279-1303) of the C++ facade,
pinned by the point groups bcc Fe fixes (tests/cpp/symmetry_check.cpp): O_h's 48 signed permutation
matrices for a perfect-crystal environment, C3v's 6 beside a vacancy, the identity alone for a distorted
one; a <111> hop expands to 8 directions, a <100> hop to 6, and under C3v the hop into the vacancy to
itself while a hop along another <111> expands to its orbit of 3.  Host code only (no GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_symmetry_program(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    exe = os.path.join(tmp_path, "symmetry_check")
    src = os.path.join(ROOT, "tests", "cpp", "symmetry_check.cpp")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", exe, src], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "symmetry ok" in r.stdout
