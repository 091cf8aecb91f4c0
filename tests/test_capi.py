"""The C-ABI library loads and exports every function include/flygpu.h declares (CPU, no device calls)."""
import ctypes
import os
import re

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "flygpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(flyg_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    names = declared()
    assert len(names) >= 40
    lib = ctypes.CDLL(fa.flygpu.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(fa.declared_symbols()) == declared()


def test_library_loads_without_device():
    fa.lib()
    # Creating a context on a machine without a HIP device must fail loudly, never fall back to CPU.
    try:
        fa.Context(0)
    except fa.FlyError as e:
        assert "device" in str(e).lower() or "hip" in str(e).lower()


def test_setfl_writer_roundtrip():
    # The synthetic potential file is read back by the oracle's setfl parser with the recorded forms.
    import numpy as np
    path = systems.potential_path(("Fe", "X"))
    d = po.DataEAM(path)
    assert d.ntypes == 2 and abs(d.r_cut - 5.3) < 1e-12
    r = np.array([2.0, 2.4725, 3.5, 4.9])
    S = systems.smooth_cutoff(r)
    assert np.allclose(d.eval("phi", 0, 0, r), np.exp(-2 * (r - 2.48)) * S, rtol=1e-7, atol=1e-9)
    assert np.allclose(d.eval("phi", 1, 0, r), np.exp(-2 * (r - 1.8)) * S, rtol=1e-7, atol=1e-9)
    assert np.allclose(d.eval("V", 0, 0, r), systems.morse(r, 0.4, 2.5) * S, rtol=1e-6, atol=1e-8)
    assert np.allclose(d.eval("V", 1, 0, r), systems.morse(r, 0.1, 1.8) * S, rtol=1e-6, atol=1e-8)
    rho = np.array([1.0, 9.0, 40.0])
    assert np.allclose(d.eval("F", 0, 0, rho), -np.sqrt(rho), rtol=1e-6)


def test_workload_generators():
    import numpy as np
    c1 = systems.fe_vacancy_16(1)
    assert c1.n == 8191 and abs(c1.basis[0, 0] - 16 * 2.855) < 1e-12
    c2 = systems.fe_x_alloy(2, cells=10)
    assert c2.n == 2000 - 20 + int(round(0.02 * 1980))
    assert set(np.unique(c2.types)) == {0, 1}
    c3 = systems.fe_defects_20(3)
    assert c3.n == 16000 - 20 + 40
    assert len(systems.defect_adjacent_atoms(c3)) > 100
    tet = systems.tetra_sites(1.0, 1)
    assert len(np.unique(np.round(tet, 6), axis=0)) == 12


def test_bench_refuses_several_gpus():
    # bench.py measures one GPU: --gpus above 1 exits with a message before touching a device.
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
