"""flyg_eam_gradient_noalloc (the allocation-free EAM::gradient) against flyg_eam_gradient, called through
the C-ABI with device pointers: bit-identical gradients on two calls in a row, the second on a smaller
system, in both precisions; sizes beyond the array are refused with an error before any work."""
import ctypes as C

import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _fn():
    f = fa.lib().flyg_eam_gradient_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    f.restype = C.c_int
    return f


def device_system(torch, s, dev):
    t = torch.tensor(s.types.astype(np.int64), dtype=torch.int64).to(torch.uint32).to(dev)
    f = torch.tensor(s.frozen, dtype=torch.uint8, device=dev)
    return t, f


@pytest.mark.parametrize("precision", [fa.FP64, fa.MIXED])
def test_noalloc_gradient_equals_gradient(ctx, precision):
    torch = pytest.importorskip("torch")
    dev = torch.device(f"cuda:{ctx.device}")
    pot = fa.EAM(ctx, systems.potential_path(("Fe", "X")))
    pot.set_precision(precision)
    noalloc = _fn()
    # Two calls in a row, the second with fewer atoms (the shared array keeps the first call's contents).
    for cells in (24, 11):
        s = systems.fe_x_alloy(seed=3, cells=cells)
        s.frozen[::9] = 1
        print(f"case noalloc fe_x_alloy({cells}^3) n={s.n} precision={precision}", flush=True)
        nl = fa.List(ctx, fa.Box(s.basis), 5.3 + 0.6)
        nl.rebuild(s.pos)
        t, f = device_system(torch, s, dev)
        g_ref = torch.zeros((s.n, 3), dtype=torch.float64, device=dev)
        g_new = torch.full((s.n, 3), float("nan"), dtype=torch.float64, device=dev)
        pot.gradient(t, f, nl, g_ref)
        rc = noalloc(pot.h, nl.h, C.c_void_p(t.data_ptr()), C.c_void_p(f.data_ptr()), s.n, C.c_void_p(g_new.data_ptr()))
        assert rc == 0, fa.lib().flyg_ctx_last_error(ctx.h)
        ctx.synchronize()
        a, b = g_ref.cpu().numpy(), g_new.cpu().numpy()
        assert np.array_equal(a, b)
        assert np.all(b[s.frozen == 1] == 0.0) and np.abs(b).max() > 0


def test_noalloc_gradient_refuses_large_systems(ctx):
    torch = pytest.importorskip("torch")
    dev = torch.device(f"cuda:{ctx.device}")
    pot = fa.EAM(ctx, H.WEN)
    s = systems.defect_system(41, 0, 0, 1)   # 137,842 atoms > 131,072
    nl = fa.List(ctx, fa.Box(s.basis), pot.r_cut())
    nl.rebuild(s.pos)
    t, f = device_system(torch, s, dev)
    g = torch.zeros((s.n, 3), dtype=torch.float64, device=dev)
    rc = _fn()(pot.h, nl.h, C.c_void_p(t.data_ptr()), C.c_void_p(f.data_ptr()), s.n, C.c_void_p(g.data_ptr()))
    assert rc == 1   # FLYG_INVALID_ARG
    assert b"131072" in fa.lib().flyg_ctx_last_error(ctx.h)
    ctx.synchronize()
    assert float(g.abs().max()) == 0.0   # nothing ran
