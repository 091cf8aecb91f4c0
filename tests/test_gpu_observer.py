"""Step observers of the plain minimiser and dimer (flyg_lbfgs_set_observer / flyg_dimer_set_observer): the
reference's Options::fout dump, one frame of positions at the start of every iteration before the
convergence test (lbfgs.cpp:75-77, dimer.cpp:88-90).

Bar: a run converging at iteration k reports k + 1 frames, iterations 0..k in order; the first frame is the
start, the last the returned positions (bit-equal); an observer asking to stop ends the run with
FLYG_STOPPED after exactly the frames it saw; an exception in the observer comes back from the call; with the
observer removed the result is bit-equal to the observed run's."""
import numpy as np
import pytest

import openfly_amd as fa
from oracle import pyoracle as po
from tests import helpers as H
from tests.test_gpu_saddle import small_vacancy_system

pytestmark = pytest.mark.gpu

A = 2.855324


@pytest.fixture(scope="module")
def wen(ctx):
    return fa.EAM(ctx, H.WEN), po.DataEAM(H.WEN)


def test_lbfgs_observer_frames(ctx, wen):
    pot, _ = wen
    pos, basis = small_vacancy_system(cells=5, seed=0, a=A)
    n = len(pos)
    t, f = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    m = fa.LBFGS(ctx, fa.Box(basis))
    frames = []
    m.set_observer(lambda it, x: frames.append((it, x)))
    print("case lbfgs observer, 249 atoms", flush=True)
    r, g, failed, it = m.minimise(pos, t, f, pot)
    assert not failed
    assert [k for k, _ in frames] == list(range(it + 1))
    assert np.array_equal(frames[0][1], pos)
    assert np.array_equal(frames[-1][1], r)
    m.set_observer(None)
    r2, g2, failed2, it2 = m.minimise(pos, t, f, pot)
    assert (it2, failed2) == (it, failed) and np.array_equal(r2, r) and np.array_equal(g2, g)

    seen = []
    m.set_observer(lambda k, x: seen.append(k) or k == 3)
    with pytest.raises(fa.Stopped):
        m.minimise(pos, t, f, pot)
    assert seen == [0, 1, 2, 3]

    def bad(k, x):
        if k == 2:
            raise ValueError("observer failed at 2")

    m.set_observer(bad)
    with pytest.raises(ValueError, match="observer failed at 2"):
        m.minimise(pos, t, f, pot)
    m.set_observer(None)


def test_dimer_observer_frames(ctx, wen):
    pot, data = wen
    pos, basis = small_vacancy_system(cells=5, seed=3, a=A, noise=0.0)
    n = len(pos)
    t, f = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    box_o = po.Box(basis)
    mn, _, failed, _ = po.minimise(box_o, data, pos, t, f, threads=4)
    assert not failed
    d = mn - 0.0
    d -= np.diag(basis) * np.round(d / np.diag(basis))
    centre = int(np.argmin(np.einsum("ij,ij->i", d, d)))
    nl_p = po.List(box_o, 6.0)
    nl_p.rebuild(mn, 2)
    start, axis, _ = po.perturb([2, 3, 3, 4], 0.6, 4.0, nl_p, centre, mn, f)
    dim = fa.Dimer(ctx, fa.Box(basis))
    print("case dimer observer, 249 atoms", flush=True)
    plain = dim.find_sp(start, axis, t, f, mn, pot)
    frames = []
    dim.set_observer(lambda it, x: frames.append((it, x)))
    res = dim.find_sp(start, axis, t, f, mn, pot)
    dim.set_observer(None)
    print(f"  exit {res['exit']} steps {res['steps']}, {len(frames)} frames")
    assert res["exit"] == plain["exit"] and res["steps"] == plain["steps"]
    assert np.array_equal(res["pos"], plain["pos"])
    assert np.array_equal(frames[0][1], start)
    if res["exit"] == 0:
        assert [k for k, _ in frames] == list(range(res["steps"] + 1))
        assert np.array_equal(frames[-1][1], res["pos"])
    else:
        assert [k for k, _ in frames] == list(range(len(frames)))
