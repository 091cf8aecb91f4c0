"""Master::perturb (find.cpp:1147-1214) in the library (flyg_perturb, host code) against the oracle's, whose
rules and statistics tests/test_oracle_saddle.py pins: the same seeds give bit-identical starts, axes and
standard deviations, and the generator state continues the same stream.  No GPU is touched."""
import numpy as np
import pytest

import openfly_amd as fa
from oracle import pyoracle as po
from openfly_amd import systems


def both(basis, periodic, pos, frozen, centre, seed, stddev=0.6, r_pert=4.0):
    nl = po.List(po.Box(basis, periodic), max(r_pert, 1e-3) + 0.5)
    nl.rebuild(pos, 2)
    xo, ao, so = po.perturb(seed, stddev, r_pert, nl, centre, pos, frozen)
    st = np.array(seed, dtype=np.uint64)
    xg, ag, sg = fa.perturb(fa.Box(basis, periodic), pos, frozen, centre, st, stddev, r_pert)
    return (xo, ao, so), (xg, ag, sg), st


@pytest.mark.parametrize("case", range(6))
def test_perturb_bit_equal_to_oracle(case):
    rng = np.random.default_rng(100 + case)
    a, cells = 2.855324, 5
    pos = systems.bcc_sites(a, cells)
    pos = np.delete(pos, rng.integers(len(pos)), axis=0) + rng.normal(0, 0.05, (len(pos) - 1, 3))
    pos += rng.uniform(-20, 20, 3)   # not canonical: the library wraps as the neighbour list does
    basis = np.diag([cells * a] * 3)
    periodic = (1, 1, 1) if case < 4 else (1, 0, 1)
    n = len(pos)
    frozen = (rng.random(n) < 0.3).astype(np.uint8)
    centre = int(rng.integers(n))
    frozen[centre] = 0
    seed = [int(v) for v in rng.integers(1, 2 ** 63, 4, dtype=np.uint64)]
    (xo, ao, so), (xg, ag, sg), st = both(basis, periodic, pos, frozen, centre, seed)
    assert so == sg
    assert np.array_equal(xo, xg) and np.array_equal(ao, ag)


def test_perturb_stream_continues():
    # Two calls from one state equal the oracle's generator advanced by the first call's draws: the second
    # start differs from the first and matches a fresh oracle call seeded with the advanced state.
    a, cells = 2.855324, 4
    pos = systems.bcc_sites(a, cells)
    basis = np.diag([cells * a] * 3)
    n = len(pos)
    f = np.zeros(n, np.uint8)
    st = np.array([11, 22, 33, 44], dtype=np.uint64)
    x1, a1, _ = fa.perturb(fa.Box(basis), pos, f, 5, st)
    mid = [int(v) for v in st]
    x2, a2, s2 = fa.perturb(fa.Box(basis), pos, f, 5, st)
    nl = po.List(po.Box(basis), 4.5)
    nl.rebuild(pos, 2)
    xo, ao, so = po.perturb(mid, 0.6, 4.0, nl, 5, pos, f)
    assert not np.array_equal(x1, x2)
    assert np.array_equal(x2, xo) and np.array_equal(a2, ao) and s2 == so


def test_perturb_small_box_images():
    # A box narrower than 2 r_pert: an atom within r_pert through two images is displaced once per image.
    rng = np.random.default_rng(7)
    basis = np.diag([7.0, 12.0, 12.0])
    pos = (basis @ rng.random((3, 60))).T
    f = np.zeros(60, np.uint8)
    seed = [5, 6, 7, 8]
    (xo, ao, so), (xg, ag, sg), _ = both(basis, (1, 1, 1), pos, f, 3, seed)
    assert np.array_equal(xo, xg) and np.array_equal(ao, ag) and so == sg


def test_perturb_rejects_frozen_centre():
    pos = systems.bcc_sites(2.855324, 4)
    f = np.zeros(len(pos), np.uint8)
    f[2] = 1
    with pytest.raises(fa.FlyError):
        fa.perturb(fa.Box(np.diag([4 * 2.855324] * 3)), pos, f, 2, np.array([1, 2, 3, 4], np.uint64))


def test_theta_tol_schedule():
    # ((30 - 2.5) exp(-0.02 tot) + 2.5) degrees (find.cpp:692).
    for tot in (0, 1, 10, 100, 1000):
        want = ((30 - 2.5) * np.exp(-0.02 * tot) + 2.5) * np.pi / 180
        assert abs(fa.theta_tol(tot) - want) < 1e-15
