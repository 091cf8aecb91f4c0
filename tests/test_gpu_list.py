"""GPU neighbour list vs the FP64 oracle and brute force: pair sets (i, j, image shift) must be equal."""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu


def gpu_pairs(ctx, basis, periodic, pos, rc, r_query=None):
    nl = fa.List(ctx, fa.Box(basis, periodic), rc)
    nl.rebuild(pos)
    return nl, nl.pairs(r_query)


def oracle_pairs(basis, periodic, pos, rc, r_query=None):
    nl = po.List(po.Box(basis, periodic), rc)
    nl.rebuild(pos, 4)
    return nl, nl.pairs(r_query)


@pytest.mark.parametrize("seed", range(12))
def test_random_boxes_equal_pair_sets(ctx, seed):
    rng = np.random.default_rng(1000 + seed)
    basis, periodic, pos = H.random_box(rng)
    rc = H.min_width(basis) / 3
    print(f"case seed={seed} n={len(pos)} periodic={periodic.tolist()}", flush=True)
    _, (go, gj, gs, gdr) = gpu_pairs(ctx, basis, periodic, pos, rc)
    _, (oo, oj, os_, odr) = oracle_pairs(basis, periodic, pos, rc)
    assert np.array_equal(H.pair_set(go, gj, gs), H.pair_set(oo, oj, os_))
    want = H.brute_pairs(basis, periodic, pos, rc)
    i = np.repeat(np.arange(len(pos)), np.diff(go))
    assert len(want) == len(i)
    for a, b, d in zip(i, gj, gdr):
        assert np.linalg.norm(want[(int(a), int(b))] - d) < 1e-10


def test_bcc_config_pair_sets_and_counts(ctx):
    s = systems.fe_vacancy_16(seed=11)
    rc = 5.3 + 2e-3 + (1.1 ** (1 / 3) - 1) * (5.3 + 2e-3)   # dimer list cut-off
    print(f"case fe_vacancy_16 n={s.n}", flush=True)
    gnl, (go, gj, gs, gdr) = gpu_pairs(ctx, s.basis, s.periodic, s.pos, rc)
    onl, (oo, oj, os_, odr) = oracle_pairs(s.basis, s.periodic, s.pos, rc)
    assert np.array_equal(go, oo)
    assert np.array_equal(H.pair_set(go, gj, gs), H.pair_set(oo, oj, os_))
    assert gnl.num_ghosts() == onl.num_ghosts()
    # sub-cut-off query (the potential's 5.3 A) on the same list
    _, (go2, gj2, gs2, _) = (gnl, gnl.pairs(5.3))
    oo2, oj2, os2, _ = onl.pairs(5.3)
    assert np.array_equal(H.pair_set(go2, gj2, gs2), H.pair_set(oo2, oj2, os2))
    mean = np.diff(go2).mean()
    assert 56 <= mean <= 60   # bcc shells within 5.3 A hold 58 neighbours


def test_small_box_multiple_images(ctx):
    # Box narrower than 2 r_cut along x: an atom can see two images of the same neighbour.
    rng = np.random.default_rng(5)
    basis = np.diag([7.0, 12.0, 13.0])
    pos = rng.random((60, 3)) * np.diag(basis)
    rc = 4.0
    _, (go, gj, gs, _) = gpu_pairs(ctx, basis, np.ones(3, np.int32), pos, rc)
    _, (oo, oj, os_, _) = oracle_pairs(basis, np.ones(3, np.int32), pos, rc)
    assert np.array_equal(H.pair_set(go, gj, gs), H.pair_set(oo, oj, os_))


def test_update_matches_brute_force(ctx):
    rng = np.random.default_rng(77)
    for _ in range(4):
        basis, periodic, pos = H.random_box(rng, full_periodic=True)
        rc = H.min_width(basis) / 3
        skin = 0.1
        nl = fa.List(ctx, fa.Box(basis, periodic), rc + 2 * skin)
        nl.rebuild(pos)
        d = rng.random(pos.shape)
        d = d / np.linalg.norm(d, axis=1, keepdims=True) * (skin * 0.99)
        nl.update(d)
        want = H.brute_pairs(basis, periodic, pos - d, rc)
        off, j, s, dr = nl.pairs(rc)
        i = np.repeat(np.arange(len(pos)), np.diff(off))
        assert len(i) == len(want)
        for a, b, x in zip(i, j, dr):
            assert np.linalg.norm(want[(int(a), int(b))] - x) < 1e-10


def test_positions_are_canonical(ctx):
    s = systems.fe_vacancy_16(seed=2)
    shifted = s.pos + np.array([100.0, -45.7, 3.0])
    nl = fa.List(ctx, fa.Box(s.basis), 5.5)
    nl.rebuild(shifted)
    onl = po.Box(s.basis)
    assert np.array_equal(nl.positions(), onl.canon(shifted).reshape(-1, 3))


def test_errors(ctx):
    with pytest.raises(fa.FlyError):
        fa.List(ctx, fa.Box(np.diag([3.0, 3.0, 3.0])), 5.0)   # r_cut too big for the box
    nl = fa.List(ctx, fa.Box(np.diag([20.0, 20.0, 20.0])), 5.0)
    with pytest.raises(fa.FlyError):
        nl.update(np.zeros((3, 3)))   # before rebuild
    nl.rebuild(np.random.default_rng(0).random((50, 3)) * 20)
    with pytest.raises(fa.FlyError):
        nl.pairs(6.0)   # larger than the list cut-off (list.hpp:141-143)
