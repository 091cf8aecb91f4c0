"""Pins the FP64 oracle (oracle/) against the reference's known answers, brute force and closed forms.

CPU only.  These are the checks that make the oracle trustworthy as the parity reference for the GPU
path; see oracle/oracle.hpp for which check pins which function.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import helpers as H


@pytest.fixture(scope="module")
def wen():
    return po.DataEAM(H.WEN)


@pytest.fixture(scope="module")
def kat(wen):
    pos, basis = H.kat_lattice()
    box = po.Box(basis)
    nl = po.List(box, wen.r_cut)
    nl.rebuild(pos, 4)
    n = len(pos)
    return dict(pos=pos, basis=basis, box=box, nl=nl, types=np.zeros(n, np.uint32), frozen=np.zeros(n, np.uint8))


# ---- DataEAM / Spline -----------------------------------------------------------------------------

def test_setfl_parser_kat(wen):
    # tests/potential/EAM/data.cpp:33-43
    assert wen.ntypes == 2
    F = wen.eval("F", 0, 0, [0.0, 1.9999999999999999e-02])
    assert abs(F[0] - H.KAT_F0[0]) < 1e-14
    assert abs(F[1] - H.KAT_F0[1]) < 1e-14
    assert wen.r_cut == pytest.approx(5.3, abs=1e-12)
    assert wen.masses[0] == pytest.approx(55.847)


def test_spline_constant():
    # tests/utility/spline.cpp:27-36
    sp = po.Spline([1, 1, 1, 1], 1.0)
    assert sp.eval(0.0, 0)[0] == pytest.approx(1.0)
    assert abs(sp.eval(0.0, 1)[0]) < 1e-15
    assert abs(sp.eval(0.0, 2)[0]) < 1e-15


def test_spline_reproduces_cubics():
    # tests/utility/spline.cpp:38-83: cubic tables are reproduced away from the ends.
    rng = np.random.default_rng(1334)
    for _ in range(20):
        c = rng.uniform(-1, 1, 4)
        f = lambda x: ((c[0] * x + c[1]) * x + c[2]) * x + c[3]
        fp = lambda x: (3 * c[0] * x + 2 * c[1]) * x + c[2]
        fpp = lambda x: 6 * c[0] * x + 2 * c[1]
        dx = 1.0 / 100
        sp = po.Spline(f(np.arange(100) * dx), dx)
        x = rng.uniform(0.1, 0.9, 200)
        assert np.allclose(sp.eval(x, 0), f(x), rtol=1e-2, atol=1e-5)
        assert np.allclose(sp.eval(x, 1), fp(x), rtol=1e-2, atol=1e-5)
        assert np.allclose(sp.eval(x, 2), fpp(x), rtol=1e-2, atol=1e-5)


def test_spline_natural_interpolates_knots():
    y = np.sin(np.linspace(0, 3, 50))
    sp = po.Spline(y, 0.1)
    k = np.arange(50) * 0.1
    assert np.allclose(sp.eval(k, 0), y, atol=1e-14)
    assert abs(sp.eval(0.0, 2)[0]) < 1e-12   # natural end condition


def test_spline_out_of_bounds():
    sp = po.Spline(np.ones(10), 0.5)   # 10 values + 5 pads -> 14 segments -> x < 7
    sp.eval(6.99)
    with pytest.raises(po.OracleOutOfBounds):
        sp.eval(7.0)


# ---- EAM -----------------------------------------------------------------------------------------

def test_eam_energy_kat(wen, kat):
    E = po.energy(wen, kat["nl"], kat["types"], kat["frozen"], 4)
    assert abs(E - H.KAT_E0) <= 1e-10 * abs(H.KAT_E0)


def test_eam_gradient_kat(wen, kat):
    g = po.gradient(wen, kat["nl"], kat["types"], kat["frozen"], 4)
    assert np.linalg.norm(g[0] - H.KAT_G0) < 1e-10
    assert np.linalg.norm(g.sum(axis=0)) < 1e-9   # translation invariance


def test_eam_hessian_kat(wen, kat):
    Hm = po.hessian(wen, kat["nl"], kat["types"], kat["frozen"], 4)
    assert np.linalg.norm(Hm[:3, :3] - H.KAT_H00) < 1e-10


def test_eam_perfect_lattice_minimum(wen):
    # tests/potential/EAM/eam.cpp:149-188: zero gradient and exactly three zero modes.
    pos, basis = H.kat_lattice(erase=False)
    box = po.Box(basis)
    nl = po.List(box, wen.r_cut)
    nl.rebuild(pos, 4)
    n = len(pos)
    t, f = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    assert np.linalg.norm(po.gradient(wen, nl, t, f, 4)) < 1e-7
    Hm = po.hessian(wen, nl, t, f, 4)
    Hs = np.tril(Hm) + np.tril(Hm, -1).T
    ev = np.linalg.eigvalsh(Hs)
    near0 = np.abs(ev) < 1e-6
    assert near0.sum() == 3 and np.all(near0[:3])


def test_eam_gradient_is_energy_derivative(wen):
    rng = np.random.default_rng(7)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.05, pos.shape)
    box = po.Box(basis)
    n = len(pos)
    t = (rng.random(n) < 0.1).astype(np.uint32)   # a few H atoms: exercises phi(a,b) != phi(b,a)
    f = np.zeros(n, np.uint8)
    nl = po.List(box, wen.r_cut)
    nl.rebuild(pos)
    g = po.gradient(wen, nl, t, f, 4)
    h = 1e-5
    for a in (0, 17, 300):
        for k in range(3):
            e = []
            for s in (+1, -1):
                p = pos.copy()
                p[a, k] += s * h
                nl2 = po.List(box, wen.r_cut)
                nl2.rebuild(p)
                e.append(po.energy(wen, nl2, t, f, 4))
            assert (e[0] - e[1]) / (2 * h) == pytest.approx(g[a, k], abs=2e-7)


def test_eam_frozen_atoms(wen, kat):
    f = kat["frozen"].copy()
    f[::3] = 1
    g = po.gradient(wen, kat["nl"], kat["types"], f, 4)
    g0 = po.gradient(wen, kat["nl"], kat["types"], kat["frozen"], 4)
    assert np.all(g[::3] == 0.0)
    m = f == 0
    assert np.array_equal(g[m], g0[m])
    # Energy includes frozen atoms (eam.cpp:33-51).
    assert po.energy(wen, kat["nl"], kat["types"], f) == pytest.approx(po.energy(wen, kat["nl"], kat["types"], kat["frozen"]), rel=1e-13)


def test_eam_hessian_is_gradient_derivative(wen):
    rng = np.random.default_rng(3)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.04, pos.shape)
    box = po.Box(basis)
    n = len(pos)
    t = np.zeros(n, np.uint32)
    f = np.zeros(n, np.uint8)
    nl = po.List(box, wen.r_cut)
    nl.rebuild(pos)
    v = rng.normal(size=(n, 3))
    hv = po.hvp(wen, nl, t, f, v, 4)
    eps = 1e-5
    gs = []
    for s in (+1, -1):
        nl2 = po.List(box, wen.r_cut)
        nl2.rebuild(pos + s * eps * v)
        gs.append(po.gradient(wen, nl2, t, f, 4))
    fd = (gs[0] - gs[1]) / (2 * eps)
    assert np.max(np.abs(fd - hv)) < 1e-5 * np.max(np.abs(hv))


# ---- Neighbour list -------------------------------------------------------------------------------

def _oracle_pairs(basis, periodic, pos, rc):
    box = po.Box(basis, periodic)
    nl = po.List(box, rc)
    nl.rebuild(pos, 2)
    off, j, s, dr = nl.pairs()
    i = np.repeat(np.arange(len(pos)), np.diff(off))
    return {(int(a), int(b)): d for a, b, d in zip(i, j, dr)}, len(i)


def test_list_fuzz_vs_brute_force():
    # tests/neigh/list.cpp:167-208 (random triclinic / orthorhombic boxes, mixed periodicity).
    rng = np.random.default_rng(1334)
    for _ in range(25):
        basis, periodic, pos = H.random_box(rng)
        rc = H.min_width(basis) / 3
        want = H.brute_pairs(basis, periodic, pos, rc)
        got, count = _oracle_pairs(basis, periodic, pos, rc)
        assert count == len(got) == len(want)
        for k, d in want.items():
            assert np.linalg.norm(got[k] - d) < 1e-10


def test_list_update():
    # tests/neigh/list.cpp:210-251: displacements below the skin keep every pair.
    rng = np.random.default_rng(1111)
    for _ in range(5):
        basis, periodic, pos = H.random_box(rng, full_periodic=True)
        rc = H.min_width(basis) / 3
        skin = 0.1
        box = po.Box(basis, periodic)
        nl = po.List(box, rc + 2 * skin)
        nl.rebuild(pos, 2)
        d = rng.random(pos.shape)
        d = d / np.linalg.norm(d, axis=1, keepdims=True) * (skin * 0.99)
        nl.update(d)
        want = H.brute_pairs(basis, periodic, pos - d, rc)
        off, j, s, dr = nl.pairs(rc)
        i = np.repeat(np.arange(len(pos)), np.diff(off))
        got = {(int(a), int(b)): x for a, b, x in zip(i, j, dr)}
        assert len(got) == len(want)
        for k, x in want.items():
            assert np.linalg.norm(got[k] - x) < 1e-10


def test_grid_shape_and_cells():
    # HyperGrid: shape = 2 + floor(L / rc), cell = L / floor(L / rc) (hypergrid.hpp:43-53).
    box = po.Box(np.diag([10.0, 7.0, 5.5]))
    shape, cell = box.grid(2.0)
    assert list(shape) == [7, 5, 4]
    assert np.allclose(cell, [10 / 5, 7 / 3, 5.5 / 2])
    idx = box.cell_idx(2.0, np.array([[0.0, 0.0, 0.0], [-0.5, 6.9, 5.4], [9.99, 0.1, 100.0]]))
    # x fastest: idx = ix + 7 (iy + 5 iz), clamped into [0, shape - 1]
    assert list(idx) == [1 + 7 * (1 + 5 * 1), 0 + 7 * (3 + 5 * 2), 5 + 7 * (1 + 5 * 3)]


# ---- L-BFGS ---------------------------------------------------------------------------------------

def test_steplbfgs_first_step_is_identity():
    s = po.StepLBFGS(5)
    g = np.arange(6.0)
    assert np.array_equal(s.newton_step(np.zeros(6), g), g)


def test_steplbfgs_quadratic_exact():
    # On f = 1/2 x^T A x with exact line search L-BFGS (m >= n) reaches the minimiser in n + 1 steps;
    # with unit steps x <- x - q the two-loop reproduces A^-1 g once n independent pairs are stored.
    rng = np.random.default_rng(0)
    n = 6
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    A = Q @ np.diag(np.linspace(1, 4, n)) @ Q.T
    s = po.StepLBFGS(n + 2)
    x = rng.normal(size=n)
    for _ in range(40):
        g = A @ x
        q = s.newton_step(x, g)
        x = x - q
    assert np.linalg.norm(x) < 1e-8


def test_minimise_perfect_lattice_stays(wen):
    pos, basis = H.kat_lattice(erase=False)
    rng = np.random.default_rng(5)
    p = pos + rng.normal(0, 0.02, pos.shape)
    box = po.Box(basis)
    n = len(pos)
    r, g, failed, it = po.minimise(box, wen, p, np.zeros(n, np.uint32), np.zeros(n, np.uint8), threads=4)
    assert not failed
    assert np.linalg.norm(g) < 1e-6
    # back on the lattice up to a rigid translation
    d = r - pos
    d -= d.mean(axis=0)
    assert np.max(np.abs(d)) < 1e-4


# ---- Gaussian sampler / PRNG ----------------------------------------------------------------------

def test_xoshiro_reference_stream():
    # Xoshiro256** (Blackman & Vigna): first outputs from the state {1, 2, 3, 4}.
    out = po.xoshiro([1, 2, 3, 4], 4)
    assert list(out[:2]) == [11520, 0]
    assert out[2] == 1509978240


def test_normal_moments():
    z = po.normals([1, 3, 3, 4], 200000)
    assert abs(z.mean()) < 0.01
    assert abs(z.std() - 1) < 0.01
