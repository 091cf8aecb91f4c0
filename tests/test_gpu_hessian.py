"""Mass-weighted Hessian spectrum of the non-frozen atoms on the device (EAM::hessian eam.cpp:122-233,
Hessian::mass_weight hessian.hpp:136-143, Hessian::eigenvalues hessian.hpp:86-93) against numpy's eigvalsh of
the FP64 oracle's dense Hessian (pinned by the reference's H(0,0) known answer), restricted to the same block
and weighted by the oracle's masses: eigenvalues within 1e-9 of the largest in magnitude.  At the vacancy-hop
saddle the spectrum has exactly one negative mode and the three translation zero modes (find.cpp:849-866)."""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

A_WEN = 2.855324


@pytest.fixture(scope="module")
def wen(ctx):
    return fa.EAM(ctx, H.WEN), po.DataEAM(H.WEN)


def oracle_block(data, basis, pos, types, frozen):
    nl = po.List(po.Box(basis), data.r_cut)
    nl.rebuild(pos, 4)
    Hl = np.tril(po.hessian(data, nl, types, frozen, 4))   # the lower triangle holds the Hessian
    Hs = Hl + np.tril(Hl, -1).T
    dof = np.nonzero(np.repeat(frozen == 0, 3))[0]
    m = np.repeat(data.masses[types], 3)[dof]
    B = Hs[np.ix_(dof, dof)] / np.sqrt(np.outer(m, m))
    return np.linalg.eigvalsh(B)


def vacancy_cell(cells):
    pos = np.delete(systems.bcc_sites(A_WEN, cells), 0, axis=0)
    return pos, np.diag([cells * A_WEN] * 3)


@pytest.mark.parametrize("frozen_radius", [5.0, None])
def test_hessian_spectrum_vs_oracle(ctx, wen, frozen_radius):
    pot, data = wen
    rng = np.random.default_rng(4)
    pos, basis = vacancy_cell(5)
    pos = pos + rng.normal(0, 0.05, pos.shape)
    n = len(pos)
    t = np.zeros(n, np.uint32)
    f = np.zeros(n, np.uint8)
    if frozen_radius:
        L = np.diag(basis)
        d = pos - pos[0]
        d -= L * np.round(d / L)
        f = (np.einsum("ij,ij->i", d, d) > frozen_radius ** 2).astype(np.uint8)
    print(f"case hessian n={n} mobile={int((f == 0).sum())}", flush=True)
    nl = fa.List(ctx, fa.Box(basis), pot.r_cut())
    nl.rebuild(pos)
    w = pot.hessian_eigenvalues(t, f, nl)
    wo = oracle_block(data, basis, pos, t, f)
    assert len(w) == 3 * int((f == 0).sum())
    err = np.max(np.abs(w - wo))
    print(f"max |dlambda| {err:.3e} of max |lambda| {np.abs(wo).max():.3e}")
    assert err <= 1e-9 * np.abs(wo).max()
    if frozen_radius is None:   # translation invariance: three zero modes
        assert np.sum(np.abs(w) < 1e-6) == 3


def test_saddle_spectrum_and_prefactor(ctx, wen):
    # The vacancy hop's saddle (found by the oracle's dimer from a nudge towards the vacancy) and the relaxed
    # minimum: one negative mode at the saddle, none at the minimum, three zero modes each; the prefactor
    # sqrt(prod lambda_min / prod lambda_sp+) / sqrt(2 pi amu) lands in the THz range.
    pot, data = wen
    pos, basis = vacancy_cell(5)
    n = len(pos)
    t = np.zeros(n, np.uint32)
    f = np.zeros(n, np.uint8)
    box = po.Box(basis)
    mn, _, failed, _ = po.minimise(box, data, pos, t, f, threads=4)
    assert not failed
    L = np.diag(basis)
    d = mn - systems.bcc_sites(A_WEN, 5)[0]
    d -= L * np.round(d / L)
    hop = int(np.argmin(np.einsum("ij,ij->i", d, d)))
    u = -d[hop] / np.linalg.norm(d[hop])
    x = mn.copy()
    x[hop] += 0.35 * A_WEN * np.sqrt(3) / 2 * u
    ax = np.zeros_like(x)
    ax[hop] = u
    o = po.find_sp(box, data, x, ax, t, f, mn, threads=4)
    assert o["exit"] == 0
    nl = fa.List(ctx, fa.Box(basis), pot.r_cut())
    out = {}
    for name, p in (("min", mn), ("sp", o["pos"])):
        nl.rebuild(p)
        w = pot.hessian_eigenvalues(t, f, nl)
        wo = oracle_block(data, basis, p, t, f)
        assert np.max(np.abs(w - wo)) <= 1e-9 * np.abs(wo).max()
        out[name] = fa.log_prod_eigen(w, 1e-6)
        print(f"{name}: lowest {w[:5]}, (log prod, zero, negative) = {out[name]}")
    assert out["min"][1] == 3 and out["min"][2] == 0
    assert out["sp"][1] == 3 and out["sp"][2] == 1
    A = fa.prefactor(out["min"][0], out["sp"][0])
    print(f"prefactor {A:.3e} Hz")
    assert 1e11 < A < 1e15
