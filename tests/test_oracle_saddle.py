"""Pins of the oracle's saddle search (Rotor::eff_gradient rotor.cpp:49-164, Dimer::find_sp dimer.cpp:28-171,
Master::perturb find.cpp:1147-1214) by what the mathematics fixes, independent of the restatement itself:

* quadratic surfaces E = 1/2 (x - c)^T A (x - c): the finite-difference curvature is exactly the Rayleigh
  quotient n^T A n, repeated rotations converge to the lowest eigenvector with C = lambda_min, the effective
  gradient is the reflection g - 2 (g.n) n, the dimer walks to the only stationary point c when exactly one
  eigenvalue is negative and stops `convex` when none is;
* the Mueller-Brown surface (Theor. Chim. Acta 53 (1979) 75): converged searches end on its two published
  saddle points;
* the EAM vacancy hop in bcc Fe (wen.eam.fs, whose EAM energy, gradient and Hessian are pinned by the
  reference's known answers): the converged point is stationary, has exactly one negative Hessian
  eigenvalue, its curvature and axis are those of the Hessian, and the bcc inversion symmetry of the hop
  maps the saddle onto a second saddle of the same energy;
* perturb's rules and the statistics of its Gaussians.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from openfly_amd import systems
from tests import helpers as H


def random_sym(rng, n3, eigs):
    q, _ = np.linalg.qr(rng.normal(size=(n3, n3)))
    return (q * eigs) @ q.T


# ---------------------------------------------------------------------------------------------------
# Rotor on quadratics
# ---------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(4))
def test_rotor_curvature_is_rayleigh_quotient(seed):
    rng = np.random.default_rng(seed)
    n3 = 12
    eigs = np.concatenate([[-1.5], rng.uniform(0.5, 4.0, n3 - 1)])
    A = random_sym(rng, n3, eigs)
    c = rng.normal(size=n3)
    x = rng.normal(size=n3)
    ax = rng.normal(size=n3)
    ax /= np.linalg.norm(ax)
    C, ax2, eff, grads = po.rotor_quadratic(A, c, x, ax, count_frozen=1)
    # The rotated axis stays a unit vector (it is rotated, not renormalised, rotor.cpp:124-126).
    assert abs(np.linalg.norm(ax2) - 1) < 1e-12
    # (g(x + dr n) - g(x)) / dr = A n exactly on a quadratic: the curvature is n^T A n.
    assert abs(C - ax2 @ A @ ax2) < 1e-8 * np.abs(eigs).max()
    assert lambda_ok(C, eigs)
    # g_eff = g0 - 2 (g0 . n) n (rotor.cpp:156-160) with g0 = A (x - c).
    g0 = A @ (x - c)
    assert np.max(np.abs(eff - (g0 - 2 * (g0 @ ax2) * ax2))) < 1e-11 * np.abs(g0).max()
    assert grads >= 2


def lambda_ok(C, eigs):
    return eigs.min() - 1e-9 <= C <= eigs.max() + 1e-9


def test_rotor_converges_to_lowest_mode():
    rng = np.random.default_rng(7)
    n3 = 15
    eigs = np.concatenate([[-0.8, 0.3], rng.uniform(1.0, 5.0, n3 - 2)])
    A = random_sym(rng, n3, eigs)
    w, V = np.linalg.eigh(A)
    vmin = V[:, 0]
    c = np.zeros(n3)
    x = rng.normal(size=n3)
    ax = rng.normal(size=n3)
    ax /= np.linalg.norm(ax)
    rot = dict(theta_tol=1e-7, iter_max_rot=200)
    for _ in range(5):   # the dimer calls the rotor again from the axis it left (same x here)
        C, ax, eff, _ = po.rotor_quadratic(A, c, x, ax, count_frozen=1, rotor=rot)
    assert abs(abs(ax @ vmin) - 1) < 1e-10
    assert abs(C - w[0]) < 1e-9


def test_rotor_projections():
    rng = np.random.default_rng(11)
    n3 = 9
    eigs = np.array([0.5, 1, 2, 3, 4, 5, 6, 7, 8.0])
    A = random_sym(rng, n3, eigs)
    c = np.zeros(n3)
    x = rng.normal(size=n3)
    ax = rng.normal(size=n3)
    ax /= np.linalg.norm(ax)
    g0 = A @ x
    # Positive curvature with relax_in_convex = false: g_eff = -(g0 . n) n (rotor.cpp:156-158).
    C, ax2, eff, _ = po.rotor_quadratic(A, c, x, ax, count_frozen=1, rotor=dict(relax_in_convex=0))
    assert C > 0
    assert np.max(np.abs(eff - (-(g0 @ ax2) * ax2))) < 1e-11 * np.abs(g0).max()
    # No frozen atoms: the effective gradient has its atom-mean removed (trans_proj, rotor.cpp:33-47).
    C, ax3, eff, _ = po.rotor_quadratic(A, c, x, ax, count_frozen=0)
    e3 = eff.reshape(-1, 3)
    assert np.max(np.abs(e3.mean(axis=0))) < 1e-13 * np.abs(g0).max()
    raw = (g0 - 2 * (g0 @ ax3) * ax3).reshape(-1, 3)
    shift = raw - e3
    assert np.max(np.abs(shift - shift[0])) < 1e-12 * np.abs(g0).max()


# ---------------------------------------------------------------------------------------------------
# Dimer on quadratics and Mueller-Brown
# ---------------------------------------------------------------------------------------------------

def test_dimer_quadratic_finds_the_saddle():
    rng = np.random.default_rng(3)
    n3 = 12
    eigs = np.concatenate([[-2.0], rng.uniform(0.5, 3.0, n3 - 1)])
    A = random_sym(rng, n3, eigs)
    c = rng.normal(size=n3)
    x = c + rng.normal(0, 0.3, n3)
    ax = rng.normal(size=n3)
    ax /= np.linalg.norm(ax)
    o = po.dimer_quadratic(A, c, x, ax, in_min=x, dimer=dict(use_history=0))
    assert o["exit"] == 0
    g = A @ (o["pos"] - c)
    assert np.linalg.norm(g) < 1e-5          # f2norm
    assert np.linalg.norm(o["pos"] - c) < 1e-5 / 0.5
    n = o["axis"]
    assert abs(o["curv"] - n @ A @ n) < 1e-8 * np.abs(eigs).max()
    # The rotor stops within theta_tol of the lowest mode: C - lambda_min <= spread * sin^2(theta_tol).
    assert 0 <= o["curv"] - eigs.min() <= (eigs.max() - eigs.min()) * np.sin(np.pi / 360) ** 2 + 1e-12


def test_dimer_quadratic_convex_exit():
    # No negative eigenvalue: every curvature is positive, so the search stops `convex` after
    # convex_max translation steps (dimer.cpp:95-97).
    rng = np.random.default_rng(4)
    n3 = 9
    A = random_sym(rng, n3, rng.uniform(0.5, 3.0, n3))
    x = rng.normal(size=n3)
    ax = rng.normal(size=n3)
    ax /= np.linalg.norm(ax)
    o = po.dimer_quadratic(A, np.zeros(n3), x, ax, in_min=x, dimer=dict(use_history=0))
    assert o["exit"] == 1 and o["steps"] == 3 and o["curv"] > 0


def test_dimer_history_collision():
    # A cached saddle in the direction the search is heading: `collision` at the first history check
    # (step 0, dimer.cpp:102-131).
    rng = np.random.default_rng(5)
    n3 = 12
    eigs = np.concatenate([[-2.0], rng.uniform(0.5, 3.0, n3 - 1)])
    A = random_sym(rng, n3, eigs)
    c = rng.normal(size=n3)
    x_min = c + rng.normal(0, 1.0, n3)
    x = x_min + 0.3 * (c - x_min) + rng.normal(0, 0.01, n3)
    ax = rng.normal(size=n3)
    ax /= np.linalg.norm(ax)
    d1, d2 = x - x_min, c - x_min
    angle = np.arccos(d1 @ d2 / np.linalg.norm(d1) / np.linalg.norm(d2))
    o = po.dimer_quadratic(A, c, x, ax, in_min=x_min, hist=[c], theta_tol=angle * 1.01)
    assert o["exit"] == 3 and o["steps"] == 0
    # A tighter tolerance: the direction to the search turns towards the saddle as it converges, so the
    # collision comes at a later history check (every hist_check_freq = 5 steps).
    o = po.dimer_quadratic(A, c, x, ax, in_min=x_min, hist=[c], theta_tol=angle * 0.5)
    assert o["exit"] == 3 and o["steps"] > 0 and o["steps"] % 5 == 0
    o = po.dimer_quadratic(A, c, x, ax, in_min=x_min, hist=[c], theta_tol=1e-9)
    assert o["exit"] == 0
    o = po.dimer_quadratic(A, c, x, ax, in_min=x_min, hist=[c], theta_tol=angle * 1.01,
                           dimer=dict(use_history=0))
    assert o["exit"] == 0


MB_SADDLES = np.array([[-0.82200156, 0.6243128], [0.21248658, 0.29298833]])   # E = -40.665, -72.249
MB_MINIMA = np.array([[-0.55822363, 1.44172584], [0.6234994, 0.02803776], [-0.05001082, 0.4666941]])


def mb_hessian(x, h=1e-5):
    Hm = np.zeros((3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        Hm[:, k] = (po.mueller_brown(x + e)[1] - po.mueller_brown(x - e)[1]) / (2 * h)
    return 0.5 * (Hm + Hm.T)


def test_mueller_brown_published_values():
    # Published stationary points (energies to the digits the literature gives).
    for p, e in zip(MB_SADDLES, (-40.665, -72.249)):
        E, g = po.mueller_brown(np.r_[p, 0.0])
        assert abs(E - e) < 1e-3 and np.linalg.norm(g) < 1e-4
    for p, e in zip(MB_MINIMA, (-146.700, -108.167, -80.768)):
        E, g = po.mueller_brown(np.r_[p, 0.0])
        assert abs(E - e) < 1e-3 and np.linalg.norm(g) < 1e-4


def test_dimer_mueller_brown_reaches_published_saddles():
    rng = np.random.default_rng(0)
    found = 0
    for k in range(30):
        m = np.r_[MB_MINIMA[k % 3], 0.0]
        x = m + np.r_[rng.normal(0, 0.05, 2), 0.0]
        ax = np.r_[rng.normal(size=2), 0.0]
        ax /= np.linalg.norm(ax)
        o = po.dimer_mueller_brown(x, ax, m, dimer=dict(min_trust=0.01, max_trust=0.1, convex_max=1000))
        if o["exit"] != 0:
            continue
        found += 1
        p = o["pos"]
        assert np.min(np.linalg.norm(MB_SADDLES - p[:2], axis=1)) < 1e-6
        assert np.linalg.norm(po.mueller_brown(p)[1]) < 1e-5
        w = np.linalg.eigvalsh(mb_hessian(p))
        assert (w < 0).sum() == 1
        assert abs(o["curv"] - w[0]) < 5e-3 * abs(w[0])   # forward difference over dr = 1e-3: O(dr)
    assert found >= 10


# ---------------------------------------------------------------------------------------------------
# EAM vacancy hop (wen.eam.fs)
# ---------------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def vacancy_hop():
    a, cells = 2.855324, 5
    L = cells * a
    sites = systems.bcc_sites(a, cells)
    pos0 = np.delete(sites, 0, axis=0)   # vacancy at the origin site
    basis = np.diag([L] * 3)
    n = len(pos0)
    t, f = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    data = po.DataEAM(H.WEN)
    box = po.Box(basis)
    pmin, _, failed, _ = po.minimise(box, data, pos0, t, f, threads=4)
    assert not failed
    d = sites[1:] - sites[0]
    d -= L * np.round(d / L)
    hop = int(np.argmin(np.abs(np.linalg.norm(d, axis=1) - a * np.sqrt(3) / 2)))
    u = -d[hop] / np.linalg.norm(d[hop])
    x = pmin.copy()
    x[hop] += 0.35 * a * np.sqrt(3) / 2 * u
    ax = np.zeros_like(x)
    ax[hop] = u
    o = po.find_sp(box, data, x, ax, t, f, pmin, threads=4)
    return dict(box=box, data=data, L=L, t=t, f=f, pmin=pmin, hop=hop, o=o)


def eam_eval(v, x):
    nl = po.List(v["box"], v["data"].r_cut)
    nl.rebuild(x, 4)
    # The Hessian is meaningful in its lower triangle (the reference's storage, hessian.hpp:42-148).
    Hl = np.tril(po.hessian(v["data"], nl, v["t"], v["f"], 4))
    return (po.energy(v["data"], nl, v["t"], v["f"], 4), po.gradient(v["data"], nl, v["t"], v["f"], 4),
            Hl + np.tril(Hl, -1).T)


def test_vacancy_hop_saddle_is_first_order(vacancy_hop):
    v = vacancy_hop
    o = v["o"]
    assert o["exit"] == 0
    sp = o["pos"]
    E, g, Hm = eam_eval(v, sp)
    # Stationary: |g_eff| = |g| (a reflection), so |grad E| < f2norm = 1e-5.
    assert np.linalg.norm(g) < 1e-5
    w, V = np.linalg.eigh(Hm)
    # Exactly one negative eigenvalue; the next three are the zero translation modes.
    assert (w < -1e-6).sum() == 1
    assert np.all(np.abs(w[1:4]) < 1e-6) and w[4] > 1e-3
    n = o["axis"].ravel()
    # Curvature = n^T H n within the finite-difference error O(dr); axis = lowest eigenvector.
    assert abs(o["curv"] - n @ Hm @ n) < 5e-3 * abs(w[0])
    # The rotor accepts an axis once its estimated rotation (from gradients dr = 1e-3 apart) is below
    # theta_tol; measured |cos| = 0.9983 against the exact lowest mode here.
    assert abs(abs(n @ V[:, 0]) / np.linalg.norm(n)) > 0.995
    assert abs(o["curv"] - w[0]) < 1e-2 * abs(w[0])


def test_vacancy_hop_inversion_symmetry(vacancy_hop):
    # bcc is symmetric under inversion through the midpoint of a nearest-neighbour bond, which swaps the
    # vacancy and the hopping atom.  With this potential the hop is double-humped: relaxing forward from
    # the saddle ends in the split-vacancy minimum, which must be inversion-symmetric about the hopping atom,
    # and the saddle's image under that inversion must be a second saddle of the same energy.
    v = vacancy_hop
    L, hop, o = v["L"], v["hop"], v["o"]
    sp = o["pos"]
    axv = o["axis"] / np.linalg.norm(o["axis"])
    E_sp, _, _ = eam_eval(v, sp)
    mids = []
    for sgn in (1, -1):
        pm, _, failed, _ = po.minimise(v["box"], v["data"], sp + sgn * 0.05 * axv, v["t"], v["f"], threads=4)
        assert not failed
        mids.append(pm)
    d0 = [np.linalg.norm(np.mod(m - v["pmin"] + L / 2, L) - L / 2) for m in mids]
    split = mids[int(np.argmax(d0))]
    back = mids[int(np.argmin(d0))]
    # One side is the initial minimum, up to the uniform drift of the start (no atom is frozen).
    db = np.mod(back - v["pmin"] + L / 2, L) - L / 2
    assert np.max(np.abs(db - db.mean(axis=0))) < 1e-4

    def inverted(x, centre):
        y = 2 * centre - x
        return y - L * np.floor(y / L)

    def mismatch(x, y):
        xc = x - L * np.floor(x / L)
        dd = xc[:, None, :] - y[None, :, :]
        dd -= L * np.round(dd / L)
        return np.max(np.min(np.linalg.norm(dd, axis=2), axis=1))

    c = split[hop]
    assert mismatch(split, inverted(split, c)) < 1e-4
    sp2 = inverted(sp, c)
    E2, g2, _ = eam_eval(v, sp2)
    assert abs(E2 - E_sp) < 1e-9 * abs(E_sp)
    assert np.linalg.norm(g2) < 1e-5


# ---------------------------------------------------------------------------------------------------
# perturb
# ---------------------------------------------------------------------------------------------------

def test_perturb_rules():
    a, cells = 2.855324, 4
    L = cells * a
    pos = np.delete(systems.bcc_sites(a, cells), 0, axis=0)
    n = len(pos)
    frozen = np.zeros(n, np.uint8)
    frozen[::7] = 1
    centre = 3
    frozen[centre] = 0
    r_pert = 4.0
    nl = po.List(po.Box(np.diag([L] * 3)), 6.0)
    nl.rebuild(pos, 2)
    d = pos - pos[centre]
    d -= L * np.round(d / L)
    r = np.linalg.norm(d, axis=1)
    inside = (r < r_pert) & (frozen == 0)
    for seed in range(5):
        x, ax, sd = po.perturb([seed + 1, 3, 3, 4], 0.6, r_pert, nl, centre, pos, frozen)
        assert sd > 0
        moved = np.any(x != pos, axis=1)
        assert np.array_equal(moved, inside)                  # the non-frozen atoms within r_pert (+ the centre)
        assert np.array_equal(np.any(ax != 0, axis=1), inside)
        assert abs(np.linalg.norm(ax) - 1) < 1e-14            # unit axis
    with pytest.raises(po.OracleError):
        fz = frozen.copy()
        fz[centre] = 1
        po.perturb([1, 3, 3, 4], 0.6, r_pert, nl, centre, pos, fz)


def test_perturb_statistics():
    # Displacements are N(0, sd) scaled by (1 - r / r_pert), the centre's unscaled; sd ~ N(0.6, 0.2)
    # truncated at 0 (find.cpp:1154-1178).  Checked through the moments of many draws.
    a, cells = 2.855324, 4
    L = cells * a
    pos = systems.bcc_sites(a, cells)
    n = len(pos)
    frozen = np.zeros(n, np.uint8)
    centre = 0
    r_pert = 4.0
    nl = po.List(po.Box(np.diag([L] * 3)), 6.0)
    nl.rebuild(pos, 2)
    d = pos - pos[centre]
    d -= L * np.round(d / L)
    r = np.linalg.norm(d, axis=1)
    scale = np.where(np.arange(n) == centre, 1.0, 1 - r / r_pert)
    inside = r < r_pert
    z, sds = [], []
    # Full-entropy seeds, as the reference seeds its generators from std::random_device (find.cpp:186-193):
    # seeds that differ in one word only give Xoshiro256** the same first output.
    seeds = np.random.default_rng(0).integers(1, 2 ** 63, size=(400, 4), dtype=np.uint64)
    for seed in seeds:
        x, ax, sd = po.perturb(seed, 0.6, r_pert, nl, centre, pos, frozen)
        sds.append(sd)
        z.append(((x - pos)[inside] / scale[inside, None] / sd).ravel())
    z = np.concatenate(z)
    sds = np.array(sds)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    # Truncated N(0.6, 0.2): the truncation at 0 is 3 sigma away, so the moments are those of the normal.
    assert abs(sds.mean() - 0.6) < 3 * 0.2 / np.sqrt(len(sds)) + 1e-3
    assert abs(sds.std() - 0.2) < 0.02
