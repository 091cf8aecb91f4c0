"""Master::find_batch semantics on the device (find.cpp:684-755) against the FP64 oracle:

* the saddle-history test (Dimer::find_sp's hist_sp with find_batch's theta_tol, dimer.cpp:102-131) ends
  the same searches as `collision` at the same step, in the batch and in the plain dimer;
* searches queued with flyg_batch_submit_perturbed start from Master::perturb (bit-equal to explicit
  starts made by flyg_perturb, which equals the oracle's perturb);
* a search whose start evaluates a spline outside its table is re-perturbed from the same stream and run
  again, at most three times (find.cpp:706-723); the failed starts raise Spline::OutOfBounds in the oracle.
Potential: the reference's wen.eam.fs; for the retry test a copy whose F(rho) table ends at rho = 30
(the perfect crystal sits at 26.3), so compressed starts leave it.
"""
import os

import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

A_WEN = 2.855324
RADIUS = 6.0
LONG = 100   # oracle steps beyond which a history search's exit is reported, not compared


def vacancy_cell(cells=7):
    pos = systems.bcc_sites(A_WEN, cells)
    vac = pos[0].copy()
    pos = np.delete(pos, 0, axis=0)
    basis = np.diag([cells * A_WEN] * 3)
    return pos, basis, vac


@pytest.fixture(scope="module")
def setup(ctx):
    pos, basis, vac = vacancy_cell()
    n = len(pos)
    t = np.zeros(n, np.uint32)
    data = po.DataEAM(H.WEN)
    box = po.Box(basis)
    mn, _, failed, _ = po.minimise(box, data, pos, t, np.zeros(n, np.uint8), threads=8)
    assert not failed
    L = np.diag(basis)
    # canonical, so explicit (dense) and library-perturbed (sparse) start states coincide bit for bit
    mn = mn - L * np.floor(mn / L)
    mn = np.where(mn >= L, mn - L, mn)
    d = mn - vac
    d -= L * np.round(d / L)
    near = np.argsort(np.einsum("ij,ij->i", d, d))[:8]
    pot = fa.EAM(ctx, H.WEN)
    return dict(mn=mn, basis=basis, n=n, t=t, data=data, box=box, near=near, pot=pot, L=L)


def frozen_mask(S, centre, radius=RADIUS):
    d = S["mn"] - S["mn"][centre]
    d -= S["L"] * np.round(d / S["L"])
    return (np.einsum("ij,ij->i", d, d) > radius * radius).astype(np.uint8)


def seeds(k, base=0):
    return np.random.default_rng(1234 + base).integers(1, 2 ** 63, size=(k, 4), dtype=np.uint64)


def explicit_starts(S, centres, prng, attempts=1):
    """flyg_perturb starts (the attempts-th draw from each stream), with the batch's frozen masks."""
    out = []
    for c, g in zip(centres, prng):
        st = g.copy()
        f = frozen_mask(S, c)
        for _ in range(attempts):
            x, ax, _ = fa.perturb(fa.Box(S["basis"]), S["mn"], f, c, st)
        out.append((x, ax, f))
    return out


def test_submit_perturbed_equals_explicit_starts(ctx, setup):
    S = setup
    centres = [int(S["near"][k % 8]) for k in range(6)]
    prng = seeds(6)
    print("case submit_perturbed vs explicit, 6 searches", flush=True)
    b1 = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], precision=fa.FP64)
    g = prng.copy()
    ids1 = b1.submit_perturbed(centres, [RADIUS] * 6, g)
    b1.run()
    st = explicit_starts(S, centres, prng)
    b2 = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], precision=fa.FP64)
    ids2 = b2.submit(centres, [RADIUS] * 6, np.stack([x for x, _, _ in st]), np.stack([a for _, a, _ in st]))
    b2.run()
    for i, j, (x, ax, f), c in zip(ids1, ids2, st, centres):
        r1, r2 = b1.result(i, positions=True), b2.result(j, positions=True)
        assert r1["exit"] == r2["exit"] and r1["steps"] == r2["steps"] and r1["barrier"] == r2["barrier"]
        assert np.array_equal(r1["pos"], r2["pos"])
        o = po.find_sp(S["box"], S["data"], x, ax, S["t"], f, S["mn"], threads=8)
        print(f"  centre {c}: gpu exit {r1['exit']} steps {r1['steps']} | oracle exit {o['exit']} steps {o['steps']}")
        assert r1["exit"] == o["exit"]
    # the generator states advanced by one perturbation each
    for k in range(6):
        s2 = prng[k].copy()
        fa.perturb(fa.Box(S["basis"]), S["mn"], frozen_mask(S, centres[k]), centres[k], s2)
        assert np.array_equal(g[k], s2)


def test_history_collision_matches_oracle(ctx, setup):
    # Converged saddles of a first batch become the history of a second batch from new starts at the same
    # sites; every search the oracle ends within LONG steps must end as the oracle's does with the same
    # history and theta_tol, and collisions at the same step.  Longer searches wander through convex
    # regions where trajectories part at the rounding level (the device's own result for one such start
    # changes with the region a relaunch continues in), so they are reported, not compared.
    S = setup
    centres = [int(S["near"][k % 8]) for k in range(8)]
    b0 = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], precision=fa.FP64)
    ids0 = b0.submit_perturbed(centres, [RADIUS] * 8, seeds(8, 1))
    b0.run()
    hist = [b0.result(i, positions=True)["pos"] for i in ids0 if b0.result(i)["exit"] == 0]
    assert len(hist) >= 2
    prng = seeds(8, 2)
    st = explicit_starts(S, centres, prng)
    collisions = compared = 0
    for tot in (0, 40):
        theta = fa.theta_tol(tot)
        print(f"case history: {len(hist)} saddles, theta_tol {np.degrees(theta):.2f} deg", flush=True)
        b = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], precision=fa.FP64)
        ids = b.submit(centres, [RADIUS] * 8, np.stack([x for x, _, _ in st]), np.stack([a for _, a, _ in st]),
                       theta_tol=theta, hist=hist)
        b.run()
        for i, (x, ax, f), c in zip(ids, st, centres):
            r = b.result(i)
            o = po.find_sp(S["box"], S["data"], x, ax, S["t"], f, S["mn"], hist=hist, theta_tol=theta, threads=8)
            print(f"  centre {c}: gpu exit {r['exit']} steps {r['steps']} launches {r['launches']} | oracle exit "
                  f"{o['exit']} steps {o['steps']}")
            if o["steps"] > LONG:
                continue
            assert r["exit"] == o["exit"]
            compared += 1
            if o["exit"] == 3:
                assert r["steps"] == o["steps"]
                collisions += 1
    assert collisions >= 1
    assert compared >= 12


def test_plain_dimer_history_collision(ctx, setup):
    # The plain form (flyg_dimer_find_sp with hist_sp) ends with the oracle's exit and step count.
    S = setup
    c = int(S["near"][0])
    f = np.zeros(S["n"], np.uint8)
    x0, a0, _ = fa.perturb(fa.Box(S["basis"]), S["mn"], f, c, seeds(1, 3)[0].copy())
    o0 = po.find_sp(S["box"], S["data"], x0, a0, S["t"], f, S["mn"], threads=8)
    assert o0["exit"] == 0
    x1, a1, _ = fa.perturb(fa.Box(S["basis"]), S["mn"], f, c, seeds(1, 4)[0].copy())
    theta = fa.theta_tol(0)
    o = po.find_sp(S["box"], S["data"], x1, a1, S["t"], f, S["mn"], hist=[o0["pos"]], theta_tol=theta, threads=8)
    dim = fa.Dimer(ctx, fa.Box(S["basis"]))
    r = dim.find_sp(x1, a1, S["t"], f, S["mn"], S["pot"], hist_sp=[o0["pos"]], theta_tol=theta)
    print(f"plain: gpu exit {r['exit']} steps {r['steps']} | oracle exit {o['exit']} steps {o['steps']}")
    assert r["exit"] == o["exit"] and r["steps"] == o["steps"]


def truncated_wen(path_out, rho_max):
    """wen.eam.fs with every F(rho) table cut to rho < rho_max (same spacing)."""
    lines = open(H.WEN).read().split("\n")
    head = lines[:5]
    toks = " ".join(lines[5:]).split()
    nrho, drho, nr, dr, rc = head[4].split()
    nrho, nr = int(nrho), int(nr)
    ntypes = int(head[3].split()[0])
    keep = int(rho_max / float(drho))
    out = head[:4] + [f"{keep} {drho} {nr} {dr} {rc}"]
    p = 0

    def chunk(vals):
        return [" ".join(vals[i:i + 5]) for i in range(0, len(vals), 5)]

    for _ in range(ntypes):
        out.append(" ".join(toks[p:p + 4]))
        p += 4
        out += chunk(toks[p:p + keep])
        p += nrho
        out += chunk(toks[p:p + ntypes * nr])
        p += ntypes * nr
    out += chunk(toks[p:])
    open(path_out, "w").write("\n".join(out) + "\n")


def test_spline_oob_retries(ctx, setup, tmp_path):
    S = setup
    path = os.path.join(tmp_path, "wen_rho30.eam.fs")
    truncated_wen(path, 30.0)
    pot = fa.EAM(ctx, path)
    data = po.DataEAM(path)
    k = 16
    centres = [int(S["near"][q % 8]) for q in range(k)]
    prng = seeds(k, 5)
    print("case spline OOB retries, 16 searches, F table to rho = 30", flush=True)
    b = fa.Batch(ctx, pot, fa.Box(S["basis"]), S["mn"], S["t"], precision=fa.FP64)
    ids = b.submit_perturbed(centres, [RADIUS] * k, prng.copy(), stddev=0.6)
    b.run()
    retried = 0
    for q, i in enumerate(ids):
        r = b.result(i)
        print(f"  search {q}: exit {r['exit']} retries {r['retries']}")
        assert 0 <= r["retries"] <= 3
        assert r["exit"] != 6 or r["retries"] == 3
        retried += r["retries"] > 0
        # every earlier start fails in the oracle too; the last one is the result
        for att in range(1, r["retries"] + 2):
            (x, ax, f), = explicit_starts(S, [centres[q]], prng[q:q + 1], attempts=att)
            if att <= r["retries"] or r["exit"] == 6:
                with pytest.raises(po.OracleOutOfBounds):
                    po.find_sp(S["box"], data, x, ax, S["t"], f, S["mn"], threads=8)
            else:
                o = po.find_sp(S["box"], data, x, ax, S["t"], f, S["mn"], threads=8)
                assert o["exit"] == r["exit"]
    assert retried >= 1
