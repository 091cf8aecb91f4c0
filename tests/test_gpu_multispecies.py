"""Batched searches on a two-species system, which run the search kernel's multi-species form (per-atom
types, phi_{a->g} and phi_{g->a} tables of unlike pairs, V of the pair): bcc Fe (7^3 cells) with one
vacancy and one H atom on a tetrahedral site next to it, with the reference's Fe-H potential wen.eam.fs,
relaxed by the oracle.

Bar: FP64 batched searches end with the oracle's exit code (searches the oracle ends within 100 steps;
longer ones wander and are reported) and converged barriers within 1e-4 eV of the oracle's; mixed
precision: every converged barrier (FP64 local energy) equals the oracle's full-system FP64 energy
difference at the returned saddle within 1e-6 eV.
"""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

A_WEN = 2.855324
RADIUS = 6.0
LONG = 100


@pytest.fixture(scope="module")
def fe_h(ctx):
    cells = 7
    pos = systems.bcc_sites(A_WEN, cells)
    vac = pos[0].copy()
    pos = np.delete(pos, 0, axis=0)
    h = vac + A_WEN * np.array([0.5, 0.25, 0.0])   # tetrahedral site beside the vacancy
    pos = np.vstack([pos, h])
    n = len(pos)
    t = np.zeros(n, np.uint32)
    t[-1] = 1                                       # H, the potential's second species
    basis = np.diag([cells * A_WEN] * 3)
    data = po.DataEAM(H.WEN)
    box = po.Box(basis)
    mn, _, failed, _ = po.minimise(box, data, pos, t, np.zeros(n, np.uint8), threads=8)
    assert not failed
    L = np.diag(basis)
    d = mn - vac
    d -= L * np.round(d / L)
    near = list(np.argsort(np.einsum("ij,ij->i", d, d))[:5])
    if n - 1 not in near:
        near[-1] = n - 1                            # the H atom is one of the centres
    nl_p = po.List(box, 6.0)
    nl_p.rebuild(mn, 4)
    return dict(mn=mn, basis=basis, n=n, t=t, data=data, box=box, near=[int(c) for c in near], nl_p=nl_p,
                L=L, pot=fa.EAM(ctx, H.WEN))


def frozen_mask(S, centre):
    d = S["mn"] - S["mn"][centre]
    d -= S["L"] * np.round(d / S["L"])
    return (np.einsum("ij,ij->i", d, d) > RADIUS * RADIUS).astype(np.uint8)


def starts(S, k):
    out = []
    for s in range(k):
        c = S["near"][s % len(S["near"])]
        f = frozen_mask(S, c)
        x, ax, _ = po.perturb([11 + s, 5, 5, 5], 0.6, 4.0, S["nl_p"], c, S["mn"], f)
        out.append((c, f, x, ax))
    return out


def energy(S, pos):
    nl = po.List(S["box"], S["data"].r_cut)
    nl.rebuild(pos, 4)
    return po.energy(S["data"], nl, S["t"], np.zeros(S["n"], np.uint8), 4)


def run(ctx, S, st, precision, f2norm):
    b = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], dimer=dict(f2norm=f2norm), precision=precision)
    ids = b.submit([c for c, _, _, _ in st], [RADIUS] * len(st), np.stack([x for _, _, x, _ in st]),
                   np.stack([a for _, _, _, a in st]))
    b.run()
    return [b.result(i, positions=True) for i in ids]


def test_fp64_two_species_batch_matches_oracle(ctx, fe_h):
    S = fe_h
    st = starts(S, 8)
    print("case Fe + H batch fp64, 8 searches", flush=True)
    res = run(ctx, S, st, fa.FP64, 1e-5)
    E0 = energy(S, S["mn"])
    compared = converged = 0
    for (c, f, x, ax), r in zip(st, res):
        o = po.find_sp(S["box"], S["data"], x, ax, S["t"], f, S["mn"], threads=8)
        print(f"  centre {c} (type {S['t'][c]}): gpu exit {r['exit']} steps {r['steps']} | oracle exit {o['exit']} "
              f"steps {o['steps']}")
        if o["steps"] > LONG:
            continue
        assert r["exit"] == o["exit"]
        compared += 1
        if o["exit"] == 0:
            Eo = energy(S, o["pos"]) - E0
            print(f"    barrier gpu {r['barrier']:.6f} oracle {Eo:.6f}")
            assert abs(r["barrier"] - Eo) < 1e-4
            converged += 1
    assert compared >= 5 and converged >= 1


def test_mixed_two_species_barriers_are_fp64_energies(ctx, fe_h):
    S = fe_h
    st = starts(S, 8)
    print("case Fe + H batch mixed, 8 searches", flush=True)
    res = run(ctx, S, st, fa.MIXED, 1e-3)
    E0 = energy(S, S["mn"])
    converged = 0
    for (c, _, _, _), r in zip(st, res):
        if r["exit"] != 0:
            continue
        Eg = energy(S, r["pos"]) - E0
        print(f"  centre {c}: barrier {r['barrier']:.8f}, oracle energy at the saddle {Eg:.8f}")
        assert abs(r["barrier"] - Eg) < 1e-6
        converged += 1
    assert converged >= 1
