"""Batched on-device dimer searches vs the FP64 oracle's full-system searches with the same frozen mask.

Bar: FP64 batched searches end with the oracle's exit code and a barrier within 1e-4 eV; mixed precision
(FP32 pair arithmetic, FP64 accumulation / curvature / dots / energies, |g| < 1e-3 eV/A as in the batched
workload) within 1e-3 eV; a search gives bit-identical results alone and inside a larger batch.
Potential: the reference's wen.eam.fs (Fe vacancy hop, barrier ~0.64 eV).
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


@pytest.fixture(scope="module")
def setup(ctx):
    cells = 7
    pos = systems.bcc_sites(A_WEN, cells)
    vac = pos[0].copy()
    pos = np.delete(pos, 0, axis=0)
    basis = np.diag([cells * A_WEN] * 3)
    n = len(pos)
    t = np.zeros(n, np.uint32)
    data = po.DataEAM(H.WEN)
    box = po.Box(basis)
    mn, _, failed, _ = po.minimise(box, data, pos, t, np.zeros(n, np.uint8), threads=8)
    assert not failed
    L = np.diag(basis)
    d = mn - vac
    d -= L * np.round(d / L)
    near = np.argsort(np.einsum("ij,ij->i", d, d))[:8]   # the 8 first neighbours of the vacancy
    nl_p = po.List(box, 6.0)
    nl_p.rebuild(mn, 4)
    pot = fa.EAM(ctx, H.WEN)
    return dict(mn=mn, basis=basis, n=n, t=t, data=data, box=box, near=near, nl_p=nl_p, pot=pot, L=L)


def frozen_mask(S, centre, radius=RADIUS):
    d = S["mn"] - S["mn"][centre]
    d -= S["L"] * np.round(d / S["L"])
    return (np.einsum("ij,ij->i", d, d) > radius * radius).astype(np.uint8)


def starts(S, k, seed0=1):
    out = []
    for s in range(k):
        centre = int(S["near"][s % len(S["near"])])
        f = frozen_mask(S, centre)
        x, ax, _ = po.perturb([seed0 + s, 7, 7, 7], 0.6, 4.0, S["nl_p"], centre, S["mn"], f)
        out.append((centre, f, x, ax))
    return out


def energy_orc(S, pos):
    nl = po.List(S["box"], S["data"].r_cut)
    nl.rebuild(pos, 4)
    return po.energy(S["data"], nl, S["t"], np.zeros(S["n"], np.uint8), 4)


def run_batch(ctx, S, st, precision, f2norm):
    b = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], dimer=dict(f2norm=f2norm), precision=precision)
    ids = b.submit([c for c, _, _, _ in st], [RADIUS] * len(st), np.stack([x for _, _, x, _ in st]),
                   np.stack([a for _, _, _, a in st]))
    b.run()
    return b, [b.result(i, positions=True) for i in ids]


def test_fp64_batch_matches_oracle(ctx, setup):
    S = setup
    st = starts(S, 6)
    print("case batch fp64, 6 searches", flush=True)
    b, res = run_batch(ctx, S, st, fa.FP64, 1e-5)
    E0 = energy_orc(S, S["mn"])
    ok = 0
    for (centre, f, x, ax), r in zip(st, res):
        o = po.find_sp(S["box"], S["data"], x, ax, S["t"], f, S["mn"], threads=8)
        print(f"  centre {centre} mobile {r['n_mobile']} region {r['n_region']}: gpu exit {r['exit']} steps "
              f"{r['steps']} barrier {r['barrier']:.6f} max_disp {r['max_disp']:.3f} launches {r['launches']} | "
              f"oracle exit {o['exit']} steps {o['steps']} max_disp {np.abs(o['pos'] - S['mn']).max():.3f}")
        assert r["n_mobile"] == int((f == 0).sum())
        assert r["exit"] == o["exit"]
        if o["exit"] == 0:
            Eo = energy_orc(S, o["pos"]) - E0
            Eg_full = energy_orc(S, r["pos"]) - E0   # the device's local-energy barrier vs a full FP64 energy
            print(f"    oracle barrier {Eo:.6f}, full energy at gpu saddle {Eg_full:.6f}")
            assert abs(r["barrier"] - Eg_full) < 1e-8
            assert abs(r["barrier"] - Eo) < 1e-4
            ok += 1
    assert ok >= 2


def test_mixed_batch_matches_oracle(ctx, setup):
    S = setup
    st = starts(S, 8, seed0=11)
    print("case batch mixed, 8 searches", flush=True)
    b, res = run_batch(ctx, S, st, fa.MIXED, 1e-3)
    E0 = energy_orc(S, S["mn"])
    both = 0
    for (centre, f, x, ax), r in zip(st, res):
        o = po.find_sp(S["box"], S["data"], x, ax, S["t"], f, S["mn"], threads=8, dimer=dict(f2norm=1e-3))
        print(f"  gpu exit {r['exit']} barrier {r['barrier']:.6f} | oracle exit {o['exit']}")
        if r["exit"] == 0:
            # every converged barrier re-evaluated in FP64 at the device's saddle
            assert abs(r["barrier"] - (energy_orc(S, r["pos"]) - E0)) < 1e-6
        if r["exit"] == 0 and o["exit"] == 0:
            Eo = energy_orc(S, o["pos"]) - E0
            print(f"    oracle barrier {Eo:.6f}")
            assert abs(r["barrier"] - Eo) < 1e-3
            both += 1
    assert both >= 2


def test_alone_equals_in_mix(ctx, setup):
    S = setup
    st = starts(S, 10, seed0=31)
    _, alone = run_batch(ctx, S, st[3:4], fa.MIXED, 1e-3)
    _, mixed = run_batch(ctx, S, st, fa.MIXED, 1e-3)
    a, m = alone[0], mixed[3]
    assert a["exit"] == m["exit"] and a["steps"] == m["steps"] and a["grads"] == m["grads"]
    assert a["barrier"] == m["barrier"]
    assert np.array_equal(a["pos"], m["pos"])


def test_results_independent_of_kernel_shape(ctx, setup):
    # Row sums use a 16-slot tree and workgroup dot products a 1024-slot tree, so the autotuner's choice of
    # (block threads, lanes per atom) cannot change any result bit.
    S = setup
    st = starts(S, 4, seed0=51)
    outs = []
    for T, L in ((256, 8), (512, 16), (1024, 8), (256, 16)):
        b = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], dimer=dict(f2norm=1e-3), precision=fa.MIXED,
                     block_threads=T, lanes_per_atom=L)
        ids = b.submit([c for c, _, _, _ in st], [RADIUS] * len(st), np.stack([x for _, _, x, _ in st]),
                       np.stack([a for _, _, _, a in st]))
        b.run()
        outs.append([b.result(i, positions=True) for i in ids])
    for other in outs[1:]:
        for a, m in zip(outs[0], other):
            assert a["exit"] == m["exit"] and a["steps"] == m["steps"] and a["barrier"] == m["barrier"]
            assert np.array_equal(a["pos"], m["pos"])


def test_tuned_shape_is_reported_and_reusable(ctx, setup):
    # flyg_batch_shape: (0, 0) before the first run, then the autotuned (block threads, lanes per atom);
    # a batch given that shape skips the tuning and reproduces the results bit for bit.
    S = setup
    st = starts(S, 3, seed0=61)
    args = ([c for c, _, _, _ in st], [RADIUS] * len(st), np.stack([x for _, _, x, _ in st]),
            np.stack([a for _, _, _, a in st]))
    b = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], dimer=dict(f2norm=1e-3), precision=fa.MIXED)
    assert b.shape() == (0, 0)
    ids = b.submit(*args)
    b.run()
    T, L = b.shape()
    print(f"tuned shape {T}x{L}", flush=True)
    assert T in (256, 512, 1024) and L in (8, 16)
    b2 = fa.Batch(ctx, S["pot"], fa.Box(S["basis"]), S["mn"], S["t"], dimer=dict(f2norm=1e-3), precision=fa.MIXED,
                  block_threads=T, lanes_per_atom=L)
    ids2 = b2.submit(*args)
    b2.run()
    assert b2.shape() == (T, L)
    for i, j in zip(ids, ids2):
        r1, r2 = b.result(i), b2.result(j)
        assert r1["exit"] == r2["exit"] and r1["steps"] == r2["steps"] and r1["barrier"] == r2["barrier"]
