"""Mechanisms on the device (Master::do_adj_min + the pathway checks of Master::saddle_2_mech).

Bar: for a converged FP64 batched search with mechanisms on, the two minima are stationary points of
the FP64 oracle's energy (|g| < 1e-5 eV/A on the mobile atoms), equal the oracle's LBFGS::minimise from
the same nudged saddle within 1e-4 A, and the reported barrier / delta equal the oracle's full-system
energy differences within 1e-6 eV; the classification follows the reference's tolerances; the server
reports the same mechanisms as the batch API; the mixed form refuses a minimiser tolerance below its
FP32 force noise.  Potential: the reference's wen.eam.fs (Fe vacancy hop).
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
    mn = mn - L * np.floor(mn / L)
    mn = np.where(mn >= L, mn - L, mn)
    d = mn - vac
    d -= L * np.round(d / L)
    near = np.argsort(np.einsum("ij,ij->i", d, d))[:8]
    s = systems.System(mn, t, basis)
    s.vacancies = vac[None]
    return dict(s=s, data=data, box=box, near=near, pot=fa.EAM(ctx, H.WEN), L=L, n=n)


def frozen_of(S, centre):
    d = S["s"].pos - S["s"].pos[centre]
    d -= S["L"] * np.round(d / S["L"])
    return (np.einsum("ij,ij->i", d, d) > RADIUS * RADIUS).astype(np.uint8)


def starts(S, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for q in range(k):
        c = int(S["near"][q % len(S["near"])])
        idx, p, a = H.library_kick(S["s"], c, RADIUS, rng.integers(1, 2 ** 63, size=4, dtype=np.uint64))
        x = S["s"].pos.copy()
        ax = np.zeros_like(x)
        x[idx] = p
        ax[idx] = a
        out.append((c, x, ax, (idx, p, a)))
    return out


def energy(S, p):
    nl = po.List(S["box"], S["data"].r_cut)
    nl.rebuild(p, 4)
    return po.energy(S["data"], nl, S["s"].types, np.zeros(S["n"], np.uint8), 4)


def test_batch_mechanisms_match_oracle(ctx, setup):
    S = setup
    st = starts(S, 8, 3)
    b = fa.Batch(ctx, S["pot"], fa.Box(S["s"].basis), S["s"].pos, S["s"].types, dimer=dict(f2norm=1e-4),
                 precision=fa.FP64)
    b.set_mechanisms()
    ids = b.submit([c for c, _, _, _ in st], [RADIUS] * len(st), np.stack([x for _, x, _, _ in st]),
                   np.stack([a for _, _, a, _ in st]))
    b.run()
    checked = 0
    for (c, _, _, _), i in zip(st, ids):
        r = b.result(i, positions=True)
        if r["exit"] != 0:
            continue
        m = b.mechanism(i, positions=True)
        print(f"case centre {c}: exit {r['exit']} barrier {r['barrier']:.6f} -> mechanism {fa.MECH_NAMES[m['code']]} "
              f"barrier {m['barrier']:.6f} delta {m['delta']:.6f} min_iters {m['min_iters']}", flush=True)
        if m["code"] == 1:
            continue
        f = frozen_of(S, c)
        rev, fwd, sp = m["rev"], m["fwd"], r["pos"]
        # Stationary in the FP64 oracle.
        for p in (rev, fwd):
            nl = po.List(S["box"], S["data"].r_cut)
            nl.rebuild(p, 4)
            g = po.gradient(S["data"], nl, S["s"].types, f, 4)
            assert np.abs(g[f == 0]).max() < 1e-5
        # Same minima as the oracle's minimiser from the same nudged saddles.
        disp = np.linalg.norm(sp - S["s"].pos)
        ax = r["axis"]
        mins = []
        for sgn in (1.0, -1.0):
            mo, _, failed, _ = po.minimise(S["box"], S["data"], sp + sgn * ax * disp * 0.01, S["s"].types, f, threads=4)
            assert not failed
            mins.append(mo)
        a, bb = mins
        if np.linalg.norm(bb - S["s"].pos) > np.linalg.norm(a - S["s"].pos):
            a, bb = bb, a      # rev = the one nearer the initial minimum
        assert np.abs(bb - rev).max() < 1e-4 and np.abs(a - fwd).max() < 1e-4
        # Energies in FP64 against the full system.
        e_rev, e_sp, e_fwd = energy(S, rev), energy(S, sp), energy(S, fwd)
        assert abs(m["barrier"] - (e_sp - e_rev)) < 1e-6
        assert abs(m["delta"] - (e_fwd - e_rev)) < 1e-6
        # Classification by the reference's tolerances.
        drift = (rev - S["s"].pos).sum(axis=0) / S["n"]
        i2r = np.linalg.norm(rev - drift - S["s"].pos)
        r2f = np.linalg.norm(fwd - rev)
        big = np.argmax(np.einsum("ij,ij->i", fwd - rev, fwd - rev))
        assert np.allclose(m["drift"], drift, atol=1e-12)
        if m["code"] == 0:
            assert i2r <= 0.05 and r2f >= 0.05 and big == c and e_sp > e_rev
            checked += 1
        elif m["code"] == 2:
            assert i2r > 0.05
        elif m["code"] == 3:
            assert r2f < 0.05
        elif m["code"] == 6:
            assert big != c
    assert checked >= 1


def test_server_mechanisms_equal_batch(ctx, setup):
    S = setup
    st = starts(S, 4, 9)
    b = fa.Batch(ctx, S["pot"], fa.Box(S["s"].basis), S["s"].pos, S["s"].types, dimer=dict(f2norm=1e-4),
                 block_threads=256, lanes_per_atom=8)
    b.set_mechanisms(minimiser=dict(f2norm=1e-4))
    ids = b.submit([c for c, _, _, _ in st], [RADIUS] * len(st), np.stack([x for _, x, _, _ in st]),
                   np.stack([a for _, _, a, _ in st]))
    b.run()
    srv = fa.Server(ctx, S["pot"], dimer=dict(f2norm=1e-4), block_threads=256, lanes_per_atom=8)
    srv.set_mechanisms(minimiser=dict(f2norm=1e-4))
    sid = srv.add_system(fa.Box(S["s"].basis), S["s"].pos, S["s"].types)
    sids = srv.submit(sid, [c for c, _, _, _ in st], [RADIUS] * len(st), [k for _, _, _, k in st])
    got = {}
    for _ in range(100000):
        for d in srv.progress(wait_ms=20.0):
            r = srv.result(d)
            got[int(d)] = (r, srv.mechanism(d))
        if len(got) == len(sids):
            break
    for i, j in zip(ids, sids):
        mb = b.mechanism(i)
        r, ms = got[int(j)]
        print(f"search {i}: exit {r['exit']} mechanism {fa.MECH_NAMES[mb['code']]}", flush=True)
        assert r["exit"] == b.result(i)["exit"]
        assert mb["code"] == ms["code"] and mb["barrier"] == ms["barrier"] and mb["delta"] == ms["delta"]
        if r["exit"] == 0:
            assert mb["code"] != 1   # the mixed relaxations reach 1e-4 eV/A


def test_mixed_mechanisms_refuse_unreachable_tolerance(ctx, setup):
    S = setup
    b = fa.Batch(ctx, S["pot"], fa.Box(S["s"].basis), S["s"].pos, S["s"].types, precision=fa.MIXED)
    with pytest.raises(fa.FlyError, match="f2norm"):
        b.set_mechanisms()                      # the reference default 1e-6 is below the FP32 noise
    b.set_mechanisms(minimiser=dict(f2norm=1e-4))
    b64 = fa.Batch(ctx, S["pot"], fa.Box(S["s"].basis), S["s"].pos, S["s"].types, precision=fa.FP64)
    b64.set_mechanisms()
    srv = fa.Server(ctx, S["pot"])
    with pytest.raises(fa.FlyError, match="f2norm"):
        srv.set_mechanisms()
