"""The benchmark workloads at their full sizes against the FP64 oracle.

* 30^3 Fe-X alloy (54,529 atoms): 300 steps of random displacement (|d| <= 0.05 A) through List::update,
  rebuilt once the accumulated largest move exceeds half the 0.6 A skin; pair sets (i, j, image) at
  r_c = 5.3 A equal the oracle's, row by row and as a multiset checksum, at every rebuild and every
  10th step.
* 20^3 Fe with 20 vacancies + 40 interstitials (16,020 atoms), 16 searches (radius U(4, 9) A, Master::perturb
  starts in the library, |g| < 1e-3 eV/A): FP64 batched searches end with the oracle's exit code, and
  converged barriers agree within 1e-4 eV; mixed-precision barriers equal FP64 energy differences at their
  own saddles within 1e-6 eV, and are compared with the oracle's from the same start.
* a served 40^3 system (128k atoms) beside a small one: served barriers equal oracle FP64 energy
  differences at the served saddles within 1e-6 eV, and a search gives the same bits alone and in the mix.
* 80^3 Fe with clustered vacancies (1,028,096 atoms), 16 concurrent searches: every converged barrier
  re-evaluated as an oracle FP64 full-system energy difference, within 1e-6 eV.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def pair_checksum(off, j, s):
    """(pair count, order-independent 64-bit checksum of the (i, j, shift) multiset)."""
    i = np.repeat(np.arange(len(off) - 1, dtype=np.uint64), np.diff(off).astype(np.int64))
    z = (i << np.uint64(32)) ^ (j.astype(np.uint64) << np.uint64(5)) ^ s.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)   # splitmix64 finaliser per key
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return len(i), int(np.sum(z, dtype=np.uint64))


def row_r2(off, dr):
    """Per-row sums of |dr|^2 (independent of the order inside a row up to rounding)."""
    r2 = np.einsum("ij,ij->i", dr, dr)
    return np.add.reduceat(r2, off[:-1]) * (np.diff(off) > 0)


def test_config2_incremental_pair_sets(ctx):
    s = systems.fe_x_alloy(seed=2, cells=30)
    n = s.n
    rc_list, skin = 5.3 + 0.6, 0.6
    print(f"case fe_x_alloy(30^3) n={n}, 300 update steps", flush=True)
    gnl = fa.List(ctx, fa.Box(s.basis), rc_list)
    onl = po.List(po.Box(s.basis), rc_list)
    pos = s.pos.copy()
    gnl.rebuild(pos)
    onl.rebuild(pos, THREADS)
    rng = np.random.default_rng(7)
    acc, checks, rebuilds = 0.0, 0, 0

    def check(tag):
        go, gj, gs, gdr = gnl.pairs(5.3)
        oo, oj, os_, odr = onl.pairs(5.3)
        assert np.array_equal(np.diff(go), np.diff(oo)), tag
        assert pair_checksum(go, gj, gs) == pair_checksum(oo, oj, os_), tag
        assert np.max(np.abs(row_r2(go, gdr) - row_r2(oo, odr))) < 1e-9, tag

    check("initial")
    for step in range(1, 301):
        d = rng.normal(size=(n, 3))
        d *= (0.05 * rng.random(n) / np.linalg.norm(d, axis=1))[:, None]
        pos += d
        acc += float(np.max(np.linalg.norm(d, axis=1)))
        if acc > 0.5 * skin:
            gnl.rebuild(pos)
            onl.rebuild(pos, THREADS)
            acc = 0.0
            rebuilds += 1
            check(f"rebuild at step {step}")
            checks += 1
        else:
            gnl.update(-d)
            onl.update(-d)
            if step % 10 == 0:
                check(f"update step {step}")
                checks += 1
    print(f"{rebuilds} rebuilds, {checks} checks")
    assert rebuilds >= 20 and checks >= 40


def relaxed(ctx, pot, s, f2norm=1e-4):
    mn, _, failed, _ = fa.LBFGS(ctx, fa.Box(s.basis), f2norm=f2norm, iter_max=5000).minimise(
        s.pos, s.types, np.zeros(s.n, np.uint8), pot)
    assert not failed
    s.pos = np.ascontiguousarray(mn)
    return s


def frozen_beyond(s, centre, radius):
    L = np.diag(s.basis)
    d = s.pos - s.pos[centre]
    d -= L * np.round(d / L)
    return (np.einsum("ij,ij->i", d, d) > radius * radius).astype(np.uint8)


def oracle_energy(data, s, pos):
    nl = po.List(po.Box(s.basis), data.r_cut)
    nl.rebuild(pos, THREADS)
    return po.energy(data, nl, s.types, np.zeros(s.n, np.uint8), THREADS)


@pytest.fixture(scope="module")
def defects20(ctx):
    pot = fa.EAM(ctx, H.WEN)
    s = relaxed(ctx, pot, systems.fe_defects_20(3))
    return s, pot, po.DataEAM(H.WEN)


def test_config3_batched_vs_oracle(ctx, defects20):
    s, pot, data = defects20
    k = 16
    rng = np.random.default_rng(33)
    centres = rng.choice(systems.defect_adjacent_atoms(s), k, replace=True)
    radii = rng.uniform(4.0, 9.0, k)
    prng = rng.integers(1, 2 ** 63, size=(k, 4), dtype=np.uint64)
    print(f"case fe_defects_20 n={s.n}, {k} searches", flush=True)
    starts = []
    for c, R, st in zip(centres, radii, prng):
        f = frozen_beyond(s, int(c), R)
        x, ax, _ = fa.perturb(fa.Box(s.basis), s.pos, f, int(c), st.copy())
        starts.append((x, ax, f))

    def oracle(q):
        x, ax, f = starts[q]
        return po.find_sp(po.Box(s.basis), data, x, ax, s.types, f, s.pos, threads=1, dimer=dict(f2norm=1e-3))

    with cf.ThreadPoolExecutor(max_workers=THREADS) as ex:
        orc = list(ex.map(oracle, range(k)))
    e0 = oracle_energy(data, s, s.pos)
    res = {}
    for name, prec in (("fp64", fa.FP64), ("mixed", fa.MIXED)):
        b = fa.Batch(ctx, pot, fa.Box(s.basis), s.pos, s.types, dimer=dict(f2norm=1e-3), precision=prec)
        ids = b.submit_perturbed(centres, radii, prng.copy())
        b.run()
        res[name] = [b.result(i, positions=True) for i in ids]
    agree = both = 0
    for q in range(k):
        o, r64, rmx = orc[q], res["fp64"][q], res["mixed"][q]
        eo = oracle_energy(data, s, o["pos"]) - e0 if o["exit"] == 0 else None
        print(f"  search {q}: M={r64['n_mobile']} oracle exit {o['exit']} steps {o['steps']} barrier {eo} | fp64 exit "
              f"{r64['exit']} steps {r64['steps']} barrier {r64['barrier']:.6f} | mixed exit {rmx['exit']} barrier "
              f"{rmx['barrier']:.6f}")
        assert r64["exit"] == o["exit"]
        if o["exit"] == 0:
            assert abs(r64["barrier"] - eo) < 1e-4
        if rmx["exit"] == 0:
            # the reported barrier is an FP64 energy difference at the mixed form's own saddle
            assert abs(rmx["barrier"] - (oracle_energy(data, s, rmx["pos"]) - e0)) < 1e-6
            if o["exit"] == 0:
                both += 1
                agree += abs(rmx["barrier"] - eo) <= 1e-3
    print(f"mixed vs oracle: {agree} of {both} converged in both within 1e-3 eV")
    assert both == 0 or agree >= both // 2


def test_config4_served_large_system(ctx):
    pot = fa.EAM(ctx, H.WEN)
    data = po.DataEAM(H.WEN)
    big = relaxed(ctx, pot, systems.defect_system(40, 320, 320, 17))
    small = relaxed(ctx, pot, systems.defect_system(10, 5, 5, 18))
    assert big.n >= 100000
    print(f"case served: {big.n} + {small.n} atoms", flush=True)
    rng = np.random.default_rng(19)
    reqs = []
    for s, k in ((big, 12), (small, 8)):
        reqs.append((s, rng.choice(systems.defect_adjacent_atoms(s), k), rng.uniform(4.0, 9.0, k),
                     rng.integers(1, 2 ** 63, size=(k, 4), dtype=np.uint64)))
    srv = fa.Server(ctx, pot, dimer=dict(f2norm=1e-3))
    sids = [srv.add_system(fa.Box(s.basis), s.pos, s.types) for s, _, _, _ in reqs]
    ids = [srv.submit_perturbed(sid, c, R, st.copy()) for sid, (_, c, R, st) in zip(sids, reqs)]
    got = {}
    while len(got) < sum(len(x) for x in ids):
        for d in srv.progress(wait_ms=20.0):
            got[int(d)] = srv.result(d, positions=True)
    for (s, _, _, _), idl in zip(reqs, ids):
        e0 = oracle_energy(data, s, s.pos)
        for i in idl:
            r = got[int(i)]
            if r["exit"] != 0:
                continue
            p = s.pos.copy()
            p[r["atoms"]] = r["pos"]
            de = oracle_energy(data, s, p) - e0
            print(f"  n={s.n} served barrier {r['barrier']:.8f} oracle FP64 {de:.8f}")
            assert abs(de - r["barrier"]) < 1e-6
    # the first large-system search again, alone on an idle server: the same bits
    s, c, R, st = reqs[0]
    srv2 = fa.Server(ctx, pot, dimer=dict(f2norm=1e-3), block_threads=srv.status()["block_threads"],
                     lanes_per_atom=srv.status()["lanes_per_atom"])
    sid2 = srv2.add_system(fa.Box(s.basis), s.pos, s.types)
    (i2,) = srv2.submit_perturbed(sid2, c[:1], R[:1], st[:1].copy())
    alone = None
    while alone is None:
        for d in srv2.progress(wait_ms=20.0):
            alone = srv2.result(d, positions=True)
    mixed = got[int(ids[0][0])]
    assert alone["exit"] == mixed["exit"] and alone["steps"] == mixed["steps"]
    assert alone["barrier"] == mixed["barrier"] and np.array_equal(alone["pos"], mixed["pos"])


def test_config5_large_system_fp64_reevaluation(ctx):
    pot = fa.EAM(ctx, H.WEN)
    data = po.DataEAM(H.WEN)
    s = relaxed(ctx, pot, systems.fe_clusters_80(5))
    print(f"case fe_clusters_80 n={s.n}, 16 searches", flush=True)
    assert s.n > 1_000_000
    rng = np.random.default_rng(55)
    k = 16
    centres = rng.choice(systems.defect_adjacent_atoms(s), k)
    radii = rng.uniform(4.0, 9.0, k)
    prng = rng.integers(1, 2 ** 63, size=(k, 4), dtype=np.uint64)
    b = fa.Batch(ctx, pot, fa.Box(s.basis), s.pos, s.types, dimer=dict(f2norm=1e-3), precision=fa.MIXED)
    ids = b.submit_perturbed(centres, radii, prng.copy())
    b.run()
    res = [b.result(i, positions=True) for i in ids]
    conv = [r for r in res if r["exit"] == 0]
    print(f"{len(conv)} of {k} converged", flush=True)
    assert len(conv) >= 3
    e0 = oracle_energy(data, s, s.pos)
    for r in conv:
        de = oracle_energy(data, s, r["pos"]) - e0
        print(f"  barrier {r['barrier']:.8f} oracle FP64 {de:.8f}")
        assert abs(de - r["barrier"]) < 1e-6
