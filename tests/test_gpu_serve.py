"""Serving (continuous batching over resident systems) on the device.

Bar: a served search gives bit-identical results to the same search run by the batch API, alone on an
idle server, and in a mix where other systems' searches were already running when it arrived; the
memory budget gates admission; converged served barriers agree with the FP64 energy at the saddle.
"""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import serve, systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

A_WEN = 2.855324
SHAPE = dict(block_threads=256, lanes_per_atom=8)


def vacancy_system(cells, drop=0):
    pos = systems.bcc_sites(A_WEN, cells)
    pos = np.delete(pos, drop, axis=0)
    return pos, np.diag([cells * A_WEN] * 3)


@pytest.fixture(scope="module")
def wen(ctx):
    pos, basis = vacancy_system(7)
    n = len(pos)
    t = np.zeros(n, np.uint32)
    data = po.DataEAM(H.WEN)
    mn, _, failed, _ = po.minimise(po.Box(basis), data, pos, t, np.zeros(n, np.uint8), threads=8)
    assert not failed
    L = np.diag(basis)
    mn = mn - L * np.floor(mn / L)   # canonical, so dense and sparse start states coincide bit for bit
    mn = np.where(mn >= L, mn - L, mn)
    s = systems.System(mn, t, basis)
    s.vacancies = systems.bcc_sites(A_WEN, 7)[:1]
    return dict(sys=s, pot=fa.EAM(ctx, H.WEN), data=data)


def requests_for(s, k, seed):
    rng = np.random.default_rng(seed)
    centres = serve.sample_centres(s, k, rng)
    radii = rng.uniform(5.0, 7.0, k)
    states = rng.integers(1, 2 ** 63, size=(k, 4), dtype=np.uint64)
    kicks = [H.library_kick(s, int(c), float(R), st) for c, R, st in zip(centres, radii, states)]
    return centres, radii, kicks


def drain(server, ids):
    """Results of `ids`; finished searches that belong to other callers are kept on the server object."""
    got = server.__dict__.setdefault("_finished", {})
    for _ in range(100000):
        if all(int(i) in got for i in ids):
            return [got[int(i)] for i in ids]
        for d in server.progress(wait_ms=20.0):
            got[int(d)] = server.result(d, positions=True)
    raise AssertionError("searches did not finish")


def same(a, b):
    return (a["exit"] == b["exit"] and a["steps"] == b["steps"] and a["grads"] == b["grads"]
            and a["barrier"] == b["barrier"] and np.array_equal(a["pos"], b["pos"]))


def test_served_equals_batch(ctx, wen):
    s = wen["sys"]
    centres, radii, kicks = requests_for(s, 6, 5)
    print("case served vs batch, 6 searches", flush=True)
    srv = fa.Server(ctx, wen["pot"], dimer=dict(f2norm=1e-3), **SHAPE)
    sid = srv.add_system(fa.Box(s.basis), s.pos, s.types)
    served = drain(srv, srv.submit(sid, centres, radii, kicks))
    starts = np.repeat(s.pos[None], len(centres), axis=0)
    axes = np.zeros_like(starts)
    for k, (idx, p, a) in enumerate(kicks):
        starts[k, idx] = p
        axes[k, idx] = a
    b = fa.Batch(ctx, wen["pot"], fa.Box(s.basis), s.pos, s.types, dimer=dict(f2norm=1e-3), **SHAPE)
    ids = b.submit(centres, radii, starts, axes)
    b.run()
    for r, i in zip(served, ids):
        rb = b.result(i, positions=True)
        full = s.pos.copy()
        full[r["atoms"]] = r["pos"]
        print(f"  exit {r['exit']} steps {r['steps']} barrier {r['barrier']:.6f} | batch exit {rb['exit']} "
              f"barrier {rb['barrier']:.6f}")
        assert r["exit"] == rb["exit"] and r["steps"] == rb["steps"] and r["barrier"] == rb["barrier"]
        assert np.array_equal(full, rb["pos"])


def test_alone_equals_in_mix(ctx, wen):
    s = wen["sys"]
    centres, radii, kicks = requests_for(s, 4, 9)
    print("case alone", flush=True)
    alone_srv = fa.Server(ctx, wen["pot"], dimer=dict(f2norm=1e-3), **SHAPE)
    sid = alone_srv.add_system(fa.Box(s.basis), s.pos, s.types)
    alone = drain(alone_srv, alone_srv.submit(sid, centres, radii, kicks))

    print("case mix: three other systems in flight, then the same request", flush=True)
    mix = fa.Server(ctx, wen["pot"], dimer=dict(f2norm=1e-3), streams=2, **SHAPE)
    others = []
    for q, cells in enumerate((6, 8, 9)):
        pos, basis = vacancy_system(cells, drop=q)
        o = systems.System(pos, np.zeros(len(pos), np.uint32), basis)
        o.vacancies = systems.bcc_sites(A_WEN, cells)[q:q + 1]
        c, R, kk = requests_for(o, 12, 20 + q)
        oid = mix.add_system(fa.Box(basis), pos, o.types)
        others.extend(mix.submit(oid, c, R, kk))
    mix.progress(wait_ms=0.0)   # the other systems' searches are launched now
    st = mix.status()
    assert st["in_flight"] > 0
    sid = mix.add_system(fa.Box(s.basis), s.pos, s.types)
    in_mix = drain(mix, mix.submit(sid, centres, radii, kicks))
    drain(mix, others)
    assert mix.status()["launches"] >= 2
    for a, m in zip(alone, in_mix):
        assert same(a, m)


def test_memory_budget_gates_admission(ctx, wen):
    s = wen["sys"]
    n = s.n
    # Room to admit one system of this size, not two (launches then run one search at a time).
    per = n * 41 + 4096 + 648 + n * 700
    srv = fa.Server(ctx, wen["pot"], dimer=dict(f2norm=1e-3, max_steps=5), mem_budget=per + 20 * n, **SHAPE)
    a = srv.add_system(fa.Box(s.basis), s.pos, s.types)
    with pytest.raises(fa.Busy):
        srv.add_system(fa.Box(s.basis), s.pos, s.types)
    centres, radii, kicks = requests_for(s, 2, 3)
    drain(srv, srv.submit(a, centres, radii, kicks))
    srv.remove_system(a)
    srv.progress()
    assert srv.status()["systems"] == 0
    b = srv.add_system(fa.Box(s.basis), s.pos, s.types)
    assert b != a


def test_served_barriers_match_fp64_energy(ctx, wen):
    s = wen["sys"]
    centres, radii, kicks = requests_for(s, 24, 13)
    srv = fa.Server(ctx, wen["pot"], dimer=dict(f2norm=1e-3))
    sid = srv.add_system(fa.Box(s.basis), s.pos, s.types)
    res = drain(srv, srv.submit(sid, centres, radii, kicks))
    box = po.Box(s.basis)

    def energy(p):
        nl = po.List(box, wen["data"].r_cut)
        nl.rebuild(p, 4)
        return po.energy(wen["data"], nl, s.types, np.zeros(s.n, np.uint8), 4)

    E0 = energy(s.pos)
    ok = 0
    for r in res:
        if r["exit"] != 0:
            continue
        full = s.pos.copy()
        full[r["atoms"]] = r["pos"]
        dE = energy(full) - E0
        print(f"  barrier {r['barrier']:.6f} FP64 full-system {dE:.6f}")
        assert abs(r["barrier"] - dE) < 1e-6
        ok += 1
    assert ok >= 2


def test_serve_loop_small(ctx, wen):
    """A short Poisson replay over three relaxed kMC systems: every request completes, systems are
    admitted once and reused, and the server is empty afterwards."""
    pot = wen["pot"]

    def relax(s):
        mn, _, failed, _ = fa.LBFGS(ctx, fa.Box(s.basis), f2norm=1e-4, iter_max=4000).minimise(
            s.pos, s.types, np.zeros(s.n, np.uint8), pot)
        assert not failed
        return mn

    pool = serve.make_pool(3, seed=1, n_min=2000, n_max=6000, relax=relax)
    reqs = serve.make_requests(pool, 8, rate=200.0, seed=2, searches=(4, 8))
    srv = fa.Server(ctx, pot, dimer=dict(f2norm=1e-3))
    out = serve.serve(srv, reqs)
    print(out)
    assert out["requests"] == 8 and out["searches"] == sum(len(r.centres) for r in reqs)
    assert out["systems_admitted"] <= 3
    assert all(r.done >= r.t for r in reqs)
    assert srv.status()["systems"] == 0
