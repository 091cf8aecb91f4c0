"""GPU EAM energy / gradient / HVP vs the reference's known answers and the FP64 oracle.

Tolerances: FP64 path within 1e-10 relative of the oracle (energies; gradients relative to the largest
component); mixed precision (FP32 pair arithmetic, FP64 accumulation) within 2e-6 relative on forces.
"""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wen_gpu(ctx):
    return fa.EAM(ctx, H.WEN)


@pytest.fixture(scope="module")
def wen_orc():
    return po.DataEAM(H.WEN)


def gpu_eval(ctx, pot, sys_pos, basis, types, frozen, rc=None, periodic=(1, 1, 1)):
    nl = fa.List(ctx, fa.Box(basis, periodic), rc or pot.r_cut())
    nl.rebuild(sys_pos)
    return nl, pot.energy(types, frozen, nl), pot.gradient(types, frozen, nl)


def orc_eval(data, sys_pos, basis, types, frozen, rc=None, periodic=(1, 1, 1)):
    nl = po.List(po.Box(basis, periodic), rc or data.r_cut)
    nl.rebuild(sys_pos, 4)
    return nl, po.energy(data, nl, types, frozen, 4), po.gradient(data, nl, types, frozen, 4)


def test_parser_kat(wen_gpu):
    assert wen_gpu.ntypes == 2 and wen_gpu.names == ["Fe", "H"]
    F = wen_gpu.table("F", 0, 0, [0.0, 1.9999999999999999e-02])
    assert abs(F[0] - H.KAT_F0[0]) < 1e-14 and abs(F[1] - H.KAT_F0[1]) < 1e-14


def test_tables_match_oracle(wen_gpu, wen_orc):
    rng = np.random.default_rng(0)
    r = rng.uniform(0, 5.3, 2000)
    rho = rng.uniform(0, 150, 2000)
    for kind, a, b, x in (("F", 0, 0, rho), ("F", 1, 1, rho), ("phi", 0, 1, r), ("phi", 1, 0, r), ("V", 1, 0, r),
                          ("V", 0, 0, r)):
        for d in (0, 1, 2):
            g = wen_gpu.table(kind, a, b, x, d)
            o = wen_orc.eval(kind, a, b, x, d)
            assert np.allclose(g, o, rtol=1e-12, atol=1e-12 * np.abs(o).max()), (kind, a, b, d)


def test_energy_gradient_kat(ctx, wen_gpu):
    pos, basis = H.kat_lattice()
    n = len(pos)
    t, f = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    print(f"case KAT 7^3 n={n}", flush=True)
    _, E, g = gpu_eval(ctx, wen_gpu, pos, basis, t, f)
    assert abs(E - H.KAT_E0) <= 1e-10 * abs(H.KAT_E0)
    assert np.linalg.norm(g[0] - H.KAT_G0) < 1e-10


def test_hvp_kat_block(ctx, wen_gpu):
    # H e_k for atom 0's three unit vectors reproduces the reference's H(0,0) block.
    pos, basis = H.kat_lattice()
    n = len(pos)
    t, f = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    nl = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut())
    nl.rebuild(pos)
    for k in range(3):
        v = np.zeros((n, 3))
        v[0, k] = 1.0
        hv = wen_gpu.hvp(t, f, nl, v)
        assert np.linalg.norm(hv[0] - H.KAT_H00[:, k]) < 1e-10


@pytest.mark.parametrize("frozen_every", [0, 4])
def test_two_species_vs_oracle(ctx, wen_gpu, wen_orc, frozen_every):
    rng = np.random.default_rng(9)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.06, pos.shape)
    n = len(pos)
    t = (rng.random(n) < 0.15).astype(np.uint32)
    f = np.zeros(n, np.uint8)
    if frozen_every:
        f[::frozen_every] = 1
    _, E, g = gpu_eval(ctx, wen_gpu, pos, basis, t, f)
    _, Eo, go = orc_eval(wen_orc, pos, basis, t, f)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))
    assert np.all(g[f == 1] == 0.0)


def test_hvp_vs_oracle_dense_hessian(ctx, wen_gpu, wen_orc):
    rng = np.random.default_rng(21)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.05, pos.shape)
    n = len(pos)
    t = (rng.random(n) < 0.1).astype(np.uint32)
    f = np.zeros(n, np.uint8)
    f[rng.random(n) < 0.2] = 1
    v = rng.normal(size=(n, 3))
    nl = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut())
    nl.rebuild(pos)
    hv = wen_gpu.hvp(t, f, nl, v)
    onl = po.List(po.Box(basis), wen_orc.r_cut)
    onl.rebuild(pos, 4)
    ho = po.hvp(wen_orc, onl, t, f, v, 4)
    assert np.max(np.abs(hv - ho)) <= 1e-10 * np.max(np.abs(ho))


def test_synthetic_fs_vacancy16_vs_oracle(ctx):
    path = systems.potential_path(("Fe",))
    pot = fa.EAM(ctx, path)
    data = po.DataEAM(path)
    s = systems.fe_vacancy_16(seed=1)
    print(f"case fe_vacancy_16 n={s.n}", flush=True)
    rc = 5.3 + 2e-3 + (1.1 ** (1 / 3) - 1) * (5.3 + 2e-3)
    _, E, g = gpu_eval(ctx, pot, s.pos, s.basis, s.types, s.frozen, rc)
    _, Eo, go = orc_eval(data, s.pos, s.basis, s.types, s.frozen, rc)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))
    # Mixed precision: FP32 pair arithmetic, FP64 accumulation.
    pot.set_precision(fa.MIXED)
    nl = fa.List(ctx, fa.Box(s.basis), rc)
    nl.rebuild(s.pos)
    gm = pot.gradient(s.types, s.frozen, nl)
    err = np.max(np.abs(gm - go)) / np.max(np.abs(go))
    print(f"mixed precision max rel force error {err:.3e}")
    assert err < 2e-6
    assert pot.energy(s.types, s.frozen, nl) == E   # energies stay FP64
    pot.set_precision(fa.FP64)


def test_fe_x_alloy_two_species_scale(ctx):
    path = systems.potential_path(("Fe", "X"))
    pot = fa.EAM(ctx, path)
    data = po.DataEAM(path)
    s = systems.fe_x_alloy(seed=2, cells=16)
    print(f"case fe_x_alloy(16^3) n={s.n}", flush=True)
    _, E, g = gpu_eval(ctx, pot, s.pos, s.basis, s.types, s.frozen)
    _, Eo, go = orc_eval(data, s.pos, s.basis, s.types, s.frozen)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))
    assert np.linalg.norm(g.sum(axis=0)) < 1e-8 * np.abs(g).max() * s.n


def test_gradient_deterministic(ctx):
    pot = fa.EAM(ctx, systems.potential_path(("Fe",)))
    s = systems.fe_vacancy_16(seed=4)
    nl = fa.List(ctx, fa.Box(s.basis), 5.3)
    nl.rebuild(s.pos)
    g1 = pot.gradient(s.types, s.frozen, nl)
    nl.rebuild(s.pos)
    g2 = pot.gradient(s.types, s.frozen, nl)
    assert np.array_equal(g1, g2)


def test_gradient_timing_counts_calls(ctx):
    # flyg_eam_set_timing / flyg_eam_timing: one event pair per gradient call while on; the timed
    # gradient equals an untimed one bit for bit.
    pot = fa.EAM(ctx, systems.potential_path(("Fe",)))
    s = systems.fe_vacancy_16(seed=5)
    nl = fa.List(ctx, fa.Box(s.basis), 5.3)
    nl.rebuild(s.pos)
    g0 = pot.gradient(s.types, s.frozen, nl)
    pot.set_timing(True)
    assert pot.timing() == (0.0, 0)
    for _ in range(3):
        g = pot.gradient(s.types, s.frozen, nl)
    ms, calls = pot.timing()
    pot.set_timing(False)
    pot.gradient(s.types, s.frozen, nl)
    assert calls == 3 and 0.0 < ms < 1e3
    assert pot.timing()[1] == 3          # calls while off are not recorded
    assert np.array_equal(g, g0)


def test_spline_out_of_bounds_raises(ctx, wen_gpu):
    # 40 atoms within 2 A: the density exceeds the F table (rho > (Nrho + 4) drho ~ 200), which must raise
    # Spline::OutOfBounds (spline.hpp:141-143); phi_Fe(0.05 A) ~ 185 in wen.eam.fs.
    basis = np.diag([30.0, 30.0, 30.0])
    pos = np.array([[10.0, 10.0, 10.0]] + [[10.0 + 0.05 * k, 10.0, 10.0] for k in range(1, 40)])
    n = len(pos)
    nl = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut())
    nl.rebuild(pos)
    with pytest.raises(fa.SplineOutOfBounds):
        wen_gpu.gradient(np.zeros(n, np.uint32), np.zeros(n, np.uint8), nl)
    # The context stays usable after the error.
    pos2, basis2 = H.kat_lattice()
    _, E, _ = gpu_eval(ctx, wen_gpu, pos2, basis2, np.zeros(len(pos2), np.uint32), np.zeros(len(pos2), np.uint8))
    assert abs(E - H.KAT_E0) <= 1e-10 * abs(H.KAT_E0)


def test_list_cutoff_smaller_than_potential_raises(ctx, wen_gpu):
    pos, basis = H.kat_lattice()
    nl = fa.List(ctx, fa.Box(basis), 5.0)
    nl.rebuild(pos)
    with pytest.raises(fa.FlyError):
        wen_gpu.energy(np.zeros(len(pos), np.uint32), np.zeros(len(pos), np.uint8), nl)


@pytest.mark.parametrize("precision", [fa.FP64, fa.MIXED])
def test_gradient_independent_of_lanes(ctx, precision):
    # Row sums use a fixed 16-slot tree, so 4, 8 and 16 lanes per atom give bitwise identical gradients.
    pot = fa.EAM(ctx, systems.potential_path(("Fe", "X")))
    pot.set_precision(precision)
    s = systems.fe_x_alloy(seed=5, cells=10)
    nl = fa.List(ctx, fa.Box(s.basis), 5.3)
    nl.rebuild(s.pos)
    out = []
    for L in (4, 8, 16):
        pot.set_lanes(L)
        out.append(pot.gradient(s.types, s.frozen, nl))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[1], out[2])


@pytest.fixture(scope="module")
def alloy30():
    """The 30^3 Fe-X alloy of the gradient benchmark (54,529 atoms) with its FP64 oracle gradient at the
    benchmark's list cutoff (r_c + 0.6 A skin)."""
    path = systems.potential_path(("Fe", "X"))
    s = systems.fe_x_alloy(seed=2, cells=30)
    data = po.DataEAM(path)
    rc = 5.3 + 0.6
    _, Eo, go = orc_eval(data, s.pos, s.basis, s.types, s.frozen, rc)
    return path, s, rc, Eo, go


# Mixed form (FP32 pair distance, splines and force magnitude; FP64 sums): the largest force error
# against the FP64 oracle, relative to the largest force component.  Measured 3e-7 on the 30^3 alloy.
MIXED_FORCE_RTOL = 2e-6


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_fe_x_alloy_mixed_vs_oracle_30(ctx, alloy30, lanes):
    # The kernels the gradient benchmark times: two species (MULTI) in the mixed form.
    path, s, rc, Eo, go = alloy30
    print(f"case fe_x_alloy(30^3) n={s.n} mixed lanes={lanes}", flush=True)
    pot = fa.EAM(ctx, path)
    pot.set_precision(fa.MIXED)
    pot.set_lanes(lanes)
    nl = fa.List(ctx, fa.Box(s.basis), rc)
    nl.rebuild(s.pos)
    g = pot.gradient(s.types, s.frozen, nl)
    err = np.max(np.abs(g - go)) / np.max(np.abs(go))
    rms = np.sqrt(np.mean((g - go) ** 2)) / np.sqrt(np.mean(go ** 2))
    print(f"mixed two-species max rel force error {err:.3e}, rms rel {rms:.3e}")
    assert err < MIXED_FORCE_RTOL
    E = pot.energy(s.types, s.frozen, nl)   # energies stay FP64
    assert abs(E - Eo) <= 1e-10 * abs(Eo)


def test_fe_x_alloy_fp64_vs_oracle_30(ctx, alloy30):
    path, s, rc, Eo, go = alloy30
    pot = fa.EAM(ctx, path)
    nl = fa.List(ctx, fa.Box(s.basis), rc)
    nl.rebuild(s.pos)
    g = pot.gradient(s.types, s.frozen, nl)
    assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))


def test_rejected_types_leave_no_stale_staging(ctx, wen_gpu, wen_orc):
    # A call rejected for an out-of-range TypeID must not leave the staged types (or the one-species
    # shortcut) of that call behind: the next call with the earlier types matches the oracle.
    rng = np.random.default_rng(31)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.05, pos.shape)
    n = len(pos)
    t1 = (rng.random(n) < 0.2).astype(np.uint32)
    f = np.zeros(n, np.uint8)
    nl = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut())
    nl.rebuild(pos)
    g_first = wen_gpu.gradient(t1, f, nl)
    # A larger system (reallocates the staging buffers) whose last type is out of range.
    big = np.concatenate([pos, pos + np.array([basis[0, 0], 0, 0])])
    bigbox = basis.copy()
    bigbox[0, 0] *= 2
    nl2 = fa.List(ctx, fa.Box(bigbox), wen_gpu.r_cut())
    nl2.rebuild(big)
    bad = np.zeros(2 * n, np.uint32)
    bad[-1] = 7
    with pytest.raises(fa.FlyError):
        wen_gpu.gradient(bad, np.zeros(2 * n, np.uint8), nl2)
    # One species but one bad entry (the one-species shortcut must not stick either).
    bad1 = np.zeros(n, np.uint32)
    bad1[-1] = 9
    with pytest.raises(fa.FlyError):
        wen_gpu.gradient(bad1, f, nl)
    g = wen_gpu.gradient(t1, f, nl)
    onl = po.List(po.Box(basis), wen_orc.r_cut)
    onl.rebuild(pos, 4)
    go = po.gradient(wen_orc, onl, t1, f, 4)
    assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))
    assert np.array_equal(g, g_first)


def test_device_arrays_checked(ctx, wen_gpu):
    torch = pytest.importorskip("torch")
    pos, basis = H.kat_lattice()
    n = len(pos)
    dev = torch.device(f"cuda:{ctx.device}")
    nl = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut())
    nl.rebuild(pos)
    t = torch.zeros(n, dtype=torch.uint32, device=dev)
    f = torch.zeros(n, dtype=torch.uint8, device=dev)
    g = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    wen_gpu.gradient(t, f, nl, g)
    torch.cuda.synchronize()
    assert np.linalg.norm(g[0].cpu().numpy() - H.KAT_G0) < 1e-10
    with pytest.raises(fa.FlyError):   # FP32 output: would receive 3n doubles
        wen_gpu.gradient(t, f, nl, torch.zeros((n, 3), dtype=torch.float32, device=dev))
    with pytest.raises(fa.FlyError):   # non-contiguous view
        wen_gpu.gradient(t, f, nl, torch.zeros((3, n), dtype=torch.float64, device=dev).t())
    with pytest.raises(fa.FlyError):   # too few elements
        wen_gpu.gradient(t, f, nl, torch.zeros((n - 1, 3), dtype=torch.float64, device=dev))
    with pytest.raises(fa.FlyError):   # int64 types
        wen_gpu.gradient(t.to(torch.int64), f, nl, g)
    # Out-of-range device TypeIDs are caught by the kernels and reported at the next synchronisation.
    tb = t.clone()
    tb[5] = 11
    with pytest.raises(fa.FlyError):
        wen_gpu.energy(tb, f, nl)
    E = wen_gpu.energy(t, f, nl)
    assert abs(E - H.KAT_E0) <= 1e-10 * abs(H.KAT_E0)


def test_slot_layout_follows_types_positions_and_rebuilds(ctx, wen_gpu, wen_orc):
    # The gradient passes run in the list's cell-order layout with the types gathered into it and into the
    # w slot of its positions (cached per list build for staged host types).  Every change the cache must
    # see: new host types on the same build, an update (positions by slot, species kept), a rebuild,
    # device types in between, two lists in turn, and two potential handles on one list.
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(41)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.05, pos.shape)
    n = len(pos)
    f = np.zeros(n, np.uint8)
    f[::7] = 1
    t1 = (rng.random(n) < 0.2).astype(np.uint32)
    t2 = (rng.random(n) < 0.3).astype(np.uint32)

    def oracle(p, t):
        onl = po.List(po.Box(basis), wen_orc.r_cut)
        onl.rebuild(p, 4)
        return po.gradient(wen_orc, onl, t, f, 4)

    def check(g, p, t):
        go = oracle(p, t)
        assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))

    nl = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut() + 0.5)
    nl.rebuild(pos)
    check(wen_gpu.gradient(t1, f, nl), pos, t1)
    check(wen_gpu.gradient(t2, f, nl), pos, t2)             # new staged types, same build
    d = rng.normal(0, 0.02, pos.shape)                      # |d| well inside half the 0.5 A skin
    nl.update(d)                                            # positions move by -d (List::update)
    moved = pos - d
    check(wen_gpu.gradient(t2, f, nl), moved, t2)
    other = moved + rng.normal(0, 0.1, pos.shape)
    nl.rebuild(other)                                       # a new layout
    check(wen_gpu.gradient(t2, f, nl), other, t2)
    dev = torch.device(f"cuda:{ctx.device}")
    td = torch.from_numpy(t1).to(dev)
    fd = torch.from_numpy(f).to(dev)
    gd = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    wen_gpu.gradient(td, fd, nl, gd)                         # device types overwrite the slot-order copy
    torch.cuda.synchronize()
    check(gd.cpu().numpy(), other, t1)
    check(wen_gpu.gradient(t2, f, nl), other, t2)            # staged types again: gathered again
    nl_b = fa.List(ctx, fa.Box(basis), wen_gpu.r_cut() + 0.5)
    nl_b.rebuild(pos)
    for lst, p in ((nl_b, pos), (nl, other), (nl_b, pos)):
        check(wen_gpu.gradient(t2, f, lst), p, t2)
    # Two potential handles with their own staged types on one list: the list's per-slot species copy
    # was last written by the other handle on every call.
    wen_b = fa.EAM(ctx, H.WEN)
    for pot_, t_ in ((wen_gpu, t1), (wen_b, t2), (wen_gpu, t1), (wen_b, t2)):
        check(pot_.gradient(t_, f, nl), other, t_)


def test_staged_atoms_follow_stage_calls_and_rebuilds(ctx, wen_orc):
    # flyg_eam_stage_atoms keeps species and frozen flags in the handle; calls given NULL for both use
    # them (the list-layout gather is then cached per build, as for host arrays).  Every result must follow
    # the last staging: from device tensors, from host arrays, after an update and a rebuild; the staged
    # form gives the bits of the per-call forms; a size mismatch or nothing staged is refused.
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(43)
    pos, basis = H.kat_lattice()
    pos = pos + rng.normal(0, 0.05, pos.shape)
    n = len(pos)
    f = np.zeros(n, np.uint8)
    f[::5] = 1
    t1 = (rng.random(n) < 0.2).astype(np.uint32)
    t2 = (rng.random(n) < 0.4).astype(np.uint32)
    dev = torch.device(f"cuda:{ctx.device}")

    def oracle(p, t):
        onl = po.List(po.Box(basis), wen_orc.r_cut)
        onl.rebuild(p, 4)
        return po.gradient(wen_orc, onl, t, f, 4)

    def check(g, p, t):
        go = oracle(p, t)
        assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))

    pot = fa.EAM(ctx, H.WEN)
    nl = fa.List(ctx, fa.Box(basis), pot.r_cut() + 0.5)
    nl.rebuild(pos)
    with pytest.raises(fa.FlyError, match="staged"):
        pot.gradient(None, None, nl)
    pot.stage(torch.from_numpy(t1).to(dev), torch.from_numpy(f).to(dev))
    g1 = pot.gradient(None, None, nl)
    check(g1, pos, t1)
    assert np.array_equal(g1, pot.gradient(None, None, nl))          # cached gather, same bits
    gd = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    pot.gradient(torch.from_numpy(t1).to(dev), torch.from_numpy(f).to(dev), nl, gd)
    torch.cuda.synchronize()
    assert np.array_equal(g1, gd.cpu().numpy())                      # per-call device arrays: same bits
    pot.stage(torch.from_numpy(t2).to(dev), torch.from_numpy(f).to(dev))
    check(pot.gradient(None, None, nl), pos, t2)
    d = rng.normal(0, 0.02, pos.shape)
    nl.update(d)
    check(pot.gradient(None, None, nl), pos - d, t2)
    other = pos - d + rng.normal(0, 0.1, pos.shape)
    nl.rebuild(other)
    check(pot.gradient(None, None, nl), other, t2)
    pot.stage(t1, f)                                                 # host arrays
    check(pot.gradient(None, None, nl), other, t1)
    assert np.array_equal(pot.gradient(t1, f, nl), pot.gradient(None, None, nl))
    pot.stage(t1[:-1], f[:-1])
    with pytest.raises(fa.FlyError, match="staged"):
        pot.gradient(None, None, nl)
