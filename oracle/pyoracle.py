"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker or the timed CPU baseline, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same sources (the CPU baseline's -march=native copy).
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_l = C.POINTER(C.c_long)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
_u64 = C.POINTER(C.c_uint64)
_vp = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_box_new": (_vp, [_d, _i]),
            "orc_box_free": (None, [_vp]),
            "orc_box_info": (C.c_int, [_vp, _i, _d]),
            "orc_box_canon": (C.c_int, [_vp, _d, C.c_long, _d]),
            "orc_grid_info": (C.c_int, [_vp, C.c_double, _i, _d]),
            "orc_grid_cell_idx": (C.c_int, [_vp, C.c_double, _d, C.c_long, _i]),
            "orc_list_new": (_vp, [_vp, C.c_double]),
            "orc_list_free": (None, [_vp]),
            "orc_list_rebuild": (C.c_int, [_vp, _d, C.c_long, C.c_int]),
            "orc_list_update": (C.c_int, [_vp, _d, C.c_long]),
            "orc_list_num_ghosts": (C.c_long, [_vp]),
            "orc_list_count": (C.c_long, [_vp, C.c_double]),
            "orc_list_export": (C.c_int, [_vp, C.c_double, _l, _i, _i, _d]),
            "orc_spline_new": (_vp, [_d, C.c_long, C.c_double]),
            "orc_spline_free": (None, [_vp]),
            "orc_spline_eval": (C.c_int, [_vp, _d, C.c_long, C.c_int, _d]),
            "orc_spline_coeffs": (C.c_long, [_vp, _d]),
            "orc_data_new": (_vp, [C.c_char_p]),
            "orc_data_free": (None, [_vp]),
            "orc_data_info": (C.c_int, [_vp, _i, _d, _d]),
            "orc_data_eval": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _d, C.c_long, C.c_int, _d]),
            "orc_energy": (C.c_int, [_vp, _vp, _u32, _u8, C.c_long, C.c_int, _d]),
            "orc_gradient": (C.c_int, [_vp, _vp, _u32, _u8, C.c_long, C.c_int, _d]),
            "orc_hessian": (C.c_int, [_vp, _vp, _u32, _u8, C.c_long, C.c_int, _d]),
            "orc_hvp": (C.c_int, [_vp, _vp, _u32, _u8, C.c_long, C.c_int, _d, _d]),
            "orc_steplbfgs_new": (_vp, [C.c_int]),
            "orc_steplbfgs_free": (None, [_vp]),
            "orc_steplbfgs_clear": (None, [_vp]),
            "orc_steplbfgs_step": (C.c_int, [_vp, _d, _d, C.c_long, _d]),
            "orc_minimise": (C.c_int, [_vp, _vp, _d, _d, _u32, _u8, C.c_long, C.c_int, _d, _i]),
            "orc_find_sp": (C.c_int, [_vp, _vp, _d, _d, _u32, _u8, C.c_long, _d, _d, C.c_int, C.c_double,
                                      C.c_int, C.c_int, _d, _d, _i, _l, _d]),
            "orc_find_sp_many": (C.c_int, [_vp, _vp, _d, _d, _u32, _u8, C.c_long, C.c_int, _d, C.c_int, C.c_int,
                                           _d, _d, _i, _l]),
            "orc_perturb": (C.c_int, [_u64, C.c_double, C.c_double, _vp, C.c_long, _d, _u8, C.c_long, _d, _d, _d]),
            "orc_xoshiro": (C.c_int, [_u64, C.c_long, C.c_int, _u64]),
            "orc_normals": (C.c_int, [_u64, C.c_long, C.c_double, C.c_double, _d]),
            "orc_rotor_eff_gradient": (C.c_int, [_vp, _vp, _d, _d, _u32, _u8, C.c_long, C.c_int, C.c_int, _d, _d,
                                                 _d, _l]),
            "orc_rotor_quadratic": (C.c_int, [C.c_long, _d, _d, _d, _d, C.c_int, _d, _d, _d, _l]),
            "orc_dimer_quadratic": (C.c_int, [C.c_long, _d, _d, _d, _d, _d, _d, C.c_int, C.c_double, C.c_int, _d,
                                              _d, _i, _l, _d]),
            "orc_dimer_mueller_brown": (C.c_int, [C.c_double, _d, _d, _d, _d, C.c_int, C.c_double, _d, _d, _i, _l,
                                                  _d]),
            "orc_mueller_brown": (C.c_int, [_d, C.c_double, _d, _d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


class OracleOutOfBounds(OracleError):
    pass


def _check(rc):
    if rc == -2:
        raise OracleOutOfBounds(lib().orc_last_error().decode())
    if rc < 0:
        raise OracleError(lib().orc_last_error().decode())
    return rc


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Box:
    def __init__(self, basis, periodic=(True, True, True)):
        self.basis = _f64(basis).reshape(3, 3)
        self.periodic = np.ascontiguousarray(np.asarray(periodic, dtype=np.int32))
        self.h = lib().orc_box_new(_p(self.basis, _d), _p(self.periodic, _i))
        if not self.h:
            raise OracleError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_box_free(self.h)

    def info(self):
        o = C.c_int()
        w = C.c_double()
        lib().orc_box_info(self.h, C.byref(o), C.byref(w))
        return bool(o.value), w.value

    def canon(self, x):
        x = _f64(x)
        out = np.empty_like(x)
        lib().orc_box_canon(self.h, _p(x, _d), x.size // 3, _p(out, _d))
        return out

    def grid(self, rc):
        shape = np.zeros(3, np.int32)
        cell = np.zeros(3)
        _check(lib().orc_grid_info(self.h, rc, _p(shape, _i), _p(cell, _d)))
        return shape, cell

    def cell_idx(self, rc, x):
        x = _f64(x)
        out = np.zeros(x.size // 3, np.int32)
        _check(lib().orc_grid_cell_idx(self.h, rc, _p(x, _d), x.size // 3, _p(out, _i)))
        return out


class List:
    def __init__(self, box: Box, r_cut: float):
        self.box = box
        self.r_cut = r_cut
        self.h = lib().orc_list_new(box.h, r_cut)
        if not self.h:
            raise OracleError(lib().orc_last_error().decode())
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_list_free(self.h)

    def rebuild(self, pos, threads=1):
        pos = _f64(pos)
        self.n = pos.size // 3
        _check(lib().orc_list_rebuild(self.h, _p(pos, _d), self.n, threads))

    def update(self, delta):
        d = _f64(delta)
        _check(lib().orc_list_update(self.h, _p(d, _d), d.size // 3))

    def num_ghosts(self):
        return lib().orc_list_num_ghosts(self.h)

    def pairs(self, r_cut=None):
        """Return (offsets, j, shift, dr) for every (i, neighbour) with r < r_cut."""
        rc = self.r_cut if r_cut is None else r_cut
        m = lib().orc_list_count(self.h, rc)
        if m < 0:
            raise OracleError(lib().orc_last_error().decode())
        off = np.zeros(self.n + 1, np.int64)
        j = np.zeros(m, np.int32)
        s = np.zeros(m, np.int32)
        dr = np.zeros((m, 3))
        _check(lib().orc_list_export(self.h, rc, _p(off, _l), _p(j, _i), _p(s, _i), _p(dr, _d)))
        return off, j, s, dr


class Spline:
    def __init__(self, y, dx):
        y = _f64(y)
        self.h = lib().orc_spline_new(_p(y, _d), y.size, dx)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_spline_free(self.h)

    def eval(self, x, deriv=0):
        x = _f64(np.atleast_1d(x))
        out = np.zeros_like(x)
        _check(lib().orc_spline_eval(self.h, _p(x, _d), x.size, deriv, _p(out, _d)))
        return out

    def coeffs(self):
        n = lib().orc_spline_coeffs(self.h, None)
        out = np.zeros((n, 4))
        lib().orc_spline_coeffs(self.h, _p(out, _d))
        return out


class DataEAM:
    def __init__(self, path):
        self.h = lib().orc_data_new(os.fsencode(path))
        if not self.h:
            raise OracleError(lib().orc_last_error().decode())
        nt = C.c_int()
        rc = C.c_double()
        lib().orc_data_info(self.h, C.byref(nt), C.byref(rc), None)
        self.ntypes = nt.value
        self.r_cut = rc.value
        self.masses = np.zeros(self.ntypes)
        lib().orc_data_info(self.h, C.byref(nt), C.byref(rc), _p(self.masses, _d))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_data_free(self.h)

    def eval(self, kind, a, b, x, deriv=0):
        x = _f64(np.atleast_1d(x))
        out = np.zeros_like(x)
        k = {"F": 0, "phi": 1, "V": 2}[kind]
        _check(lib().orc_data_eval(self.h, k, a, b, _p(x, _d), x.size, deriv, _p(out, _d)))
        return out


def _sys(types, frozen, n):
    t = np.ascontiguousarray(np.asarray(types, dtype=np.uint32).reshape(n))
    f = np.ascontiguousarray(np.asarray(frozen, dtype=np.uint8).reshape(n))
    return t, f


def energy(data, nl, types, frozen, threads=1):
    t, f = _sys(types, frozen, nl.n)
    E = C.c_double()
    _check(lib().orc_energy(data.h, nl.h, _p(t, _u32), _p(f, _u8), nl.n, threads, C.byref(E)))
    return E.value


def gradient(data, nl, types, frozen, threads=1):
    t, f = _sys(types, frozen, nl.n)
    g = np.zeros((nl.n, 3))
    _check(lib().orc_gradient(data.h, nl.h, _p(t, _u32), _p(f, _u8), nl.n, threads, _p(g, _d)))
    return g


def hessian(data, nl, types, frozen, threads=1):
    t, f = _sys(types, frozen, nl.n)
    H = np.zeros((3 * nl.n, 3 * nl.n))
    _check(lib().orc_hessian(data.h, nl.h, _p(t, _u32), _p(f, _u8), nl.n, threads, _p(H, _d)))
    return H


def hvp(data, nl, types, frozen, v, threads=1):
    t, f = _sys(types, frozen, nl.n)
    v = _f64(v)
    y = np.zeros(3 * nl.n)
    _check(lib().orc_hvp(data.h, nl.h, _p(t, _u32), _p(f, _u8), nl.n, threads, _p(v, _d), _p(y, _d)))
    return y.reshape(-1, 3)


class StepLBFGS:
    def __init__(self, m):
        self.h = lib().orc_steplbfgs_new(m)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_steplbfgs_free(self.h)

    def clear(self):
        lib().orc_steplbfgs_clear(self.h)

    def newton_step(self, r, g):
        r = _f64(r).ravel()
        g = _f64(g).ravel()
        out = np.zeros_like(r)
        _check(lib().orc_steplbfgs_step(self.h, _p(r, _d), _p(g, _d), r.size, _p(out, _d)))
        return out


LBFGS_DEFAULTS = dict(n=10, iter_max=2000, f2norm=1e-6, skin_frac=1.1, proj_tol=0.0, max_trust=0.5,
                      min_trust=0.01, grow_trust=1.5, shrink_trust=0.5)
DIMER_DEFAULTS = dict(n=10, max_steps=1000, hist_check_freq=5, f2norm=1e-5, skin_frac=1.1, proj_tol=0.0,
                      max_trust=0.5, min_trust=0.05, grow_trust=1.5, shrink_trust=0.5, convex_max=3,
                      use_history=1)
ROTOR_DEFAULTS = dict(n=6, iter_max_rot=20, delta_r=1e-3, theta_tol=np.pi / 360, relax_in_convex=1)


def _opts(defaults, over):
    d = dict(defaults)
    for k, v in (over or {}).items():
        if k not in d:
            raise KeyError(k)
        d[k] = v
    return np.array([float(v) for v in d.values()])


def minimise(box, data, pos, types, frozen, threads=1, **opts):
    n = pos.size // 3
    r = _f64(pos).copy().ravel()
    g = np.zeros(3 * n)
    t, f = _sys(types, frozen, n)
    o = _opts(LBFGS_DEFAULTS, opts)
    it = C.c_int()
    rc = _check(lib().orc_minimise(box.h, data.h, _p(r, _d), _p(g, _d), _p(t, _u32), _p(f, _u8), n, threads,
                                   _p(o, _d), C.byref(it)))
    return r.reshape(-1, 3), g.reshape(-1, 3), rc == 1, it.value


def find_sp(box, data, pos, axis, types, frozen, in_min, hist=None, theta_tol=0.0, count_frozen=None,
            threads=1, dimer=None, rotor=None):
    n = pos.size // 3
    r = _f64(pos).copy().ravel()
    ax = _f64(axis).copy().ravel()
    t, f = _sys(types, frozen, n)
    if count_frozen is None:
        count_frozen = int(f.sum())
    mn = _f64(in_min).ravel()
    hs = np.zeros(1) if hist is None or len(hist) == 0 else _f64(np.stack([np.ravel(h) for h in hist]))
    nh = 0 if hist is None else len(hist)
    do = _opts(DIMER_DEFAULTS, dimer)
    ro = _opts(ROTOR_DEFAULTS, rotor)
    ex = C.c_int()
    st = np.zeros(3, np.int64)
    curv = C.c_double()
    _check(lib().orc_find_sp(box.h, data.h, _p(r, _d), _p(ax, _d), _p(t, _u32), _p(f, _u8), n, _p(mn, _d),
                             _p(hs, _d), nh, theta_tol, count_frozen, threads, _p(do, _d), _p(ro, _d),
                             C.byref(ex), _p(st, _l), C.byref(curv)))
    return dict(pos=r.reshape(-1, 3), axis=ax.reshape(-1, 3), exit=ex.value, steps=int(st[0]), grads=int(st[1]),
                rebuilds=int(st[2]), curv=curv.value)


def find_sp_many(box, data, starts, axes, types, frozen, in_min, threads, count_frozen=None, dimer=None, rotor=None):
    k, n = starts.shape[0], starts.shape[1]
    r = _f64(starts).copy()
    ax = _f64(axes).copy()
    t, f = _sys(types, frozen, n)
    if count_frozen is None:
        count_frozen = int(f.sum())
    mn = _f64(in_min).ravel()
    do = _opts(DIMER_DEFAULTS, dimer)
    ro = _opts(ROTOR_DEFAULTS, rotor)
    ex = np.zeros(k, np.int32)
    st = np.zeros((k, 3), np.int64)
    _check(lib().orc_find_sp_many(box.h, data.h, _p(r, _d), _p(ax, _d), _p(t, _u32), _p(f, _u8), n, k, _p(mn, _d),
                                  count_frozen, threads, _p(do, _d), _p(ro, _d), _p(ex, _i), _p(st, _l)))
    return r, ax, ex, st


def perturb(seed, stddev, r_pert, nl, centre, pos, frozen):
    n = pos.size // 3
    s = np.ascontiguousarray(np.asarray(seed, dtype=np.uint64))
    r_in = _f64(pos).ravel()
    f = np.ascontiguousarray(np.asarray(frozen, dtype=np.uint8))
    r_out = np.zeros(3 * n)
    ax = np.zeros(3 * n)
    sd = C.c_double()
    _check(lib().orc_perturb(_p(s, _u64), stddev, r_pert, nl.h, centre, _p(r_in, _d), _p(f, _u8), n, _p(r_out, _d),
                             _p(ax, _d), C.byref(sd)))
    return r_out.reshape(-1, 3), ax.reshape(-1, 3), sd.value


def xoshiro(seed, m, jumps=0):
    s = np.ascontiguousarray(np.asarray(seed, dtype=np.uint64))
    out = np.zeros(m, np.uint64)
    _check(lib().orc_xoshiro(_p(s, _u64), m, jumps, _p(out, _u64)))
    return out


def normals(seed, m, mean=0.0, sd=1.0):
    s = np.ascontiguousarray(np.asarray(seed, dtype=np.uint64))
    out = np.zeros(m)
    _check(lib().orc_normals(_p(s, _u64), m, mean, sd, _p(out, _d)))
    return out


def rotor_eff_gradient(box, data, pos, axis, types, frozen, count_frozen=None, threads=1, rotor=None):
    """Rotor::eff_gradient at pos: (curvature, rotated axis, effective gradient, gradient evaluations)."""
    n = pos.size // 3
    r = _f64(pos).ravel()
    ax = _f64(axis).copy().ravel()
    t, f = _sys(types, frozen, n)
    if count_frozen is None:
        count_frozen = int(f.sum())
    out = np.zeros(3 * n)
    curv = C.c_double()
    grads = C.c_long()
    _check(lib().orc_rotor_eff_gradient(box.h, data.h, _p(r, _d), _p(ax, _d), _p(t, _u32), _p(f, _u8), n,
                                        count_frozen, threads, _p(_opts(ROTOR_DEFAULTS, rotor), _d), _p(out, _d),
                                        C.byref(curv), C.byref(grads)))
    return curv.value, ax.reshape(-1, 3), out.reshape(-1, 3), grads.value


def rotor_quadratic(A, c, x, axis, count_frozen=1, rotor=None):
    """The same rotor on E = 1/2 (x - c)^T A (x - c): (curvature, axis, effective gradient, evaluations)."""
    A = _f64(A)
    n3 = A.shape[0]
    c = _f64(c).ravel()
    x = _f64(x).ravel()
    ax = _f64(axis).copy().ravel()
    out = np.zeros(n3)
    curv = C.c_double()
    grads = C.c_long()
    _check(lib().orc_rotor_quadratic(n3, _p(A, _d), _p(c, _d), _p(x, _d), _p(ax, _d), count_frozen,
                                     _p(_opts(ROTOR_DEFAULTS, rotor), _d), _p(out, _d), C.byref(curv),
                                     C.byref(grads)))
    return curv.value, ax, out, grads.value


def _dimer_out(r, ax, ex, st, curv):
    return dict(pos=r, axis=ax, exit=ex.value, steps=int(st[0]), grads=int(st[1]), rebuilds=int(st[2]),
                curv=curv.value)


def dimer_quadratic(A, c, x, axis, in_min, hist=None, theta_tol=0.0, count_frozen=1, dimer=None, rotor=None):
    """Dimer::find_sp's loop on E = 1/2 (x - c)^T A (x - c)."""
    A = _f64(A)
    n3 = A.shape[0]
    c = _f64(c).ravel()
    r = _f64(x).copy().ravel()
    ax = _f64(axis).copy().ravel()
    mn = _f64(in_min).ravel()
    hs = np.zeros(1) if not hist else _f64(np.stack([np.ravel(h) for h in hist]))
    ex, st, curv = C.c_int(), np.zeros(3, np.int64), C.c_double()
    _check(lib().orc_dimer_quadratic(n3, _p(A, _d), _p(c, _d), _p(r, _d), _p(ax, _d), _p(mn, _d), _p(hs, _d),
                                     len(hist or []), theta_tol, count_frozen, _p(_opts(DIMER_DEFAULTS, dimer), _d),
                                     _p(_opts(ROTOR_DEFAULTS, rotor), _d), C.byref(ex), _p(st, _l), C.byref(curv)))
    return _dimer_out(r, ax, ex, st, curv)


def dimer_mueller_brown(x, axis, in_min, kz=100.0, hist=None, theta_tol=0.0, dimer=None, rotor=None):
    """Dimer::find_sp's loop for one atom on the Mueller-Brown surface (x, y) + kz z^2 / 2."""
    r = _f64(x).copy().ravel()
    ax = _f64(axis).copy().ravel()
    mn = _f64(in_min).ravel()
    hs = np.zeros(1) if not hist else _f64(np.stack([np.ravel(h) for h in hist]))
    ex, st, curv = C.c_int(), np.zeros(3, np.int64), C.c_double()
    _check(lib().orc_dimer_mueller_brown(kz, _p(r, _d), _p(ax, _d), _p(mn, _d), _p(hs, _d), len(hist or []),
                                         theta_tol, _p(_opts(DIMER_DEFAULTS, dimer), _d),
                                         _p(_opts(ROTOR_DEFAULTS, rotor), _d), C.byref(ex), _p(st, _l),
                                         C.byref(curv)))
    return _dimer_out(r, ax, ex, st, curv)


def mueller_brown(x, kz=100.0):
    """(energy, gradient) of the Mueller-Brown surface as the oracle evaluates it."""
    x = _f64(x).ravel()
    e = C.c_double()
    g = np.zeros(3)
    lib().orc_mueller_brown(_p(x, _d), kz, C.byref(e), _p(g, _d))
    return e.value, g
