"""Shared test inputs: the reference's EAM known-answer lattice and small defect systems."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEN = os.path.join(HERE, "golden", "wen.eam.fs")

# Known answers from the reference tests (tests/potential/EAM/eam.cpp:119-141, data.cpp:42-43).
KAT_E0 = -2748.5339306065985
KAT_G0 = -0.19182439085483005
KAT_H00 = np.array([
    [+9.0654715551510830, -1.0875103661087880, -1.0875103661087877],
    [-1.0875103661087880, +9.0654715551510830, -1.0875103661087873],
    [-1.0875103661087877, -1.0875103661087873, +9.0654715551510830],
])
KAT_F0 = (-0.1414216239132407, -0.2000010747904311)


def kat_lattice(erase=True):
    """make_super (tests/potential/EAM/eam.cpp:36-96): 7^3 bcc at the T=300 K lattice constant, site 1
    removed; sites ordered z, y, x, motif."""
    T = 300.0
    V = 11.64012 + T * (9.37798e-5 + T * (3.643134e-7 + T * (1.851593e-10 + T * 5.669148e-14)))
    a = (2 * V) ** (1.0 / 3)
    lat = []
    for k in range(7):
        for j in range(7):
            for i in range(7):
                for off in (0.0, 0.5):
                    lat.append((np.array([i, j, k], float) + off) * a)
    pos = np.array(lat)
    if erase:
        pos = np.delete(pos, 1, axis=0)
    return pos, np.diag([7 * a] * 3)


def random_box(rng, full_periodic=False):
    """gen_rand (tests/neigh/list.cpp:128-165): random upper-triangular (half the time diagonal) basis with
    diagonal in [1, 10] + U(0,1), 100..1000 atoms uniformly inside, random periodic axes."""
    basis = rng.random((3, 3))
    basis[np.diag_indices(3)] += 1 + 9 * rng.random(3)
    basis = np.triu(basis)
    if rng.random() < 0.5:
        basis = np.diag(np.diag(basis))
    periodic = np.ones(3, np.int32) if full_periodic else (rng.random(3) < 0.5).astype(np.int32)
    n = int(rng.integers(100, 1001))
    pos = (basis @ rng.random((3, n))).T
    return basis, periodic, np.ascontiguousarray(pos)


def min_width(basis):
    if np.allclose(basis, np.diag(np.diag(basis)), atol=1e-4):
        return float(np.min(np.diag(basis)))
    w = []
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        nrm = np.cross(basis[:, j], basis[:, k])
        nrm /= np.linalg.norm(nrm)
        w.append(abs(nrm @ basis[:, i]))
    return float(min(w))


def brute_pairs(basis, periodic, pos, rc):
    """slow_neigh_list (tests/neigh/list.cpp:42-97): O(N^2) minimum image over +-1 cell shifts; returns
    {(i, j): dr} with dr = (x_j + shift) - x_i."""
    shifts = []
    for s2 in (-1, 0, 1):
        for s1 in (-1, 0, 1):
            for s0 in (-1, 0, 1):
                s = np.array([s0, s1, s2])
                if np.any((s != 0) & (periodic == 0)):
                    continue
                shifts.append(basis @ s)
    shifts = np.array(shifts)
    n = len(pos)
    out = {}
    for i in range(n):
        d = pos[None, :, :] + shifts[:, None, :] - pos[i][None, None, :]   # (S, n, 3)
        r2 = np.einsum("snk,snk->sn", d, d)
        best = np.argmin(r2, axis=0)
        bd = d[best, np.arange(n)]
        br = r2[best, np.arange(n)]
        for j in np.nonzero(br < rc * rc)[0]:
            if j != i:
                out[(i, int(j))] = bd[j]
    return out


def pair_set(off, j, s):
    """Multiset of (i, j, shift code) from CSR export."""
    n = len(off) - 1
    i = np.repeat(np.arange(n), np.diff(off))
    keys = np.stack([i, j.astype(np.int64), s.astype(np.int64)], axis=1)
    return keys[np.lexsort(keys.T[::-1])]


def library_kick(s, centre, radius, state):
    """Master::perturb start of system s (library flyg_perturb) around `centre` with atoms beyond `radius`
    frozen, in the sparse form the serving API takes: (atoms, positions, axis) of the displaced atoms.
    `state` (4 uint64) advances."""
    import openfly_amd as fa
    L = np.diag(s.basis)
    d = s.pos - s.pos[centre]
    d -= L * np.round(d / L)
    frozen = (np.einsum("ij,ij->i", d, d) > radius * radius).astype(np.uint8)
    x, ax, _ = fa.perturb(fa.Box(s.basis), s.pos, frozen, int(centre), state)
    idx = np.nonzero(np.any(ax != 0, axis=1))[0].astype(np.int64)
    return idx, x[idx], ax[idx]
