"""Local-environment catalogue: the request source of the serving loop (cache misses of a kMC run).

Follows env::Catalogue::rebuild (catalogue.hpp:170-206, catalogue.cpp:66-200): every atom's local
environment is the atoms within r_env of it (colour = 2 * type + frozen, env/geometry.hpp:391-407),
centred on its centroid; environments are bucketed by a discrete key, matched within a bucket by their
fingerprints (the sorted centre distances r_0j and the sorted neighbour-pair distances r_ij,
heuristics.hpp:65-122, Fingerprint::equiv heuristics.cpp:36-56 with delta = min(0.4 r_min, delta_max),
catalogue.hpp:425-427), and the atoms whose environment is not yet known are returned, ascending (of
several equal new environments only the first).  Three steps differ from the reference, which needs
nauty: the discrete key is a permutation-invariant invariant of the coloured graph joining atoms closer
than r_edge (the environment size, the centre's colour and the sorted (colour, degree) multiset) instead
of the canonical graph hash; the fingerprint sorts r_0j within each colour and r_ij within each colour
pair (the reference's sorts are colour-blind and its colours enter through the canonical hash and
permute_onto, which keeps colour); and there is no final permute_onto: environments whose fingerprints
agree within delta are equal.  With the default delta_max = 0.1 A, distinct environments of a relaxed
crystal stay apart.

The environments come from a neighbour list's rows (the library's List exported at r_env, or a k-d tree
on the host); atoms are processed in chunks of equal environment size, each chunk's distances and sorts
computed at once (on a torch device when one is given, FP64 either way) and matched against a bucket's
known environments at once.  Only unmatched atoms go through the sequential insertion, which gives the
same new atoms as an atom-by-atom pass (environments of different keys never match).
"""
from __future__ import annotations

import numpy as np

_SQRT2 = np.sqrt(2.0)


def host_pairs(pos, basis, r_env, atoms=None):
    """(offsets, neighbour index, displacement) of every pair closer than r_env (orthorhombic periodic box,
    minimum image), one row per atom of `atoms` (all atoms when None), rows in ascending neighbour index."""
    from scipy.spatial import cKDTree
    pos = np.asarray(pos, dtype=np.float64)
    L = np.diag(np.asarray(basis, dtype=np.float64))
    w = pos - L * np.floor(pos / L)
    w = np.where(w >= L, w - L, w)
    centres = np.arange(len(w)) if atoms is None else np.asarray(atoms, dtype=np.int64)
    rows = cKDTree(w, boxsize=L).query_ball_point(w[centres], r_env)
    off = np.zeros(len(centres) + 1, np.int64)
    js = []
    for r, (i, nb) in enumerate(zip(centres, rows)):
        nb = sorted(x for x in nb if x != i)
        js.append(nb)
        off[r + 1] = off[r] + len(nb)
    j = np.fromiter((x for nb in js for x in nb), dtype=np.int64, count=int(off[-1]))
    i_of = np.repeat(centres, np.diff(off))
    d = w[j] - w[i_of]
    d -= L * np.round(d / L)
    return off, j, d


class Catalogue:
    def __init__(self, r_env=5.2, r_edge=3.0, delta_max=0.1, chunk=2048):
        self.r_env = float(r_env)
        self.r_edge = float(r_edge)
        self.delta_max = float(delta_max)
        self.chunk = int(chunk)
        self.buckets = {}   # key -> list of (r_0j, r_ij, r_min)
        self.size = 0
        self._dev = {}      # device copies of known environments
        self._stack = {}    # key -> (R0 (K, n), RIJ (K, m), r_min (K,)) of the bucket, for the insertion test

    # Distances are sorted within colour classes by sorting class * _CLASS + r (r < 2 r_env < _CLASS).
    _CLASS = 64.0

    def _fingerprints(self, rel, col, device):
        """r_0j (B, n) sorted by the neighbour's colour, then distance; r_ij (B, n(n-1)/2) sorted by the
        pair's colours, then distance; the smallest distance (B,); degrees (B, n+1) in the r_edge graph
        (host array).  Torch tensors on `device` when one is given, numpy arrays otherwise."""
        n = rel.shape[1] - 1
        nc = int(col.max()) + 1
        c0 = col[:, 1:].astype(np.float64)
        cmin = np.minimum(col[:, 1:, None], col[:, None, 1:])
        cmax = np.maximum(col[:, 1:, None], col[:, None, 1:])
        cp = (cmin * nc + cmax).astype(np.float64)
        if device is not None:
            import torch
            t = torch.from_numpy(rel).to(device)
            full = torch.cdist(t, t, compute_mode="donot_use_mm_for_euclid_dist")
            iu = torch.triu_indices(n, n, 1, device=device)
            k0 = torch.from_numpy(c0).to(device) * self._CLASS
            kij = torch.from_numpy(cp).to(device)[:, iu[0], iu[1]] * self._CLASS
            r0 = torch.sort(full[:, 0, 1:] + k0, dim=1).values - torch.sort(k0, dim=1).values
            rij = torch.sort(full[:, 1:, 1:][:, iu[0], iu[1]] + kij, dim=1).values - torch.sort(kij, dim=1).values
            rmin = torch.minimum(r0.amin(dim=1), rij.amin(dim=1)) if n > 1 else r0.amin(dim=1)
            deg = ((full < self.r_edge).sum(dim=2) - 1).cpu().numpy()
            return r0, rij, rmin, deg
        full = np.sqrt(((rel[:, :, None, :] - rel[:, None, :, :]) ** 2).sum(axis=3))
        iu = np.triu_indices(n, 1)
        k0 = c0 * self._CLASS
        kij = cp[:, iu[0], iu[1]] * self._CLASS
        r0 = np.sort(full[:, 0, 1:] + k0, axis=1) - np.sort(k0, axis=1)
        rij = np.sort(full[:, 1:, 1:][:, iu[0], iu[1]] + kij, axis=1) - np.sort(kij, axis=1)
        rmin = np.minimum(r0.min(axis=1), rij.min(axis=1)) if n > 1 else r0.min(axis=1)
        deg = (full < self.r_edge).sum(axis=2) - 1
        return r0, rij, rmin, deg

    def _groups(self, off, j, d, colour, centres, device):
        """Yields (key, atom indices, R0, RIJ, r_min) per chunk of atoms with n neighbours and one key
        (row r of the pairs is atom centres[r])."""
        cnt = np.diff(off)
        for n in np.unique(cnt):
            if n < 1:
                continue
            atoms = np.nonzero(cnt == n)[0]
            for c0 in range(0, len(atoms), self.chunk):
                idx = atoms[c0:c0 + self.chunk]
                B = len(idx)
                sel = off[idx][:, None] + np.arange(n)[None, :]
                rel = np.concatenate([np.zeros((B, 1, 3)), d[sel]], axis=1)
                rel -= rel.mean(axis=1, keepdims=True)   # Geometry::centre: origin at the centroid
                col = np.concatenate([colour[centres[idx]][:, None], colour[j[sel]]], axis=1)
                r0, rij, rmin, deg = self._fingerprints(rel, col, device)
                code = np.sort(col * 4096 + deg, axis=1)
                groups = {}
                for b in range(B):
                    groups.setdefault((int(col[b, 0]), int(n) + 1, code[b].tobytes()), []).append(b)
                for k, bs in groups.items():
                    if device is not None:
                        import torch
                        rows = torch.as_tensor(bs, device=device)
                        yield k, centres[idx[bs]], r0[rows], rij[rows], rmin[rows]
                    else:
                        yield k, centres[idx[bs]], r0[bs], rij[bs], rmin[bs]

    @staticmethod
    def _match(R0, RIJ, rmin, ref, delta_max):
        """Fingerprint::equiv of each row against one known environment (numpy or torch rows, ref alike)."""
        if isinstance(R0, np.ndarray):
            tol = np.minimum(0.4 * np.minimum(rmin, ref[2]), delta_max) * _SQRT2
            return (np.abs(R0 - ref[0]) <= tol[:, None]).all(axis=1) & (np.abs(RIJ - ref[1]) <= tol[:, None]).all(axis=1)
        import torch
        tol = torch.clamp(0.4 * torch.clamp(rmin, max=ref[2]), max=delta_max) * _SQRT2
        return (torch.abs(R0 - ref[0]) <= tol[:, None]).all(dim=1) & (torch.abs(RIJ - ref[1]) <= tol[:, None]).all(dim=1)

    def dump(self, file):
        """Catalogue::dump (catalogue.hpp:341-348): the options and every known environment (per bucket
        key, the fingerprints in insertion order) to a numpy .npz archive holding plain arrays only."""
        arrays = {"options": np.array([self.r_env, self.r_edge, self.delta_max]),
                  "counts": np.array([len(self._stack), self.size], np.int64)}
        for b, (key, st) in enumerate(self._stack.items()):
            arrays[f"k{b}_head"] = np.array([key[0], key[1]], np.int64)
            arrays[f"k{b}_code"] = np.frombuffer(key[2], np.uint8)
            arrays[f"k{b}_r0"], arrays[f"k{b}_rij"], arrays[f"k{b}_rmin"] = st
        np.savez(file, **arrays)

    @classmethod
    def load(cls, file, r_env=5.2, r_edge=3.0, delta_max=0.1, chunk=2048):
        """Catalogue(opt, fin) (catalogue.hpp:314-339): a catalogue written by dump.  Its options must be
        the given ones (FlyError otherwise, as the reference's verify), and a file that is not such an
        archive raises FlyError("Catalogue is a bad binary: ...")."""
        import zipfile

        from .flygpu import FlyError
        c = cls(r_env, r_edge, delta_max, chunk)
        try:
            with np.load(file, allow_pickle=False) as z:
                opt = z["options"]
                nb, size = (int(v) for v in z["counts"])
                stacks = []
                for b in range(nb):
                    head = z[f"k{b}_head"]
                    key = (int(head[0]), int(head[1]), z[f"k{b}_code"].astype(np.uint8).tobytes())
                    r0, rij, rmin = (np.asarray(z[f"k{b}_{x}"], dtype=np.float64) for x in ("r0", "rij", "rmin"))
                    if r0.ndim != 2 or rij.ndim != 2 or rmin.shape != (len(r0),) or len(rij) != len(r0):
                        raise ValueError(f"bucket {b} has inconsistent shapes")
                    stacks.append((key, (r0, rij, rmin)))
        except (KeyError, ValueError, OSError, EOFError, zipfile.BadZipFile) as e:
            raise FlyError(f"Catalogue is a bad binary: {e}") from None
        for name, mine, theirs in (("r_env", c.r_env, opt[0]), ("r_edge", c.r_edge, opt[1]),
                                   ("delta_max", c.delta_max, opt[2])):
            if mine != float(theirs):
                raise FlyError(f"Loading a catalogue with contradicting .{name}")
        for key, st in stacks:
            c._stack[key] = st
            c.buckets[key] = [(st[0][i], st[1][i], float(st[2][i])) for i in range(len(st[2]))]
        c.size = size
        if c.size != sum(len(v) for v in c.buckets.values()):
            raise FlyError("Catalogue is a bad binary: environment count does not match its buckets")
        return c

    def _device_ref(self, ref, device):
        import torch
        key = (id(ref), str(device))
        if key not in self._dev:
            self._dev[key] = (torch.from_numpy(ref[0]).to(device), torch.from_numpy(ref[1]).to(device), ref[2])
        return self._dev[key]

    def rebuild(self, pos, types, frozen, basis, pairs=None, device=None, atoms=None):
        """Catalogue::rebuild: indices of the atoms whose environment was new (now inserted), ascending.
        pairs: (offsets, j, displacement) rows within r_env (e.g. List.pairs(r_env) of the library),
        built on the host when absent; device: a torch device for the distances and sorts; atoms: only
        these atoms' environments are looked up (the ones a kMC step changed), pairs then holding their
        rows in that order."""
        centres = np.arange(len(pos)) if atoms is None else np.asarray(atoms, dtype=np.int64)
        off, j, d = pairs if pairs is not None else host_pairs(pos, basis, self.r_env, atoms)
        off = np.asarray(off, dtype=np.int64)
        j = np.asarray(j, dtype=np.int64)
        d = np.asarray(d, dtype=np.float64)
        colour = 2 * np.asarray(types, dtype=np.int64) + (np.asarray(frozen) != 0)
        pending = []   # unmatched environments, inserted in ascending atom order below
        for key, idx, R0, RIJ, rmin in self._groups(off, j, d, colour, centres, device):
            if device is None:
                left = np.ones(len(idx), dtype=bool)
                for ref in self.buckets.get(key, ()):
                    left &= ~self._match(R0, RIJ, rmin, ref, self.delta_max)
                    if not left.any():
                        break
                rows = np.nonzero(left)[0]
                R0h, RIJh, rminh = R0[rows], RIJ[rows], rmin[rows]
            else:
                import torch
                left = torch.ones(len(idx), dtype=torch.bool, device=device)
                for ref in self.buckets.get(key, ()):
                    left &= ~self._match(R0, RIJ, rmin, self._device_ref(ref, device), self.delta_max)
                rows = np.nonzero(left.cpu().numpy())[0]
                sel = torch.as_tensor(rows, device=device)
                R0h, RIJh, rminh = R0[sel].cpu().numpy(), RIJ[sel].cpu().numpy(), rmin[sel].cpu().numpy()
            for q, b in enumerate(rows):
                pending.append((int(idx[b]), key, R0h[q], RIJh[q], float(rminh[q])))
        new = []
        for i, key, r0, rij, rmin in sorted(pending, key=lambda t: t[0]):
            st = self._stack.get(key)
            if st is not None:   # the bucket's environments, one vectorised Fingerprint::equiv
                tol = np.minimum(0.4 * np.minimum(rmin, st[2]), self.delta_max) * _SQRT2
                if ((np.abs(st[0] - r0) <= tol[:, None]).all(axis=1)
                        & (np.abs(st[1] - rij) <= tol[:, None]).all(axis=1)).any():
                    continue
            self.buckets.setdefault(key, []).append((r0, rij, rmin))
            self._stack[key] = ((r0[None], rij[None], np.array([rmin])) if st is None else
                                (np.vstack([st[0], r0]), np.vstack([st[1], rij]), np.append(st[2], rmin)))
            self.size += 1
            new.append(i)
        return new
