"""Serving loop: saddle-search requests from a simulated kMC loop, served with continuous batching.

A kMC loop (SKMC::skmc, skmc.hpp:240-262) asks Master::find_mechs (find.cpp:211-312) for new mechanisms
whenever its catalogue misses an environment; each miss becomes a request of 8-64 dimer searches on the
current system.  Here the requests arrive as a Poisson process; each carries its own system (bcc Fe,
N log-uniform in [2000, 250000] atoms, 0.5-2 % vacancies + interstitials), which the server admits
under its device memory budget, keeps resident while the request's searches run, and frees afterwards.
Searches of requests that arrive while others run join the next launch (flyg_server_*).

Latency of a request = time from its arrival to the completion of its last search.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import flygpu as fg
from . import systems


@dataclasses.dataclass
class Request:
    t: float                 # arrival time (s after the start)
    system: systems.System   # the kMC system it came from (shared by that loop's requests)
    centres: np.ndarray
    radii: np.ndarray
    prng: np.ndarray         # per search: Xoshiro256** state of its Master::perturb start (k x 4 uint64)
    sys_id: int = -1
    remaining: int = 0
    admitted: float = -1.0
    done: float = -1.0
    results: dict = dataclasses.field(default_factory=dict)


class Near:
    """Periodic fixed-radius neighbour queries on a system's positions (k-d tree over the box)."""

    def __init__(self, s: systems.System):
        from scipy.spatial import cKDTree
        self.L = np.diag(s.basis).copy()
        w = s.pos - self.L * np.floor(s.pos / self.L)
        w = np.where(w >= self.L, w - self.L, w)
        self.tree = cKDTree(w, boxsize=self.L)
        self.pos = s.pos

    def within(self, p, r):
        """Sorted indices of the atoms within r of point p (minimum image)."""
        q = np.asarray(p, float) - self.L * np.floor(np.asarray(p, float) / self.L)
        q = np.where(q >= self.L, q - self.L, q)
        return np.array(sorted(self.tree.query_ball_point(q, r)), dtype=np.int64)


def _near_of(s):
    nr = getattr(s, "_near", None)
    if nr is None:
        nr = Near(s)
        s._near = nr
    return nr


def sample_centres(s: systems.System, k, rng, r=3.0):
    """k random defect-adjacent atoms: pick a random defect (vacancy site or interstitial), then a random
    atom within r of it."""
    pts = []
    if getattr(s, "vacancies", None) is not None and len(s.vacancies):
        pts.append(np.asarray(s.vacancies))
    if getattr(s, "interstitial_ids", None) is not None and len(s.interstitial_ids):
        pts.append(s.pos[s.interstitial_ids])
    pts = np.concatenate(pts) if pts else s.pos
    nr = _near_of(s)
    out = np.empty(k, np.int64)
    for q in range(k):
        near = nr.within(pts[rng.integers(len(pts))], r)
        out[q] = rng.choice(near) if len(near) else rng.integers(s.n)
    return out


def make_pool(count, seed=4, n_min=2000, n_max=250000, defects=(0.005, 0.02), relax=None):
    """`count` kMC systems: bcc Fe with N log-uniform in [n_min, n_max] atoms and a defect fraction uniform
    in `defects` (half vacancies, half tetrahedral interstitials).  relax(system) -> positions, if given,
    turns each into its relaxed minimum (the state a kMC loop searches from)."""
    rng = np.random.default_rng(seed)
    pool = []
    for q in range(count):
        n_target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        cells = max(10, int(round((n_target / 2.0) ** (1.0 / 3.0))))
        frac = rng.uniform(*defects)
        n_def = max(2, int(round(frac * 2 * cells ** 3)))
        s = systems.defect_system(cells, n_def // 2, n_def - n_def // 2, int(rng.integers(1 << 30)))
        if relax is not None:
            s.pos = relax(s)
        s.pool_index = q
        pool.append(s)
    return pool


def make_requests(pool, count, rate=200.0, seed=5, searches=(8, 64), radii=(4.0, 9.0), one_system_each=False,
                  misses=None):
    """Poisson arrivals at `rate`/s, each from a uniformly chosen system of `pool`, asking for a uniform
    number of searches in `searches` around defect-adjacent atoms with radii uniform in `radii`; each search
    starts from Master::perturb (in the library) with its own generator state.  one_system_each: request q
    comes from pool[q] (a kMC loop whose system changed since its last request) instead of a random one.

    misses(system) -> atom indices: the request source of a kMC loop (Master::find_mechs over the
    catalogue's new environments, find.cpp:211-312): the request's centres are the atoms whose local
    environment the catalogue has not seen (a shared catalogue.Catalogue across requests), as many searches
    as the request asks for spread over them in turn; a system with no new environment sends no request."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    reqs = []
    for q in range(count):
        s = pool[q % len(pool)] if one_system_each else pool[int(rng.integers(len(pool)))]
        k = int(rng.integers(searches[0], searches[1] + 1))
        if misses is not None:
            new = np.asarray(misses(s), dtype=np.int64)
            if len(new) == 0:
                continue
            if len(new) > k:
                new = np.sort(rng.choice(new, k, replace=False))
            centres = np.resize(new, k)   # several searches per environment when there are fewer than k
        else:
            centres = sample_centres(s, k, rng)
        rad = rng.uniform(radii[0], radii[1], k)
        prng = rng.integers(1, 2 ** 63, size=(k, 4), dtype=np.uint64)
        reqs.append(Request(float(t[q]), s, centres, rad, prng))
    return reqs


def serve(server: fg.Server, reqs, poll_ms=2.0, keep_positions=False, max_seconds=600.0, report=None):
    """Replays `reqs` against `server` in real time (arrival times relative to the call).  A kMC system is
    admitted on its first request and stays resident for later ones; admission is FIFO, and when the
    memory budget is full the least recently used idle system is evicted, else the request waits for
    running ones to finish.  Returns a summary dict; per-search results land in each request's
    `results`.  `report(msg)`, if given, gets a progress line about every 10 s."""
    pending = collections.deque()
    owner = {}
    resident = collections.OrderedDict()   # id(system) -> [sys_id, live requests], LRU order
    i = 0
    t0 = time.perf_counter()
    live = 0
    finished = 0
    admitted_systems = 0
    next_report = 10.0

    def admit(r):
        nonlocal admitted_systems
        key = id(r.system)
        if key in resident:
            resident.move_to_end(key)
            return resident[key]
        while True:
            try:
                sid = server.add_system(fg.Box(r.system.basis), r.system.pos, r.system.types)
                break
            except fg.Busy:
                idle = [k for k, v in resident.items() if v[1] == 0]
                if not idle:
                    return None
                server.remove_system(resident.pop(idle[0])[0])
                server.progress()
        admitted_systems += 1
        resident[key] = [sid, 0]
        return resident[key]

    while i < len(reqs) or pending or live:
        now = time.perf_counter() - t0
        if report is not None and now >= next_report:
            report(f"{now:.0f} s: {i} arrived, {finished} requests done, {live} running, {len(pending)} waiting")
            next_report += 10.0
        if now > max_seconds:
            raise fg.FlyError(f"serving did not finish within {max_seconds} s")
        while i < len(reqs) and reqs[i].t <= now:
            pending.append(reqs[i])
            i += 1
        while pending:
            r = pending[0]
            slot = admit(r)
            if slot is None:
                break
            slot[1] += 1
            r.sys_id = slot[0]
            ids = server.submit_perturbed(r.sys_id, r.centres, r.radii, r.prng.copy())
            r.remaining = len(ids)
            r.admitted = time.perf_counter() - t0
            for sid in ids:
                owner[int(sid)] = r
            live += 1
            pending.popleft()
        wait = poll_ms
        if i < len(reqs):
            wait = min(wait, max(0.0, (reqs[i].t - (time.perf_counter() - t0)) * 1e3))
        for sid in server.progress(wait_ms=wait):
            r = owner.pop(int(sid))
            r.results[int(sid)] = server.result(sid, positions=keep_positions, release=True)
            r.remaining -= 1
            if r.remaining == 0:
                r.done = time.perf_counter() - t0
                resident[id(r.system)][1] -= 1
                live -= 1
                finished += 1
    wall = time.perf_counter() - t0
    for sid, _ in resident.values():
        server.remove_system(sid)
    server.progress()
    lat = np.array([r.done - r.t for r in reqs])
    res = [x for r in reqs for x in r.results.values()]
    return {
        "requests": len(reqs),
        "searches": len(res),
        "converged": int(sum(1 for x in res if x["exit"] == 0)),
        "wall_s": wall,
        "searches_per_s": len(res) / wall,
        "converged_per_s": sum(1 for x in res if x["exit"] == 0) / wall,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "mean_queue_wait_s": float(np.mean([r.admitted - r.t for r in reqs])),
        "offered_requests_per_s": len(reqs) / max(reqs[-1].t, 1e-9),
        "mean_atoms": float(np.mean([r.system.n for r in reqs])),
        "systems_admitted": admitted_systems,
        "mean_searches_per_request": float(np.mean([len(r.centres) for r in reqs])),
    }
