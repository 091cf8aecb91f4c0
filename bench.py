"""Benchmark of the saddle-search hot path on one MI355X.

Workload (one "step"): the incremental neighbour-list + EAM gradient pass of the 30^3-cell bcc Fe system
(54,000 sites, 1% vacancies, a second species at 2 per 100 Fe on tetrahedral sites, thermal noise 0.08 A):
every step displaces every atom by a random vector of norm <= 0.05 A (pre-generated, resident in HBM),
updates the neighbour list (or rebuilds it once the accumulated displacement exceeds half the skin)
and evaluates the EAM gradient with FP32 pair arithmetic and FP64 accumulation.

Prints ONE JSON line: value = EAM atom-gradients per second over the timed steps, with the roofline of
the EAM gradient kernels (HIP events on the library's stream) and the FP64 CPU oracle timed on the host
cores for the same system.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The serving leg runs this many launches at once, each on a stream of its own.  HIP maps a process's
# streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless set), and streams that share a queue run their
# kernels one after the other, so the queues are raised to match before anything initialises HIP
# (32 requests at 216/s: p50 latency 0.34 s with 4 streams on 4 queues, 0.28 s with 8 on 8).
SERVE_STREAMS = 8
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < SERVE_STREAMS:
    os.environ["GPU_MAX_HW_QUEUES"] = str(SERVE_STREAMS)

METRIC = "saddle searches/s (ΔE≤1e-3 eV) + EAM atom-grads/s; p50/p99 search latency"
HBM_PEAK_GBS = 8000.0
R_CUT = 5.3
SKIN = (1.1 ** (1.0 / 3.0) - 1.0) * R_CUT   # LBFGS::minimise skin policy (lbfgs.cpp:39), CPU baseline list
LIST_SKIN = 0.6   # Verlet skin of the timed update/rebuild loop (rebuild once the accumulated move > skin / 2)
MECH_F2NORM = 1e-4   # |g| tolerance of the mixed form's mechanism relaxations (its FP32 force noise is ~1e-5 eV/A)

# HBM traffic of one gradient (density + force pass) at the default workload, as measured with rocprofv3
# (profiles/eam_gradient_fe30_x_alloy.md): FETCH_SIZE and WRITE_SIZE per dispatch in KiB, each scaled by the
# reading for a copy kernel of known size (python bench.py --copy-calibration: a 64 MiB copy reads
# FETCH_SIZE 32780.9 KiB and WRITE_SIZE 65536 KiB).
GRADIENT_TRAFFIC = {"fetch_kib": 12217.7 + 12747.6, "write_kib": 1704.0 + 1490.2,
                    "fetch_scale": 65536.0 / 32780.9, "write_scale": 1.0}


_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(counts, pos_bytes=8):
    """B_atom = 8 n_i + 6 s_p + 56 per atom-gradient (two passes): list entries (4 B) read once per
    pass, own position read once per pass, type x2, row offset x2, F' write + read, gradient write."""
    return float(np.sum(8.0 * counts + 6 * pos_bytes + 56.0))


def native_oracle():
    """Builds the oracle with the reference's release flags (-O3 -march=native, CMakePresets.json:43) on
    this host and points the oracle binding at it; returns a note on the build used."""
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                       capture_output=True, timeout=120)
        os.environ["ORACLE_LIB"] = os.path.join(ROOT, "oracle", "_build_native", "liboracle.so")
        return "g++ -O3 -march=native -fopenmp"
    except (subprocess.SubprocessError, OSError) as e:
        log(f"native oracle build failed ({e}); timing the portable x86-64-v3 build")
        return "g++ -O3 -march=x86-64-v3 -fopenmp (native build failed)"


def cpu_threads():
    """Host cores this process may use (the affinity mask, else the CPU count)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(sys_, path, build, repeats=5, block_s=2.0):
    """FP64 oracle (OpenMP over the host cores, static schedule as eam.cpp:81,115): one rebuild, one
    warm-up evaluation, then the median of `repeats` timed blocks of about block_s seconds each."""
    from oracle import pyoracle as po
    threads = cpu_threads()
    data = po.DataEAM(path)
    nl = po.List(po.Box(sys_.basis), R_CUT + SKIN)
    nl.rebuild(sys_.pos, threads)
    po.gradient(data, nl, sys_.types, sys_.frozen, threads)   # warm-up
    rates, evals = [], 0
    for _ in range(repeats):
        reps = 0
        t0 = time.perf_counter()
        while True:
            po.gradient(data, nl, sys_.types, sys_.frozen, threads)
            reps += 1
            el = time.perf_counter() - t0
            if el >= block_s:
                break
        rates.append(sys_.n * reps / el)
        evals += reps
    return {"value": float(np.median(rates)), "unit": "atom-grads/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle EAM gradient of the same {sys_.n}-atom system after one rebuild: one warm-up, "
                      f"median of {repeats} blocks of ~{block_s:g} s ({evals} evaluations), OpenMP static schedule "
                      f"over {threads} threads, built {build}",
            "repeats": [round(r, 1) for r in rates]}


def explicit_starts(fa, s, centres, radii, prng):
    """The starts flyg_batch_submit_perturbed draws (flyg_perturb with each search's generator state and
    the atoms beyond its radius frozen), as full arrays for the plain form and the CPU oracle."""
    L = np.diag(s.basis)
    starts, axes, frozen = [], [], []
    for c, R, st in zip(centres, radii, prng):
        d = s.pos - s.pos[c]
        d -= L * np.round(d / L)
        f = (np.einsum("ij,ij->i", d, d) > R * R).astype(np.uint8)
        x, ax, _ = fa.perturb(fa.Box(s.basis), s.pos, f, int(c), st.copy())
        starts.append(x)
        axes.append(ax)
        frozen.append(f)
    return starts, axes, frozen


WEN = os.path.join(ROOT, "tests", "golden", "wen.eam.fs")   # the reference's Fe potential (test fixture)


def relax(ctx, fa, pot, s):
    """min0: the system relaxed by the device L-BFGS (plain FP64 form) to |g| < 1e-4 eV/A."""
    mn, g, failed, it = fa.LBFGS(ctx, fa.Box(s.basis), f2norm=1e-4, iter_max=5000).minimise(
        s.pos, s.types, np.zeros(s.n, np.uint8), pot)
    if failed:
        raise RuntimeError("relaxation of the base system failed")
    return mn


def search_bench(ctx, fa, systems, n_search=256, plain_sample=16, quick=False):
    """256 dimer searches on the 20^3-cell bcc Fe system with 20 vacancies and 40 interstitials
    (relaxed: min0), active radius U(4, 9) A around defect-adjacent atoms, Master::perturb starts
    (kick 0.6 A) drawn in the library, |g| < 1e-3 eV/A; the batched kernel in both precisions, with and
    without mechanism checks, and the plain form (host-driven dimer, full-system FP64 gradients) on the
    first plain_sample starts; potential wen.eam.fs.  quick: the mixed batch only (its autotune run and
    one timed run), for profiling."""
    s = systems.fe_defects_20(3)
    pot = fa.EAM(ctx, WEN)
    t0 = time.perf_counter()
    s.pos = relax(ctx, fa, pot, s)
    t_relax = time.perf_counter() - t0
    rng = np.random.default_rng(33)
    cand = systems.defect_adjacent_atoms(s)
    centres = rng.choice(cand, n_search, replace=True)
    radii = rng.uniform(4.0, 9.0, n_search)
    prng = np.random.default_rng(34).integers(1, 2 ** 63, size=(n_search, 4), dtype=np.uint64)
    box = fa.Box(s.basis)
    dimer = dict(f2norm=1e-3)
    out = {"workload": f"bcc Fe 20^3 + 20 vacancies + 40 interstitials ({s.n} atoms) relaxed to min0, "
                       f"{n_search} searches, radius U(4,9) A, Master::perturb starts (stddev 0.6 A, r_pert 4 A), "
                       "|g| < 1e-3 eV/A, potential wen.eam.fs",
           "relax_seconds": t_relax}
    res = {}
    legs = (("batched_mixed", fa.MIXED),) if quick else (("batched_mixed", fa.MIXED), ("batched_fp64", fa.FP64))
    for name, prec in legs:
        log(f"search leg: {name} (autotune run + timed run of {n_search} searches)")
        b = fa.Batch(ctx, pot, box, s.pos, s.types, dimer=dimer, precision=prec)
        ids = b.submit_perturbed(centres, radii, prng.copy())
        t0 = time.perf_counter()
        b.run()   # untimed: the first run of a batch includes the shape autotuner
        first = time.perf_counter() - t0
        T, L = b.shape()
        b2 = fa.Batch(ctx, pot, box, s.pos, s.types, dimer=dimer, precision=prec, block_threads=T, lanes_per_atom=L)
        ids2 = b2.submit_perturbed(centres, radii, prng.copy())
        t0 = time.perf_counter()
        b2.run()
        el = time.perf_counter() - t0
        r = [b2.result(i) for i in ids2]
        res[name] = r
        ok = sum(1 for x in r if x["exit"] == 0)
        ms = np.array([x["device_ms"] for x in r])
        grads = sum(x["grads"] for x in r)
        out[name] = {"seconds": el, "first_run_seconds_with_autotune": first, "shape": [T, L],
                     # device time per region gradient, and the part of it inside the gradient passes (the
                     # rest is the rotor / L-BFGS algebra and barriers of the search)
                     "device_us_per_gradient": 1e3 * float(ms.sum()) / grads,
                     "gradient_pass_us_per_gradient": 1e3 * sum(x["device_grad_ms"] for x in r) / grads,
                     "converged": ok, "searches_per_s": n_search / el,
                     "converged_per_s": ok / el, "exits": np.bincount([x["exit"] for x in r], minlength=7).tolist(),
                     "mean_grads": float(np.mean([x["grads"] for x in r])),
                     "mean_mobile": float(np.mean([x["n_mobile"] for x in r])),
                     "mean_region": float(np.mean([x["n_region"] for x in r])),
                     "relaunched": int(sum(1 for x in r if x["launches"] > 1)),
                     "reperturbed": int(sum(1 for x in r if x["retries"] > 0)),
                     "search_device_ms_p50": float(np.percentile(ms, 50)),
                     "search_device_ms_p99": float(np.percentile(ms, 99)),
                     "search_device_ms_max": float(ms.max())}
    if quick:
        return out, None
    # The same mixed batch with every converged saddle turned into a checked min->sp->min mechanism.
    log("search leg: batched_mixed with mechanisms")
    T, L = out["batched_mixed"]["shape"]
    bm = fa.Batch(ctx, pot, box, s.pos, s.types, dimer=dimer, precision=fa.MIXED, block_threads=T, lanes_per_atom=L)
    bm.set_mechanisms(minimiser=dict(f2norm=MECH_F2NORM))
    idm = bm.submit_perturbed(centres, radii, prng.copy())
    t0 = time.perf_counter()
    bm.run()
    el = time.perf_counter() - t0
    rm = [bm.result(i) for i in idm]
    mm = [bm.mechanism(i) for i, r in zip(idm, rm) if r["exit"] == 0]
    codes = np.bincount([m["code"] for m in mm if m["code"] >= 0], minlength=8).tolist()
    out["batched_mixed_mechanisms"] = {"seconds": el, "searches_per_s": n_search / el,
                                       "converged": sum(1 for r in rm if r["exit"] == 0), "pathways": codes[0],
                                       "pathways_per_s": codes[0] / el, "mechanism_codes": codes,
                                       "minimiser_f2norm": MECH_F2NORM}
    # Plain form (one search at a time, full-system FP64 gradients, host-driven dimer) on the first starts.
    log(f"search leg: plain FP64 form on the first {plain_sample} starts")
    dim = fa.Dimer(ctx, box, dimer=dimer)
    starts, axes, frozen = explicit_starts(fa, s, centres[:plain_sample], radii[:plain_sample], prng[:plain_sample])
    t0 = time.perf_counter()
    plain = [dim.find_sp(starts[k], axes[k], s.types, frozen[k], s.pos, pot) for k in range(plain_sample)]
    el = time.perf_counter() - t0
    # The plain form reports no barrier: FP64 energies at its saddles, after the timed run.
    nl = fa.List(ctx, box, pot.r_cut())
    nl.rebuild(s.pos)
    fz0 = np.zeros(s.n, np.uint8)
    e0 = pot.energy(s.types, fz0, nl)
    for r in plain:
        if r["exit"] == 0:
            nl.rebuild(r["pos"])
            r["barrier"] = pot.energy(s.types, fz0, nl) - e0
    out["plain_fp64_gpu"] = {"sample": plain_sample, "seconds": el, "searches_per_s": plain_sample / el,
                             "converged": int(sum(1 for r in plain if r["exit"] == 0)),
                             "exits": np.bincount([r["exit"] for r in plain], minlength=7).tolist()}
    return out, dict(s=s, centres=centres, radii=radii, prng=prng, dimer=dimer, res=res, plain=plain)


def qualify(out, S, orc):
    """A search counts only if it converges and its barrier is within 1e-3 eV of the FP64 oracle's barrier
    from the same start (the oracle must converge too).  Fills searches_per_s_dE_1e-3 and the ratios."""
    ob = orc["barriers"]
    def count(results, idx):
        good, diffs = 0, []
        for q, r in zip(idx, results):
            if r["exit"] == 0 and ob[q] is not None:
                d = abs(r["barrier"] - ob[q])
                diffs.append(d)
                good += int(d <= 1e-3)
        return good, diffs
    n_all = len(S["centres"])
    for name in ("batched_mixed", "batched_fp64"):
        good, diffs = count(S["res"][name], range(n_all))
        o = out[name]
        o["within_1e-3eV_of_oracle"] = good
        o["searches_per_s_dE_1e-3"] = good / o["seconds"]
        o["max_abs_dE_vs_oracle"] = float(max(diffs)) if diffs else None
        o["both_converged"] = len(diffs)
    k = len(S["plain"])
    good, diffs = count(S["plain"], range(k))
    p = out["plain_fp64_gpu"]
    p["within_1e-3eV_of_oracle"] = good
    p["searches_per_s_dE_1e-3"] = good / p["seconds"]
    cpu = out["cpu_baseline"]
    mixed = out["batched_mixed"]["searches_per_s_dE_1e-3"]
    out["speedups_dE_1e-3"] = {
        "batched_mixed_vs_plain_gpu": mixed / p["searches_per_s_dE_1e-3"] if good else None,
        "batched_mixed_vs_cpu": mixed / cpu["searches_per_s_dE_1e-3"] if cpu["searches_per_s_dE_1e-3"] else None,
        "note": "qualified searches/s: converged and within 1e-3 eV of the FP64 oracle's barrier from the same "
                "start; batched over all 256 starts, plain GPU form and CPU over the first 16 of them"}


def cpu_search_leg(S, build, sample=16, repeats=5):
    """FP64 oracle, one search per thread over the host cores (find.cpp:697-732 with threads = 1 in each
    search, find.cpp:990): (1) all 256 starts once, whose barriers qualify the device's searches;
    (2) the CPU rate: one warm-up search, then the median of `repeats` runs over the first `sample` starts."""
    from oracle import pyoracle as po
    import concurrent.futures as cf
    import openfly_amd as fa
    s = S["s"]
    threads = cpu_threads()
    data = po.DataEAM(WEN)
    box = po.Box(s.basis)
    starts, axes, frozen = explicit_starts(fa, s, S["centres"], S["radii"], S["prng"])

    def one(i):
        return po.find_sp(box, data, starts[i], axes[i], s.types, frozen[i], s.pos, threads=1, dimer=S["dimer"])

    def run(idx):
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            return list(ex.map(one, idx))

    n_all = len(starts)
    log(f"search leg: FP64 oracle on all {n_all} starts ({threads} threads)")
    t0 = time.perf_counter()
    full = run(range(n_all))
    el_full = time.perf_counter() - t0

    def energy(p):
        nl = po.List(box, data.r_cut)
        nl.rebuild(p, threads)
        return po.energy(data, nl, s.types, np.zeros(s.n, np.uint8), threads)

    e0 = energy(s.pos)
    barriers = [energy(r["pos"]) - e0 if r["exit"] == 0 else None for r in full]
    log(f"search leg: CPU rate, warm-up + median of {repeats} runs over {sample} starts")
    one(0)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        run(range(sample))
        times.append(time.perf_counter() - t0)
    t_med = float(np.median(times))
    conv_sample = sum(1 for r in full[:sample] if r["exit"] == 0)
    base = {"value": sample / t_med, "unit": "searches/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle Dimer::find_sp on the first {sample} starts, one search per thread: one warm-up, "
                      f"median of {repeats} runs ({t_med:.2f} s); built {build}",
            "converged_in_sample": conv_sample, "searches_per_s_dE_1e-3": conv_sample / t_med,
            "repeat_seconds": [round(t, 3) for t in times],
            "all_starts": {"searches": n_all, "seconds": el_full, "searches_per_s": n_all / el_full,
                           "converged": int(sum(1 for r in full if r["exit"] == 0)),
                           "exits": np.bincount([r["exit"] for r in full], minlength=7).tolist()}}
    return base, {"barriers": barriers}


def serve_bench(ctx, fa, systems, n_req=64, rate=200.0, streams=SERVE_STREAMS):
    """Serving: Poisson arrivals at 200/s of requests from a simulated kMC loop, each on its own system
    (bcc Fe, N log-uniform in [2000, 250000], 0.5-2 % vacancies + interstitials, relaxed), 8-64 searches
    each centred on the system's catalogue misses (radius U(4, 9) A, Master::perturb starts in the
    library), served with continuous batching; latency = arrival to last search done.  Then the same systems at 0.9x the throughput sustained under
    that overload (the highest rate the server keeps up with), for latency below saturation."""
    from openfly_amd import serve
    from openfly_amd.catalogue import Catalogue
    import torch
    pot = fa.EAM(ctx, WEN)
    log(f"serving leg: building and relaxing {n_req} kMC systems (one per request)")
    t0 = time.perf_counter()
    pool = serve.make_pool(n_req, seed=41, relax=lambda sy: relax(ctx, fa, pot, sy))
    t_pool = time.perf_counter() - t0
    srv = fa.Server(ctx, pot, dimer=dict(f2norm=1e-3), streams=streams)
    log("serving leg: warm-up")
    serve.serve(srv, serve.make_requests(pool[:2], 4, rate=rate, seed=40, searches=(8, 16)))   # + shape autotune
    # Request source: each system is one kMC state whose last hop was at one of its defects; the catalogue
    # shared by the run looks at the environments within 2 r_env of that site (the ones the hop changed)
    # and the request's searches go to the new ones (cache misses), computed once per system before the
    # replays.
    log("serving leg: catalogue of local environments (cache misses per system)")
    cat = Catalogue()
    dev = torch.device("cuda", 0)
    found = {}
    rng_hop = np.random.default_rng(44)
    t0 = time.perf_counter()
    for q, sy in enumerate(pool):
        sites = [np.asarray(sy.vacancies)] if getattr(sy, "vacancies", None) is not None and len(sy.vacancies) else []
        if getattr(sy, "interstitial_ids", None) is not None and len(sy.interstitial_ids):
            sites.append(sy.pos[sy.interstitial_ids])
        sites = np.concatenate(sites)
        hop = sites[rng_hop.integers(len(sites))]
        near = serve.Near(sy).within(hop, 2.0 * cat.r_env)
        found[id(sy)] = cat.rebuild(sy.pos, sy.types, sy.frozen, sy.basis, device=dev, atoms=near)
        if q % 16 == 15:
            log(f"serving leg: catalogue {q + 1}/{len(pool)} systems, {cat.size} environments")
    t_cat = time.perf_counter() - t0
    misses = [len(found[id(sy)]) for sy in pool]
    reqs = serve.make_requests(pool, n_req, rate=rate, seed=42, one_system_each=True,
                               misses=lambda sy: found[id(sy)])
    log(f"serving leg: replay of {n_req} requests at {rate:g}/s")
    out = serve.serve(srv, reqs, report=log)
    sustained = out["requests"] / out["wall_s"]
    rate2 = 0.9 * sustained
    log(f"serving leg: replay at {rate2:.1f} requests/s (0.9x the sustained {sustained:.1f}/s)")
    reqs2 = serve.make_requests(pool, n_req, rate=rate2, seed=43, one_system_each=True,
                                misses=lambda sy: found[id(sy)])
    light = serve.serve(srv, reqs2, report=log)
    keys = ("requests", "searches", "converged", "wall_s", "searches_per_s", "latency_p50_s", "latency_p99_s",
            "latency_max_s", "mean_queue_wait_s", "offered_requests_per_s", "mean_searches_per_request")
    out["sustained_requests_per_s"] = sustained
    out["at_sustainable_rate"] = {k: light[k] for k in keys}
    st = srv.status()
    out.update(workload=f"{n_req} kMC systems (bcc Fe, N log-uniform [2000, 250000], 0.5-2% vacancies + "
                        f"interstitials, relaxed), Poisson {rate:g}/s; each system sends one request for 8-64 searches "
                        "centred on its local environments that a shared catalogue had not seen (cache misses; "
                        "systems with none send none), radius U(4,9) A, Master::perturb starts, |g| < 1e-3 eV/A, "
                        "FP32 pair / FP64 accumulation, wen.eam.fs",
               catalogue={"environments": cat.size, "seconds": t_cat, "misses_per_system_min": int(min(misses)),
                          "misses_per_system_max": int(max(misses)),
                          "misses_per_system_mean": float(np.mean(misses))},
               pool_relax_seconds=t_pool, launches=st["launches"], kernel_ms=st["kernel_ms"], host_ms=st["host_ms"],
               streams=streams, hw_queues=os.environ.get("GPU_MAX_HW_QUEUES"),
               block_threads=st["block_threads"], lanes_per_atom=st["lanes_per_atom"],
               system_atoms_min=int(min(p.n for p in pool)), system_atoms_max=int(max(p.n for p in pool)))
    return out


def large_bench(ctx, fa, systems, n_search=1024):
    """Large system: bcc Fe 80^3 cells (1,028,096 atoms) with 0.1 % vacancies in clusters of 1-6 and 0.5 %
    interstitials, relaxed (min0); 1024 concurrent searches (radius U(4, 9) A, kick 0.6 A, |g| < 1e-3 eV/A)
    in one launch of the autotuned persistent kernel; every converged barrier re-evaluated as a full-system
    FP64 energy difference E(sp) - E(min0)."""
    from openfly_amd import serve
    t0 = time.perf_counter()
    s = systems.fe_clusters_80(5)
    pot = fa.EAM(ctx, WEN)
    s.pos = relax(ctx, fa, pot, s)
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(55)
    centres = serve.sample_centres(s, n_search, rng)
    radii = rng.uniform(4.0, 9.0, n_search)
    prng = rng.integers(1, 2 ** 63, size=(n_search, 4), dtype=np.uint64)
    srv = fa.Server(ctx, pot, dimer=dict(f2norm=1e-3), max_batch=n_search, streams=1)
    box = fa.Box(s.basis)
    t0 = time.perf_counter()
    sysid = srv.add_system(box, s.pos, s.types)
    t_admit = time.perf_counter() - t0

    def drain(ids):
        want, got = set(int(i) for i in ids), {}
        while len(got) < len(want):
            for d in srv.progress(wait_ms=50.0):
                got[int(d)] = srv.result(d, positions=True, release=True)
        return [got[int(i)] for i in ids]

    log("1M-atom leg: autotune on 64 searches")
    drain(srv.submit_perturbed(sysid, centres[:64], radii[:64], prng[:64].copy()))
    log(f"1M-atom leg: {n_search} searches")
    t0 = time.perf_counter()
    res = drain(srv.submit_perturbed(sysid, centres, radii, prng.copy()))
    el = time.perf_counter() - t0
    st = srv.status()
    conv = [r for r in res if r["exit"] == 0]
    # FP64 re-evaluation of every converged barrier (full system, fresh list each time).
    log(f"1M-atom leg: FP64 re-evaluation of all {len(conv)} converged barriers")
    t0 = time.perf_counter()
    nl = fa.List(ctx, box, pot.r_cut())
    frozen = np.zeros(s.n, np.uint8)
    nl.rebuild(s.pos)
    e0 = pot.energy(s.types, frozen, nl)
    L = np.diag(s.basis)
    errs, pairs = [], []
    p = s.pos.copy()
    for r in conv:
        p[r["atoms"]] = r["pos"]
        nl.rebuild(p)
        de = pot.energy(s.types, frozen, nl) - e0
        p[r["atoms"]] = s.pos[r["atoms"]]
        errs.append(abs(de - r["barrier"]))
        d = r["pos"] - s.pos[r["atoms"]]
        d -= L * np.round(d / L)   # minimum image
        pairs.append((r["barrier"], de, float(np.sqrt(np.einsum("ij,ij->i", d, d)).max())))
    t_reeval = time.perf_counter() - t0
    ms = np.array([r["device_ms"] for r in res])
    return {"workload": f"bcc Fe 80^3 cells ({s.n} atoms), 0.1% vacancies in clusters of 1-6, 0.5% interstitials, "
                        f"relaxed; {n_search} concurrent searches, radius U(4,9) A, kick 0.6 A, |g| < 1e-3 eV/A, "
                        "FP32 pair / FP64 accumulation, wen.eam.fs",
            "atoms": int(s.n), "setup_seconds": t_setup, "admit_seconds": t_admit, "seconds": el,
            "searches_per_s": n_search / el, "converged": len(conv), "converged_per_s": len(conv) / el,
            "exits": np.bincount([r["exit"] for r in res], minlength=7).tolist(),
            "block_threads": st["block_threads"], "lanes_per_atom": st["lanes_per_atom"],
            "search_device_ms_p50": float(np.percentile(ms, 50)), "search_device_ms_p99": float(np.percentile(ms, 99)),
            "fp64_reevaluated": len(errs), "fp64_reevaluation_seconds": t_reeval,
            "max_abs_barrier_error_eV": float(max(errs)) if errs else None,
            "barrier_error_eV_p50": float(np.median(errs)) if errs else None,
            "first_barrier_fp64_maxdisp": [[round(a, 6), round(b, 6), round(c, 3)] for a, b, c in pairs[:5]]}


def gradient_leg(ctx, fa, torch, s, path, precision, steps, warmup, skin, pool_n=16, device_types=False):
    """The incremental list + EAM gradient loop on system `s`: each step moves every atom by a random vector
    of norm <= 0.05 A (pre-generated, resident in HBM), updates the list (a rebuild once the accumulated
    displacement exceeds skin / 2) and evaluates the gradient.  Returns the wall time of the timed steps,
    the kernel time per gradient (the library's HIP event pair on its stream) and the algorithmic bytes.
    Species and frozen flags are staged in the potential handle once (flyg_eam_stage_atoms, from HBM) and
    every call uses that copy, which the library gathers into the list's layout once per list build.
    device_types=True passes the device tensors to every call instead: their contents may change between
    calls, so each call gathers the species again."""
    pot = fa.EAM(ctx, path)
    pot.set_precision(fa.MIXED if precision == "mixed" else fa.FP64)
    dev = torch.device("cuda", 0)
    # Inputs resident in HBM before timing.
    pos = torch.from_numpy(s.pos.copy()).to(dev)
    types = torch.from_numpy(s.types.astype(np.uint32)).to(dev)
    frozen = torch.from_numpy(s.frozen).to(dev)
    if not device_types:
        pot.stage(types, frozen)
        types = frozen = None
    grad = torch.zeros((s.n, 3), dtype=torch.float64, device=dev)
    scratch = torch.zeros((s.n, 3), dtype=torch.float64, device=dev)
    rng = np.random.default_rng(12)
    pool, pool_neg, pool_max = [], [], []
    for _ in range(pool_n):
        d = rng.normal(size=(s.n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d *= 0.05 * rng.random((s.n, 1)) ** (1.0 / 3.0)   # uniform in the 0.05 A ball
        pool.append(torch.from_numpy(d).to(dev))
        pool_neg.append(torch.from_numpy(-d).to(dev))
        pool_max.append(float(np.max(np.linalg.norm(d, axis=1))))
    torch.cuda.synchronize()

    nl = fa.List(ctx, fa.Box(s.basis), R_CUT + skin)
    nl.rebuild(pos)
    off, _, _, _ = nl.pairs(R_CUT + skin)   # stored rows (n_i = stored list length)
    counts = np.diff(off).astype(np.float64)

    state = {"acc": 0.0, "k": 0, "rebuilds": 0}
    def step():
        # Steps 2j and 2j+1 apply +d and -d of pool entry j, so positions oscillate about the thermal
        # configuration instead of drifting; every step still moves every atom by up to 0.05 A.
        k = (state["k"] // 2) % pool_n
        sign = 1.0 if state["k"] % 2 == 0 else -1.0
        state["k"] += 1
        nl.update(pool[k] if sign > 0 else pool_neg[k])
        state["acc"] += pool_max[k]
        if state["acc"] > 0.5 * skin:
            fa.lib().flyg_list_positions(nl.h, fa.flygpu.C.c_void_p(scratch.data_ptr()), fa.DEVICE)
            nl.rebuild(scratch)
            state["acc"] = 0.0
            state["rebuilds"] += 1
        pot.gradient(types, frozen, nl, out=grad)

    for _ in range(warmup):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    state["rebuilds"] = 0
    pot.set_timing(True)   # the library's own HIP event pair around each call's two kernels
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib_ms, lib_calls = pot.timing()
    pot.set_timing(False)
    # Kernel time per gradient from the event pair the library records on its stream right before the
    # density launch and right after the force launch (an event pair recorded around the Python call
    # would also hold the host time of the call whenever the stream had drained before the launch).
    kern_ms = lib_ms / lib_calls
    b = algorithmic_bytes(counts, 8)
    return {"atoms": int(s.n), "precision": precision, "steps": steps, "seconds": el,
            "atom_grads_per_s": s.n * steps / el, "ms_per_step": el * 1e3 / steps, "kernel_ms": kern_ms,
            "algorithmic_bytes": b, "achieved_GBs": b / (kern_ms * 1e-3) / 1e9, "rebuilds": state["rebuilds"],
            "mean_neighbours": float(counts.mean()), "counts": counts}


def extra_gradient_legs(ctx, fa, torch, systems, steps, warmup, skin):
    """The same loop at reference precision (FP64) on the alloy, on the 16^3 single-vacancy cell (FP64,
    Fe-only synthetic potential) and on the 1,028,096-atom 80^3 cluster system (mixed and FP64)."""
    legs = {}
    cases = (("fe_x_alloy_30_fp64", lambda: systems.fe_x_alloy(seed=2, cells=30), ("Fe", "X"), "fp64"),
             ("fe_vacancy_16_fp64", lambda: systems.fe_vacancy_16(seed=1), ("Fe",), "fp64"),
             ("fe_clusters_80_mixed", lambda: systems.fe_clusters_80(5), ("Fe",), "mixed"),
             ("fe_clusters_80_fp64", lambda: systems.fe_clusters_80(5), ("Fe",), "fp64"))
    cache = {}
    for name, make, species, prec in cases:
        key = name.rsplit("_", 1)[0]
        if key not in cache:
            cache.clear()
            cache[key] = make()
        s = cache[key]
        log(f"gradient leg {name}: {s.n} atoms")
        g = gradient_leg(ctx, fa, torch, s, systems.potential_path(species), prec, steps, warmup, skin)
        g.pop("counts")
        legs[name] = g
    return legs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cells", type=int, default=30)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--skin", type=float, default=LIST_SKIN, help="Verlet skin of the gradient leg's list (A)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the FP64, 16^3 and 80^3 gradient legs beside the headline one")
    ap.add_argument("--no-search", action="store_true", help="skip the batched saddle-search leg")
    ap.add_argument("--search-quick", action="store_true",
                    help="search leg: the mixed batch only, no FP64 batch, mechanisms or plain form (profiling)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served saddle-search leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 1M-atom, 1024-search leg")
    ap.add_argument("--serve-streams", type=int, default=SERVE_STREAMS,
                    help="concurrent launches of the serving leg (set GPU_MAX_HW_QUEUES to match)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 64 MiB device copies (the known-size kernel that scales FETCH_SIZE / "
                         "WRITE_SIZE readings to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": a.numel() * 8, "dispatches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the saddle-search path runs on a "
              "single MI355X and is not sharded across GPUs.", file=sys.stderr)
        sys.exit(2)

    build = native_oracle() if not args.no_cpu else None
    import torch
    import openfly_amd as fa
    from openfly_amd import systems

    s = systems.fe_x_alloy(seed=2, cells=args.cells)
    path = systems.potential_path(("Fe", "X"))
    ctx = fa.Context(0)
    log(f"gradient leg: {s.n} atoms, {args.warmup} warm-up + {args.steps} timed steps")
    g = gradient_leg(ctx, fa, torch, s, path, args.precision, args.steps, args.warmup, args.skin)
    el, kern_ms, bytes_per_grad, counts = g["seconds"], g["kernel_ms"], g["algorithmic_bytes"], g["counts"]
    achieved = bytes_per_grad / (kern_ms * 1e-3) / 1e9
    value = s.n * args.steps / el
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-grads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 accumulation" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"bcc Fe {args.cells}^3 cells + 1% vacancies + 2% X tetrahedral interstitials, "
                        "random displacement <= 0.05 A per step, list update/rebuild, EAM gradient "
                        "(species and frozen flags staged in the potential once, from HBM)",
            "atoms": s.n,
            "mean_neighbours": float(counts.mean()),
            "rebuilds_in_timed_steps": g["rebuilds"],
            "list_cutoff": R_CUT + args.skin,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (1024.0 * (GRADIENT_TRAFFIC["fetch_kib"] * GRADIENT_TRAFFIC["fetch_scale"]
                                  + GRADIENT_TRAFFIC["write_kib"] * GRADIENT_TRAFFIC["write_scale"])
                        if (args.cells == 30 and args.skin == LIST_SKIN and args.precision == "mixed") else None),
            "kernel": "EAM gradient (density pass + force pass)",
            "kernel_ms": kern_ms,
            "algorithmic_bytes_per_launch": bytes_per_grad,
        },
    }
    if not args.no_extra:
        out["gradient_legs"] = extra_gradient_legs(ctx, fa, torch, systems, min(args.steps, 100),
                                                   min(args.warmup, 10), args.skin)
        log("extra gradient legs done")
    sargs = None
    if not args.no_search:
        sr, sargs = search_bench(ctx, fa, systems, quick=args.search_quick)
        log("search leg done")
        out["searches"] = sr
    if not args.no_serve:
        out["serving"] = serve_bench(ctx, fa, systems, streams=args.serve_streams)
        log("serving leg done")
    if not args.no_large:
        log("1M-atom leg: building and relaxing the system")
        out["large"] = large_bench(ctx, fa, systems)
        log("1M-atom leg done")
    if not args.no_cpu:
        log("gradient leg: CPU oracle baseline")
        out["cpu_baseline"] = cpu_baseline(s, path, build)
        if sargs is not None:
            sr["cpu_baseline"], orc = cpu_search_leg(sargs, build)
            qualify(sr, sargs, orc)
            log("search leg: CPU oracle done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
