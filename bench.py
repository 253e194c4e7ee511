"""Throughput of the exact tetrahedron mass deposit on one MI355X (tets/s), one JSON line.

Workload (the heavy-tailed batched configuration, the largest that fits one GPU): a 256^3-particle
lattice (257^3 unwrapped vertices, 100,663,296 tets in the Kuhn decomposition, tet mass 1/6 so a
particle carries mass 1) with a seeded k^-2 Gaussian displacement field at rms 4 lattice spacings
plus 64 seeded spherical infall centres (radius 2-8 Mpc/h) that pull tets into caustics; the
reference's adaptive octree (max depth 6, split above 50,000 estimated tets) with 16^3-cell leaves
(`value`) and, as an extra field, 32^3-cell leaves.  A step is the per-leaf binning of the lattice
plus the deposit of every lattice tet into every leaf it overlaps, from positions resident in HBM to
FP64 leaf grids in HBM (the octree is built once, outside the timed region, and timed separately).
tets/s = 6 N^3 / step time; each tet counts once however many leaves it touches.
`--workload c2` runs the same measurement on the 128^3 lattice (rms 3 spacings, no infall).

Extra fields: the plain one-tet-per-thread per-leaf-launch form measured the same way, the
hipGraph replay, the whole one-rank workflow, the total-mass drift, the roofline of the dominant
kernel (HIP events on its launch stream; algorithmic FLOPs and bytes from the oracle's counts in
tests/golden/c{2,3}_flops.json) and a CPU baseline: the compiled reference CPU path (oracle/_ref,
kind "reference") or the oracle restatement (kind "port") timed on a bounded sample of the same
leaves on this host's cores, and on one core.

Studies, one JSON line each instead of the headline: --heavy-tail (256^3 lattice with infall
centres, batched vs plain), --precision (precision modes and near-degenerate tets), --serve (ROI
requests, continuous batching vs serial plain, with the step kernel's roofline from the server's
step-kernel time and tests/golden/c5_flops.json), --program (the whole program from a snapshot file
to leaf files, SORTED_BY_ID and DARK_SKY, beside the reference program) and --boundary (the drop-in
deposit_tets called per leaf, plain vs batched form).
"""

from __future__ import annotations

import argparse
import concurrent.futures
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import adeco_amd as A  # noqa: E402
from adeco_amd.lattice import make_lattice  # noqa: E402

# the two adaptive workloads: lattice cubes per axis, rms displacement (spacings), seed, infall
# centres, flops fixture
WORKLOADS = {
    "c3": {"n": 256, "rms": 4.0, "seed": 3, "infall": 64, "flops": "c3_flops.json"},
    "c2": {"n": 128, "rms": 3.0, "seed": 2, "infall": 0, "flops": "c2_flops.json"},
}
N = 128
RMS = 3.0
SEED = 2
MAX_TETS = 50000
MAX_LEVEL = 6
DIMS = 16
DIMS_FINE = 32
TET_MASS = 1.0 / 6.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak (the roof of an FP64 VALU-bound kernel)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of k_deposit_waves<FP64, false> at 16^3-cell leaves: rocprofv3 FETCH_SIZE
# and WRITE_SIZE (KiB per dispatch) scaled by what each reads for a 1 GiB copy (2.0 and 1.0),
# profiles/deposit_waves_c3_pmc.md (256^3 infall workload) and profiles/deposit_waves_c2_pmc.md
# (128^3)
TRAFFIC_BYTES = {"c3": 22706736.9375 * 2 * 1024 + 18270762.125 * 1024, "c2": 2575602.0 * 2 * 1024 + 1951831.875 * 1024}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_steps(tree, steps, warmup, mode, precision, decomp, dims=DIMS):
    """One step = the per-leaf tet binning of the lattice plus the deposit of every leaf (the
    octree is built once: its refinement is reported separately as octree_ms)."""
    for _ in range(warmup):
        tree.bin(decomp)
        tree.deposit(dims, decomposition=decomp, tet_mass=TET_MASS, mode=mode, precision=precision)
    wall, main_ms, tl = [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        tree.bin(decomp)
        tree.deposit(dims, decomposition=decomp, tet_mass=TET_MASS, mode=mode, precision=precision)
        wall.append(time.perf_counter() - t0)
        tl = A.Octree.last_timings()
        main_ms.append(tl["main_ms"])
    return float(np.sum(wall)), float(np.mean(main_ms)), tl


def max_rel_err(a, b, floor):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def _stdout_to_stderr(fn):
    """Runs fn with file descriptor 1 pointed at stderr: the reference library logs to stdout, and
    stdout must carry the JSON line alone."""
    import ctypes
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn()
    finally:
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def host_cpu_info():
    """This host's CPUs: nproc, the model name (lscpu), the affinity set and the cgroup CPU quota
    (cpu.max), and the threads a CPU baseline may use (the smaller of affinity and quota)."""
    import subprocess
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return {"nproc": nproc, "cpu_model": model, "affinity_cpus": aff, "cgroup_cpu_quota": quota, "threads": threads}


def cpu_baseline(lat, dims, budget_s=20.0, one_core_s=10.0):
    """Times the reference CPU path (or the oracle port) on this host's cores; returns a dict.

    The compiled reference deposits leaves as its worker threads do (leaf gather +
    DepositWrapper::deposit_tets into a float grid per thread, REF_EFFECTIVE tets), one leaf per
    thread at a time, in a seeded random leaf order, until every leaf is done or the time budget
    has passed; a partial run is extrapolated to the whole job by the meta-patch share of the leaves
    it finished.  Run once on all the threads this host gives the process and once on one thread."""
    import threading
    from tests import _oracle as O  # checker side only
    from tests._octree_io import ref_octree
    host = host_cpu_info()
    threads = host["threads"]
    ref = O.load_ref()
    if ref is not None:
        t_tree = time.perf_counter()
        d = _stdout_to_stderr(lambda: ref_octree(lat, MAX_TETS, MAX_LEVEL))
        t_tree = time.perf_counter() - t_tree
        dfs = list(d["dfs"])
        n_meta = np.diff(d["mc_offsets"])
        meta_of = {int(node): int(n_meta[i]) for i, node in enumerate(dfs)}
        order = [int(x) for x in np.random.default_rng(0).permutation(np.asarray(d["order"]))]

        def run(nthreads, budget):
            lock = threading.Lock()
            todo = iter(order)
            done = []
            t0 = time.perf_counter()

            def worker():
                g = np.zeros(dims ** 3, np.float32)
                while time.perf_counter() - t0 < budget:
                    with lock:
                        node = next(todo, None)
                    if node is None:
                        return
                    g[:] = 0
                    assert ref.ref_octree_deposit_leaf(node, dims, O.P(g), None, None, None) == 0
                    with lock:
                        done.append(node)

            pool = [threading.Thread(target=worker) for _ in range(nthreads)]
            for th in pool:
                th.start()
            for th in pool:
                th.join()
            wall = time.perf_counter() - t0
            frac = sum(meta_of[n] for n in done) / float(sum(meta_of.values()))
            return wall, len(done), frac

        wall, ndone, frac = _stdout_to_stderr(lambda: run(threads, budget_s))
        wall1, ndone1, frac1 = _stdout_to_stderr(lambda: run(1, one_core_s))
        whole = ndone == len(order)
        return {"value": lat.num_tets / (wall / frac), "unit": "tets/s", "cores": threads, "kind": "reference",
                "one_core_value": lat.num_tets / (wall1 / frac1), **{k: host[k] for k in host if k != "threads"},
                "sample": (f"{'all' if whole else ndone} of {len(order)} leaves"
                           f"{'' if whole else f' ({frac:.3f} of the meta-patch work, whole-job time extrapolated by that share)'}"
                           f" in a seeded random order, deposited at {dims}^3 cells by the compiled reference (its leaf "
                           f"gather + DepositWrapper::deposit_tets, REF_EFFECTIVE tets, float grids) on {threads} threads "
                           f"in {wall:.2f} s; one thread: {ndone1} leaves ({frac1:.4f} of the work) in {wall1:.2f} s; the "
                           f"reference octree (generate_tree, one thread) took {t_tree:.2f} s before it, not included"),
                "octree_s": t_tree}
    # fallback: the oracle port, FP64, on a fixed tet sample of the whole lattice
    box = np.asarray(A.cubic_root_box(lat.bbox()))
    ids = np.arange(0, lat.num_tets, 97, dtype=np.int32)
    t0 = time.perf_counter()
    O.deposit_lattice(lat, box, dims * 8, decomp=2, tet_mass=TET_MASS, tet_ids=ids, precision=1, nthreads=threads)
    wall = time.perf_counter() - t0
    return {"value": ids.size / wall, "unit": "tets/s", "cores": threads, "kind": "port",
            **{k: host[k] for k in host if k != "threads"},
            "sample": f"{ids.size} tets (every 97th) of the lattice on one grid of {dims * 8}^3 cells over the root box, "
                      f"FP64, {threads} threads, {wall:.2f} s"}


def algorithmic_work(workload):
    """Reference-algorithm FP64 operations and nonzero (tet, cell) contributions of one step,
    counted by the oracle's counting instantiation (tests/golden/make_flops.py) and committed as
    tests/golden/c{2,3}_flops.json; None when the fixture is absent.  Returns {dims: {"flops",
    "nonzero", "nonzero_per_pair_by_level"}}."""
    path = os.path.join(ROOT, "tests", "golden", WORKLOADS[workload]["flops"])
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if "dims" not in d:  # every leaf counted, 16^3 cells
        return {16: {"flops": float(d["flops"]), "nonzero": d.get("nonzero_contributions"), "by_level": None,
                     "flops_rel_std_err": 0.0, "pairs": int(d["pairs"])}}
    return {int(k): {"flops": float(v["flops"]), "nonzero": float(v["nonzero_contributions"]),
                     "by_level": {int(l): r["nonzero_per_pair"] for l, r in v["per_level"].items()},
                     "flops_rel_std_err": float(v["flops_rel_std_err"]), "pairs": int(d["pairs"])}
            for k, v in d["dims"].items()}


def roofline_of(work, dims, kernel_ms, lat, n_leaves, counts=None, levels=None, traffic=None):
    """The dominant kernel's roofline: algorithmic FLOPs / its HIP-event duration against the FP64
    vector peak, and algorithmic bytes (12 B per lattice vertex, 8 B per nonzero (tet, cell)
    contribution - the FP64 accumulation - and the leaf grids once) against the HBM peak.  With the
    run's per-leaf pair counts and leaf levels, the nonzero contributions are the fixture's rate per
    pair at each octree level times this run's pairs at that level."""
    w = work.get(dims) if work else None
    if w is None:
        return None
    achieved = w["flops"] / (kernel_ms * 1e-3) / 1e12
    nonzero = w["nonzero"]
    if w["by_level"] and counts is not None and levels is not None:
        nonzero = float(sum(w["by_level"].get(int(l), 0.0) * float(counts[levels == l].sum()) for l in np.unique(levels)))
    nv = float(np.prod(lat.dims))
    alg_bytes = None if nonzero is None else 12.0 * nv + 8.0 * nonzero + 8.0 * n_leaves * dims ** 3
    rec = {"bound": "valu", "pipe": "VALU FP64 (clipping geometry; no matrix work)",
           "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
           "frac_of_fp32_peak": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
           "flops_per_launch": w["flops"], "flops_rel_std_err": w["flops_rel_std_err"], "kernel_ms": kernel_ms,
           "kernel": "k_deposit_waves<FP64>", "leaf_dims": dims}
    if alg_bytes is not None:
        rec.update({"algorithmic_bytes": alg_bytes, "nonzero_contributions": nonzero,
                    "hbm_achieved_gbs": alg_bytes / (kernel_ms * 1e-3) / 1e9,
                    "hbm_frac": alg_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
    return rec


SERVE_N = 256
SERVE_RES = (64, 128, 256, 512)


def serve_requests(count, seed):
    """A request set for the Poisson streams: unit-rate exponential gaps (divided by the rate),
    ROI boxes (side U[5,50] Mpc/h, centre U[0,100)^3) and resolutions drawn from 64^3..512^3."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0, count)
    gaps[0] = 0.0
    side = rng.uniform(5.0, 50.0, count)
    centre = rng.uniform(0.0, 100.0, (count, 3))
    dims = rng.choice(SERVE_RES, count)
    rois = np.empty((count, 6))
    rois[:, 0::2] = centre - side[:, None] / 2
    rois[:, 1::2] = centre + side[:, None] / 2
    return np.cumsum(gaps), rois, dims


def percentiles(lat):
    a = np.asarray(lat)
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


def serve_study(rates, count, seed=5, max_inflight=16):
    """Serving: the resident 256^3 lattice (the 128^3 workload's field type, rms 3 spacings) answers
    Poisson streams of ROI requests; p50/p99 latency and tets/s of the continuous-batching server
    (real-time arrivals) against serial plain-GPU processing of the same requests (one at a time,
    arrival order: FIFO over the measured service times)."""
    import threading
    t0 = time.perf_counter()
    lat = make_lattice(SERVE_N, RMS, seed=seed)
    log(f"[serve] lattice {SERVE_N}^3 generated in {time.perf_counter() - t0:.1f}s")
    unit_arr, rois, dims = serve_requests(count, seed)
    out = {"lattice": SERVE_N, "tets": lat.num_tets, "requests": count, "max_inflight": max_inflight, "rates": []}
    # the reference algorithm's FLOPs of this request set (tests/golden/c5_flops.json: 60 requests, seed 5)
    serve_flops = None
    fx = os.path.join(ROOT, "tests", "golden", "c5_flops.json")
    if count == 60 and seed == 5 and os.path.isfile(fx):
        with open(fx) as f:
            serve_flops = float(json.load(f)["flops"])
    with A.ResidentLattice.from_lattice(lat) as rl:
        # serial plain GPU service times (independent of the arrival rate)
        service, plain_keep, n_tets = [], {}, []
        for i in range(count):
            s0 = time.perf_counter()
            g, n = A.roi_deposit_plain(rl, rois[i], int(dims[i]), A.DECOMP_KUHN, TET_MASS)
            service.append(time.perf_counter() - s0)
            n_tets.append(n)
            if i < 4:
                plain_keep[i] = g
        tets = float(np.sum(n_tets))
        log(f"[serve] serial plain: {np.sum(service):.2f} s for {count} requests ({tets / np.sum(service):.3e} tets/s)")
        out["plain_serial_busy_s"] = float(np.sum(service))
        out["plain_serial_tets_per_s_busy"] = tets / float(np.sum(service))
        for rate in rates:
            arr = unit_arr / rate
            lat_s, err = [], 0.0
            with A.Server(rl, A.DECOMP_KUHN, TET_MASS, max_inflight=max_inflight) as srv:
                ids = [None] * count
                start = time.perf_counter()

                def submitter():
                    for i in range(count):
                        dt = arr[i] - (time.perf_counter() - start)
                        if dt > 0:
                            time.sleep(dt)
                        ids[i] = srv.submit(rois[i], int(dims[i]))

                th = threading.Thread(target=submitter)
                th.start()
                th.join()
                for i in range(count):
                    g, info = srv.wait(ids[i])
                    lat_s.append(info["latency"])
                    if i in plain_keep:
                        cell_vol = np.prod((rois[i][1::2] - rois[i][0::2]) / dims[i])
                        floor = 1e-8 * TET_MASS / cell_vol
                        want = plain_keep[i].astype(np.float64)
                        err = max(err, float(np.max(np.abs(g.astype(np.float64) - want) / np.maximum(np.abs(want), floor))))
                makespan = time.perf_counter() - start
                st = srv.stats()
            p50, p99 = percentiles(lat_s)
            finish, lat_p = 0.0, []
            for i in range(count):
                finish = max(finish, arr[i]) + service[i]
                lat_p.append(finish - arr[i])
            q50, q99 = percentiles(lat_p)
            by_res = {}
            for d in SERVE_RES:
                sel = [i for i in range(count) if int(dims[i]) == d]
                if sel:
                    b50, b99 = percentiles([lat_s[i] for i in sel])
                    s50, s99 = percentiles([lat_p[i] for i in sel])
                    by_res[f"{d}^3"] = {"requests": len(sel), "batched_p50_s": b50, "batched_p99_s": b99,
                                        "plain_serial_p50_s": s50, "plain_serial_p99_s": s99}
            rec = {"rate_req_s": rate, "batched_p50_s": p50, "batched_p99_s": p99,
                   "batched_tets_per_s": tets / makespan, "batched_makespan_s": makespan, "steps": int(st["steps"]),
                   "step_tiles": int(st["tiles"]), "worker_busy_s": float(st["busy_s"]),
                   "step_kernel_s": float(st["step_kernel_s"]),
                   "plain_serial_p50_s": q50, "plain_serial_p99_s": q99, "plain_serial_tets_per_s": tets / finish,
                   "plain_serial_makespan_s": finish, "batched_vs_plain_max_rel_err": err, "by_resolution": by_res}
            if serve_flops is not None:
                ach = serve_flops / float(st["step_kernel_s"]) / 1e12
                rec["roofline"] = {"bound": "valu", "kernel": "k_deposit_step<FP64>", "achieved": ach,
                                   "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                                   "flops": serve_flops, "kernel_s": float(st["step_kernel_s"])}
            log(f"[serve] {rate:g} req/s: batched p50 {p50 * 1e3:.1f} ms p99 {p99 * 1e3:.1f} ms "
                f"({tets / makespan:.3e} tets/s); serial plain p50 {q50 * 1e3:.1f} ms p99 {q99 * 1e3:.1f} ms "
                f"({tets / finish:.3e} tets/s); max rel diff {err:.2e}")
            out["rates"].append(rec)
    return out


def grid_errors(got, ref, floor):
    """Per-cell max relative error (floor for near-empty cells), L1 error relative to the total, and
    the total-mass drift of `got` against `ref`."""
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    tot = float(np.sum(np.abs(ref)))
    return {"max_rel": float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))),
            "l1_rel": float(np.sum(np.abs(got - ref)) / tot), "mass_drift": float(abs(got.sum() - ref.sum()) / tot)}


def precision_study(n_degenerate=1_000_000, seed=4):
    """Precision modes on the 128^3 workload's leaves and on near-degenerate tets (edge ratios
    1e3..1e6, volumes down to ~1e-12 cell^3) deposited as a soup on a 128^3 grid over the box:
    FP64; FP32 clip with FP64 orientation/volume and FP64 recompute where the FP32 error bound
    fails (MIXED); pure FP32 clip.  Errors are against the FP64 form, itself held to the
    extended-precision oracle by the GPU parity tests."""
    from adeco_amd.lattice import make_degenerate_tets
    out = {"leaves": {}, "degenerate": {}}
    floor = 1e-8 * TET_MASS
    lat = make_lattice(N, RMS, seed=SEED)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, MAX_TETS, MAX_LEVEL) as tree:
        tree.bin(A.DECOMP_KUHN)
        ref = None
        for name, prec in (("fp64", A.PREC_FP64), ("mixed", A.PREC_MIXED), ("fp32", A.PREC_FP32)):
            tree.deposit(DIMS, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED, precision=prec)
            tl = A.Octree.last_timings()
            g = tree.leaf_grids()
            rec = {"kernel_ms": tl["main_ms"], "partial_cells": tl["partials"],
                   "fp64_recompute_fraction": tl["fallbacks"] / max(tl["partials"], 1)}
            if ref is None:
                ref = g
                rec["mass_drift_vs_exact"] = abs(g.sum() - N ** 3) / N ** 3
            else:
                rec.update(grid_errors(g, ref, floor))
            out["leaves"][name] = rec
            log(f"[precision] leaves {name}: {rec}")
    tets = make_degenerate_tets(n_degenerate, seed=seed)
    box = np.array([0, 100, 0, 100, 0, 100], np.float32)
    ref = A.deposit_soup(tets, box, 128, TET_MASS, mode=A.MODE_BATCHED, precision=A.PREC_FP64)
    tl = A.Octree.last_timings()
    out["degenerate"]["fp64"] = {"kernel_ms": tl["main_ms"], "partial_cells": tl["partials"]}
    plain = A.deposit_soup(tets, box, 128, TET_MASS, mode=A.MODE_PLAIN)
    out["degenerate"]["fp64_plain_kernel"] = grid_errors(plain, ref, floor)
    for name, prec in (("mixed", A.PREC_MIXED), ("fp32", A.PREC_FP32)):
        g = A.deposit_soup(tets, box, 128, TET_MASS, mode=A.MODE_BATCHED, precision=prec)
        tl = A.Octree.last_timings()
        rec = {"kernel_ms": tl["main_ms"], "partial_cells": tl["partials"],
               "fp64_recompute_fraction": tl["fallbacks"] / max(tl["partials"], 1)}
        rec.update(grid_errors(g, ref, floor))
        out["degenerate"][name] = rec
        log(f"[precision] degenerate {name}: {rec}")
    out["degenerate"]["tets"] = int(n_degenerate)
    return out


def heavy_tail_study(dims_list=(16, 32), n=256, steps=3):
    """The heavy-tailed batched workload: a 256^3 lattice (100,663,296 Kuhn tets) at rms 4 spacings
    with 64 seeded spherical infall centres (radius 2-8 Mpc/h), the reference octree (depth <= 6,
    split above 50,000 estimated tets) and 16^3 or 32^3-cell leaves.  The batched form (autotuned
    launch shape; eager and hipGraph replay) against the plain per-leaf launches."""
    t0 = time.perf_counter()
    log(f"[c3] generating the {n}^3 lattice with 64 infall centres")
    lat = make_lattice(n, 4.0, seed=3, infall_centres=64)
    log(f"[c3] lattice {n}^3 with infall generated in {time.perf_counter() - t0:.1f}s")
    out = {"lattice": n, "tets": lat.num_tets, "runs": []}
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, MAX_TETS, MAX_LEVEL) as tree:
        counts = tree.bin(A.DECOMP_KUHN)
        nz = counts[counts > 0]
        out["leaves"] = int(tree.num_leaves)
        out["binned_pairs"] = int(counts.sum())
        out["leaf_tets_min_median_max"] = [int(nz.min()), int(np.median(nz)), int(nz.max())]
        out["leaf_tets_p99_over_median"] = float(np.percentile(nz, 99) / np.median(nz))
        log(f"[c3] {tree.num_leaves} leaves, pairs {counts.sum()}, per-leaf tets min/median/max "
            f"{out['leaf_tets_min_median_max']}")
        inside = None
        for dims in dims_list:
            shape = tree.autotune(dims, A.DECOMP_KUHN, TET_MASS, A.PREC_FP64, reps=1)
            wall, main_ms, tl = run_steps(tree, steps, 1, A.MODE_BATCHED, A.PREC_FP64, A.DECOMP_KUHN, dims)
            g = tree.leaf_grids()
            total = float(g.sum())
            del g
            with A.DepositGraph(tree, dims, A.DECOMP_KUHN, TET_MASS, A.PREC_FP64) as gr:
                gr.launch()
                gms = float(np.mean([gr.launch() for _ in range(steps)]))
            wall_p, main_p, _ = run_steps(tree, 1, 0, A.MODE_PLAIN, A.PREC_FP64, A.DECOMP_KUHN, dims)
            rec = {"leaf_dims": dims, "launch_shape": shape, "batched_ms_per_step": wall / steps * 1e3,
                   "batched_kernel_ms": main_ms, "batched_tets_per_s": lat.num_tets / (wall / steps),
                   "graph_replay_ms": gms, "graph_tets_per_s": lat.num_tets / (gms * 1e-3),
                   "plain_ms_per_step": wall_p * 1e3, "plain_tets_per_s": lat.num_tets / wall_p,
                   "batched_over_plain": wall_p / (wall / steps), "candidates": tl["candidates"],
                   "partials": tl["partials"], "deposited_mass": total}
            out["runs"].append(rec)
            log(f"[c3] {dims}^3 leaves: batched {rec['batched_ms_per_step']:.1f} ms/step "
                f"({rec['batched_tets_per_s']:.3e} tets/s), graph {gms:.1f} ms, plain {wall_p * 1e3:.0f} ms "
                f"-> {rec['batched_over_plain']:.1f}x; mass {total:.6f}")
    return out


def program_study(n=N, runs=3, with_reference=True):
    """The reference program's single-rank workflow end to end on one GPU (adeco_deposit_workflow,
    the body of adeco_resample): a periodic n^3-particle SORTED_BY_ID snapshot (k^-2 field at rms
    3 spacings, wrapped into the box) read from a file, uploaded, periodic correction, octree
    (depth <= 6, split above 50,000 estimated tets), batched FP64 leaf deposit of the reference's
    tets (mass 1) at 16^3 cells per leaf, and one write_patch file per leaf.  One untimed run,
    then `runs` timed runs; step times are the program's own.  The same snapshot as a DARK_SKY
    file (shuffled 32-byte records, Morton ids placed on the GPU) is timed the same way.
    with_reference: also time the reference program on the SORTED_BY_ID input
    (reference_program_time)."""
    import tempfile
    from adeco_amd.lattice import make_snapshot, write_dark_sky, write_sorted_by_id
    from adeco_amd.workflow import AMD_Parameters, deposit_workflow
    t0 = time.perf_counter()
    snap = make_snapshot(n, RMS, seed=SEED)
    log(f"[program] snapshot {n}^3 generated in {time.perf_counter() - t0:.1f}s")
    out = {"particles": n ** 3, "tets": 6 * (n - 1) ** 3, "runs": [], "dark_sky_runs": []}
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "snapshot.bin")
        write_sorted_by_id(path, snap)
        ds_path = os.path.join(tmp, "snapshot.sdf")
        write_dark_sky(ds_path, snap)
        for fmt, fpath, key in (("SORTED_BY_ID", path, "runs"), ("DARK_SKY", ds_path, "dark_sky_runs")):
            p = AMD_Parameters(input_file=fpath, id_type=fmt, max_refinement_level=MAX_LEVEL,
                               max_tets_per_octree_node=MAX_TETS, linear_patch_resolution=DIMS,
                               periodic_boundaries=True)
            for r in range(runs + 1):
                p.output_directory = os.path.join(tmp, f"out_{key}{r}")
                os.mkdir(p.output_directory)
                t0 = time.perf_counter()
                res = deposit_workflow(p)
                wall = time.perf_counter() - t0
                rec = {"wall_s": wall,
                       **{k: res[k] for k in ("load_s", "octree_s", "deposit_s", "write_s", "kernel_ms")},
                       "leaves": res["num_leaves"], "total_mass": res["total_mass"]}
                log(f"[program] {fmt} run {r}: {wall * 1e3:.1f} ms (load {res['load_s'] * 1e3:.1f}, octree "
                    f"{res['octree_s'] * 1e3:.1f}, deposit {res['deposit_s'] * 1e3:.1f} [gpu {res['kernel_ms']:.2f}], "
                    f"files {res['write_s'] * 1e3:.1f}); {res['num_leaves']} leaves, mass {res['total_mass']:.6f}")
                if r:
                    out[key].append(rec)
        if with_reference:
            p.input_file, p.id_type = path, "SORTED_BY_ID"
            out["cpu_baseline"] = reference_program_time(path, p, n, tmp)
            if out["cpu_baseline"].get("value"):
                ref_dir = os.path.join(tmp, "ref_out")
                out["vs_reference_files"] = compare_leaf_files(os.path.join(tmp, f"out_runs{runs}"), ref_dir, DIMS)
                # the same in the reference's geometry (float leaf boxes, 1/dx scaling), which
                # separates the geometry's share of the differences from the float error's
                p.output_directory = os.path.join(tmp, "out_reference_geometry")
                os.mkdir(p.output_directory)
                saved = A.get_cell_scaling()
                A.set_cell_scaling(A.SCALING_REFERENCE)
                try:
                    deposit_workflow(p)
                finally:
                    A.set_cell_scaling(saved)
                out["vs_reference_files_reference_geometry"] = compare_leaf_files(p.output_directory, ref_dir, DIMS)
    out["dark_sky_median_wall_s"] = float(np.median([r["wall_s"] for r in out["dark_sky_runs"]]))
    w = float(np.median([r["wall_s"] for r in out["runs"]]))
    out["median_wall_s"] = w
    out["tets_per_s_end_to_end"] = out["tets"] / w
    if out.get("cpu_baseline", {}).get("value"):
        out["cpu_baseline"]["gpu_speedup"] = out["cpu_baseline"]["wall_s"] / w
    return out


def compare_leaf_files(ours_dir, ref_dir, dims):
    """Checker leg of the program study: this program's leaf files against the reference
    program's on the same snapshot - the same names, and per file the largest relative density
    difference (floor: 1e-3 of the file's largest density) and the mass difference.  The
    reference accumulates in float, so its own error sets these figures."""
    ours, ref = sorted(os.listdir(ours_dir)), sorted(os.listdir(ref_dir))
    worst, worst_mass = 0.0, 0.0
    total_ours = total_ref = 0.0
    for f in ours:
        if f not in ref:
            continue
        a = np.fromfile(os.path.join(ours_dir, f), np.float32).astype(np.float64)
        b = np.fromfile(os.path.join(ref_dir, f), np.float32).astype(np.float64)
        floor = 1e-3 * max(float(np.abs(b).max()), 1e-30)
        worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))))
        sa, sb = float(a.sum()), float(b.sum())
        total_ours, total_ref = total_ours + sa, total_ref + sb
        worst_mass = max(worst_mass, abs(sa - sb) / max(abs(sb), 1e-30))
    res = {"same_file_names": ours == ref, "files": len(ours), "max_rel_density_diff": worst,
           "max_rel_leaf_mass_diff": worst_mass, "total_rel_diff": abs(total_ours - total_ref) / max(total_ref, 1e-30)}
    log(f"[program] vs the reference program's files: {res}")
    return res


def reference_program_time(snapshot_path, params, params_n, tmp):
    """CPU baseline of the program study: the reference program itself (main.cpp, compiled from
    the reference sources into oracle/_ref by oracle/build_ref.sh) on the same snapshot and
    parameters, with one worker thread per core (at most 16), in a child process (it finalises
    MPI).  Wall time of the child, which includes process and MPI start-up."""
    import subprocess
    ref = os.path.join(ROOT, "oracle", "_ref", "libadeco_ref.so")
    if not os.path.exists(ref):
        return {"value": None, "note": "oracle/_ref not built"}
    threads = min(16, os.cpu_count() or 1)
    outdir = os.path.join(tmp, "ref_out")
    os.mkdir(outdir)
    pfile = os.path.join(tmp, "ref_params.txt")
    with open(pfile, "w") as f:
        f.write(f"input_file               ={snapshot_path}\noutput_directory         ={outdir}\n"
                f"file_format              = SORTED_BY_ID\nuse_region_of_interest   = false\n"
                f"max_refinement_level     = {params.max_refinement_level}\nstride                   = 1\n"
                f"max_tets_per_octree_node = {params.max_tets_per_octree_node}\n"
                f"linear_patch_resolution  = {params.linear_patch_resolution}\n"
                f"num_mpi_ranks_per_node   = 1\nnum_pthreads             = {threads}\n"
                f"periodic_boundaries      = true\n")
    code = ("import sys; sys.path.insert(0, sys.argv[2]); from tests import _oracle as O; "
            "sys.exit(O.load_ref().ref_cli_main(sys.argv[1].encode()))")
    log(f"[program] reference program on {threads} threads")
    t0 = time.perf_counter()
    r = subprocess.run([sys.executable, "-c", code, pfile, ROOT], capture_output=True, text=True, timeout=900)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        return {"value": None, "note": f"reference program failed: {r.stderr[-300:]}"}
    total = None
    for line in r.stdout.splitlines():
        if "total mass" in line:
            total = float(line.split("=")[-1])
    files = len(os.listdir(outdir))
    log(f"[program] reference program: {wall:.2f} s, {files} leaf files, total mass {total}")
    tets = 6 * (params_n - 1) ** 3
    return {"value": tets / wall, "unit": "tets/s", "cores": threads, "wall_s": wall, "leaf_files": files,
            "total_mass": total, "kind": "reference", "sample": "the whole program on the same snapshot"}


def boundary_study(reps=2):
    """The drop-in boundary (DepositWrapper::deposit_tets, adeco_deposit_patches) on the 128^3
    workload, called as the reference's worker calls it (AdDynamicLoadBalancing.h:200-205): once
    per octree leaf, with the leaf's gathered meta-patches in host memory (its meta-cell list, the
    reference's gather order), the leaf's float box, 16^3 cells, the reference's tets of mass 1,
    into a zeroed float grid.  The plain form (one thread per tet, the reference kernel's
    structure) against the batched form; host-to-device copies of the patches and the grid
    readback are inside the timed loop."""
    from adeco_amd.lattice import meta_patches
    lat = make_lattice(N, RMS, seed=SEED)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, MAX_TETS, MAX_LEVEL) as tree:
        lv = tree.leaves()
        boxes = tree.leaf_boxes()
    t0 = time.perf_counter()
    patches = [meta_patches(lat, lv["mc_ids"][lv["mc_offsets"][l]:lv["mc_offsets"][l + 1]]) for l in range(boxes.shape[0])]
    log(f"[boundary] gathered {sum(p.size for p in patches) * 4 / 1e6:.0f} MB of meta-patches for "
        f"{len(patches)} leaves in {time.perf_counter() - t0:.1f}s")
    out = {"leaves": len(patches), "patch_bytes": int(sum(p.size for p in patches) * 4), "forms": {}}
    grids = {}
    for name, mode in (("plain", A.MODE_PLAIN), ("batched", A.MODE_BATCHED)):
        prev = A.set_boundary_form(mode=mode)
        try:
            best = None
            for _ in range(reps + 1):
                res = []
                t0 = time.perf_counter()
                for p, box in zip(patches, boxes):
                    g = np.zeros(DIMS ** 3, np.float32)
                    if p.size:
                        A.DepositWrapper.deposit_tets(True, 0, 0, p.size // 1536, p, p.size, 7, box, DIMS, g)
                    res.append(g)
                wall = time.perf_counter() - t0
                best = wall if best is None else min(best, wall)
            grids[name] = res
        finally:
            A.set_boundary_form(**prev)
        mass = float(sum(float(g.sum(dtype=np.float64)) for g in grids[name]))
        out["forms"][name] = {"s": best, "tets_per_s": 6 * (N ** 3) / best, "total_mass": mass}
        log(f"[boundary] {name}: {best * 1e3:.1f} ms for {len(patches)} calls ({6 * N ** 3 / best:.3e} tets/s), "
            f"mass {mass:.6f}")
    err = max_rel_err(np.concatenate(grids["batched"]), np.concatenate(grids["plain"]), 1e-8)
    out["batched_over_plain"] = out["forms"]["plain"]["s"] / out["forms"]["batched"]["s"]
    out["batched_vs_plain_max_rel_err"] = err
    log(f"[boundary] batched / plain: {out['batched_over_plain']:.1f}x, per-cell max rel err {err:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="skip the 32^3-leaf, plain-kernel and reference-decomposition comparison runs")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c3",
                    help="c3: 256^3 lattice with infall (default); c2: 128^3 lattice")
    ap.add_argument("--shape", default="", help="launch shape CHUNK,WORKGROUPS_PER_CU instead of the autotuner's")
    ap.add_argument("--serve", action="store_true", help="run only the ROI serving study (256^3 lattice)")
    ap.add_argument("--precision", action="store_true", help="run only the precision-mode study")
    ap.add_argument("--heavy-tail", action="store_true", help="run only the 256^3 heavy-tailed batched study")
    ap.add_argument("--boundary", action="store_true",
                    help="run only the drop-in boundary study (per-leaf deposit_tets calls, plain vs batched)")
    ap.add_argument("--program", action="store_true",
                    help="run only the whole single-rank program (file in, leaf files out) on a 128^3 snapshot")
    ap.add_argument("--serve-rates", default="20,50,100,200")
    ap.add_argument("--serve-requests", type=int, default=60)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one MI355X; run with --gpus 1"}))
        sys.exit(2)
    if args.heavy_tail:
        res = heavy_tail_study()
        print(json.dumps({"metric": "tets/s exact deposit, batched vs plain per-leaf launches", "unit": "tets/s",
                          "higher_is_better": True, "n_gpus": 1, "dtype": "f64",
                          "data": "synthetic (256^3 lattice, rms 4 spacings + 64 infall centres)", **res}))
        return
    if args.boundary:
        res = boundary_study()
        print(json.dumps({"metric": "tets/s through DepositWrapper::deposit_tets (per-leaf calls, PCIe included)",
                          "unit": "tets/s", "value": res["forms"]["batched"]["tets_per_s"], "higher_is_better": True,
                          "n_gpus": 1, "dtype": "f64", "data": "synthetic (128^3 lattice, rms 3 spacings)", **res}))
        return
    if args.program:
        res = program_study(with_reference=not args.no_cpu_baseline)
        print(json.dumps({"metric": "tets/s end to end: snapshot file to leaf files", "unit": "tets/s",
                          "value": res["tets_per_s_end_to_end"], "higher_is_better": True, "n_gpus": 1,
                          "dtype": "f64", "data": "synthetic (128^3 periodic snapshot, rms 3 spacings)", **res}))
        return
    if args.precision:
        res = precision_study()
        print(json.dumps({"metric": "per-cell max rel error, L1 error and mass drift of the precision modes vs FP64",
                          "unit": "relative", "higher_is_better": False, "n_gpus": 1, "dtype": "f64/f32",
                          "data": "synthetic (128^3 lattice leaves; 1M near-degenerate tets)", **res}))
        return
    if args.serve:
        res = serve_study([float(r) for r in args.serve_rates.split(",")], args.serve_requests)
        print(json.dumps({"metric": "p50/p99 ROI latency (s) and tets/s, continuous batching vs serial plain GPU",
                          "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
                          "data": "synthetic (256^3 lattice, k^-2 field at rms 3 spacings; Poisson ROI requests)",
                          "config": {"workload": "ROI serving, side U[5,50] Mpc/h, resolutions 64^3..512^3, fp32 density"},
                          **res}))
        return

    headline(args)


def headline(args):
    wl = WORKLOADS[args.workload]
    n = wl["n"]
    t0 = time.perf_counter()
    lat = make_lattice(n, wl["rms"], seed=wl["seed"], infall_centres=wl["infall"])
    log(f"[bench] lattice {n}^3 ({args.workload}) generated in {time.perf_counter() - t0:.1f}s")
    decomp = A.DECOMP_KUHN
    work = algorithmic_work(args.workload)
    if work is None:
        log(f"[bench] roofline skipped: tests/golden/{wl['flops']} is missing")
    with A.ResidentLattice.from_lattice(lat) as rl:
        t0 = time.perf_counter()
        tree = A.Octree(rl, MAX_TETS, MAX_LEVEL)
        t_oct = time.perf_counter() - t0
        t0 = time.perf_counter()
        counts = tree.bin(decomp)
        t_bin = time.perf_counter() - t0
        leaves = tree.leaves()["dfs"]
        levels = tree.nodes()["level"][leaves]
        nz = counts[counts > 0]
        log(f"[bench] octree {tree.num_nodes} nodes / {tree.num_leaves} leaves in {t_oct*1e3:.1f} ms; "
            f"binning {counts.sum()} pairs in {t_bin*1e3:.1f} ms; per-leaf tets min/median/max "
            f"{nz.min()}/{int(np.median(nz))}/{nz.max()}")

        # launch shape by the autotuner (untimed) or --shape, then the headline: batched persistent
        # form, FP64
        if args.shape:
            chunk, blocks = (int(v) for v in args.shape.split(","))
            A.set_launch_shape(chunk, blocks)
            shape = {"chunk": chunk, "blocks_per_cu": blocks}
            log(f"[bench] launch shape {shape}")
        else:
            t0 = time.perf_counter()
            shape = tree.autotune(DIMS, decomp, TET_MASS, A.PREC_FP64, reps=1)
            log(f"[bench] autotune: {shape} in {time.perf_counter() - t0:.1f}s")
        wall, main_ms, tl = run_steps(tree, args.steps, args.warmup, A.MODE_BATCHED, A.PREC_FP64, decomp)
        step_ms = wall / args.steps * 1e3
        value = lat.num_tets / (wall / args.steps)
        grids64 = tree.leaf_grids()
        total = float(grids64.sum())
        drift = abs(total - n ** 3) / n ** 3
        log(f"[bench] batched fp64 {DIMS}^3 leaves: {step_ms:.2f} ms/step, kernel {main_ms:.2f} ms, {value:.3e} tets/s, "
            f"candidates {tl['candidates']}, partials {tl['partials']}, exact recomputes {tl['fallbacks']}, "
            f"mass drift {drift:.2e}")
        roofline = roofline_of(work, DIMS, main_ms, lat, tree.num_leaves, counts, levels, TRAFFIC_BYTES[args.workload])

        extra = {"exact_recomputes": tl["fallbacks"], "partial_candidates": tl["partials"], "candidates": tl["candidates"]}
        # the same deposit captured as a hipGraph and replayed (candidate space built once)
        with A.DepositGraph(tree, DIMS, decomp, TET_MASS, A.PREC_FP64) as g:
            g.launch()
            gms = [g.launch() for _ in range(max(1, args.steps))]
            gerr = max_rel_err(g.leaf_grids(), grids64, 1e-8 * TET_MASS)
        extra["graph_replay_ms"] = float(np.mean(gms))
        extra["graph_replay_tets_per_s"] = lat.num_tets / (float(np.mean(gms)) * 1e-3)
        extra["graph_vs_eager_max_rel_err"] = gerr
        log(f"[bench] graph replay: {np.mean(gms):.2f} ms ({extra['graph_replay_tets_per_s']:.3e} tets/s), err {gerr:.1e}")
        # the whole step (binning, pair setup, scans, tile map, deposit) as one graph with device-side
        # sizes, replayed: the eager step's work without its host round trips
        with A.DepositGraph(tree, DIMS, decomp, TET_MASS, A.PREC_FP64, step=True) as g:
            g.launch()
            sms = [g.launch() for _ in range(max(1, args.steps))]
            serr = max_rel_err(g.leaf_grids(), grids64, 1e-8 * TET_MASS)
        extra["step_graph_ms"] = float(np.mean(sms))
        extra["step_graph_tets_per_s"] = lat.num_tets / (float(np.mean(sms)) * 1e-3)
        extra["step_graph_vs_eager_max_rel_err"] = serr
        log(f"[bench] step graph replay: {np.mean(sms):.2f} ms ({extra['step_graph_tets_per_s']:.3e} tets/s), err {serr:.1e}")
        del grids64
        # the whole one-rank workflow per step: octree refinement + binning + deposit + leaf grids to host
        wf = []
        for _ in range(3):
            t0 = time.perf_counter()
            with A.Octree(rl, MAX_TETS, MAX_LEVEL) as t2:
                t2.bin(decomp)
                t2.deposit(DIMS, decomposition=decomp, tet_mass=TET_MASS, mode=A.MODE_BATCHED, precision=A.PREC_FP64)
                t2.leaf_grids()
            wf.append(time.perf_counter() - t0)
        extra["workflow_ms_per_step"] = float(np.min(wf)) * 1e3
        extra["workflow_tets_per_s"] = lat.num_tets / float(np.min(wf))
        log(f"[bench] workflow (octree + binning + deposit + leaf grids to host): {extra['workflow_ms_per_step']:.2f} ms "
            f"({extra['workflow_tets_per_s']:.3e} tets/s)")
        if not args.quick:
            # 32^3-cell leaves: the same lattice and octree, its own launch shape
            fine_shape = shape if args.shape else tree.autotune(DIMS_FINE, decomp, TET_MASS, A.PREC_FP64, reps=1)
            steps_f = max(2, args.steps // 3)
            wall_f, main_f, tlf = run_steps(tree, steps_f, 1, A.MODE_BATCHED, A.PREC_FP64, decomp, DIMS_FINE)
            fine = tree.leaf_grids()
            extra["fine_leaves"] = {
                "leaf_dims": DIMS_FINE, "tets_per_s": lat.num_tets / (wall_f / steps_f), "ms_per_step": wall_f / steps_f * 1e3,
                "kernel_ms": main_f, "launch_shape": fine_shape, "mass_drift": abs(float(fine.sum()) - n ** 3) / n ** 3,
                "candidates": tlf["candidates"], "partial_candidates": tlf["partials"],
                "roofline": roofline_of(work, DIMS_FINE, main_f, lat, tree.num_leaves, counts, levels)}
            del fine
            log(f"[bench] batched fp64 {DIMS_FINE}^3 leaves: {wall_f / steps_f * 1e3:.2f} ms/step, kernel {main_f:.2f} ms "
                f"({extra['fine_leaves']['tets_per_s']:.3e} tets/s)")
            A.set_launch_shape(shape["chunk"], shape["blocks_per_cu"])
            # the plain one-tet-per-thread form, one launch per leaf
            steps_p = 1 if n > 128 else max(2, args.steps // 2)
            wall_p, main_p, _ = run_steps(tree, steps_p, 0 if n > 128 else 1, A.MODE_PLAIN, A.PREC_FP64, decomp)
            plain_tets = lat.num_tets / (wall_p / steps_p)
            plain_grids = tree.leaf_grids()
            tree.deposit(DIMS, decomposition=decomp, tet_mass=TET_MASS, mode=A.MODE_BATCHED, precision=A.PREC_FP64)
            extra["plain_tets_per_s"] = plain_tets
            extra["batched_over_plain"] = value / plain_tets
            extra["plain_vs_batched_max_rel_err"] = max_rel_err(plain_grids, tree.leaf_grids(), 1e-8 * TET_MASS)
            del plain_grids
            log(f"[bench] plain fp64: {wall_p / steps_p * 1e3:.2f} ms/step, {plain_tets:.3e} tets/s, "
                f"batched/plain {extra['batched_over_plain']:.1f}x, max rel diff {extra['plain_vs_batched_max_rel_err']:.1e}")
            # the reference's own (non-tiling) decomposition, like for like with the CPU baseline
            tree.bin(A.DECOMP_REF_EFFECTIVE)
            wall_r, _, _ = run_steps(tree, max(2, args.steps // 2), 1, A.MODE_BATCHED, A.PREC_FP64, A.DECOMP_REF_EFFECTIVE)
            extra["ref_effective_tets_per_s"] = lat.num_tets / (wall_r / max(2, args.steps // 2))
            tree.bin(decomp)

        cpu = None
        if not args.no_cpu_baseline:
            try:
                cpu = cpu_baseline(lat, DIMS)
                log(f"[bench] cpu baseline: {cpu['value']:.3e} tets/s ({cpu['kind']}, {cpu['cores']} cores; one core "
                    f"{cpu.get('one_core_value', float('nan')):.3e})")
            except Exception as exc:
                log(f"[bench] cpu baseline skipped: {exc}")
        tree.close()

    field = f"rms {wl['rms']:g} spacings" + (f" + {wl['infall']} infall centres" if wl["infall"] else "")
    out = {
        "metric": "tets/s exact deposit; p50/p99 ROI latency vs plain GPU at \u22641e-6 mass err",
        "value": value,
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded k^-2 Gaussian displacement field, {field})",
        "config": {"workload": f"adaptive exact deposit, {n}^3-particle lattice ({lat.num_tets:,} Kuhn tets, {field}), "
                               f"octree depth<=6, split >50000, {DIMS}^3-cell leaves, batched persistent FP64 form; a "
                               f"step = per-leaf tet binning + deposit of all leaves (octree refinement reported as "
                               f"octree_ms; {DIMS_FINE}^3-cell leaves in fine_leaves)",
                   "lattice": n, "tets": lat.num_tets, "leaves": int(leaves.size), "leaf_dims": DIMS,
                   "binned_pairs": int(counts.sum()), "leaf_tets_min_median_max": [int(nz.min()), int(np.median(nz)), int(nz.max())],
                   "octree_ms": t_oct * 1e3, "binning_ms": t_bin * 1e3, "mass_drift": drift,
                   "launch_shape": {"tiles_per_pop": shape["chunk"], "workgroups_per_cu": shape["blocks_per_cu"]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
