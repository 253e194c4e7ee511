"""Throughput of the exact tetrahedron mass deposit on one MI355X (tets/s), one JSON line.

Workload (the 128^3 plain-GPU configuration): a 128^3-particle lattice (129^3 unwrapped vertices,
12,582,912 tets in the Kuhn decomposition, tet mass 1/6 so a particle carries mass 1) with a
seeded k^-2 Gaussian displacement field at rms 3 lattice spacings; the reference's adaptive
octree (max depth 6, split above 50,000 estimated tets) with 16^3-cell leaves.  A step is one
deposit of every lattice tet into every leaf it overlaps, from positions resident in HBM to FP64
leaf grids in HBM (octree and binning are built once, outside the timed region, and timed
separately).  tets/s = 6 N^3 / step time; each tet counts once however many leaves it touches.

Extra fields: the plain one-tet-per-thread per-leaf-launch form measured the same way, the
mixed-precision form with its error against FP64, the total-mass drift, the roofline of the
dominant kernel (HIP events on its launch stream) and a CPU baseline: the compiled reference
CPU path (oracle/_ref, kind "reference") or the oracle restatement (kind "port") timed on a
bounded sample of the same leaves on this host's cores.
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

N = 128
RMS = 3.0
SEED = 2
MAX_TETS = 50000
MAX_LEVEL = 6
DIMS = 16
TET_MASS = 1.0 / 6.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak (the roof of an FP64 VALU-bound kernel)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_steps(tree, steps, warmup, mode, precision, decomp):
    for _ in range(warmup):
        tree.deposit(DIMS, decomposition=decomp, tet_mass=TET_MASS, mode=mode, precision=precision)
    wall, main_ms, tl = [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        tree.deposit(DIMS, decomposition=decomp, tet_mass=TET_MASS, mode=mode, precision=precision)
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


def cpu_baseline(lat, tree_leaves_order, budget_s=20.0):
    """Times the reference CPU path (or the oracle port) on this host's cores; returns a dict.

    The reference deposits the leaves in its processing order, one leaf per worker thread at a
    time, each thread into its own float grid, until every leaf is done or budget_s has passed;
    a partial run is extrapolated by the meta-patch share of the leaves it finished."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from tests import _oracle as O  # checker side only
    from tests._octree_io import ref_octree
    threads = max(1, min(16, os.cpu_count() or 1))
    ref = O.load_ref()
    if ref is not None:
        d = _stdout_to_stderr(lambda: ref_octree(lat, MAX_TETS, MAX_LEVEL))
        order = [int(x) for x in d["order"]]
        dfs = list(d["dfs"])
        n_meta = np.diff(d["mc_offsets"])
        meta_of = {int(node): int(n_meta[i]) for i, node in enumerate(dfs)}
        lock = threading.Lock()
        todo = iter(order)
        done = []
        t0 = time.perf_counter()

        def worker():
            g = np.zeros(DIMS ** 3, np.float32)
            while time.perf_counter() - t0 < budget_s:
                with lock:
                    node = next(todo, None)
                if node is None:
                    return
                g[:] = 0
                assert ref.ref_octree_deposit_leaf(node, DIMS, O.P(g), None, None, None) == 0
                with lock:
                    done.append(node)

        pool = [threading.Thread(target=worker) for _ in range(threads)]
        for th in pool:
            th.start()
        for th in pool:
            th.join()
        wall = time.perf_counter() - t0
        frac = sum(meta_of[n] for n in done) / float(sum(meta_of.values()))
        est_wall = wall / frac
        whole = len(done) == len(order)
        return {"value": lat.num_tets / est_wall, "unit": "tets/s", "cores": threads, "kind": "reference",
                "sample": (f"{'all' if whole else len(done)} of {len(order)} leaves"
                           f"{'' if whole else f' ({frac:.3f} of the meta-patch work, whole-job time extrapolated by that share)'}"
                           f" deposited by the compiled reference DepositWrapper::deposit_tets (REF_EFFECTIVE tets, "
                           f"float grids, leaf gather included) on {threads} threads in {wall:.2f} s")}
    # fallback: the oracle port, FP64, on a fixed tet sample of the whole lattice
    box = np.asarray(A.cubic_root_box(lat.bbox()))
    ids = np.arange(0, lat.num_tets, 97, dtype=np.int32)
    t0 = time.perf_counter()
    O.deposit_lattice(lat, box, DIMS * 8, decomp=2, tet_mass=TET_MASS, tet_ids=ids, precision=1, nthreads=threads)
    wall = time.perf_counter() - t0
    return {"value": ids.size / wall, "unit": "tets/s", "cores": threads, "kind": "port",
            "sample": f"{ids.size} tets (every 97th) of the lattice on one 128^3 grid over the root box, FP64, "
                      f"{threads} threads, {wall:.2f} s"}


def algorithmic_flops(lat, tree, leaves, decomp, n_sample=24):
    """Reference-algorithm FP64 operation count per (leaf, tet) pair, from the oracle's counting
    mode on a sample of leaves; returns (flops per pair, pairs per step)."""
    from tests import _oracle as O  # checker side only
    counts = tree.bin(decomp)
    boxes = tree.leaf_boxes()
    pick = np.linspace(0, leaves.size - 1, n_sample).astype(int)
    flops = 0
    pairs = 0
    for l in pick:
        ids = tree.leaf_tets(int(l))
        _, st = O.deposit_lattice(lat, boxes[l], DIMS, decomp=decomp, tet_mass=TET_MASS, tet_ids=ids, precision=2,
                                  stats=True)
        flops += int(st[12])
        pairs += ids.size
    return flops / max(pairs, 1), int(counts.sum())


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
            rec = {"rate_req_s": rate, "batched_p50_s": p50, "batched_p99_s": p99,
                   "batched_tets_per_s": tets / makespan, "batched_makespan_s": makespan, "steps": int(st["steps"]),
                   "plain_serial_p50_s": q50, "plain_serial_p99_s": q99, "plain_serial_tets_per_s": tets / finish,
                   "plain_serial_makespan_s": finish, "batched_vs_plain_max_rel_err": err}
            log(f"[serve] {rate:g} req/s: batched p50 {p50 * 1e3:.1f} ms p99 {p99 * 1e3:.1f} ms "
                f"({tets / makespan:.3e} tets/s); serial plain p50 {q50 * 1e3:.1f} ms p99 {q99 * 1e3:.1f} ms "
                f"({tets / finish:.3e} tets/s); max rel diff {err:.2e}")
            out["rates"].append(rec)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the plain and mixed comparison runs")
    ap.add_argument("--serve", action="store_true", help="run only the ROI serving study (256^3 lattice)")
    ap.add_argument("--serve-rates", default="20,50,100,200")
    ap.add_argument("--serve-requests", type=int, default=60)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one MI355X; run with --gpus 1"}))
        sys.exit(2)
    if args.serve:
        res = serve_study([float(r) for r in args.serve_rates.split(",")], args.serve_requests)
        print(json.dumps({"metric": "p50/p99 ROI latency (s) and tets/s, continuous batching vs serial plain GPU",
                          "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
                          "data": "synthetic (256^3 lattice, k^-2 field at rms 3 spacings; Poisson ROI requests)",
                          "config": {"workload": "ROI serving, side U[5,50] Mpc/h, resolutions 64^3..512^3, fp32 density"},
                          **res}))
        return

    t0 = time.perf_counter()
    lat = make_lattice(N, RMS, seed=SEED)
    log(f"[bench] lattice {N}^3 generated in {time.perf_counter() - t0:.1f}s")
    decomp = A.DECOMP_KUHN
    with A.ResidentLattice.from_lattice(lat) as rl:
        t0 = time.perf_counter()
        tree = A.Octree(rl, MAX_TETS, MAX_LEVEL)
        t_oct = time.perf_counter() - t0
        t0 = time.perf_counter()
        counts = tree.bin(decomp)
        t_bin = time.perf_counter() - t0
        leaves = tree.leaves()["dfs"]
        log(f"[bench] octree {tree.num_nodes} nodes / {tree.num_leaves} leaves in {t_oct*1e3:.1f} ms; "
            f"binning {counts.sum()} pairs in {t_bin*1e3:.1f} ms")

        # launch shape by the autotuner (untimed), then the headline: batched persistent form, FP64
        t0 = time.perf_counter()
        shape = tree.autotune(DIMS, decomp, TET_MASS, A.PREC_FP64, reps=1)
        log(f"[bench] autotune: {shape} in {time.perf_counter() - t0:.1f}s")
        wall, main_ms, tl = run_steps(tree, args.steps, args.warmup, A.MODE_BATCHED, A.PREC_FP64, decomp)
        step_ms = wall / args.steps * 1e3
        value = lat.num_tets / (wall / args.steps)
        grids64 = tree.leaf_grids()
        total = float(grids64.sum())
        drift = abs(total - N ** 3) / N ** 3
        log(f"[bench] batched fp64: {step_ms:.2f} ms/step, kernel {main_ms:.2f} ms, {value:.3e} tets/s, "
            f"candidates {tl['candidates']}, partials {tl['partials']}, mass drift {drift:.2e}")

        extra = {}
        # the same deposit captured as a hipGraph and replayed (candidate space built once)
        with A.DepositGraph(tree, DIMS, decomp, TET_MASS, A.PREC_FP64) as g:
            g.launch()
            gms = [g.launch() for _ in range(max(1, args.steps))]
            gerr = max_rel_err(g.leaf_grids(), grids64, 1e-8 * TET_MASS)
        extra["graph_replay_ms"] = float(np.mean(gms))
        extra["graph_replay_tets_per_s"] = lat.num_tets / (float(np.mean(gms)) * 1e-3)
        extra["graph_vs_eager_max_rel_err"] = gerr
        log(f"[bench] graph replay: {np.mean(gms):.2f} ms ({extra['graph_replay_tets_per_s']:.3e} tets/s), err {gerr:.1e}")
        if not args.quick:
            wall_p, main_p, _ = run_steps(tree, max(2, args.steps // 2), 1, A.MODE_PLAIN, A.PREC_FP64, decomp)
            plain_tets = lat.num_tets / (wall_p / max(2, args.steps // 2))
            plain_grids = tree.leaf_grids()
            extra["plain_tets_per_s"] = plain_tets
            extra["batched_over_plain"] = value / plain_tets
            extra["plain_vs_batched_max_rel_err"] = max_rel_err(plain_grids, grids64, 1e-8 * TET_MASS)
            log(f"[bench] plain fp64: {wall_p / max(2, args.steps // 2) * 1e3:.2f} ms/step, {plain_tets:.3e} tets/s")
            wall_m, main_m, tlm = run_steps(tree, max(2, args.steps // 2), 1, A.MODE_BATCHED, A.PREC_MIXED, decomp)
            mixed_grids = tree.leaf_grids()
            extra["mixed_tets_per_s"] = lat.num_tets / (wall_m / max(2, args.steps // 2))
            extra["mixed_max_rel_err_vs_fp64"] = max_rel_err(mixed_grids, grids64, 1e-8 * TET_MASS)
            extra["mixed_fp64_recompute_fraction"] = tlm["fallbacks"] / max(tlm["partials"], 1)
            log(f"[bench] mixed: {extra['mixed_tets_per_s']:.3e} tets/s, err {extra['mixed_max_rel_err_vs_fp64']:.2e}, "
                f"recompute {extra['mixed_fp64_recompute_fraction']:.3f}")
            # the reference's own (non-tiling) decomposition, like for like with the CPU baseline
            wall_r, _, _ = run_steps(tree, max(2, args.steps // 2), 1, A.MODE_BATCHED, A.PREC_FP64, A.DECOMP_REF_EFFECTIVE)
            extra["ref_effective_tets_per_s"] = lat.num_tets / (wall_r / max(2, args.steps // 2))
            tree.bin(decomp)

        roofline = None
        try:
            flops_per_pair, n_pairs = algorithmic_flops(lat, tree, leaves, decomp)
            achieved = flops_per_pair * n_pairs / (main_ms * 1e-3) / 1e12
            roofline = {"bound": "mfma", "pipe": "VALU FP64 (clipping geometry; no matrix work)",
                        "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                        "flops_per_pair": flops_per_pair, "pairs_per_launch": n_pairs,
                        "kernel_ms": main_ms, "kernel": "k_deposit_waves<FP64>"}
        except Exception as exc:  # the oracle is absent: report the roofline as unmeasured
            log(f"[bench] roofline skipped: {exc}")

        cpu = None
        if not args.no_cpu_baseline:
            try:
                cpu = cpu_baseline(lat, tree.leaves()["order"])
                log(f"[bench] cpu baseline: {cpu['value']:.3e} tets/s ({cpu['kind']}, {cpu['cores']} cores)")
            except Exception as exc:
                log(f"[bench] cpu baseline skipped: {exc}")
        tree.close()

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
        "data": "synthetic (seeded k^-2 Gaussian displacement field, rms 3 spacings)",
        "config": {"workload": "adaptive exact deposit, 128^3-particle lattice (12,582,912 Kuhn tets), octree depth<=6, "
                               "split >50000, 16^3-cell leaves, batched persistent FP64 form",
                   "lattice": N, "tets": lat.num_tets, "leaves": int(leaves.size), "leaf_dims": DIMS,
                   "octree_ms": t_oct * 1e3, "binning_ms": t_bin * 1e3, "mass_drift": drift,
                   "launch_shape": {"tiles_per_pop": shape["chunk"], "workgroups_per_cu": shape["blocks_per_cu"]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
