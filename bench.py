"""Benchmark: 10-NN join of 200k synthetic nuclei against 1000 vessels on one
MI355X (mixed FP32/FP64 form, result lists identical to the FP64 form).

One step = one full `tdbase join -q nn --knn 10` over the resident tiles:
octree candidates, device filter, and the refinement rounds at LODs
20/40/60/80/100. The tiles are generated (seeded) and uploaded before the
timed region. `value` is triangle-pair distances per second over the whole
join (wall clock); `roofline` is for the dominant kernel of the refinement
(k_mx_mark in the mixed form), whose device time is taken from HIP events
recorded around its launches on the stream it runs on; its
`effective_refinement` member states the whole refinement round's rate in
TriDist-equivalent operations.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Static FP32 operation count of one triangle pair on the canonical full path
# of TriDist (9 edge-edge cases + 2 vertex-face cases, no early exit) plus the
# bound update; add/sub/mul/div/sqrt/compare = 1. Counted by running the
# oracle's TriDist with a counting number type (oracle/flop_count.cpp, checked
# in tests/test_cpu.py).
F_TRI = 997
# Static FP32 operation count of the mixed form's per-pair disc test, the
# part of k_mx_mark every triangle pair it tests goes through (mark_pair()
# through the comparison, with the voxel pair's margin; the breakdown is
# beside mark_pair()). 39 in the round that folds d alone, 44 in the rounds
# that fold lo and hi; the smaller count is used, so the figure is a lower
# bound of the kernel's work.
F_MARK = 39
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per launch on the default workload (seed 1, 200k nuclei x 1000
# vessels, 10-NN, mixed form, default tuning), from rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE runs of this script, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (the
# factors at which the counters read a plain 1 GiB device copy,
# --calibrate-copy 1024, as 1 GiB): k_mx_mark (mean over the 5 rounds), and
# one refinement round (all its kernels), per tuning the autotuner picks there
# (the default tuning, and rows per wave iteration capped at 4). Measured in
# the runs summarised in TRAFFIC_SOURCE; reported only for that workload and
# those tunings.
DEFAULT_TUNING = {"small_pairs": 8, "max_rows": 64, "drain": 0, "drain_blocks": 0,
                  "eval_blocks": 4096, "graph": 0, "row_cull": 1}
MEASURED_TRAFFIC = {  # max_rows -> (k_mx_mark bytes per launch, bytes per round)
    64: (11.91e9, 25.85e9),
    4: (11.97e9, 25.95e9),
}
TRAFFIC_SOURCE = "profiles/knn10_200k_nuclei_1000_vessels_mixed_traffic_and_valu.md"
REFINE_STAGES = ["reset", "bounds", "scan", "probe", "mark", "expand", "eval", "fold", "direct",
                 "graph"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(tiles_dir, n_targets_total, n_sub, query, knn, tau, budget_s=15.0):
    """The oracle (the C restatement, FP64; candidates, filter and refinement
    folds in parallel over targets on all host threads) on the first targets
    of the same workload, grown until it has run for about budget_s / 3 or
    reaches n_sub targets; the full join's time extrapolated linearly in the
    targets."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    t1 = orc.OracleTile(os.path.join(tiles_dir, "nuclei.dt"))
    t2 = orc.OracleTile(os.path.join(tiles_dir, "vessels.dt"))
    n_cap = min(n_targets_total, n_sub)
    m = 20
    while True:
        t0 = time.perf_counter()
        _, _, _, _, st = orc.join(t1, t2, query, k=knn, tau=tau, precision="f64",
                                  max_objects1=m)
        dt = time.perf_counter() - t0
        if dt >= budget_s / 3 or m >= n_cap:
            break
        m = min(n_cap, int(m * max(2.0, budget_s / 3 / max(dt, 1e-3))))
    return {"value": st["tri_pairs"] / dt, "unit": "tri-pair dist/s", "cores": orc.threads(),
            "kind": "port",
            "sample": f"first {m} of {n_targets_total} nuclei vs all vessels, "
                      f"{'%d-NN' % knn if query == 'nn' else 'within %g' % tau}, FP64 oracle "
                      f"(candidates, filter and refinement on {orc.threads()} threads), "
                      f"{dt:.1f} s, {st['tri_pairs']} triangle pairs",
            "sample_targets": m, "sample_s": dt,
            "extrapolated_full_join_s": dt * n_targets_total / m,
            "extrapolation": "linear in the targets (the sample's nuclei lie around the first "
                             "vessels; every target's join is independent)"}


def list_agreement(a, b):
    """How far result lists a agree with b (b the reference): entries in one
    and not the other, targets whose lists differ, and the relative error of
    a's distance ranges on the entries both hold."""
    import numpy as np
    ka = a.target.astype(np.int64) * (1 << 32) + a.obj2.astype(np.int64)
    kb = b.target.astype(np.int64) * (1 << 32) + b.obj2.astype(np.int64)
    common, ia, ib = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
    only_a = np.setdiff1d(ka, kb, assume_unique=True)
    only_b = np.setdiff1d(kb, ka, assume_unique=True)
    bad_t = np.union1d(only_a >> 32, only_b >> 32)
    out = {"entries": int(len(ka)), "reference_entries": int(len(kb)),
           "missing": int(len(only_b)), "extra": int(len(only_a)),
           "targets_differing": int(len(bad_t)),
           "identical": bool(len(only_a) == 0 and len(only_b) == 0 and np.array_equal(a.target, b.target)
                             and np.array_equal(a.obj2, b.obj2))}
    for name, x, y in (("rmin", a.rmin[ia], b.rmin[ib]), ("rmax", a.rmax[ia], b.rmax[ib])):
        fin = np.isfinite(y) & (y > 0)
        rel = np.abs(x[fin] - y[fin]) / y[fin]
        out[f"{name}_rel_err_max"] = float(rel.max()) if rel.size else 0.0
        out[f"{name}_rel_err_median"] = float(np.median(rel)) if rel.size else 0.0
        out[f"{name}_bit_identical"] = bool(np.array_equal(x.view(np.uint64), y.view(np.uint64)))
    return out


def stream_bench(tdp, n_nuclei, cube, budget_gib, knn, seed, tuning, log_fn, lazy=True,
                 workers=0):
    """Chunked streaming 10-NN self-join of n_nuclei nuclei (radii 6-14,
    uniform in the cube) within a device budget (tdb_stream_join): workers
    side by side (four by default), each decoding a chunk's voxel and object arrays
    into pinned buffers and joining it; lazily (the default) only the records
    of the voxels each LOD round reads are decoded and shipped, otherwise
    every record of every chunk object."""
    t0 = time.perf_counter()
    src = tdp.Source.generate(seed=seed, n_vessels=0, nuclei_per_vessel=n_nuclei, cube=cube,
                              nucleus_r_min=6.0, nucleus_r_max=14.0)
    eng = tdp.Engine(0)
    if tuning:
        eng.set_tuning(**tuning)
    budget = int(budget_gib * (1 << 30))
    r = eng.stream_join(src, None, query="nn", knn=knn, precision="mixed", hbm_budget=budget,
                        lazy=lazy, workers=workers)
    wall = time.perf_counter() - t0
    eng.close()
    st = r.stream
    h2d = st["h2d_bytes"]
    log_fn(f"stream ({'lazy' if lazy else 'every record'}): {n_nuclei} nuclei, "
           f"{st['chunks']:.0f} chunks, {wall:.1f} s, h2d {h2d / 1e9:.1f} GB "
           f"(records {st['record_bytes_shipped'] / 1e9:.1f} of {st['record_bytes_all'] / 1e9:.1f} GB), "
           f"overlap {st['overlap']:.2f}, peak {st['peak_device_bytes'] / 2**30:.1f} GiB")
    return {
        "workload": f"{knn}-NN self-join of {n_nuclei} nuclei (radii 6-14) in a {cube:g}^3 cube, "
                    f"LODs 20-100, mixed precision, device budget {budget_gib:g} GiB",
        "records": "per LOD round, the voxels it reads" if lazy else "every record",
        "workers": int(st["workers"]),
        "seed": seed, "wall_s": wall, "setup_s": wall - st["wall_ms"] / 1e3,
        "chunks": int(st["chunks"]), "objects_decoded": int(st["objects_decoded"]),
        "h2d_bytes": int(h2d),
        "h2d_record_bytes": int(st["record_bytes_shipped"]),
        "h2d_meta_bytes": int(st["meta_bytes"]),
        "record_bytes_of_chunk_objects": int(st["record_bytes_all"]),
        "record_share_shipped": (st["record_bytes_shipped"] / st["record_bytes_all"]
                                 if st["record_bytes_all"] else None),
        "pcie_gbps": h2d / (st["copy_ms"] * 1e6) if st["copy_ms"] else None,
        "end_to_end_gbps": h2d / (st["wall_ms"] * 1e6) if st["wall_ms"] else None,
        "decode_gbps": h2d / (st["decode_ms"] * 1e6) if st["decode_ms"] else None,
        "stage_ms": {"index": st["index_ms"], "candidates": st["candidates_ms"],
                     "plan": st["plan_ms"], "workers_setup": st["workers_setup_ms"],
                     "decode": st["decode_ms"], "copy": st["copy_ms"], "join": st["join_ms"],
                     "pipeline_wall": st["wall_ms"]},
        "candidate_targets_from_host_octree": int(st["candidate_host_targets"]),
        "overlap_fraction": st["overlap"],
        "peak_device_bytes": int(st["peak_device_bytes"]),
        "budget_bytes": budget, "chunk_splits": int(st["splits"]),
        "triangle_pairs": r.stats["tri_pairs"], "results": int(len(r.target)),
        "tri_pair_dist_per_s": r.stats["tri_pairs"] / wall,
    }


def serve_bench(tdp, vessels, gen_kw, n_queries, rates, seed, log_fn):
    """Serving benchmark: `vessels` resident; n_queries fresh nucleus
    tiles of log-uniform size 100..50000 (seeded), each kNN with k in
    {1, 5, 10, 20} or within with tau U(50, 500), arriving as a Poisson
    process (unit-rate gaps scaled to each of `rates` queries/s). At each rate
    run on the same schedule: the batched mixed server (records per LOD
    round, batches joined side by side), the same uploading every record
    with its tile, and one-query-at-a-time plain FP64 (every record
    uploaded). Returns, per rate, latency percentiles, throughput, bytes sent
    to the device and whether every query's results agree."""
    import threading
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(100), np.log(50000), n_queries)).astype(int)
    kinds = []
    for q in range(n_queries):
        if rng.random() < 0.5:
            kinds.append(("nn", int(rng.choice([1, 5, 10, 20])), 0.0))
        else:
            kinds.append(("within", 1, float(rng.uniform(50, 500))))
    unit_gaps = rng.exponential(1.0, n_queries)
    t0 = time.perf_counter()
    nv = gen_kw["n_vessels"]
    tiles = []
    for q in range(n_queries):
        kw = dict(gen_kw)
        kw.update(nuclei_per_vessel=max(1, -(-int(sizes[q]) // nv)), max_nuclei=int(sizes[q]),
                  nuclei_seed=10_000 + q, emit_vessels=0)
        tiles.append(tdp.generate(**kw)[0])
    n_nuclei = int(sum(t.n_obj for t in tiles))
    log_fn(f"serving: {n_queries} query tiles ({n_nuclei} nuclei) generated in "
           f"{time.perf_counter() - t0:.1f} s")

    def run(precision, max_batch, overlap, gaps, lazy=True):
        srv = tdp.Server(vessels, max_batch=max_batch, overlap=overlap, precision=precision,
                         lazy=lazy)
        copies = [t.clone() for t in tiles]  # each server owns what it is given
        tickets = [None] * n_queries
        start = time.perf_counter()

        def feed():
            t = start
            for q in range(n_queries):
                t += gaps[q]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[q] = srv.submit(copies[q], query=kinds[q][0], knn=kinds[q][1],
                                        within_dist=kinds[q][2])
        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res, lat = [], []
        for q in range(n_queries):
            r, l = srv.wait(tickets[q])
            res.append(r)
            lat.append(l)
        makespan = time.perf_counter() - start
        st = srv.stats()
        srv.close()
        lat = np.array(lat)
        return res, {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "mean_ms": float(lat.mean()), "queries_per_s": n_queries / makespan,
                     "nuclei_per_s": n_nuclei / makespan, "batches": st["batches"],
                     "max_batch": st["max_batch"], "peak_device_gb": st["peak_device_bytes"] / 1e9,
                     "executors": st["executors"],
                     "most_batches_at_once": st["most_batches_at_once"],
                     "mixed_kind_batches": st["mixed_kind_batches"],
                     "h2d_bytes": st["h2d_meta_bytes"] + st["h2d_record_bytes"],
                     "h2d_record_bytes": st["h2d_record_bytes"],
                     "record_bytes_of_queries": st["record_bytes_of_queries"],
                     "host_stage_ms": st["stage_ms"], "makespan_s": makespan}

    out = {"queries": n_queries, "nuclei": n_nuclei,
           "sizes": "log-uniform 100..50000 nuclei",
           "mix": "kNN k in {1,5,10,20} / within tau U(50,500)", "by_rate": []}
    for rate in rates:
        gaps = unit_gaps / rate
        res_b, batched = run("mixed", 64, True, gaps)
        log_fn(f"serving {rate:g} q/s, batched mixed: p50 {batched['p50_ms']:.1f} ms, p99 "
               f"{batched['p99_ms']:.1f} ms, {batched['queries_per_s']:.1f} q/s, "
               f"{batched['batches']} batches (largest {batched['max_batch']}), "
               f"{batched['most_batches_at_once']} at once, h2d {batched['h2d_bytes'] / 1e9:.1f} GB")
        res_e, every = run("mixed", 64, True, gaps, lazy=False)
        log_fn(f"serving {rate:g} q/s, batched mixed, every record uploaded: p50 "
               f"{every['p50_ms']:.1f} ms, p99 {every['p99_ms']:.1f} ms, "
               f"{every['queries_per_s']:.1f} q/s, h2d {every['h2d_bytes'] / 1e9:.1f} GB")
        res_s, single = run("fp64", 1, False, gaps, lazy=False)
        log_fn(f"serving {rate:g} q/s, one-at-a-time FP64: p50 {single['p50_ms']:.1f} ms, p99 "
               f"{single['p99_ms']:.1f} ms, {single['queries_per_s']:.1f} q/s")
        same = all(np.array_equal(a.target, b.target) and np.array_equal(a.obj2, b.obj2)
                   and np.array_equal(a.rmin, b.rmin) and np.array_equal(a.rmax, b.rmax)
                   for r in (res_b, res_e) for a, b in zip(r, res_s))
        out["by_rate"].append({
            "arrival_rate_per_s": rate, "batched_mixed": batched,
            "batched_mixed_every_record": every, "one_at_a_time_fp64": single,
            "p50_speedup": single["p50_ms"] / batched["p50_ms"],
            "p99_speedup": single["p99_ms"] / batched["p99_ms"],
            "throughput_speedup": batched["queries_per_s"] / single["queries_per_s"],
            "identical_results": bool(same)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--vessels", type=int, default=1000)
    ap.add_argument("--nuclei", type=int, default=200, help="nuclei per vessel")
    ap.add_argument("--cube", type=float, default=20000.0)
    ap.add_argument("--knn", type=int, default=10)
    ap.add_argument("--query", default="nn", choices=["nn", "within"])
    ap.add_argument("--tau", type=float, default=400.0, help="within distance")
    ap.add_argument("--rings-min", type=int, default=12, help="vessel rings per segment, low")
    ap.add_argument("--rings-max", type=int, default=12, help="vessel rings per segment, high")
    ap.add_argument("--precision", default="mixed", choices=["ref32", "fp64", "mixed"])
    ap.add_argument("--compare-fp64", type=int, default=1,
                    help="also time one plain FP64 join and check identical lists")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--compare-ref32", type=int, default=1,
                    help="also run the REF32 form and state its list agreement and distance "
                         "error against FP64")
    ap.add_argument("--cold-join", type=int, default=1,
                    help="also time one join on a fresh engine (tile upload included)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--serve-queries", type=int, default=24,
                    help="serving run (resident vessels, Poisson arrivals) after the join bench; 0: off")
    ap.add_argument("--serve-rates", default="8,32",
                    help="query arrival rates (per second) of the serving runs")
    ap.add_argument("--tune", default="", help="kernel tuning, e.g. drain=0,max_rows=4")
    ap.add_argument("--autotune", type=int, default=40000,
                    help="autotune on this many targets before timing (0: off)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many MiB three times (the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--stream-nuclei", type=int, default=2000000,
                    help="streaming self-join of this many nuclei after the join bench; 0: off")
    ap.add_argument("--only-stream", type=int, default=0,
                    help="run only the streaming self-join and print its figures")
    ap.add_argument("--only-serve", type=int, default=0,
                    help="run only the serving benchmark and print its figures")
    ap.add_argument("--stream-cube", type=float, default=50000.0)
    ap.add_argument("--stream-budget-gib", type=float, default=24.0)
    ap.add_argument("--stream-workers", type=str, default="0",
                    help="streaming workers side by side (0 = the library's default; a comma "
                         "list with --only-stream runs each)")
    ap.add_argument("--stream-every-record", type=int, default=0,
                    help="also run the streaming self-join shipping every record of every "
                         "chunk object, for comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the join shards by "
              f"target objects across GPUs but multi-GPU runs are not supported by this benchmark",
              file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy > 0:
        import torch
        n = args.calibrate_copy << 20
        x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}),
              flush=True)
        return

    import importlib
    import tempfile
    import numpy as np
    tdp = importlib.import_module("3dpipe_amd")
    if args.only_serve:
        gen_kw = dict(seed=args.seed, n_vessels=args.vessels, nuclei_per_vessel=args.nuclei,
                      cube=args.cube, rings_min=args.rings_min, rings_max=args.rings_max)
        _, ves = tdp.generate(**dict(gen_kw, max_nuclei=1))
        print(json.dumps(serve_bench(tdp, ves, gen_kw, args.serve_queries,
                                     [float(x) for x in args.serve_rates.split(",")],
                                     args.seed + 100, log)), flush=True)
        return
    if args.only_stream:
        ws = [int(x) for x in args.stream_workers.split(",")]
        out = stream_bench(tdp, args.stream_nuclei, args.stream_cube, args.stream_budget_gib,
                           args.knn, args.seed, None, log, workers=ws[0])
        for w in ws[1:]:
            out[f"workers_{w}"] = stream_bench(tdp, args.stream_nuclei, args.stream_cube,
                                               args.stream_budget_gib, args.knn, args.seed, None,
                                               log, workers=w)
        if args.stream_every_record:
            out["every_record"] = stream_bench(tdp, args.stream_nuclei, args.stream_cube,
                                               args.stream_budget_gib, args.knn, args.seed, None,
                                               log, lazy=False)
        print(json.dumps(out), flush=True)
        return

    t0 = time.perf_counter()
    gen_kw = dict(seed=args.seed, n_vessels=args.vessels, nuclei_per_vessel=args.nuclei,
                  cube=args.cube, rings_min=args.rings_min, rings_max=args.rings_max)
    nuc, ves = tdp.generate(**gen_kw)
    log(f"generated {nuc.n_obj} nuclei ({nuc.n_tri} triangle records) and {ves.n_obj} vessels "
        f"({ves.n_tri}) in {time.perf_counter() - t0:.1f} s")
    eng = tdp.Engine(0)
    tuned = None
    if args.tune:
        eng.set_tuning(**{k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(","))})
    elif args.autotune:
        t0 = time.perf_counter()
        tuned = eng.autotune(nuc, ves, query=args.query, knn=args.knn, within_dist=args.tau,
                             sample_targets=args.autotune, precision=args.precision)
        tuned["seconds"] = time.perf_counter() - t0
        log(f"autotune ({args.autotune} targets, {len(tuned['tried'])} tunings, "
            f"{tuned['seconds']:.1f} s): {tuned['best']}")
    name, cus = eng.device()
    log(f"device {name}, {cus} CUs")

    def run(precision):
        return eng.join(nuc, ves, query=args.query, knn=args.knn, within_dist=args.tau,
                        precision=precision)

    for w in range(args.warmup):
        r = run(args.precision)
        log(f"warmup {w}: {r.timing_ms['total']:.1f} ms, {r.stats['tri_pairs']} triangle pairs")
    times, results = [], []
    for s in range(args.steps):
        t0 = time.perf_counter()
        r = run(args.precision)
        times.append(time.perf_counter() - t0)
        results.append(r)
        tm = r.timing_ms
        log(f"step {s}: {times[-1] * 1e3:.1f} ms (candidates {tm['candidates']:.1f}, set-up "
            f"{tm['setup']:.1f}, filter {tm['filter']:.1f}, refine {tm['refine']:.1f}, refine "
            f"kernels {tm['refine_kernels']:.1f}, download {tm['download']:.1f}, assembly "
            f"{tm['assembly']:.1f})")
    r = results[-1]
    ms = 1e3 * sum(times) / len(times)
    tri_pairs = r.stats["tri_pairs"]
    refine_kernel_ms = sum(x.timing_ms["refine_kernels"] for x in results) / len(results)
    rounds = max(1, r.stats["rounds"])
    stage_ms = {k: sum(x.stage_ms[k] for x in results) / len(results) for k in REFINE_STAGES}
    stage_n = r.stage_launches
    # the measured traffic applies to the workload it was measured on
    tun = eng.tuning()
    measured = MEASURED_TRAFFIC.get(tun["max_rows"]) if all(
        tun[k] == DEFAULT_TUNING[k] for k in DEFAULT_TUNING if k != "max_rows") else None
    default_workload = (args.seed == 1 and args.vessels == 1000 and args.nuclei == 200 and
                        args.cube == 20000.0 and args.query == "nn" and args.knn == 10 and
                        args.rings_min == args.rings_max == 12 and args.precision == "mixed" and
                        measured is not None)
    # The roofline is the dominant kernel's: its stage time comes from HIP
    # events recorded around its launches on the join stream.
    dom = max((k for k in REFINE_STAGES if k not in ("scan", "reset")), key=lambda k: stage_ms[k])
    dom_kernel = {"mark": "k_mx_mark", "eval": "k_mx_eval", "bounds": "k_mx_bounds",
                  "probe": "k_mx_eval", "expand": "k_mx_expand", "fold": "k_mx_fold",
                  "direct": "k_refine", "graph": "captured round"}[dom]
    dom_launches = max(1, stage_n.get(dom, 1))
    dom_ms = stage_ms[dom] / dom_launches
    # algorithmic work per launch: the per-pair operation count of the
    # kernel's test (k_mx_mark: F_MARK; the plain forms' k_refine: F_TRI)
    # times the triangle pairs the kernel tests one by one (k_mx_mark passes
    # over whole rows ruled out against a voxel sphere, and leaves voxel pairs
    # of at most 8 triangle pairs to FP64: those pairs are not counted)
    flop_pair = {"mark": F_MARK, "direct": F_TRI}.get(dom)
    tested = r.stats.get("mark_tested", 0) if dom == "mark" else tri_pairs
    achieved = (flop_pair * tested / (stage_ms[dom] * 1e-3) / 1e12) if flop_pair else None
    roofline = {"bound": "valu-fp32" if dom != "direct" or args.precision != "fp64" else "valu-fp64",
                "kernel": dom_kernel, "achieved": achieved,
                "peak": PEAK_FP32_TFLOPS if args.precision != "fp64" else PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": None, "traffic": None,
                "launches": dom_launches, "avg_launch_ms": dom_ms, "flop_per_tri_pair": flop_pair,
                "tri_pairs_tested": tested, "tri_pairs_of_rounds": tri_pairs,
                "share_of_refine": stage_ms[dom] / refine_kernel_ms if refine_kernel_ms else None}
    if achieved is not None:
        roofline["frac"] = achieved / roofline["peak"]
    if dom == "mark" and default_workload:
        roofline["traffic"] = measured[0]
        roofline["traffic_source"] = f"{TRAFFIC_SOURCE} (max_rows {tun['max_rows']})"
    # The reference's distance kernel computes TriDist on every triangle pair
    # (F_TRI each); the rate at which this pipeline gets through that work is
    # an effective figure (most pairs are ruled out without TriDist), not a
    # measure of any one kernel.
    eff = F_TRI * tri_pairs / (refine_kernel_ms * 1e-3) / 1e12
    roofline["effective_refinement"] = {
        "kernels": "refinement round (k_mx_* pipeline)" if args.precision == "mixed"
        else "refinement round (k_refine)",
        "tridist_equivalent_tflops": eff, "vs_fp32_peak": eff / PEAK_FP32_TFLOPS,
        "rounds": rounds, "avg_round_ms": refine_kernel_ms / rounds,
        "round_traffic": measured[1] if default_workload else None,
        "stage_ms": {k: v for k, v in stage_ms.items() if v > 0}}
    out = {
        "metric": "10-NN join time (200k nuclei); tri-pair dist/s vs FP32 peak; p50/p99 query latency",
        "value": tri_pairs / (ms * 1e-3),
        "unit": "tri-pair dist/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else ("f32" if args.precision == "ref32" else "f64"),
        "data": "synthetic (seeded icosphere nuclei around tube vessels)",
        "config": {"workload": (f"{args.knn}-NN join" if args.query == "nn" else
                                f"within-{args.tau:g} join") +
                               f" nuclei->vessels, {nuc.n_obj} nuclei x {ves.n_obj} vessels "
                               f"(rings/segment {args.rings_min}-{args.rings_max}), "
                               f"{args.cube:g}^3 cube, LODs 20-100",
                   "precision": args.precision, "seed": args.seed,
                   "triangle_pairs": tri_pairs, "voxel_pairs": r.stats["voxel_pairs"],
                   "fp32_culled_tri_pairs": r.stats["fp32_culled"],
                   "fp64_tridist_evals": r.stats["fp64_evals"],
                   "fp64_probe_evals": r.stats.get("fp64_probe_evals", 0),
                   "object_pairs": r.stats["object_pairs"], "results": int(len(r.target)),
                   "tuning": eng.tuning(), "autotune": tuned,
                   "timing_ms": r.timing_ms},
        "roofline": roofline,
    }
    if args.compare_fp64 and args.precision != "fp64":
        t0 = time.perf_counter()
        rp = run("fp64")
        fp64_ms = 1e3 * (time.perf_counter() - t0)
        same = (np.array_equal(rp.target, r.target) and np.array_equal(rp.obj2, r.obj2)
                and np.array_equal(rp.rmin, r.rmin) and np.array_equal(rp.rmax, r.rmax))
        out["plain_fp64"] = {"ms": fp64_ms, "speedup": fp64_ms / ms, "identical_results": bool(same),
                             "refine_kernel_ms": rp.timing_ms["refine_kernels"],
                             "agreement": list_agreement(r, rp)}
        log(f"plain FP64: {fp64_ms:.1f} ms, identical={same}")
        if args.compare_ref32:
            # the reference's own arithmetic (FP32 on every pair) against FP64
            t0 = time.perf_counter()
            r32 = run("ref32")
            ref32_ms = 1e3 * (time.perf_counter() - t0)
            out["ref32_vs_fp64"] = dict(list_agreement(r32, rp), ms=ref32_ms)
            log(f"REF32: {ref32_ms:.1f} ms, {out['ref32_vs_fp64']}")
            del r32
        del rp
    if args.cold_join:
        # a join on a fresh engine: the tiles' upload (and bound records) included
        e2 = tdp.Engine(0)
        e2.set_tuning(**eng.tuning())
        t0 = time.perf_counter()
        rc = e2.join(nuc, ves, query=args.query, knn=args.knn, within_dist=args.tau,
                     precision=args.precision)
        cold_ms = 1e3 * (time.perf_counter() - t0)
        e2.close()
        out["cold_join"] = {"ms": cold_ms, "upload_ms": rc.timing_ms["upload"],
                            "tile_bytes_h2d": int(48 * (nuc.n_tri + ves.n_tri)),
                            "h2d_gbps": 48 * (nuc.n_tri + ves.n_tri) / (rc.timing_ms["upload"] * 1e6)
                            if rc.timing_ms["upload"] else None}
        log(f"cold join: {cold_ms:.1f} ms (upload {rc.timing_ms['upload']:.1f} ms)")
        del rc
    if args.cpu_baseline:
        with tempfile.TemporaryDirectory() as d:
            # the first nuclei of the same tile and all vessels, written for
            # the oracle
            n_sub = min(nuc.n_obj, 20000)
            sub = tdp.Source.of_tile(nuc).decode(np.arange(n_sub))
            sub.save(os.path.join(d, "nuclei.dt"))
            ves.save(os.path.join(d, "vessels.dt"))
            del sub
            out["cpu_baseline"] = cpu_baseline(d, nuc.n_obj, n_sub, args.query, args.knn,
                                               args.tau)
            out["cpu_baseline"]["vs_gpu"] = out["value"] / out["cpu_baseline"]["value"]
    tuning_used = eng.tuning()
    eng.close()
    if args.stream_nuclei > 0:
        out["streaming"] = stream_bench(tdp, args.stream_nuclei, args.stream_cube,
                                        args.stream_budget_gib, args.knn, args.seed,
                                        tuning_used, log)
    if args.serve_queries > 0:
        del nuc, results, r
        out["serving"] = serve_bench(tdp, ves, gen_kw, args.serve_queries,
                                     [float(x) for x in args.serve_rates.split(",")],
                                     args.seed + 100, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
