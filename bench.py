"""Benchmark: 10-NN join of 200k synthetic nuclei against 1000 vessels on one
MI355X (mixed FP32/FP64 form, result lists identical to the FP64 form).

One step = one full `tdbase join -q nn --knn 10` over the resident tiles:
octree candidates, device filter, and the refinement rounds at LODs
20/40/60/80/100. The tiles are generated (seeded) and uploaded before the
timed region. `value` is triangle-pair distances per second over the whole
join (wall clock); `roofline` is for the refinement kernel, whose device time
is taken from HIP events recorded on the stream it runs on.

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
# part of k_mx_mark every triangle pair goes through (disc_interval() and the
# candidate comparison; the breakdown is beside disc_interval()).
F_MARK = 89
PEAK_FP32_TFLOPS = 157.3
REFINE_STAGES = ["reset", "bounds", "scan", "probe", "mark", "expand", "eval", "fold", "direct"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(tiles_dir, n_targets_total, n_sub, query, knn, tau, budget_s=15.0):
    """The oracle (single-threaded C restatement, FP64) on the first targets
    of the same workload, grown until it has run for about budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    t1 = orc.OracleTile(os.path.join(tiles_dir, "nuclei.dt"))
    t2 = orc.OracleTile(os.path.join(tiles_dir, "vessels.dt"))
    n_targets_total = min(n_targets_total, n_sub)
    m = 20
    while True:
        t0 = time.perf_counter()
        _, _, _, _, st = orc.join(t1, t2, query, k=knn, tau=tau, precision="f64",
                                  max_objects1=m)
        dt = time.perf_counter() - t0
        if dt >= budget_s / 3 or m >= n_targets_total:
            break
        m = min(n_targets_total, int(m * max(2.0, budget_s / 3 / max(dt, 1e-3))))
    return {"value": st["tri_pairs"] / dt, "unit": "tri-pair dist/s", "cores": 1, "kind": "port",
            "sample": f"first {m} of {n_targets_total} nuclei vs all vessels, "
                      f"{'%d-NN' % knn if query == 'nn' else 'within %g' % tau}, FP64 oracle, "
                      f"{dt:.1f} s, {st['tri_pairs']} triangle pairs"}


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
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the join shards by "
              f"target objects across GPUs but multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import importlib
    import tempfile
    import numpy as np
    tdp = importlib.import_module("3dpipe_amd")

    t0 = time.perf_counter()
    gen_kw = dict(seed=args.seed, n_vessels=args.vessels, nuclei_per_vessel=args.nuclei,
                  cube=args.cube, rings_min=args.rings_min, rings_max=args.rings_max)
    nuc, ves = tdp.generate(**gen_kw)
    log(f"generated {nuc.n_obj} nuclei ({nuc.n_tri} triangle records) and {ves.n_obj} vessels "
        f"({ves.n_tri}) in {time.perf_counter() - t0:.1f} s")
    eng = tdp.Engine(0)
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
        log(f"step {s}: {times[-1] * 1e3:.1f} ms (filter {r.timing_ms['filter']:.1f}, refine "
            f"{r.timing_ms['refine']:.1f}, refine kernels {r.timing_ms['refine_kernels']:.1f})")
    r = results[-1]
    ms = 1e3 * sum(times) / len(times)
    tri_pairs = r.stats["tri_pairs"]
    refine_kernel_ms = sum(x.timing_ms["refine_kernels"] for x in results) / len(results)
    rounds = max(1, r.stats["rounds"])
    stage_ms = {k: sum(x.stage_ms[k] for x in results) / len(results) for k in REFINE_STAGES}
    stage_n = r.stage_launches
    # the refinement round (all its kernels, between two HIP events on the
    # join stream) is the measured unit: algorithmic work = F_TRI per pair
    achieved = F_TRI * tri_pairs / (refine_kernel_ms * 1e-3) / 1e12
    dom = max((k for k in REFINE_STAGES if k not in ("scan", "reset")), key=lambda k: stage_ms[k])
    dom_kernel = {"mark": "k_mx_mark", "eval": "k_mx_eval", "bounds": "k_mx_bounds1",
                  "probe": "k_mx_eval", "expand": "k_mx_expand", "fold": "k_mx_fold",
                  "direct": "k_refine"}[dom]
    dom_launches = max(1, stage_n.get(dom, 1))
    dom_obj = {"name": dom_kernel, "stage": dom, "launches": dom_launches,
               "avg_launch_ms": stage_ms[dom] / dom_launches,
               "share_of_refine": stage_ms[dom] / refine_kernel_ms if refine_kernel_ms else None}
    if dom == "mark":
        a = F_MARK * tri_pairs / (stage_ms[dom] * 1e-3) / 1e12
        dom_obj.update({"flop_per_tri_pair": F_MARK, "achieved": a, "frac": a / PEAK_FP32_TFLOPS})
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
                   "object_pairs": r.stats["object_pairs"], "results": int(len(r.target)),
                   "timing_ms": r.timing_ms},
        "roofline": {"bound": "valu-fp32", "kernel": "refinement round (k_mx_* pipeline)"
                     if args.precision == "mixed" else "refinement round (k_refine)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "launches": rounds, "avg_launch_ms": refine_kernel_ms / rounds,
                     "flop_per_tri_pair": F_TRI, "dominant_kernel": dom_obj,
                     "stage_ms": {k: v for k, v in stage_ms.items() if v > 0}},
    }
    if args.compare_fp64 and args.precision != "fp64":
        t0 = time.perf_counter()
        rp = run("fp64")
        fp64_ms = 1e3 * (time.perf_counter() - t0)
        same = (np.array_equal(rp.target, r.target) and np.array_equal(rp.obj2, r.obj2)
                and np.array_equal(rp.rmin, r.rmin) and np.array_equal(rp.rmax, r.rmax))
        out["plain_fp64"] = {"ms": fp64_ms, "speedup": fp64_ms / ms, "identical_results": bool(same),
                             "refine_kernel_ms": rp.timing_ms["refine_kernels"]}
        log(f"plain FP64: {fp64_ms:.1f} ms, identical={same}")
    if args.cpu_baseline:
        with tempfile.TemporaryDirectory() as d:
            # the first nuclei of the same workload (identical objects) and
            # all vessels, written for the oracle
            n_sub = min(nuc.n_obj, 2000)
            sub, _ = tdp.generate(max_nuclei=n_sub, **gen_kw)
            sub.save(os.path.join(d, "nuclei.dt"))
            ves.save(os.path.join(d, "vessels.dt"))
            del sub
            out["cpu_baseline"] = cpu_baseline(d, nuc.n_obj, n_sub, args.query, args.knn,
                                               args.tau)
            out["cpu_baseline"]["vs_gpu"] = out["value"] / out["cpu_baseline"]["value"]
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
