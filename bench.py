#!/usr/bin/env python3
"""Benchmark of the signed triangle counting hot path on one MI355X.

Workload (default): RMAT scale 22, edge factor 16, (a,b,c,d) =
(0.57,0.19,0.19,0.05), signs -1 with probability 0.25 -- 67,108,864 generated
lines, 64,155,555 oriented edges after dedupe.  The graph is generated on the
host, uploaded and turned into the oriented CSR on the device once (reported
as build_ms); a step is one pass of the counting phase (binning + every
intersection kernel) over that device-resident CSR.

The JSON line carries:
  value       oriented edges / s of the counting phase (wall clock over the
              timed steps, bracketed by device synchronisation)
  roofline    the dominant kernel's algorithmic bytes (4(d+(u)+d+(v)) per
              oriented edge it counted) / its mean hipEvent duration, against
              the 8 TB/s HBM peak; traffic = FETCH_SIZE + WRITE_SIZE bytes per
              launch from the committed rocprofv3 PMC summary, if present
  cpu_baseline the CPU oracle's merge count (a restatement of the reference's
              OpenMP merge path) timed on this host over a bounded sample
  learned_vs_fixed  counting time of the fixed-threshold dispatch over the
              learned dispatch on the hub-dominated graph (18M vertices)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
KERNELS = ["merge_lane", "merge_wave", "hash_wave", "hash", "bitmap_wave", "bitmap",
           "pull_wave", "pull"]
WORKLOADS = {1: "er_10k_100k", 2: "grid_960x960_diag", 3: "rmat22_ef16", 4: "hub_18m"}


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--dispatch", default="learned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hub", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="measure the HBM read rate for sequential and random-segment reads")
    ap.add_argument("--batch-only", action="store_true",
                    help="count only the 4,096-graph batch (for profiling) and print its record")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--traffic-from", default=None,
                    help="write profiles/traffic.json from a rocprofv3 PMC directory and exit")
    return ap.parse_args()


def goldens():
    p = os.path.join(ROOT, "tests", "golden", "goldens.json")
    try:
        with open(p) as f:
            return json.load(f)
    except OSError:
        return {}


def traffic_for(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 PMC summary."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
        rec = t.get(workload, {}).get(kernel)
        return rec["bytes_per_launch"] if rec else None
    except (OSError, ValueError, KeyError):
        return None


def cpu_model():
    """The host CPU's model name, as lscpu prints it (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(g, budget_s):
    """The oracle's merge count (tc_merge_sign_enu.cpp:131-145 restated), the
    median of 3 timed runs over every stride-th source vertex; a first pass
    over 1/64 of the vertices sets the stride so each run takes about a third
    of budget_s seconds (the whole graph when that fits).  Returns oriented
    edges per second."""
    import oracle
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.time()
    csr = oracle.Csr(g.n, g.u, g.v, g.sign, threads)
    build_s = time.time() - t0
    probe = csr.merge_sample(64, 7, threads)
    est_full = probe["seconds"] * 64.0
    stride = max(1, int(round(est_full / (budget_s / 3.0))))
    runs = [csr.merge_sample(stride, (1 + k) % stride, threads) for k in range(3)]
    rates = sorted(r["edges"] / r["seconds"] for r in runs if r["seconds"] > 0)
    mid = runs[1]
    return {
        "value": rates[len(rates) // 2] if rates else None,
        "unit": "oriented edges/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "kind": "port",
        "runs": [round(r, 1) for r in rates],
        "sample": "oracle merge count (OpenMP, %d threads), median of 3 runs over source "
                  "vertices v %% %d == phase: %d of %d oriented edges in %.1f s per run; oracle "
                  "CSR build (%.1f s) not included"
                  % (threads, stride, mid["edges"], csr.m, mid["seconds"], build_s),
    }


def traffic_from(pmc_dir, workload, out_path):
    """Per-launch HBM traffic of each kernel from rocprofv3 PMC runs:
    pmc_dir/{fetch,write,copy_fetch,copy_write}/run_counter_collection.csv.
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are each scaled by what the
    same counter reads for a device copy of known size (copy_*: 1 GiB)."""
    import csv
    import collections

    def read(sub):
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        with open(os.path.join(pmc_dir, sub, "run_counter_collection.csv")) as f:
            for row in csv.DictReader(f):
                name = row["Kernel_Name"]
                agg[name][row["Counter_Name"]].append(float(row["Counter_Value"]))
        return agg

    def copy_factor(sub, counter):
        for name, d in read(sub).items():
            if "copyBuffer" in name and counter in d:
                v = sorted(d[counter])[len(d[counter]) // 2]
                return (1 << 20) / v  # 1 GiB = 2^20 KiB moved each way
        return None

    ff = copy_factor("copy_fetch", "FETCH_SIZE")
    fw = copy_factor("copy_write", "WRITE_SIZE")
    fetch, write = read("fetch"), read("write")
    names = {"k_merge_lane": "merge_lane", "k_merge_wave": "merge_wave",
             "k_hash_wave": "hash_wave", "k_hash<": "hash",
             "k_bitmap_wave<false, false>": "bitmap_wave", "k_bitmap<false>": "bitmap",
             "k_bitmap_wave<false, true>": "pull_wave",
             # a two-class count's workgroup and wave pull chunks share one
             # two-phase launch (the batch kernel, as a batch of one graph)
             "k_two_phase": "pull"}
    out = {}
    try:
        with open(out_path) as f:
            out = json.load(f)
    except (OSError, ValueError):
        pass
    rec = {}
    for name, d in fetch.items():
        key = next((v for k, v in names.items() if k in name), None)
        if key is None or "FETCH_SIZE" not in d:
            continue
        fk = sum(d["FETCH_SIZE"]) / len(d["FETCH_SIZE"])
        wk = 0.0
        for wn, wd in write.items():
            if wn == name and "WRITE_SIZE" in wd:
                wk = sum(wd["WRITE_SIZE"]) / len(wd["WRITE_SIZE"])
        rec[key] = {"bytes_per_launch": round(1024.0 * (fk * ff + wk * fw)),
                    "fetch_kib_raw": fk, "write_kib_raw": wk}
    out[workload] = dict(rec, _calibration={"fetch_factor": ff, "write_factor": fw})
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out[workload], indent=1))


def roofline_of(st, wname):
    """Roofline of the dominant kernel: the bytes its strategy must read
    (alg_bytes_kernel: staged lists once per item, the streamed list of every
    counted edge, and the metadata of every edge and item visited) over its
    median hipEvent duration on the library's stream.  The merge-model B_alg
    of the whole counting phase, which charges every edge with both lists
    whatever the strategy, is reported beside it for information."""
    k = max(range(len(KERNELS)), key=lambda i: st["ms_kernel"][i])
    kname = KERNELS[k]
    kms = st["ms_kernel"][k]
    alg = st["alg_bytes_kernel"][k]
    achieved = alg / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = traffic_for(kname, wname)
    phase = st["b_alg"] / (st["ms_count"] * 1e-3) / 1e9 if st["ms_count"] > 0 else 0.0
    return {
        "bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
        "traffic": traffic,
        "traffic_over_algorithmic": round(traffic / alg, 3) if traffic and alg else None,
        "kernel_ms": round(kms, 4), "algorithmic_bytes": alg,
        "count_phase_merge_model": {
            "bytes": st["b_alg"], "achieved": round(phase, 1),
            "frac": round(phase / HBM_PEAK_GBPS, 4),
            "note": "B_alg charges 4(d+(u)+d+(v)) per oriented edge whatever the strategy; "
                    "informational, not a physical rate"},
    }


def hbm_probe(device):
    """HBM read rate (TB/s) for a sequential read and for random 128-byte-aligned
    segments of 128..1024 bytes, from the bench-only probe library (not part of
    libstc.so)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "signed_triangle_counting_amd", "libstc_bench_probe.so"))
    lib.bench_probe_read_bandwidth.argtypes = [C.c_int, C.POINTER(C.c_double)]
    out = (C.c_double * 5)()
    rc = lib.bench_probe_read_bandwidth(device, out)
    if rc != 0:
        raise RuntimeError("bench_probe_read_bandwidth returned %d" % rc)
    keys = ("sequential", "segment_128B", "segment_256B", "segment_512B", "segment_1024B")
    return {k: round(float(v), 3) for k, v in zip(keys, out)}


def hub_learned_vs_fixed(ctx, stc, steps):
    g = stc.gen_workload(4)
    ctx.prepare(g)
    out = {"workload": WORKLOADS[4], "m_oriented": ctx.stats()["m_oriented"]}
    res = {}
    for mode in ("fixed", "learned"):
        ctx.count_prepared(dispatch=mode, repeat=2)
        res[mode] = ctx.count_prepared(dispatch=mode, repeat=steps)
        out[mode + "_ms"] = round(ctx.stats()["ms_count"], 4)
    out["exact_equal"] = res["fixed"] == res["learned"]
    out["speedup"] = round(out["fixed_ms"] / out["learned_ms"], 3) if out["learned_ms"] else None
    return out


def batch_workload(ctx, stc, steps):
    """4,096 independent graphs counted in one persistent launch per step,
    replayed from a captured hipGraph; per-graph counts against the fixture."""
    b = stc.gen_workload_batch(4096)
    t0 = time.perf_counter()
    got = ctx.count_batch(b, dispatch="learned", repeat=steps)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    per = goldens().get("batch", {}).get("per_graph")
    exact = None if per is None else all(list(got[i]) == per[i][1:3] for i in range(b.k))
    return {"graphs": b.k, "lines": b.total_m, "m_oriented": st["m_oriented"],
            "ms_per_replay": round(st["ms_count"], 4),
            "oriented_edges_per_s": st["m_oriented"] / (st["ms_count"] * 1e-3),
            "build_ms": round(st["ms_build"], 3), "upload_ms": round(st["ms_upload"], 3),
            "call_wall_s": round(wall, 3), "exact_vs_reference": exact}


def main():
    args = parse_args()
    if args.traffic_from:
        traffic_from(args.traffic_from, WORKLOADS[args.workload],
                     os.path.join(ROOT, "profiles", "traffic.json"))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build counts one graph (or one batch) per "
              "GPU and has no multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import signed_triangle_counting_amd as stc

    if args.hbm_probe:
        print(json.dumps({"hbm_read_tbps": hbm_probe(0),
                          "peak_tbps": HBM_PEAK_GBPS / 1000.0}), flush=True)
        return
    if args.batch_only:
        with stc.Counter(0) as ctx:
            print(json.dumps({"batch": batch_workload(ctx, stc, args.steps)}), flush=True)
        return
    wname = WORKLOADS[args.workload]
    g = stc.gen_workload(args.workload)
    ctx = stc.Counter(0)
    ctx.prepare(g)
    st0 = ctx.stats()
    if args.warmup > 0:
        ctx.count_prepared(dispatch=args.dispatch, repeat=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counts = ctx.count_prepared(dispatch=args.dispatch, repeat=args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ms_per_step = 1e3 * wall / args.steps
    m_plus = st["m_oriented"]
    gold = goldens().get("workloads", {}).get(str(args.workload))
    exact = None if gold is None else (counts == (gold["balanced"], gold["unbalanced"]))

    roofline = roofline_of(st, wname)
    line = {
        "metric": "oriented_edges_per_s", "value": m_plus / (ms_per_step * 1e-3),
        "unit": "oriented edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": wname, "n": st["n"], "lines": st["m_in"],
                   "m_oriented": m_plus, "dispatch": args.dispatch,
                   "parallelism": "single GPU"},
        "counts": {"balanced": counts[0], "unbalanced": counts[1], "exact_vs_reference": exact},
        "ms_count_event_median": round(st["ms_count"], 4), "ms_bin": round(st["ms_bin"], 4),
        "ms_kernels": dict(zip(KERNELS, [round(x, 4) for x in st["ms_kernel"]])),
        "build_ms": round(st0["ms_build"], 3), "upload_ms": round(st0["ms_upload"], 3),
        "b_alg_bytes": st["b_alg"],
        "roofline": roofline,
        "edges_by_strategy": {"merge": st["edges_merge"], "hash": st["edges_hash"],
                              "bitmap": st["edges_bitmap"], "pull": st["edges_pull"],
                              "skipped": st["edges_skipped"]},
        "model_loaded": ctx.model_status == 0,
    }
    if not args.no_hub:
        line["learned_vs_fixed"] = hub_learned_vs_fixed(ctx, stc, max(3, args.steps // 2))
    if not args.no_batch:
        line["batch"] = batch_workload(ctx, stc, 3)
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
