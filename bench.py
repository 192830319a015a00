#!/usr/bin/env python3
"""Benchmark of the signed triangle counting hot path on one MI355X.

Headline workload: RMAT scale 22, edge factor 16, (a,b,c,d) =
(0.57,0.19,0.19,0.05), signs -1 with probability 0.25 -- 67,108,864 generated
lines, 64,155,555 oriented edges after dedupe.  The graph is generated on the
host, uploaded and turned into the oriented CSR (and the transposed index the
pull kernels read) on the device -- twice: build_ms / upload_ms are the second
preparation's (device buffers and pinned staging reused), build_ms_first /
upload_ms_first the first's (allocations included); a step is one
pass of the counting phase (binning + every intersection kernel) over that
device-resident CSR -- the reference's "Triangle Counting Execution" boundary
(Multicore Implementations/tc_merge_sign_enu.cpp:131-145), which also leaves
the CSR build out.

The JSON line carries:
  value         oriented edges / s of the counting phase (wall clock over the
                timed steps, bracketed by device synchronisation)
  roofline      the dominant kernel's algorithmic bytes -- what its strategy
                must read (alg_bytes_kernel) -- over its median hipEvent
                duration, against the 8 TB/s HBM peak; traffic = FETCH_SIZE +
                WRITE_SIZE bytes per launch from the committed rocprofv3 PMC
                summary (profiles/traffic.json); the merge-model B_alg of the
                whole phase beside it, for information
  cpu_baseline  the oracle's merge count (the reference's OpenMP merge path
                restated) on this host, OMP threads = nproc, bounded sample
  configs       one record per configuration (ER 10k/100k, grid 960^2, RMAT-22,
                hub 18M, the 4,096-graph batch): event-timed count (median of
                the timed repeats after warm-up), oriented edges/s, dominant
                kernel roofline, counter traffic, every dispatch mode's time,
                learned / tuned-fixed and learned / best-forced, the reference
                goldens, and the CPU merge and hash paths (oracle) on this host
  learned_vs_fixed  the hub-dominated graph's record, repeated at top level
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
KERNELS = ["merge_lane", "merge_wave", "hash_wave", "hash", "bitmap_wave", "bitmap",
           "pull_wave", "pull"]
WORKLOADS = {1: "er_10k_100k", 2: "grid_960x960_diag", 3: "rmat22_ef16", 4: "hub_18m",
             5: "batch4096"}
MODES = ("learned", "fixed", "merge", "hash", "bitmap", "pull")
FORCED = ("merge", "hash", "bitmap", "pull")
TIMED_REGION = ("binning + intersection kernels over the device-resident oriented CSR and its "
                "transposed index, both built once by stc_prepare (build_ms); the reference's "
                "counting phase (tc_merge_sign_enu.cpp:131-145) likewise excludes its CSR build")


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dispatch", default="learned")
    ap.add_argument("--reps", type=int, default=7,
                    help="event-timed repeats per dispatch mode in the config records")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="only the headline RMAT-22 line (no per-config records)")
    ap.add_argument("--profile-config", type=int, default=0,
                    help="count only config N (1-5) with learned dispatch, for rocprofv3")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 1 GiB device-to-device copy (calibrates FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="measure the HBM read rate for sequential and random-segment reads")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--traffic-from", default=None,
                    help="write profiles/traffic.json from a rocprofv3 PMC directory and exit")
    ap.add_argument("--workload", type=int, default=3,
                    help="with --traffic-from: the config the PMC runs counted")
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


def cpu_threads():
    """OMP threads as the reference runs them: OMP_NUM_THREADS, else $(nproc)
    (README.md:8)."""
    nproc = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return int(env) if env and env.isdigit() and int(env) > 0 else nproc, nproc


def median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2] if xs else None


def cpu_baseline(g, budget_s):
    """The oracle's merge count (tc_merge_sign_enu.cpp:131-145 restated), the
    median of 3 timed runs over every stride-th source vertex; a first pass
    over 1/64 of the vertices sets the stride so each run takes about a third
    of budget_s seconds (the whole graph when that fits).  Returns oriented
    edges per second (and the oracle's CSR, for the hash path's timing)."""
    import oracle
    threads, nproc = cpu_threads()
    t0 = time.time()
    csr = oracle.Csr(g.n, g.u, g.v, g.sign, threads)
    build_s = time.time() - t0
    probe = csr.merge_sample(64, 7, threads)
    est_full = probe["seconds"] * 64.0
    stride = max(1, int(round(est_full / (budget_s / 3.0))))
    runs = [csr.merge_sample(stride, (1 + k) % stride, threads) for k in range(3)]
    rates = sorted(r["edges"] / r["seconds"] for r in runs if r["seconds"] > 0)
    mid = runs[1]
    return csr, {
        "value": rates[len(rates) // 2] if rates else None,
        "unit": "oriented edges/s",
        "cores": threads,
        "nproc": nproc,
        "cpu_model": cpu_model(),
        "kind": "port",
        "runs": [round(r, 1) for r in rates],
        "sample": "oracle merge count (OpenMP, %d threads), median of 3 runs over source "
                  "vertices v %% %d == phase: %d of %d oriented edges in %.1f s per run; oracle "
                  "CSR build (%.1f s) not included"
                  % (threads, stride, mid["edges"], csr.m, mid["seconds"], build_s),
    }


def cpu_paths(graphs):
    """The reference's two OpenMP counting paths restated by the oracle, on
    this host with OMP threads = nproc: the merge path's counting loop
    (tc_merge_sign_enu.cpp:131-145) and the hash path's (tc_hash_sign_enu.cpp:
    124-165, its table build timed apart), median of 3 runs each, summed over
    the given graphs; the oracle's CSR build reported apart."""
    import oracle
    threads, nproc = cpu_threads()
    build = merge_s = hash_s = hash_build = 0.0
    m = 0
    counts = [0, 0]
    for g in graphs:
        t0 = time.time()
        csr = oracle.Csr(g.n, g.u, g.v, g.sign, threads)
        build += time.time() - t0
        m += csr.m
        mt, hb, hc = [], [], []
        for _ in range(3):
            t, b, u = csr.time_merge(threads)
            mt.append(t)
            th, bs, b2, u2 = csr.time_hash(threads)
            hc.append(th)
            hb.append(bs)
            if (b, u) != (b2, u2):
                raise RuntimeError("oracle merge and hash paths disagree")
        counts[0] += b
        counts[1] += u
        merge_s += median(mt)
        hash_s += median(hc)
        hash_build += median(hb)
        del csr
    return {"threads": threads, "nproc": nproc, "m_oriented": m,
            "merge_s": round(merge_s, 4),
            "merge_edges_per_s": m / merge_s if merge_s > 0 else None,
            "hash_s": round(hash_s, 4), "hash_table_build_s": round(hash_build, 4),
            "hash_edges_per_s": m / hash_s if hash_s > 0 else None,
            "csr_build_s": round(build, 3), "counts": counts}


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
             # one graph's workgroup and wave pull chunks, or a whole batch,
             # share one two-phase launch (reported in the last slot)
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
        prev = rec.get(key)
        # the launch that did the work (the others of the slot exit at once)
        if prev is None or fk > prev["fetch_kib_raw"]:
            short = re.search(r"(k_\w+(<[^>]*>)?)", name)
            rec[key] = {"kernel": short.group(1) if short else name.split("(")[0],
                        "bytes_per_launch":
                        round(1024.0 * (fk * ff + wk * fw)),
                        "fetch_bytes": round(1024.0 * fk * ff), "write_bytes": round(1024.0 * wk * fw),
                        "fetch_kib_raw": fk, "write_kib_raw": wk}
    out[workload] = dict(rec, _calibration={"fetch_factor": ff, "write_factor": fw})
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out[workload], indent=1))


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
        "traffic_rate_frac": (round(traffic / (kms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
                              if traffic and kms > 0 else None),
        "kernel_ms": round(kms, 4), "algorithmic_bytes": alg,
        "count_phase_merge_model": {
            "bytes": st["b_alg"], "achieved": round(phase, 1),
            "frac": round(phase / HBM_PEAK_GBPS, 4),
            "note": "B_alg charges 4(d+(u)+d+(v)) per oriented edge whatever the strategy; "
                    "informational, not a physical rate"},
    }


def time_modes(ctx, reps, warmup, batch=False):
    """Event-timed counting phase under every dispatch mode: {mode: (ms, counts)}
    (single graph: median of `reps` runs after `warmup`; batch: median of
    `reps` hipGraph replays after the eager runs)."""
    out = {}
    for mode in MODES:
        if batch:
            cnt = tuple(map(tuple, ctx.count_batch_prepared(dispatch=mode, repeat=reps)))
        else:
            if warmup:
                ctx.count_prepared(dispatch=mode, repeat=warmup)
            cnt = ctx.count_prepared(dispatch=mode, repeat=reps)
        out[mode] = (ctx.stats()["ms_count"], cnt)
    return out


def prepare_twice(ctx, prep):
    """Prepare (upload + device CSR build) twice: the first call also
    allocates the device buffers and pins the staging memory, the second
    reuses them.  Returns the second call's statistics, with the first call's
    times as ms_build_first / ms_upload_first."""
    prep()
    first = ctx.stats()
    prep()
    st = ctx.stats()
    st["ms_build_first"] = first["ms_build"]
    st["ms_upload_first"] = first["ms_upload"]
    return st


def config_record(ctx, wid, g, reps, warmup, batch=False):
    """One configuration: prepare, time every dispatch mode, then the learned
    dispatch's statistics and roofline."""
    wname = WORKLOADS[wid]
    prep = (lambda: ctx.prepare_batch(g)) if batch else (lambda: ctx.prepare(g))
    st0 = prepare_twice(ctx, prep)
    res = time_modes(ctx, reps, warmup, batch)
    counts = {v[1] for v in res.values()}
    if len(counts) != 1:
        raise RuntimeError("dispatch modes disagree on %s" % wname)
    # learned last, so the statistics below are the learned dispatch's
    if batch:
        got = ctx.count_batch_prepared(dispatch="learned", repeat=reps)
    else:
        got = ctx.count_prepared(dispatch="learned", repeat=reps)
    st = ctx.stats()
    ms = {m: round(v[0], 4) for m, v in res.items()}
    best = min(FORCED, key=lambda m: res[m][0])
    gold = goldens()
    if batch:
        per = gold.get("batch", {}).get("per_graph")
        exact = None if per is None else all(list(got[i]) == per[i][1:3] for i in range(len(got)))
        bal, unb = sum(x[0] for x in got), sum(x[1] for x in got)
    else:
        gw = gold.get("workloads", {}).get(str(wid))
        exact = None if gw is None else (got == (gw["balanced"], gw["unbalanced"]))
        bal, unb = got
    t = res["learned"][0]
    return {
        "workload": wname, "n": st0["n"], "lines": st0["m_in"], "m_oriented": st0["m_oriented"],
        "graphs": len(got) if batch else 1,
        "counts": {"balanced": bal, "unbalanced": unb, "exact_vs_reference": exact},
        "t_count_ms": round(t, 4), "oriented_edges_per_s": st0["m_oriented"] / (t * 1e-3),
        "ms_by_dispatch": ms,
        "learned_vs_fixed": round(res["fixed"][0] / t, 3),
        "learned_vs_best_forced": round(res[best][0] / t, 3), "best_forced": best,
        "ms_bin": round(st["ms_bin"], 4),
        "roofline": roofline_of(st, wname),
        "edges_by_strategy": {"merge": st["edges_merge"], "hash": st["edges_hash"],
                              "bitmap": st["edges_bitmap"], "pull": st["edges_pull"],
                              "skipped": st["edges_skipped"]},
        "build_ms": round(st0["ms_build"], 3), "upload_ms": round(st0["ms_upload"], 3),
        "build_ms_first": round(st0["ms_build_first"], 3),
        "upload_ms_first": round(st0["ms_upload_first"], 3),
        "timing": ("hipEvent median of %d %s" % (reps, "hipGraph replays after 2 eager runs"
                                                 if batch else "runs after %d warm-up" % warmup)),
    }


def profile_config(wid, steps, warmup):
    """Count only config wid under learned dispatch (for rocprofv3 runs)."""
    import signed_triangle_counting_amd as stc
    with stc.Counter(0) as ctx:
        if wid == 5:
            b = stc.gen_workload_batch(4096)
            ctx.prepare_batch(b)
            got = ctx.count_batch_prepared(dispatch="learned", repeat=steps)
            print(json.dumps({"workload": WORKLOADS[5], "graphs": len(got),
                              "ms_per_replay": ctx.stats()["ms_count"]}), flush=True)
        else:
            g = stc.gen_workload(wid)
            ctx.prepare(g)
            ctx.count_prepared(repeat=max(warmup, 1))
            got = ctx.count_prepared(repeat=steps)
            print(json.dumps({"workload": WORKLOADS[wid], "counts": got,
                              "ms_count": ctx.stats()["ms_count"]}), flush=True)


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

    if args.copy_calibration:
        a = torch.full((1 << 28,), 7, dtype=torch.int32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": 1 << 30, "ok": bool((b[:16] == 7).all())}), flush=True)
        return
    if args.hbm_probe:
        print(json.dumps({"hbm_read_tbps": hbm_probe(0),
                          "peak_tbps": HBM_PEAK_GBPS / 1000.0}), flush=True)
        return
    if args.profile_config:
        profile_config(args.profile_config, args.steps, args.warmup)
        return
    wname = WORKLOADS[3]
    g = stc.gen_workload(3)
    ctx = stc.Counter(0)
    if ctx.model_status != 0:
        raise RuntimeError("dispatcher model not loaded: " + ctx.last_error())
    st0 = prepare_twice(ctx, lambda: ctx.prepare(g))
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
    gold = goldens().get("workloads", {}).get("3")
    exact = None if gold is None else (counts == (gold["balanced"], gold["unbalanced"]))
    line = {
        "metric": "oriented_edges_per_s", "value": m_plus / (ms_per_step * 1e-3),
        "unit": "oriented edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": wname, "n": st["n"], "lines": st["m_in"],
                   "m_oriented": m_plus, "dispatch": args.dispatch,
                   "parallelism": "single GPU"},
        "timed_region": TIMED_REGION,
        "counts": {"balanced": counts[0], "unbalanced": counts[1], "exact_vs_reference": exact},
        "ms_count_event_median": round(st["ms_count"], 4), "ms_bin": round(st["ms_bin"], 4),
        "ms_kernels": dict(zip(KERNELS, [round(x, 4) for x in st["ms_kernel"]])),
        "build_ms": round(st0["ms_build"], 3), "upload_ms": round(st0["ms_upload"], 3),
        "build_ms_first": round(st0["ms_build_first"], 3),
        "upload_ms_first": round(st0["ms_upload_first"], 3),
        "roofline": roofline_of(st, wname),
        "edges_by_strategy": {"merge": st["edges_merge"], "hash": st["edges_hash"],
                              "bitmap": st["edges_bitmap"], "pull": st["edges_pull"],
                              "skipped": st["edges_skipped"]},
        "model_loaded": ctx.model_status == 0,
    }
    configs = []
    graphs = {3: g}
    if not args.no_configs:
        for wid in (1, 2, 3, 4):
            gw = graphs.get(wid) or stc.gen_workload(wid)
            graphs[wid] = gw
            rec = config_record(ctx, wid, gw, args.reps, 2)
            configs.append(rec)
            print("config %s: %.4f ms, learned/fixed %.3f" % (rec["workload"], rec["t_count_ms"],
                                                            rec["learned_vs_fixed"]),
                  file=sys.stderr, flush=True)
        batch = stc.gen_workload_batch(4096)
        rec = config_record(ctx, 5, batch, 5, 0, batch=True)
        configs.append(rec)
        print("config batch4096: %.4f ms" % rec["t_count_ms"], file=sys.stderr, flush=True)
        hub = configs[3]
        line["learned_vs_fixed"] = {
            "workload": hub["workload"], "fixed_ms": hub["ms_by_dispatch"]["fixed"],
            "learned_ms": hub["ms_by_dispatch"]["learned"], "speedup": hub["learned_vs_fixed"],
            "learned_vs_best_forced": hub["learned_vs_best_forced"],
            "best_forced": hub["best_forced"],
            "exact_equal": True,  # config_record raises when dispatch modes disagree
            "fixed_rule": "tuned (models/tune.txt; autotune.py on the dispatcher's training set)"}
    ctx.close()
    if not args.no_cpu:
        csr, line["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds)
        if configs:
            for rec, wid in zip(configs, (1, 2, 3, 4)):
                if wid == 3:
                    continue  # the headline sample above covers RMAT-22's merge path
                rec["cpu"] = cpu_paths([graphs[wid]])
            # RMAT-22: one full hash count beside the headline merge sample
            threads, nproc = cpu_threads()
            th, tb, _, _ = csr.time_hash(threads)
            cb = line["cpu_baseline"]
            configs[2]["cpu"] = {"threads": threads, "nproc": nproc,
                                 "merge_edges_per_s": cb["value"],
                                 "merge_s": round(m_plus / cb["value"], 3) if cb["value"] else None,
                                 "hash_s": round(th, 3), "hash_table_build_s": round(tb, 3),
                                 "hash_edges_per_s": m_plus / th if th > 0 else None,
                                 "note": "merge: the headline cpu_baseline sample; hash: one run"}
            # batch: every 16th graph (the sample the reference goldens pin)
            sample = [batch.graph(i) for i in range(0, batch.k, 16)]
            cpu = cpu_paths(sample)
            cpu["sample"] = "every 16th graph: %d of %d" % (len(sample), batch.k)
            configs[4]["cpu"] = cpu
            for rec in configs:
                c = rec["cpu"]
                if c.get("merge_edges_per_s"):
                    c["gpu_over_cpu_merge"] = round(rec["oriented_edges_per_s"] /
                                                    c["merge_edges_per_s"], 1)
        del csr
    else:
        line["cpu_baseline"] = None
    if configs:
        line["configs"] = configs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
