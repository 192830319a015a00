"""Autotuner for the counter's launch parameters and fixed rule; writes tune.txt.

    python -m signed_triangle_counting_amd.autotune --out DIR [--only SECTIONS]

The reference fixes its tile sizes, lanes per item and hash-table sizes at
compile time (SharedMEM/signed_tc_hash_shmem.hxx:14-19, signed_tc_shmem.hxx:
15-19, GlobalMEM/signed_tc.hxx:503-504).  Here they are runtime parameters,
tuned by timing the counting phase on synthetic graphs whose seeds differ from
the benchmark workloads:

  merge     merge_wave_min_log2: the list length from which one wave64 per
            edge (tiled merge) replaces one lane per edge -- one global value,
            then one value per shorter-side class;
  bitmap    bitmap_span_log2 (rank window of the workgroup LDS bitmap: LDS
            size and so workgroups per CU) and wave_window_log2 (per-wave
            window for source lists of at most 64 entries);
  hash      hash_max_log2: the longest source list given an LDS hash table;
  pull      pull_block_chunk_log2 / pull_wave_chunk_log2: in-edges per
            workgroup / wave pull item; wave_pop: items a wave takes per queue
            pop;
  wavepull  pull_wave_max_log2 / batch_pull_wave_max_log2: the longest
            target list (over a span within the wave window) counted by a
            wave instead of a workgroup, in single graphs / in batches;
  batch     lane_item_log2: merge-lane edges per workgroup item of the batched
            kernel (timed on small-graph batches);
  fixed     the fixed rule's four thresholds (fixed_merge_log2,
            fixed_pull_log2, fixed_skew_log2, fixed_hub_log2), by a full grid
            search on the learned dispatcher's own training graphs and batches
            (train_dispatcher.py), minimising the mean over graphs of the
            rule's count time relative to the graph's best single strategy.

Every candidate must reproduce the counts of the default parameters.  The
values each section sweeps are CANDIDATES, which the GPU tests also count
under one by one.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import signed_triangle_counting_amd as stc  # noqa: E402

KEYS = ["merge_wave_min_log2", "hash_max_log2", "bitmap_span_log2", "wave_window_log2",
        "fixed_merge_log2", "fixed_pull_log2", "fixed_skew_log2", "fixed_hub_log2",
        "pull_block_chunk_log2", "pull_wave_chunk_log2", "pull_wave_max_log2",
        "batch_pull_wave_max_log2", "wave_pop", "lane_item_log2"]

DEFAULT = {"merge_wave_min_log2": [5] * 16, "hash_max_log2": 10, "bitmap_span_log2": 17,
           "wave_window_log2": 14, "fixed_merge_log2": 3, "fixed_pull_log2": 0,
           "fixed_skew_log2": 2, "fixed_hub_log2": 9, "pull_block_chunk_log2": 11,
           "pull_wave_chunk_log2": 8, "pull_wave_max_log2": 6, "batch_pull_wave_max_log2": 6,
           "wave_pop": 8, "lane_item_log2": 9}

# Values each parameter is swept over (merge_wave_min_log2: as one global value).
CANDIDATES = {
    "merge_wave_min_log2": [3, 4, 5, 6, 7, 8],
    "bitmap_span_log2": [16, 17, 18, 19],
    "wave_window_log2": [13, 14, 15, 16],
    "hash_max_log2": [9, 10, 11],
    "pull_block_chunk_log2": [9, 10, 11, 12, 13],
    "pull_wave_chunk_log2": [5, 6, 7, 8, 9, 10],
    "pull_wave_max_log2": [6, 7, 8, 9],
    "batch_pull_wave_max_log2": [6, 7, 8, 9],
    "wave_pop": [2, 4, 8],
    "lane_item_log2": [9, 10, 11],
    # the fixed rule: 16 switches the pull, skew and hub branches off, -1 the
    # merge branch; a negative pull threshold pulls longer suffixes too
    "fixed_merge_log2": [-1, 2, 3, 4, 5, 6, 7],
    "fixed_pull_log2": [-6, -4, -2, -1, 0, 1, 2, 16],
    "fixed_skew_log2": [1, 2, 3, 4, 16],
    "fixed_hub_log2": [5, 7, 9, 11, 16],
}

FIXED_KEYS = ["fixed_merge_log2", "fixed_pull_log2", "fixed_skew_log2", "fixed_hub_log2"]


def tune_text(p):
    lines = ["stc-tune v1"]
    for k in KEYS:
        v = p[k]
        lines.append("%s %s" % (k, " ".join(str(x) for x in v) if isinstance(v, list) else v))
    return "\n".join(lines) + "\n"


def read_tune(path):
    """The parameters of a tune.txt (defaults for keys it does not name)."""
    p = {k: (list(v) if isinstance(v, list) else v) for k, v in DEFAULT.items()}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] in p:
                vals = [int(x) for x in parts[1:]]
                p[parts[0]] = vals if parts[0] == "merge_wave_min_log2" else vals[0]
    return p


def tuning_graphs():
    return [
        ("rmat20_ef16", lambda: stc.gen_rmat(20, 16, 0.57, 0.19, 0.19, 0.25, 201)),
        ("rmat21_ef16", lambda: stc.gen_rmat(21, 16, 0.57, 0.19, 0.19, 0.25, 202)),
        ("er_1e6_8e6", lambda: stc.gen_er(1000000, 8000000, 0.3, 203)),
        ("er_2e4_2e6", lambda: stc.gen_er(20000, 2000000, 0.3, 204)),
        ("hub_3m", lambda: stc.gen_hub(3000000, 3200000, 8, 1.2, 0.6, 30000, 0.1, 206)),
        ("grid_700", lambda: stc.gen_grid(700, 700, 0.5, 0.2, 207)),
    ]


def tuning_batches():
    return [("batch_512", lambda: stc.gen_batch(512, seed=221)),
            ("batch_256_dense", lambda: stc.gen_batch(256, d_lo=16.0, d_hi=64.0, seed=222))]


class Timer:
    """Counts prepared graphs under tune candidates; checks every count."""

    def __init__(self, ctx, reps):
        self.ctx, self.reps, self.want = ctx, reps, {}

    def time(self, name, p, mode, batch=False):
        self.ctx.set_tune(tune_text(p))
        if batch:  # the call's warm eager run precedes the timed replays
            got = self.ctx.count_batch_prepared(dispatch=mode, repeat=self.reps)
            got = (sum(x[0] for x in got), sum(x[1] for x in got))
        else:
            self.ctx.count_prepared(dispatch=mode, repeat=1)
            got = self.ctx.count_prepared(dispatch=mode, repeat=self.reps)
        w = self.want.setdefault(name, got)
        if got != w:
            raise RuntimeError("tune candidate changed a count on %s: %r != %r" % (name, got, w))
        return float(self.ctx.stats()["ms_count"])


def sweep(timer, items, best, key, mode, log, batch=False):
    """Time each candidate of `key` (others at `best`) over the prepared items
    (name, prepare thunk); returns {value: total ms}."""
    res = {}
    for val in CANDIDATES[key]:
        res[val] = 0.0
    for name, prep in items:
        prep()
        for val in CANDIDATES[key]:
            p = dict(best)
            p[key] = [val] * 16 if key == "merge_wave_min_log2" else val
            res[val] += timer.time(name, p, mode, batch)
    for val in CANDIDATES[key]:
        log("  %s=%s (%s): %.3f ms" % (key, val, mode, res[val]))
    return res


def fixed_table(hist, f):
    """The fixed rule on the non-empty buckets: 0 merge, 1 wave merge (skew
    branch), 2 pull, 3 push bitmap -- the strategies the device rule returns
    before its merge variant is chosen."""
    out = []
    for b in np.nonzero(hist)[0]:
        la, lb = int(b) // 16, int(b) % 16
        if max(la, lb) <= f["fixed_merge_log2"]:
            s = 0
        elif lb - la >= f["fixed_pull_log2"]:
            s = 2
        elif abs(la - lb) >= f["fixed_skew_log2"]:
            s = 1
        elif la >= f["fixed_hub_log2"]:
            s = 3
        else:
            s = 0
        out.append(s)
    return tuple(out)


def tune_fixed(ctx, timer, best, log):
    """Grid search of the fixed rule on the dispatcher's training set: per
    graph, every grid point's table is derived from the bucket histogram and
    each distinct table is timed once (fixed dispatch under that point)."""
    from signed_triangle_counting_amd import train_dispatcher as td
    grid = [dict(zip(FIXED_KEYS, vals)) for vals in
            itertools.product(*[CANDIDATES[k] for k in FIXED_KEYS])]
    per_graph = []
    # keyed apart from the tuning graphs (a training graph may share a name)
    items = [("train:" + n, th, False) for n, th in td.training_graphs()] + \
            [("train:" + n, th, True) for n, th in td.training_batches()]
    for name, thunk, is_batch in items:
        g = thunk()
        if is_batch:
            ctx.prepare_batch(g)
        else:
            ctx.prepare(g)
        hist = ctx.bucket_histogram()
        forced = {}
        for s in ("merge", "hash", "bitmap", "pull", "learned"):
            p = dict(best)
            forced[s] = timer.time(name, p, s, is_batch)
        best_forced = min(forced[s] for s in ("merge", "hash", "bitmap", "pull"))
        sig_time = {}
        times = []
        for f in grid:
            sig = fixed_table(hist, f)
            if sig not in sig_time:
                sig_time[sig] = timer.time(name, dict(best, **f), "fixed", is_batch)
                if len(sig_time) % 100 == 0:
                    log("    %s: %d tables timed" % (name, len(sig_time)))
            times.append(sig_time[sig])
        per_graph.append({"graph": name, "best_forced_ms": best_forced, "forced_ms": forced,
                          "times": times, "distinct_tables": len(sig_time)})
        log("  %-18s best forced %.3f ms, learned %.3f ms, %d distinct fixed tables" % (
            name, best_forced, forced["learned"], len(sig_time)))
        del g
    rel = np.array([[t / r["best_forced_ms"] for t in r["times"]] for r in per_graph])
    score = rel.mean(axis=0)
    k = int(np.argmin(score))
    chosen = grid[k]
    default = dict((key, DEFAULT[key]) for key in FIXED_KEYS)
    kd = grid.index(default)
    summary = {"chosen": chosen, "mean_rel_time": float(score[k]),
               "default": default, "default_mean_rel_time": float(score[kd]),
               "graphs": [{"graph": r["graph"], "best_forced_ms": round(r["best_forced_ms"], 4),
                           "forced_ms": {s: round(v, 4) for s, v in r["forced_ms"].items()},
                           "tuned_fixed_ms": round(r["times"][k], 4),
                           "default_fixed_ms": round(r["times"][kd], 4),
                           "distinct_tables": r["distinct_tables"]} for r in per_graph]}
    log("  fixed rule %s: mean time / best single strategy %.3f (default %s: %.3f)" % (
        chosen, score[k], default, score[kd]))
    return chosen, summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--only", default="merge,bitmap,hash,pull,wavepull,batch,fixed")
    ap.add_argument("--start", default=None, help="tune.txt to start from (default: defaults)")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    sections = set(args.only.split(","))
    ctx = stc.Counter(0)
    best = read_tune(args.start) if args.start else dict(DEFAULT)
    timer = Timer(ctx, args.reps)
    log_lines = []

    def log(s):
        print(s, flush=True)
        log_lines.append(s)

    t0 = time.time()
    graphs = [(name, thunk()) for name, thunk in tuning_graphs()]

    def items(gs=graphs):
        return [(name, (lambda g=g: ctx.prepare(g))) for name, g in gs]

    report = {}
    if "merge" in sections:
        log("merge lane -> wave threshold")
        res = sweep(timer, items(), best, "merge_wave_min_log2", "merge", log)
        best["merge_wave_min_log2"] = [min(res, key=res.get)] * 16
        # per shorter-side class: each class present, the others at the
        # global value (a bucket's class is the log2 of its shorter side)
        tot = {}
        for name, prep in items():
            prep()
            hist = ctx.bucket_histogram()
            classes = sorted({min(int(b) // 16, int(b) % 16) for b in np.nonzero(hist)[0]})
            for lmin in classes:
                for t in CANDIDATES["merge_wave_min_log2"]:
                    p = dict(best, merge_wave_min_log2=list(best["merge_wave_min_log2"]))
                    p["merge_wave_min_log2"][lmin] = t
                    tot.setdefault(lmin, {})
                    tot[lmin][t] = tot[lmin].get(t, 0.0) + timer.time(name, p, "merge")
        for lmin, res in sorted(tot.items()):
            best["merge_wave_min_log2"][lmin] = min(res, key=res.get)
            log("  shorter side class %d: threshold %d (%s)" % (
                lmin, best["merge_wave_min_log2"][lmin],
                ", ".join("%d: %.3f" % kv for kv in sorted(res.items()))))
        report["merge_wave_min_log2"] = best["merge_wave_min_log2"]
    if "bitmap" in sections:
        log("workgroup bitmap window")
        res = sweep(timer, items(), best, "bitmap_span_log2", "bitmap", log)
        best["bitmap_span_log2"] = min(res, key=res.get)
        log("per-wave bitmap window")
        res = sweep(timer, items(), best, "wave_window_log2", "bitmap", log)
        best["wave_window_log2"] = min(res, key=res.get)
    if "hash" in sections:
        log("hash table cap")
        res = sweep(timer, items(), best, "hash_max_log2", "hash", log)
        best["hash_max_log2"] = min(res, key=res.get)
    batches = None
    if "pull" in sections or "wavepull" in sections or "batch" in sections:
        batches = [(name, thunk()) for name, thunk in tuning_batches()]
    bitems = [(name, (lambda b=b: ctx.prepare_batch(b))) for name, b in (batches or [])]
    if "pull" in sections:
        for key in ("pull_block_chunk_log2", "pull_wave_chunk_log2", "wave_pop"):
            log(key)
            res = sweep(timer, items(), best, key, "pull", log)
            resb = sweep(timer, bitems, best, key, "pull", log, batch=True)
            tot = {v: res[v] + resb[v] for v in res}
            best[key] = min(tot, key=tot.get)
    if "wavepull" in sections:
        # a candidate replaces the starting value only when at least 1% faster
        # (about the run-to-run spread of these sums)
        def pick(res, cur):
            b = min(res, key=res.get)
            return b if res[b] < 0.99 * res.get(cur, float("inf")) else cur
        log("pull_wave_max_log2 (single graphs)")
        res = sweep(timer, items(), best, "pull_wave_max_log2", "pull", log)
        best["pull_wave_max_log2"] = pick(res, best["pull_wave_max_log2"])
        log("batch_pull_wave_max_log2 (batches)")
        res = sweep(timer, bitems, best, "batch_pull_wave_max_log2", "pull", log, batch=True)
        best["batch_pull_wave_max_log2"] = pick(res, best["batch_pull_wave_max_log2"])
    if "batch" in sections:
        log("batched merge-lane item")
        res = sweep(timer, bitems, best, "lane_item_log2", "merge", log, batch=True)
        best["lane_item_log2"] = min(res, key=res.get)
    del graphs, batches
    if "fixed" in sections:
        log("fixed rule (dispatcher training graphs and batches)")
        chosen, summary = tune_fixed(ctx, timer, best, log)
        best.update(chosen)
        report["fixed_rule"] = summary
    text = tune_text(best)
    with open(os.path.join(args.out, "tune.txt"), "w") as f:
        f.write(text + "# tuned on gfx950 by signed_triangle_counting_amd/autotune.py\n")
    report["tune"] = best
    report["seconds"] = round(time.time() - t0, 1)
    report["log"] = log_lines
    with open(os.path.join(args.out, "autotune_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    print(text, flush=True)


if __name__ == "__main__":
    main()
