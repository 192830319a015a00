"""Autotuner for the counter's launch parameters; writes tune.txt.

    python -m signed_triangle_counting_amd.autotune --out DIR

Tuned, by timing the counting phase on synthetic tuning graphs (seeds
distinct from the benchmark workloads):
  * merge_wave_min_log2  the list length from which one wave64 per edge
                          (tiled merge) replaces one lane per edge, per
                          shorter-side bucket;
  * bitmap_span_log2     rank window of the workgroup LDS bitmap (LDS size
                          and so workgroups per CU);
  * wave_window_log2     rank window of the per-wave LDS bitmap used for
                          source lists of at most 64 entries;
  * hash_max_log2        longest source list given an LDS hash table.
Each candidate must return the same counts as the default parameters.
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import signed_triangle_counting_amd as stc  # noqa: E402


def tune_text(p):
    return ("stc-tune v1\n"
            "merge_wave_min_log2 %s\n"
            "hash_max_log2 %d\n"
            "bitmap_span_log2 %d\n"
            "wave_window_log2 %d\n"
            "fixed_skew_log2 %d\n"
            "fixed_hub_log2 %d\n" % (" ".join(str(x) for x in p["merge_wave_min_log2"]),
                                     p["hash_max_log2"], p["bitmap_span_log2"],
                                     p["wave_window_log2"], p["fixed_skew_log2"],
                                     p["fixed_hub_log2"]))


DEFAULT = {"merge_wave_min_log2": [5] * 16, "hash_max_log2": 10, "bitmap_span_log2": 17,
           "wave_window_log2": 15, "fixed_skew_log2": 2, "fixed_hub_log2": 9}


def tuning_graphs():
    return [
        ("rmat20_ef16", stc.gen_rmat(20, 16, 0.57, 0.19, 0.19, 0.25, 201)),
        ("rmat21_ef16", stc.gen_rmat(21, 16, 0.57, 0.19, 0.19, 0.25, 202)),
        ("er_1e6_8e6", stc.gen_er(1000000, 8000000, 0.3, 203)),
        ("er_2e4_2e6", stc.gen_er(20000, 2000000, 0.3, 204)),
        ("rmat22_ef16", stc.gen_rmat(22, 16, 0.57, 0.19, 0.19, 0.25, 205)),
    ]


def timed(ctx, text, g, mode, want, reps=5):
    ctx.set_tune(text)
    ctx.prepare(g)
    got = ctx.count_prepared(dispatch=mode, repeat=2)
    if want is not None and got != want:
        raise RuntimeError("tune candidate changed a count")
    ctx.count_prepared(dispatch=mode, repeat=reps)
    return ctx.stats()["ms_count"], got


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    ctx = stc.Counter(0)
    graphs = tuning_graphs()
    best = {k: (list(v) if isinstance(v, list) else v) for k, v in DEFAULT.items()}
    want = {}
    for name, g in graphs:
        _, want[name] = timed(ctx, tune_text(best), g, "merge", None)

    def total(p, mode):
        return sum(timed(ctx, tune_text(p), g, mode, want[name])[0] for name, g in graphs)

    # 1. lane -> wave threshold for the merge kernels (global, then per shorter side)
    res = {}
    for t in (3, 4, 5, 6, 7, 8):
        p = dict(best, merge_wave_min_log2=[t] * 16)
        res[t] = total(p, "merge")
        print("merge_wave_min_log2=%d: %.3f ms" % (t, res[t]), flush=True)
    t_best = min(res, key=res.get)
    best["merge_wave_min_log2"] = [t_best] * 16
    # 2. bitmap span capacity
    res = {}
    for s in (16, 17, 18, 19):
        p = dict(best, bitmap_span_log2=s)
        res[s] = total(p, "bitmap")
        print("bitmap_span_log2=%d: %.3f ms" % (s, res[s]), flush=True)
    best["bitmap_span_log2"] = min(res, key=res.get)
    # 3. per-wave bitmap window
    res = {}
    for w in (13, 14, 15, 16):
        p = dict(best, wave_window_log2=w)
        res[w] = total(p, "bitmap")
        print("wave_window_log2=%d: %.3f ms" % (w, res[w]), flush=True)
    best["wave_window_log2"] = min(res, key=res.get)
    # 4. hash table cap
    res = {}
    for h in (9, 10, 11):
        p = dict(best, hash_max_log2=h)
        res[h] = total(p, "hash")
        print("hash_max_log2=%d: %.3f ms" % (h, res[h]), flush=True)
    best["hash_max_log2"] = min(res, key=res.get)
    text = tune_text(best)
    with open(os.path.join(args.out, "tune.txt"), "w") as f:
        f.write(text + "# tuned on gfx950 by signed_triangle_counting_amd/autotune.py\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
