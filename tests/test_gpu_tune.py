"""Counts under every tune value the autotuner sweeps (autotune.CANDIDATES),
one parameter at a time with the others at their defaults: bitmap and hash
windows, the merge threshold, pull chunk sizes, wave pop size, the batched
kernel's merge-lane item and the fixed rule's thresholds.  The LDS windows,
caps and item sizes change the kernels' control flow, so each value is
counted in every dispatch mode against the oracle (the reference's merge
path restated), single graphs and a batch."""
import numpy as np
import pytest

import oracle
import signed_triangle_counting_amd as stc
from signed_triangle_counting_amd import autotune as at

pytestmark = pytest.mark.gpu

MODES = ["learned", "fixed", "merge", "hash", "bitmap", "pull"]


def say(*a):
    print("[case]", *a, flush=True)


def _values():
    for key, vals in at.CANDIDATES.items():
        for v in vals:
            p = {k: (list(x) if isinstance(x, list) else x) for k, x in at.DEFAULT.items()}
            p[key] = [v] * 16 if key == "merge_wave_min_log2" else v
            yield key, v, at.tune_text(p)


@pytest.mark.parametrize("case", ["workload1", "rmat16", "hub"])
def test_every_tune_candidate_counts_exactly(ctx, case):
    if case == "workload1":
        g = stc.gen_workload(1)
    elif case == "rmat16":
        g = stc.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.3, 8080)
    else:
        g = stc.gen_hub(200000, 300000, 8, 1.2, 0.6, 2000, 0.1, 8081)
    want = oracle.Csr(g.n, g.u, g.v, g.sign).merge()
    ctx.prepare(g)
    try:
        for key, v, text in _values():
            ctx.set_tune(text)
            for mode in MODES:
                say("tune", case, key, v, mode)
                assert ctx.count_prepared(dispatch=mode) == want, (key, v, mode)
    finally:
        assert ctx.load_models() == 0  # the shipped tune table and model again


def test_every_tune_candidate_counts_batch_exactly(ctx):
    b = stc.gen_batch(48, n_hi=4096, seed=8082)
    graphs = [b.graph(i) for i in range(b.k)]
    want = [oracle.Csr(g.n, g.u, g.v, g.sign).merge() for g in graphs]
    ctx.prepare_batch(b)
    try:
        for key, v, text in _values():
            ctx.set_tune(text)
            for mode in ("learned", "fixed", "merge", "pull"):
                say("tune batch", key, v, mode)
                assert ctx.count_batch_prepared(dispatch=mode) == want, (key, v, mode)
    finally:
        assert ctx.load_models() == 0


def test_long_wave_pull_lists_four_classes_and_per_vertex(ctx):
    """Pull targets of up to 512 entries counted by waves (pull_wave_max_log2
    9: the list staged 64 entries at a time, its in-edges streamed as rows):
    two-class, four-class and per-vertex counts against the oracle."""
    g = stc.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.3, 8083)
    o = oracle.Csr(g.n, g.u, g.v, g.sign)
    want2, want4, want_pv = o.merge(), o.classes(), o.classes_per_vertex()
    ctx.prepare(g)
    p = {k: (list(x) if isinstance(x, list) else x) for k, x in at.DEFAULT.items()}
    p["pull_wave_max_log2"] = 9
    try:
        ctx.set_tune(at.tune_text(p))
        for mode in ("learned", "pull"):
            say("long wave pull", mode)
            assert ctx.count_prepared(dispatch=mode) == want2, mode
            ctx.count_prepared(dispatch=mode, classes=True)
            assert ctx.classes() == want4, mode
            got = ctx.count_per_vertex(dispatch=mode)
            assert (got == want_pv).all(), mode
    finally:
        assert ctx.load_models() == 0
