"""GPU parity: the HIP path, called through the C ABI, against the oracle and
against the reference programs' golden counts.  Integer work: bit-exact.
"""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
import signed_triangle_counting_amd as stc

pytestmark = pytest.mark.gpu

MODES = ["learned", "fixed", "merge", "hash", "bitmap", "pull"]


def say(*a):
    print("[case]", *a, flush=True)


def oracle_counts(n, u, v, s):
    return oracle.Csr(n, u, v, s).merge()


def random_graph(rng, n, m, p_neg=0.4, dup=0.0, loops=0.0):
    u = rng.integers(0, n, m).astype(np.uint32)
    v = rng.integers(0, n, m).astype(np.uint32)
    s = np.where(rng.random(m) < p_neg, -1, 1).astype(np.int8)
    if dup > 0:
        k = int(m * dup)
        idx = rng.integers(0, m, k)
        u2, v2 = v[idx], u[idx]
        u, v = np.concatenate([u, u2]), np.concatenate([v, v2])
        s = np.concatenate([s, np.where(rng.random(k) < 0.5, -1, 1).astype(np.int8)])
    if loops > 0:
        k = int(m * loops)
        w = rng.integers(0, n, k).astype(np.uint32)
        u, v = np.concatenate([u, w]), np.concatenate([v, w])
        s = np.concatenate([s, np.ones(k, np.int8)])
    return u, v, s


def test_small_golden_inputs(ctx, goldens):
    for name, r in goldens["small"].items():
        text = r.get("text", "")
        g = stc.parse_graph(text)
        for mode in MODES:
            say("small", name, mode)
            assert ctx.count(g, dispatch=mode) == (r["balanced"], r["unbalanced"]), (name, mode)
        if "classes" in r:
            ctx.count(g, classes=True)
            assert ctx.classes() == r["classes"], name


@pytest.mark.parametrize("wid", [1, 2, 3, 4])
def test_workloads_vs_reference_goldens(ctx, goldens, wid):
    r = goldens["workloads"][str(wid)]
    say("workload", wid, "generate")
    g = stc.gen_workload(wid)
    say("workload", wid, "prepare n=%d lines=%d" % (g.n, g.m))
    ctx.prepare(g)
    st = ctx.stats()
    assert st["m_oriented"] == r["m_oriented"]
    want = (r["balanced"], r["unbalanced"])
    for mode in MODES:
        say("workload", wid, mode)
        assert ctx.count_prepared(dispatch=mode) == want, mode
    say("workload", wid, "classes")
    ctx.count_prepared(dispatch="learned", classes=True)
    assert ctx.classes() == r["classes"]


@pytest.mark.parametrize("seed", range(8))
def test_random_graphs_vs_oracle(ctx, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(2, 3000))
    m = int(rng.integers(1, 40 * n))
    u, v, s = random_graph(rng, n, m, dup=0.2, loops=0.05)
    want = oracle_counts(n, u, v, s)
    g = stc.Graph(u=u, v=v, sign=s, n=n)
    for mode in MODES:
        say("random", seed, n, len(u), mode)
        assert ctx.count(g, dispatch=mode) == want, mode


@pytest.mark.parametrize("scale,ef", [(10, 16), (12, 32), (14, 24), (16, 16)])
def test_rmat_vs_oracle(ctx, scale, ef):
    g = stc.gen_rmat(scale, ef, 0.57, 0.19, 0.19, 0.3, 77 + scale)
    want = oracle_counts(g.n, g.u, g.v, g.sign)
    ctx.prepare(g)
    for mode in MODES:
        say("rmat", scale, ef, mode)
        assert ctx.count_prepared(dispatch=mode) == want, mode


def test_dense_graph_long_lists(ctx):
    """Out-lists beyond one wave, beyond the LDS hash cap and wide spans."""
    rng = np.random.default_rng(5)
    n = 6000
    pairs = np.array(list(itertools.combinations(range(300), 2)), np.uint32)
    u = np.concatenate([pairs[:, 0], rng.integers(0, n, 200000).astype(np.uint32)])
    v = np.concatenate([pairs[:, 1], rng.integers(0, n, 200000).astype(np.uint32)])
    s = np.where(rng.random(len(u)) < 0.35, -1, 1).astype(np.int8)
    want = oracle_counts(n, u, v, s)
    g = stc.Graph(u=u, v=v, sign=s, n=n)
    ctx.prepare(g)
    for mode in MODES:
        say("dense", mode)
        assert ctx.count_prepared(dispatch=mode) == want, mode


@pytest.mark.parametrize("n", [3, 5, 64, 200])
def test_complete_graphs(ctx, n):
    pairs = np.array(list(itertools.combinations(range(n), 2)), np.uint32)
    c3 = math.comb(n, 3)
    for sgn, want in ((1, (c3, 0)), (-1, (0, c3))):
        g = stc.Graph(u=pairs[:, 0], v=pairs[:, 1], sign=np.full(len(pairs), sgn, np.int8), n=n)
        for mode in MODES:
            say("K", n, sgn, mode)
            assert ctx.count(g, dispatch=mode) == want


def test_edge_cases(ctx):
    say("empty")
    assert ctx.count(stc.Graph(u=[], v=[], sign=[], n=0)) == (0, 0)
    assert ctx.count(stc.Graph(u=[], v=[], sign=[], n=10)) == (0, 0)
    say("loops only")
    assert ctx.count(stc.Graph(u=[1, 2, 3], v=[1, 2, 3], sign=[1, -1, 1], n=4)) == (0, 0)
    say("single edge")
    assert ctx.count(stc.Graph(u=[0], v=[1], sign=[-1], n=2)) == (0, 0)
    say("conflicting duplicates: first line wins")
    g = stc.Graph(u=[0, 1, 2, 1, 0], v=[1, 2, 0, 0, 2], sign=[1, 1, -1, -1, 1], n=3)
    assert ctx.count(g) == (0, 1)
    say("id >= n rejected")
    with pytest.raises(stc.StcError) as ei:
        ctx.count(stc.Graph(u=[0, 5], v=[1, 2], sign=[1, 1], n=3))
    assert ei.value.code == stc.STC_E_ARG
    say("context still usable after an error")
    assert ctx.count(stc.Graph(u=[0, 1, 0], v=[1, 2, 2], sign=[1, 1, 1], n=3)) == (1, 0)


@pytest.mark.parametrize("case", ["random", "rmat15"])
def test_classes_vs_oracle(ctx, case):
    """Four-class totals in every mode; the RMAT graph has targets of more than
    64 out-edges, so pull runs workgroup and wave chunks in one launch."""
    if case == "random":
        rng = np.random.default_rng(9)
        u, v, s = random_graph(rng, 2000, 60000, dup=0.1)
        g = stc.Graph(u=u, v=v, sign=s, n=2000)
    else:
        g = stc.gen_rmat(15, 16, 0.57, 0.19, 0.19, 0.3, 77)
    want = oracle.Csr(g.n, g.u, g.v, g.sign).classes()
    for mode in MODES:
        say("classes", case, mode)
        b, un = ctx.count(g, dispatch=mode, classes=True)
        assert ctx.classes() == want
        assert (b, un) == (want["t3"] + want["t1"], want["t2"] + want["t0"])


def _hip():
    """The HIP runtime libstc.so is linked against (already loaded)."""
    import ctypes as C
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def test_repeat_and_device_input(ctx):
    import ctypes as C
    g = stc.gen_workload(1)
    want = oracle_counts(g.n, g.u, g.v, g.sign)
    say("repeat")
    assert ctx.count(g, repeat=5) == want
    st = ctx.stats()
    assert st["ms_count"] > 0 and st["m_oriented"] == 100000
    say("device input")
    hip = _hip()
    ptrs = []
    for arr in (g.u, g.v, g.sign):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), arr.nbytes) == 0
        assert hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host -> device
        ptrs.append(p.value)
    try:
        got = ctx.count_device(g.n, g.m, *ptrs)
    finally:
        for p in ptrs:
            hip.hipFree(p)
    assert got == want


def test_reference_operator_shape(ctx):
    g = stc.gen_workload(1)
    res = stc.signed_tc.result_t()
    ms = stc.signed_tc.run(g, stc.signed_tc.param_t(), res, context=ctx)
    assert ms > 0
    assert (res.total_balanced, res.total_unbalanced) == oracle_counts(g.n, g.u, g.v, g.sign)


def test_cli_output(tmp_path, goldens):
    r = goldens["small"]["k4_mixed"]
    p = tmp_path / "k4.txt"
    p.write_text(r["text"])
    say("cli")
    out = subprocess.run([stc.CLI_PATH, str(p)], stdout=subprocess.PIPE, text=True,
                         timeout=120).stdout
    assert "Balanced   = %d" % r["balanced"] in out
    assert "Unbalanced = %d" % r["unbalanced"] in out
    assert "Total      = %d" % (r["balanced"] + r["unbalanced"]) in out
    assert "[Triangle Counting Execution] took" in out
    bad = subprocess.run([stc.CLI_PATH], stdout=subprocess.PIPE, text=True, timeout=60)
    assert bad.returncode == 1 and "Usage" in bad.stdout


# ---- batched counting (one persistent launch, captured in a hipGraph) -------
def test_batch_small_vs_oracle(ctx):
    rng = np.random.default_rng(77)
    graphs = []
    for i in range(40):
        n = int(rng.integers(1, 400))
        m = int(rng.integers(0, 6 * n + 1))
        if n < 2:
            m = 0
        u, v, s = random_graph(rng, max(n, 1), m, dup=0.2, loops=0.1) if m else (
            np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int8))
        graphs.append(stc.Graph(u=u, v=v, sign=s, n=n))
    graphs.append(stc.gen_rmat(13, 16, 0.57, 0.19, 0.19, 0.3, 5))
    graphs.append(stc.Graph(u=[], v=[], sign=[], n=0))
    want = [oracle_counts(g.n, g.u, g.v, g.sign) for g in graphs]
    for mode in MODES:
        say("batch small", mode)
        got = ctx.count_batch(graphs, dispatch=mode, repeat=3)
        assert got == want, mode


@pytest.mark.parametrize("seed", [11, 12])
def test_batch_mixed_strategy_tables(ctx, seed):
    """The batched kernel with tables mixing merge, hash, bitmap and pull
    bucket by bucket: every item kind of both phases (workgroup items, then
    wave items) in one launch, sharing the LDS region, counts every triangle
    once.  The batch holds small graphs, an RMAT graph with long lists and a
    hub graph whose targets span many pull chunks."""
    rng = np.random.default_rng(seed)
    b = stc.gen_batch(24, n_hi=20000, d_hi=48.0, seed=seed)
    graphs = [b.graph(i) for i in range(b.k)]
    graphs.append(stc.gen_rmat(13, 16, 0.57, 0.19, 0.19, 0.3, seed))
    graphs.append(stc.gen_hub(100000, 400000, 4, 1.2, 0.6, 1000, 0.3, seed))
    want = [oracle_counts(g.n, g.u, g.v, g.sign) for g in graphs]
    total = (sum(w[0] for w in want), sum(w[1] for w in want))
    ctx.prepare_batch(graphs)
    for _ in range(3):
        t = rng.choice([stc.TABLE[s] for s in ("merge", "hash", "bitmap", "pull")],
                       256).astype(np.uint8)
        got = ctx.count_with_table(t)
        say("batch mixed", seed, got)
        assert got == total
    say("batch mixed per graph")
    assert ctx.count_batch(graphs, dispatch="pull") == want


def test_batch_rejects_out_of_range_id(ctx):
    graphs = [stc.Graph(u=[0, 1], v=[1, 2], sign=[1, 1], n=3),
              stc.Graph(u=[0, 3], v=[1, 2], sign=[1, 1], n=3)]
    with pytest.raises(stc.StcError) as ei:
        ctx.count_batch(graphs)
    assert ei.value.code == stc.STC_E_ARG


def test_batch_workload_vs_reference_goldens(ctx, goldens):
    """4,096 graphs, per-graph counts against the oracle fixture that the
    reference programs pin on a sample (tests/golden/make_goldens.py)."""
    say("batch workload generate")
    b = stc.gen_workload_batch(4096)
    per = goldens["batch"]["per_graph"]
    assert b.total_m == goldens["batch"]["total_m"]
    for mode in ("learned", "fixed"):
        say("batch workload", mode)
        got = ctx.count_batch(b, dispatch=mode, repeat=2)
        bad = [i for i in range(b.k) if list(got[i]) != per[i][1:3]]
        assert not bad, (mode, bad[:10])


@pytest.mark.parametrize("scale,ef", [(16, 16), (18, 24)])
def test_partial_tables_strategy_invariant(ctx, scale, ef):
    """Counting any subset of buckets gives the same partial counts under
    every strategy (each triangle is counted by exactly one oriented edge),
    per bucket group of the fixed rule and for single heavy buckets."""
    g = stc.gen_rmat(scale, ef, 0.57, 0.19, 0.19, 0.3, 300 + scale)
    ctx.prepare(g)
    hist = ctx.bucket_histogram()
    nz = [int(b) for b in np.nonzero(hist)[0]]
    groups = {}
    for b in nz:
        la, lb = b // 16, b % 16
        key = "hash" if abs(la - lb) >= 2 else ("bitmap" if la >= 9 else "merge")
        groups.setdefault(key, []).append(b)
    heavy = sorted(nz, key=lambda b: -int(hist[b]))[:6]
    sets = list(groups.values()) + [[b] for b in heavy]
    total = (0, 0)
    for bs in sets:
        res = {}
        for s in ("merge", "hash", "bitmap", "pull"):
            t = np.zeros(256, np.uint8)
            t[bs] = stc.TABLE[s]
            res[s] = ctx.count_with_table(t)
        say("partial", scale, bs[:4], res)
        assert len(set(res.values())) == 1, (bs, res)
    for bs in groups.values():
        t = np.zeros(256, np.uint8)
        t[bs] = stc.TABLE["merge"]
        c = ctx.count_with_table(t)
        total = (total[0] + c[0], total[1] + c[1])
    assert total == oracle_counts(g.n, g.u, g.v, g.sign)


def test_staged_upload_partial_chunk(ctx):
    """Host edge lists of 2^20 lines and more go through the pinned staging
    buffers in chunks of 2^23 lines; a graph one partial chunk past a chunk
    boundary, and a batch whose graphs straddle chunk boundaries, count as
    the oracle says."""
    rng = np.random.default_rng(77)
    n = 3_000_000
    m = (1 << 23) + 4321
    u, v, s = random_graph(rng, n, m, p_neg=0.3)
    g = stc.Graph(u=u, v=v, sign=s, n=n)
    say("staged upload", m)
    assert ctx.count(g) == oracle_counts(n, u, v, s)
    assert ctx.stats()["m_in"] == m
    say("staged batch upload")
    graphs = []
    for i in range(9):
        mi = int(rng.integers(900_000, 1_100_000))
        ui, vi, si = random_graph(rng, 200_000, mi, p_neg=0.4)
        graphs.append(stc.Graph(u=ui, v=vi, sign=si, n=200_000))
    got = ctx.count_batch(graphs)
    for gi, c in zip(graphs, got):
        assert tuple(c) == oracle_counts(gi.n, gi.u, gi.v, gi.sign)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mixed_strategy_tables(ctx, seed):
    """A table mixing merge, hash, bitmap and pull bucket by bucket counts
    every triangle exactly once: binning and each kernel's edge test agree
    on every edge, also where push and pull edges share a vertex and where a
    target's in-edges span many pull chunks (hub graph)."""
    rng = np.random.default_rng(seed)
    graphs = [stc.gen_rmat(17, 16, 0.57, 0.19, 0.19, 0.3, 500 + seed),
              stc.gen_hub(200000, 1000000, 8, 1.2, 0.6, 2000, 0.3, 600 + seed)]
    for g in graphs:
        ctx.prepare(g)
        want = oracle_counts(g.n, g.u, g.v, g.sign)
        for _ in range(3):
            t = rng.choice([stc.TABLE[s] for s in ("merge", "hash", "bitmap", "pull")],
                           256).astype(np.uint8)
            got = ctx.count_with_table(t)
            say("mixed", seed, g.n, got)
            assert got == want


def test_kernel_statistics_account_for_every_edge(ctx):
    """Per-kernel edge counts (from the counting kernels) plus skipped edges
    (from binning) add up to the oriented edge count in every mode, the
    per-kernel merge-model bytes add up to the B_alg of the counted edges, and
    the strategy bytes of forced merge, bitmap and pull equal the byte model
    restated from the oracle's CSR mapped into rank space."""
    g = stc.gen_rmat(17, 16, 0.57, 0.19, 0.19, 0.25, 4242)
    o = oracle.Csr(g.n, g.u, g.v, g.sign)
    off, E, _, R = o.arrays()
    n = o.n
    fd = np.diff(off.astype(np.int64))
    # oriented edges in rank space, each list ascending by rank
    key = np.sort(R[np.repeat(np.arange(n), fd)].astype(np.int64) * n + R[E])
    srcr, dstr = key // n, key % n
    fdr = np.zeros(n, np.int64)
    fdr[R] = fd
    offr = np.concatenate([[0], np.cumsum(fdr)])
    da = fdr[srcr] - (np.arange(len(key)) - offr[srcr]) - 1  # entries after the target
    counted = (da > 0) & (fdr[dstr] > 0)
    words = int((fdr[srcr] + fdr[dstr])[counted].sum())
    db = fdr[dstr]
    # every list fits the LDS structures (no demotion): the byte model below
    # needs no fit test
    assert fdr.max() <= 1024
    # strategy bytes (the statistics' algorithmic bytes, restated):
    # merge: 24 + 4 (da + db) per counted edge
    want_merge = int((24 + 4 * (da + db))[counted].sum())
    # push bitmap: 4 db per counted edge; per source with a counted edge
    # 20 + 16 d+(u)
    srcs = np.unique(srcr[counted])
    want_bitmap = int((4 * db)[counted].sum()) + int((20 + 16 * fdr[srcs]).sum())
    # pull: 4 da per counted edge; per target with a counted in-edge, chunks
    # of the shipped tune's workgroup (d+ > 64) or wave chunk size in in-edges,
    # each 28 + 4 d+(v), and 8 bytes per in-edge record
    from signed_triangle_counting_amd import autotune
    tune = autotune.read_tune(os.path.join(stc.MODEL_DIR, "tune.txt"))
    tg = np.unique(dstr[counted])
    indeg = np.bincount(dstr, minlength=n)[tg]
    chunk = np.where(fdr[tg] > 64, 1 << tune["pull_block_chunk_log2"],
                     1 << tune["pull_wave_chunk_log2"])
    nchunks = (indeg + chunk - 1) // chunk
    want_pull = (int((4 * da)[counted].sum()) + int((nchunks * (28 + 4 * fdr[tg])).sum()) +
                 int((8 * indeg).sum()))
    want_alg = {"merge": want_merge, "bitmap": want_bitmap, "pull": want_pull}
    ctx.prepare(g)
    for mode in MODES:
        ctx.count_prepared(dispatch=mode)
        st = ctx.stats()
        say("stats", mode, st["edges_merge"], st["edges_hash"], st["edges_bitmap"],
            st["edges_pull"], st["edges_skipped"], sum(st["alg_bytes_kernel"]))
        assert (st["edges_merge"] + st["edges_hash"] + st["edges_bitmap"] + st["edges_pull"] +
                st["edges_skipped"]) == st["m_oriented"] == o.m
        assert st["edges_skipped"] == int((~counted).sum())
        assert sum(st["bytes_kernel"]) == 4 * words
        if mode in want_alg:
            assert sum(st["alg_bytes_kernel"]) == want_alg[mode], mode


def test_batch_two_phase_and_training_hooks(ctx):
    """stc_prepare_batch + stc_count_batch_prepared equals stc_count_batch in
    every mode; the bucket histogram and table-driven counts of a prepared
    batch cover all its edges and sum to the per-graph totals."""
    b = stc.gen_batch(64, n_hi=4096, seed=901)
    want = ctx.count_batch(b, dispatch="merge")
    ctx.prepare_batch(b)
    for mode in MODES:
        say("batch two-phase", mode)
        assert ctx.count_batch_prepared(dispatch=mode, repeat=2) == want, mode
    total = (sum(x[0] for x in want), sum(x[1] for x in want))
    hist = ctx.bucket_histogram()
    st = ctx.stats()
    assert int(hist.sum()) + st["edges_skipped"] == st["m_oriented"]
    nz = np.nonzero(hist)[0]
    for s in ("merge", "hash", "bitmap", "pull"):
        t = np.zeros(256, np.uint8)
        t[nz] = stc.TABLE[s]
        say("batch table", s)
        assert ctx.count_with_table(t) == total, s
    # prepared batch stays countable after the hooks
    assert ctx.count_batch_prepared(dispatch="learned") == want


@pytest.mark.parametrize("case", ["random", "rmat14", "rmat16", "workload1", "workload2"])
def test_per_vertex_counts_vs_oracle(ctx, case):
    """stc_count_per_vertex against the oracle's per-vertex four-class array
    (the reference's cnt[4 v + k], tc_serial.cpp:141-158), every mode; the
    rows sum to the four-class totals."""
    if case == "random":
        rng = np.random.default_rng(515)
        n = 3000
        u, v, s = random_graph(rng, n, 40000, dup=0.1, loops=0.05)
        g = stc.Graph(u=u, v=v, sign=s, n=n)
    elif case.startswith("rmat"):
        g = stc.gen_rmat(int(case[4:]), 16, 0.57, 0.19, 0.19, 0.3, 61)
    else:
        g = stc.gen_workload(int(case[-1]))
    want = oracle.Csr(g.n, g.u, g.v, g.sign).classes_per_vertex()
    ctx.prepare(g)
    for mode in MODES:
        say("per-vertex", case, mode)
        got = ctx.count_per_vertex(dispatch=mode)
        assert got.shape == want.shape
        assert (got == want).all(), mode
        tot = ctx.classes()
        assert got.sum(axis=0).tolist() == [tot["t0"], tot["t1"], tot["t2"], tot["t3"]]
    assert ctx.count_prepared() == (int(want[:, 3].sum() + want[:, 1].sum()),
                                    int(want[:, 2].sum() + want[:, 0].sum()))


def test_matrix_market_input_counts(ctx, tmp_path):
    """A graph stored as the GPU drivers store it -- MatrixMarket, 1-based,
    both directions, sign = weight > 0 -- counts the same as its text-format
    edge list and as the oracle; the CLI reads the .mtx file directly."""
    g = stc.gen_rmat(13, 16, 0.57, 0.19, 0.19, 0.3, 808)
    keep = g.u != g.v
    u, v, s = g.u[keep], g.v[keep], g.sign[keep]
    want = oracle_counts(g.n, g.u, g.v, g.sign)
    w = np.where(s < 0, -1.5, 2.0)
    lines = ["%%MatrixMarket matrix coordinate real general", "%d %d %d" % (g.n, g.n, 2 * len(u))]
    lines += ["%d %d %g" % (a + 1, b + 1, x) for a, b, x in zip(u, v, w)]
    lines += ["%d %d %g" % (b + 1, a + 1, x) for a, b, x in zip(u, v, w)]
    p = tmp_path / "g.mtx"
    p.write_text("\n".join(lines) + "\n")
    m = stc.load_graph(str(p))
    for mode in MODES:
        say("mtx", mode)
        assert ctx.count(m, dispatch=mode) == want, mode
    exe = os.path.join(os.path.dirname(stc.LIB_PATH), "stc")
    out = subprocess.run([exe, str(p)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Balanced   = %d" % want[0] in out.stdout
    assert "Unbalanced = %d" % want[1] in out.stdout
