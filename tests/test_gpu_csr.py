"""The device's oriented CSR word for word against the oracle's ForwardCSR
(the reference's build_forward_csr, Multicore Implementations/graph_io.h:
221-402, restated): dedupe, (degree, id) orientation and per-list order.  The
device keeps the CSR in rank space, so the oracle's lists are mapped through
its rank array; the transposed index and the per-rank classes the pull
kernels and binning read are checked against the same mapping."""
import numpy as np
import pytest

import oracle
import signed_triangle_counting_amd as stc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def clog2(x):
    """ceil(log2 x) for x >= 1 (0 for x <= 1), elementwise."""
    x = np.asarray(x, np.int64)
    return np.where(x > 1, np.floor(np.log2(np.maximum(x - 1, 1))).astype(np.int64) + 1, 0)


@pytest.mark.parametrize("case", ["workload1", "workload2", "workload4", "rmat16", "random"])
def test_device_csr_equals_oracle_forward_csr(ctx, case):
    if case.startswith("workload"):
        g = stc.gen_workload(int(case[-1]))
    elif case == "rmat16":
        g = stc.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.3, 9191)
    else:
        rng = np.random.default_rng(9192)
        n, m = 5000, 60000
        u = rng.integers(0, n, m).astype(np.uint32)
        v = rng.integers(0, n, m).astype(np.uint32)
        s = np.where(rng.random(m) < 0.4, -1, 1).astype(np.int8)
        # duplicates of the same sign and self-loops, where the reference is defined
        u = np.concatenate([u, v[:5000], u[:300]])
        v = np.concatenate([v, u[:5000], u[:300]])
        s = np.concatenate([s, s[:5000], np.ones(300, np.int8)])
        g = stc.Graph(u=u, v=v, sign=s, n=n)
    say("csr", case, "oracle")
    o = oracle.Csr(g.n, g.u, g.v, g.sign)
    off, E, S, R = o.arrays()
    n, m = o.n, o.m
    R = R.astype(np.int64)
    say("csr", case, "device n=%d m+=%d" % (n, m))
    ctx.prepare(g)
    st = ctx.stats()
    assert (st["n"], st["m_oriented"]) == (n, m)
    a = ctx.csr_arrays()
    # rank -> vertex id is the inverse of the oracle's (degree, id) rank
    assert (a["perm"][R] == np.arange(n)).all()
    fd = np.diff(off.astype(np.int64))
    fdr = np.zeros(n, np.int64)
    fdr[R] = fd
    want_off = np.concatenate([[0], np.cumsum(fdr)])
    assert (a["off"].astype(np.int64) == want_off).all()
    # each list: (rank << 1) | neg of the forward neighbours, ascending by rank
    src_r = np.repeat(R, fd)
    dst_r = R[E.astype(np.int64)]
    order = np.lexsort((dst_r, src_r))
    want_col = (dst_r[order] << 1) | (S[order] < 0)
    assert (a["col"].astype(np.int64) == want_col).all()
    # transposed index: per target rank, its in-edges e ascending, as
    # (e + 1, (|N+(u) after v| << 1) | neg)
    src_e, dst_e = src_r[order], dst_r[order]
    e = np.arange(m, dtype=np.int64)
    suffix = want_off[src_e + 1] - e - 1
    by_t = np.lexsort((e, dst_e))
    want_tin_off = np.concatenate([[0], np.cumsum(np.bincount(dst_e, minlength=n))])
    assert (a["tin_off"].astype(np.int64) == want_tin_off).all()
    want_dcls = np.where(fdr == 0, 0xFF, np.minimum(clog2(fdr), 15))
    want_tin = np.stack([e[by_t] + 1, (suffix[by_t] << 1) | (S[order][by_t] < 0)], axis=1)
    assert (a["tin"].astype(np.int64) == want_tin).all()
    # per-rank out-degree class and suffix-class mask (binning's two halves)
    assert (a["dcls"].astype(np.int64) == want_dcls).all()
    bits = np.where(suffix > 0, np.left_shift(1, np.minimum(clog2(suffix), 15)), 0)
    want_smask = np.zeros(n, np.int64)
    np.bitwise_or.at(want_smask, dst_e, bits)
    assert (a["smask"].astype(np.int64) == want_smask).all()
    # the bucket histogram of the binning pass: per oriented edge with a
    # non-empty suffix and a non-empty N+(v), bucket 16 min(ceil_log2 |suffix|,
    # 15) + min(ceil_log2 d+(v), 15)
    live = (suffix > 0) & (fdr[dst_e] > 0)
    want_bk = np.minimum(clog2(suffix[live]), 15) * 16 + np.minimum(clog2(fdr[dst_e][live]), 15)
    want_hist = np.bincount(want_bk, minlength=256)
    assert (ctx.bucket_histogram().astype(np.int64) == want_hist).all()
    # the merge model's bytes, 4 * sum over oriented edges of (d+(u) + d+(v))
    # + 4 per edge + 8 per vertex, and the largest out-degree
    words = int((fdr[src_r] + fdr[dst_r]).sum())
    assert st["b_alg"] == 4.0 * words + 4.0 * m + 8.0 * (n + 1)
    assert st["max_out_degree"] == (int(fdr.max()) if n else 0)
