"""Pins the CPU oracle (oracle/stc_oracle.c) without trusting its own formulas.

Each check is chosen so a plausible mistake fails: a dropped or flipped sign
term fails the colouring / global-flip / trace identities, a wrong orientation
or rank fails the CSR invariants and brute force, a parse slip fails the
hand-written fixtures.
"""
import itertools
import math

import numpy as np
import pytest

import oracle


# ---- parse (graph_io.h:72-212) ------------------------------------------
PARSE_CASES = [
    # text, expected n, expected [(u, v, sign)]
    ("0 1 1\n1 2 -1\n", 3, [(0, 1, 1), (1, 2, -1)]),
    ("# c\n% c\n3 4\n", 5, [(3, 4, 1)]),                       # comments, missing sign
    ("1\t2\t-1\r\n2 3 +1\r\n", 4, [(1, 2, -1), (2, 3, 1)]),    # tabs, CRLF, '+1'
    ("0 1 0\n1 2 -0\n2 3 7\n", 4, [(0, 1, 1), (1, 2, -1), (2, 3, 1)]),  # sign spellings
    ("5 5 1\n0 1 -1\n", 2, [(0, 1, -1)]),                      # self-loop dropped, n from kept
    (" 7 8 -1\n", 8, [(0, 7, 1)]),                             # leading space: u=0, v=7, '8' sign
    ("0 1 -1 extra words\n", 2, [(0, 1, -1)]),                 # rest of line ignored
    ("4294967297 3 1\n", 4, [(1, 3, 1)]),                      # uint32 wrap-around
    ("\n\n\r\n0 1\n", 2, [(0, 1, 1)]),                         # blank lines
    ("0 1 1", 2, [(0, 1, 1)]),                                 # no final newline
    ("3 3\n", 1, []),                                          # only a self-loop: n = 1
    ("", 0, []),                                               # empty file: n = 0
    ("%%MatrixMarket x\n3 3 2\n1 2 -1\n", 3, [(1, 2, -1)]),    # mm size line M==N is a loop
]


@pytest.mark.parametrize("text,n,edges", PARSE_CASES)
def test_parse_fixtures(text, n, edges):
    pn, u, v, s = oracle.parse(text)
    assert pn == n
    assert list(zip(u.tolist(), v.tolist(), s.tolist())) == edges


# ---- helpers ----------------------------------------------------------------
def brute(n, u, v, s):
    """O(n^3) classes over the simple graph; first line of a pair wins."""
    sign = {}
    for a, b, x in zip(u.tolist(), v.tolist(), s.tolist()):
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        if key not in sign:
            sign[key] = -1 if x < 0 else 1
    adj = {}
    for (a, b), x in sign.items():
        adj.setdefault(a, {})[b] = x
        adj.setdefault(b, {})[a] = x
    t = [0, 0, 0, 0]
    verts = sorted(adj)
    for i, a in enumerate(verts):
        for b in (x for x in adj[a] if x > a):
            for c in (x for x in adj[b] if x > b):
                if c in adj[a]:
                    neg = (adj[a][b] < 0) + (adj[b][c] < 0) + (adj[a][c] < 0)
                    t[3 - neg] += 1
    return {"t0": t[0], "t1": t[1], "t2": t[2], "t3": t[3]}


def brute_per_vertex(n, u, v, s):
    """O(n^3) per-vertex classes: each triangle is credited to its vertex of
    lowest (degree, id) -- the source of its oriented edges (tc_serial.cpp:
    141-158) -- in column 3 - (number of negative edges)."""
    sign = {}
    for a, b, x in zip(u.tolist(), v.tolist(), s.tolist()):
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        if key not in sign:
            sign[key] = -1 if x < 0 else 1
    adj = {}
    for (a, b), x in sign.items():
        adj.setdefault(a, {})[b] = x
        adj.setdefault(b, {})[a] = x
    per = np.zeros((n, 4), np.int64)
    for a, b, c in itertools.combinations(sorted(adj), 3):
        if b in adj[a] and c in adj[a] and c in adj[b]:
            neg = (adj[a][b] < 0) + (adj[b][c] < 0) + (adj[a][c] < 0)
            low = min((a, b, c), key=lambda x: (len(adj[x]), x))
            per[low, 3 - neg] += 1
    return per


def bal_unbal(t):
    return t["t3"] + t["t1"], t["t2"] + t["t0"]


def random_graph(rng, n, m, p_neg=0.4, dup=0.0, loops=0.0):
    u = rng.integers(0, n, m).astype(np.uint32)
    v = rng.integers(0, n, m).astype(np.uint32)
    s = np.where(rng.random(m) < p_neg, -1, 1).astype(np.int8)
    if dup > 0:
        k = int(m * dup)
        idx = rng.integers(0, m, k)
        u = np.concatenate([u, v[idx]])          # reversed duplicates,
        v = np.concatenate([v, u[idx]])          # possibly conflicting signs
        s = np.concatenate([s, np.where(rng.random(k) < 0.5, -1, 1).astype(np.int8)])
    if loops > 0:
        k = int(m * loops)
        w = rng.integers(0, n, k).astype(np.uint32)
        u, v = np.concatenate([u, w]), np.concatenate([v, w])
        s = np.concatenate([s, np.ones(k, np.int8)])
    return u, v, s


# ---- brute force on tiny graphs ----------------------------------------------
@pytest.mark.parametrize("seed", range(12))
def test_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 30))
    m = int(rng.integers(1, n * (n - 1) // 2 + 5))
    u, v, s = random_graph(rng, n, m, dup=0.3, loops=0.1)
    c = oracle.Csr(n, u, v, s)
    want = brute(n, u, v, s)
    assert c.classes() == want
    assert c.merge() == bal_unbal(want)
    assert c.hash() == bal_unbal(want)
    per = c.classes_per_vertex()
    assert per.sum(axis=0).tolist() == [want["t0"], want["t1"], want["t2"], want["t3"]]
    assert (per.astype(np.int64) == brute_per_vertex(n, u, v, s)).all()


# ---- closed forms --------------------------------------------------------------
@pytest.mark.parametrize("n", [3, 4, 7, 20, 45])
def test_complete_graph(n):
    pairs = np.array(list(itertools.combinations(range(n), 2)), np.uint32)
    u, v = pairs[:, 0], pairs[:, 1]
    c3 = math.comb(n, 3)
    assert oracle.Csr(n, u, v, np.ones(len(u), np.int8)).merge() == (c3, 0)
    assert oracle.Csr(n, u, v, -np.ones(len(u), np.int8)).merge() == (0, c3)
    assert oracle.Csr(n, u, v, -np.ones(len(u), np.int8)).hash() == (0, c3)
    assert oracle.Csr(n, v, u, np.ones(len(u), np.int8)).classes()["t3"] == c3


@pytest.mark.parametrize("seed", range(4))
def test_two_colouring(seed):
    """s_uv = c_u c_v makes every triangle balanced; -c_u c_v makes every one
    unbalanced (three edges, each negated)."""
    rng = np.random.default_rng(100 + seed)
    n, m = 400, 6000
    u, v, _ = random_graph(rng, n, m)
    col = np.where(rng.random(n) < 0.5, -1, 1)
    s = (col[u] * col[v]).astype(np.int8)
    c = oracle.Csr(n, u, v, s)
    b, un = c.merge()
    assert un == 0 and b > 0
    b2, u2 = oracle.Csr(n, u, v, (-s).astype(np.int8)).merge()
    assert (b2, u2) == (0, b)


def test_grid_with_diagonals():
    """Every triangle of a grid + one diagonal per chosen cell lies inside one
    cell: two triangles per diagonal; the classes come from a per-cell count."""
    rng = np.random.default_rng(7)
    R, Cc = 40, 37
    edges, sign = [], {}
    for r in range(R):
        for c in range(Cc):
            i = r * Cc + c
            if c + 1 < Cc:
                edges.append((i, i + 1))
            if r + 1 < R:
                edges.append((i, i + Cc))
    diag = {}
    for r in range(R - 1):
        for c in range(Cc - 1):
            if rng.random() < 0.5:
                i = r * Cc + c
                d = (i, i + Cc + 1) if rng.random() < 0.5 else (i + 1, i + Cc)
                diag[(r, c)] = d
                edges.append(d)
    for e in edges:
        sign[e] = -1 if rng.random() < 0.3 else 1
    t = [0, 0, 0, 0]
    for (r, c), (a, b) in diag.items():
        i = r * Cc + c
        corners = [i, i + 1, i + Cc, i + Cc + 1]
        others = [x for x in corners if x not in (a, b)]
        for x in others:
            def sg(p, q):
                return sign[(p, q)] if (p, q) in sign else sign[(q, p)]
            neg = (sg(a, b) < 0) + (sg(a, x) < 0) + (sg(b, x) < 0)
            t[3 - neg] += 1
    u = np.array([e[0] for e in edges], np.uint32)
    v = np.array([e[1] for e in edges], np.uint32)
    s = np.array([sign[e] for e in edges], np.int8)
    got = oracle.Csr(R * Cc, u, v, s).classes()
    assert sum(got.values()) == 2 * len(diag)
    assert got == {"t0": t[0], "t1": t[1], "t2": t[2], "t3": t[3]}


# ---- trace identities ------------------------------------------------------------
@pytest.mark.parametrize("seed", range(3))
def test_trace_identities(seed):
    """tr(A^3) = 6T and tr(S^3) = 6(B - U) for the +-1 signed adjacency S."""
    sp = pytest.importorskip("scipy.sparse")
    rng = np.random.default_rng(200 + seed)
    n, m = 3000, 60000
    u, v, s = random_graph(rng, n, m, dup=0.1)
    c = oracle.Csr(n, u, v, s)
    off, E, S, R = c.arrays()
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    A = sp.coo_matrix((np.ones(len(E)), (src, E)), shape=(n, n)).tocsr()
    A = A + A.T
    Sg = sp.coo_matrix((S.astype(np.float64), (src, E)), shape=(n, n)).tocsr()
    Sg = Sg + Sg.T
    trA3 = int(round((A @ A).multiply(A).sum()))
    trS3 = int(round((Sg @ Sg).multiply(Sg).sum()))
    b, un = c.merge()
    assert trA3 == 6 * (b + un)
    assert trS3 == 6 * (b - un)


# ---- invariances -----------------------------------------------------------------
def test_invariances():
    rng = np.random.default_rng(11)
    n, m = 2000, 30000
    u, v, s = random_graph(rng, n, m)
    base = oracle.Csr(n, u, v, s)
    b0, u0 = base.merge()
    cl0 = base.classes()
    # relabel vertices
    perm = rng.permutation(n).astype(np.uint32)
    assert oracle.Csr(n, perm[u], perm[v], s).merge() == (b0, u0)
    # permute lines (pairs are first-sign unique only if no duplicates conflict)
    key = np.minimum(u, v).astype(np.uint64) << 32 | np.maximum(u, v)
    _, first = np.unique(key, return_index=True)
    uu, vv, ss = u[first], v[first], s[first]
    p = rng.permutation(len(uu))
    assert oracle.Csr(n, uu[p], vv[p], ss[p]).merge() == (b0, u0)
    # switching at a vertex set leaves both counts unchanged
    flip = np.where(rng.random(n) < 0.3, -1, 1)
    s_sw = (ss * flip[uu] * flip[vv]).astype(np.int8)
    assert oracle.Csr(n, uu, vv, s_sw).merge() == (b0, u0)
    # global sign flip swaps balanced <-> unbalanced and T3<->T0, T1<->T2
    fl = oracle.Csr(n, uu, vv, (-ss).astype(np.int8))
    assert fl.merge() == (u0, b0)
    cf = fl.classes()
    assert (cf["t0"], cf["t1"], cf["t2"], cf["t3"]) == (cl0["t3"], cl0["t2"], cl0["t1"], cl0["t0"])


# ---- forward CSR invariants (graph_io.h:231-402) ------------------------------
@pytest.mark.parametrize("seed", range(3))
def test_csr_invariants(seed):
    rng = np.random.default_rng(300 + seed)
    n, m = 500, 5000
    u, v, s = random_graph(rng, n, m, dup=0.5, loops=0.05)
    c = oracle.Csr(n, u, v, s)
    off, E, S, R = c.arrays()
    keep = u != v
    lo = np.minimum(u[keep], v[keep]).astype(np.int64)
    hi = np.maximum(u[keep], v[keep]).astype(np.int64)
    pairs = np.unique(lo * n + hi)
    assert c.m == len(pairs)
    deg = np.bincount(np.concatenate([pairs // n, pairs % n]), minlength=n)
    # rank_ = position in order of (degree, id), computed independently
    order = np.lexsort((np.arange(n), deg))
    want_rank = np.empty(n, np.int64)
    want_rank[order] = np.arange(n)
    assert np.array_equal(R, want_rank)
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    assert np.all(R[E] > R[src])                     # low -> high orientation
    got = np.unique(np.minimum(src, E).astype(np.int64) * n + np.maximum(src, E))
    assert np.array_equal(got, pairs)                # each unique edge exactly once
    for x in range(n):                                # sorted lists
        seg = E[off[x]:off[x + 1]]
        assert np.all(seg[1:] > seg[:-1])
    # sign of each edge = sign of its first line
    first = {}
    for a, b, x in zip(u.tolist(), v.tolist(), s.tolist()):
        if a != b:
            first.setdefault((min(a, b), max(a, b)), x)
    for a, b, x in zip(src.tolist(), E.tolist(), S.tolist()):
        assert first[(min(a, b), max(a, b))] == x
