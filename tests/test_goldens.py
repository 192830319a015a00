"""The oracle against counts produced by the reference programs themselves.

tests/golden/goldens.json was written by tests/golden/make_goldens.py, which
compiles the reference's OpenMP programs from their sources and runs them on
this package's generated workloads.  Here the oracle must reproduce every one
of those numbers, and the generators must reproduce the same graphs
(checksums), so the same fixtures can pin the GPU path on a machine where the
reference is absent.
"""
import hashlib

import numpy as np
import pytest

import oracle
import signed_triangle_counting_amd as stc


def checksum(u, v, s):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(u, np.uint32).tobytes())
    h.update(np.ascontiguousarray(v, np.uint32).tobytes())
    h.update(np.ascontiguousarray(s, np.int8).tobytes())
    return h.hexdigest()[:32]


def test_reference_programs_agree(goldens):
    """merge / hash / split / serial agreed on every input (the reference's
    implicit cross-implementation test, MC/Tests/*.csv)."""
    recs = list(goldens["small"].values()) + list(goldens["workloads"].values())
    assert len(recs) >= 12
    for r in recs:
        vals = {(x["balanced"], x["unbalanced"]) for x in r["reference"].values()}
        assert len(vals) == 1


def test_small_inputs(goldens):
    for name, r in goldens["small"].items():
        if not r.get("text"):
            continue
        n, u, v, s = oracle.parse(r["text"])
        c = oracle.Csr(n, u, v, s)
        assert c.merge() == (r["balanced"], r["unbalanced"]), name
        assert c.hash() == (r["balanced"], r["unbalanced"]), name
        assert c.classes() == r["classes"], name


@pytest.mark.parametrize("wid", [1, 2, 4])
def test_workload_oracle(goldens, wid):
    r = goldens["workloads"][str(wid)]
    g = stc.gen_workload(wid)
    assert checksum(g.u, g.v, g.sign) == r["checksum"]
    c = oracle.Csr(g.n, g.u, g.v, g.sign)
    assert "%016x" % c.graph_hash == r["graph_hash"]
    assert c.m == r["m_oriented"]
    assert c.merge() == (r["balanced"], r["unbalanced"])
    assert c.hash() == (r["balanced"], r["unbalanced"])
    assert c.classes() == r["classes"]


@pytest.mark.slow
def test_workload_rmat22_oracle(goldens):
    r = goldens["workloads"]["3"]
    g = stc.gen_workload(3)
    assert checksum(g.u, g.v, g.sign) == r["checksum"]
    c = oracle.Csr(g.n, g.u, g.v, g.sign)
    assert c.m == r["m_oriented"]
    assert c.merge() == (r["balanced"], r["unbalanced"])


def test_batch_fixture(goldens):
    b = goldens["batch"]
    per = b["per_graph"]
    assert len(per) == b["k"] == 4096
    for gi, (bal, unb) in b["reference_subset"].items():
        assert (per[int(gi)][1], per[int(gi)][2]) == (bal, unb)
    # regenerate a few graphs and recount with the oracle
    batch = stc.gen_workload_batch(4096)
    assert batch.total_m == b["total_m"]
    for gi in (0, 1, 2, 3, 1001, 4095):
        g = batch.graph(gi)
        c = oracle.Csr(g.n, g.u, g.v, g.sign)
        assert [c.m, *c.merge()] == per[gi][:3]
