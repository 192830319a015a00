"""stc_count_prepared_static (the counting pass in the library's static device
array, no device allocation) against stc_count_prepared, through the C ABI:
bit-exact in every dispatch mode and for the four classes, on two graphs in a
row (the second smaller, so it runs over what the first left in the array),
and a graph too large for the array is rejected with STC_E_TOO_LARGE without
spoiling the context."""
import pytest

import signed_triangle_counting_amd as stc

pytestmark = pytest.mark.gpu

MODES = ["learned", "fixed", "merge", "hash", "bitmap", "pull"]


def say(*a):
    print("[case]", *a, flush=True)


def _same_counts(ctx, name):
    for mode in MODES:
        say("static", name, mode)
        want = ctx.count_prepared(dispatch=mode)
        got = ctx.count_prepared_static(dispatch=mode)
        assert got == want, (name, mode)
        got2 = ctx.count_prepared_static(dispatch=mode, repeat=2)
        assert got2 == want, (name, mode, "repeat")
    say("static", name, "classes")
    ctx.count_prepared(classes=True)
    want = ctx.classes()
    ctx.count_prepared_static(classes=True)
    assert ctx.classes() == want, name


def test_static_scratch_count_equals_allocating_count():
    with stc.Counter(0) as ctx:
        assert ctx.model_status == 0, ctx.last_error()
        big = stc.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.3, 4040)
        ctx.prepare(big)
        _same_counts(ctx, "rmat16")
        small = stc.gen_workload(1)  # ER n=10,000, 100,000 lines
        ctx.prepare(small)
        _same_counts(ctx, "er10k")
        grid = stc.gen_workload(2)  # 960 x 960 grid, 2.3 M oriented edges
        ctx.prepare(grid)
        say("static grid")
        assert ctx.count_prepared_static() == ctx.count_prepared()


def test_static_scratch_rejects_a_graph_that_does_not_fit():
    with stc.Counter(0) as ctx:
        g = stc.gen_rmat(20, 16, 0.57, 0.19, 0.19, 0.3, 4041)
        ctx.prepare(g)
        say("static too large")
        with pytest.raises(stc.StcError) as ei:
            ctx.count_prepared_static()
        assert ei.value.code == stc.STC_E_TOO_LARGE
        want = ctx.count_prepared()
        small = stc.gen_workload(1)
        ctx.prepare(small)
        say("static after rejection")
        assert ctx.count_prepared_static() == ctx.count_prepared()
        assert want[0] > 0
