"""The learned dispatcher's FP32 kernels and its fallback, through the C ABI.

  * The device training kernel (k_train_mlp, via stc_train_dispatcher) against
    an fp32 PyTorch restatement of the same loss with autograd and
    torch.optim.Adam, from the same normalised samples and initial weights.
  * The device inference fused into binning against a PyTorch forward of the
    committed model on the host's bucket features: the strategy table the
    device picked is the argmax bucket by bucket.
  * A missing, corrupt or wrongly dimensioned model: STC_E_MODEL, the fixed
    rule in use, counts equal to the reference goldens.
  * The fixed rule's table against its thresholds in the shipped tune.txt.
The reference dispatches with a fixed 4x rule only (GlobalMEM/signed_tc.hxx:
47,65); none of this has a reference counterpart.  PyTorch runs on the CPU in
a child interpreter (tests/torch_mlp_ref.py).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import signed_triangle_counting_amd as stc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1
CODE = {"merge": 1, "hash": 2, "bitmap": 3, "pull": 4}
ARGMAX_CODE = [CODE["merge"], CODE["hash"], CODE["bitmap"], CODE["pull"]]


def say(*a):
    print("[case]", *a, flush=True)


def fnv1a(text):
    h = 1469598103934665603
    for ch in text.encode():
        h ^= ch
        h = (h * 1099511628211) & MASK
    return h


def read_model(path):
    arrs = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts and parts[0] in ("mean", "inv_std", "w1", "b1", "w2", "b2"):
                arrs[parts[0]] = np.array([float(x) for x in parts[1:]], np.float32)
    arrs["w1"] = arrs["w1"].reshape(16, 7)
    arrs["w2"] = arrs["w2"].reshape(4, 16)
    return arrs


def torch_ref(mode, **arrays):
    d = os.environ.get("TMPDIR", "/tmp")
    src = os.path.join(d, "stc_mlp_in_%d.npz" % os.getpid())
    dst = os.path.join(d, "stc_mlp_out_%d.npz" % os.getpid())
    np.savez(src, **arrays)
    r = subprocess.run([sys.executable, os.path.join(HERE, "torch_mlp_ref.py"), mode, src, dst],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with np.load(dst) as z:
        out = {k: z[k] for k in z.files}
    os.remove(src)
    os.remove(dst)
    return out


def host_training_inputs(feat, cost, seed):
    """What stc_train_dispatcher hands the kernel: features normalised by
    their mean and (population) standard deviation, costs scaled to [0, 1]
    per sample, sample weights (hi - lo) / lo clamped at 1 and normalised to
    sum 1, and uniform(+-1/sqrt(fan_in)) initial weights from a splitmix64
    stream (biases zero)."""
    f64 = feat.astype(np.float64)
    mean = f64.mean(axis=0)
    sd = np.sqrt(((f64 - mean) ** 2).mean(axis=0))
    inv = 1.0 / np.maximum(sd, 1e-3)
    mean32, inv32 = mean.astype(np.float32), inv.astype(np.float32)
    x = ((feat - mean32) * inv32).astype(np.float32)
    lo, hi = cost.min(axis=1), cost.max(axis=1)
    r = ((cost - lo[:, None]) / (hi - lo + np.float32(1e-9))[:, None]).astype(np.float32)
    w = np.minimum(np.float32(1), (hi - lo) / np.where(lo > 0, lo, 1)).astype(np.float32)
    w = (w.astype(np.float64) / w.astype(np.float64).sum()).astype(np.float32)
    st = (seed ^ 0x6D6C70) & MASK

    def uni(a):
        nonlocal st
        st = (st + 0x9E3779B97F4A7C15) & MASK
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        z ^= z >> 31
        u = np.float32((z >> 11) * 2.0 ** -53)
        return np.float32(a * (np.float32(2.0) * u - np.float32(1.0)))

    a1 = np.float32(1.0) / np.sqrt(np.float32(7.0))
    a2 = np.float32(1.0) / np.sqrt(np.float32(16.0))
    params = np.zeros(16 * 7 + 16 + 4 * 16 + 4, np.float32)
    for i in range(16 * 7):
        params[i] = uni(a1)
    for i in range(4 * 16):
        params[16 * 7 + 16 + i] = uni(a2)
    return mean32, inv32, x, r, w, params


@pytest.mark.parametrize("nsamp,iters", [(48, 1), (48, 25), (200, 25)])
def test_training_kernel_matches_torch_adam(ctx, tmp_path, nsamp, iters):
    """k_train_mlp's parameters after `iters` full-batch Adam steps equal
    PyTorch's (autograd + torch.optim.Adam, fp32), within 2e-5 absolute +
    1e-4 relative: the two sum the same terms in different orders."""
    rng = np.random.default_rng(nsamp * 100 + iters)
    feat = np.column_stack([rng.integers(0, 16, nsamp), rng.integers(0, 16, nsamp),
                            rng.normal(0, 4, nsamp), rng.uniform(0, 20, nsamp),
                            rng.uniform(5, 25, nsamp), rng.uniform(5, 28, nsamp),
                            rng.uniform(0, 12, nsamp)]).astype(np.float32)
    cost = rng.uniform(0.5, 3.0, (nsamp, 4)).astype(np.float32)
    seed, lr = 7, 0.01
    path = str(tmp_path / "dispatch_mlp.txt")
    say("train", nsamp, iters)
    ctx.train_dispatcher(feat, cost, path, iters=iters, lr=lr, seed=seed)
    dev = read_model(path)
    mean, inv, x, r, w, params = host_training_inputs(feat, cost, seed)
    np.testing.assert_allclose(dev["mean"], mean, rtol=1e-6)
    np.testing.assert_allclose(dev["inv_std"], inv, rtol=1e-6)
    ref = torch_ref("train", x=x, r=r, w=w, params=params, iters=iters, lr=lr)
    for k in ("w1", "b1", "w2", "b2"):
        np.testing.assert_allclose(dev[k], ref[k].reshape(dev[k].shape), rtol=1e-4, atol=2e-5,
                                   err_msg=k)
    # the parameters moved: a kernel that skipped its update would fail here
    assert np.abs(dev["w1"].ravel() - params[:112]).max() > 0.5 * lr
    # deterministic for a seed
    path2 = str(tmp_path / "again.txt")
    ctx.train_dispatcher(feat, cost, path2, iters=iters, lr=lr, seed=seed)
    again = read_model(path2)
    for k in ("w1", "b1", "w2", "b2"):
        assert np.array_equal(again[k], dev[k]), k
    ctx.load_models()  # back to the shipped model
    assert ctx.model_status == 0


@pytest.mark.parametrize("case", ["rmat16", "hub", "batch"])
def test_learned_table_is_torch_argmax(ctx, case):
    """The strategy the device picked for every non-empty bucket is the
    argmax of a PyTorch forward of the committed model on stc_bucket_features
    of that bucket (buckets whose two best logits lie within 1e-3 are left
    out: the device takes log2 of the bucket count with the hardware's
    approximate log)."""
    assert ctx.model_status == 0, ctx.last_error()
    if case == "rmat16":
        g, k = stc.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.3, 7001), 1
        ctx.prepare(g)
        ctx.count_prepared(dispatch="learned")
    elif case == "hub":
        g, k = stc.gen_hub(400000, 500000, 8, 1.2, 0.6, 4000, 0.1, 7002), 1
        ctx.prepare(g)
        ctx.count_prepared(dispatch="learned")
    else:
        b = stc.gen_batch(128, n_hi=8192, seed=7003)
        k = b.k
        ctx.prepare_batch(b)
        ctx.count_batch_prepared(dispatch="learned")
    st = ctx.stats()
    assert st["dispatch_used"] == stc.DISPATCH["learned"]
    table = ctx.last_table()
    hist = ctx.bucket_histogram()
    nz = np.nonzero(hist)[0]
    assert (table[hist == 0] == 0).all()
    feats = np.array([stc.bucket_features(int(b_), int(hist[b_]), st["n"], st["m_oriented"], k)
                      for b_ in nz], np.float32)
    model = read_model(os.path.join(stc.MODEL_DIR, "dispatch_mlp.txt"))
    logits = torch_ref("forward", f=feats, **model)["logits"]
    top2 = np.sort(logits, axis=1)[:, -2:]
    clear = (top2[:, 1] - top2[:, 0]) >= 1e-3
    want = np.array(ARGMAX_CODE)[np.argmax(logits, axis=1)]
    say("argmax", case, len(nz), int(clear.sum()), np.bincount(want, minlength=5).tolist())
    assert clear.sum() >= 0.9 * len(nz)
    assert (table[nz][clear] == want[clear]).all(), (nz[clear][table[nz][clear] != want[clear]])


def _fixed_table(tune, hist):
    """The fixed rule of the shipped tune.txt, bucket by bucket (16 switches
    the pull, skew and hub branches off)."""
    t = np.zeros(256, np.uint8)
    for b in np.nonzero(hist)[0]:
        la, lb = int(b) // 16, int(b) % 16
        if max(la, lb) <= tune["fixed_merge_log2"]:
            s = "merge"
        elif lb - la >= tune["fixed_pull_log2"]:
            s = "pull"
        elif abs(la - lb) >= tune["fixed_skew_log2"]:
            s = "merge"  # the wave merge
        elif la >= tune["fixed_hub_log2"]:
            s = "bitmap"
        else:
            s = "merge"
        t[b] = CODE[s]
    return t


def _read_tune(path):
    out = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 2 and p[0].startswith("fixed_"):
                out[p[0]] = int(p[1])
    return out


def test_fixed_rule_table(ctx):
    g = stc.gen_rmat(16, 16, 0.57, 0.19, 0.19, 0.3, 7004)
    ctx.prepare(g)
    ctx.count_prepared(dispatch="fixed")
    table = ctx.last_table()
    hist = ctx.bucket_histogram()
    tune = _read_tune(os.path.join(stc.MODEL_DIR, "tune.txt"))
    assert set(tune) == {"fixed_merge_log2", "fixed_pull_log2", "fixed_skew_log2",
                         "fixed_hub_log2"}
    assert (table == _fixed_table(tune, hist)).all()


def _model_dir(tmp_path, name, text):
    d = tmp_path / name
    d.mkdir()
    shutil.copy(os.path.join(stc.MODEL_DIR, "tune.txt"), str(d / "tune.txt"))
    if text is not None:
        (d / "dispatch_mlp.txt").write_text(text)
    return str(d)


def test_model_fallback_to_fixed_rule(ctx, tmp_path, goldens):
    """A missing model file, a checksum mismatch, and a well-formed file whose
    dims line does not match the kernel's 7-16-4 network each leave the
    context on the fixed rule: stc_model_status is STC_E_MODEL, a learned
    count reports the fixed dispatch, uses the fixed rule's table, and
    returns the reference goldens."""
    good = open(os.path.join(stc.MODEL_DIR, "dispatch_mlp.txt")).read()
    lines = good.splitlines(keepends=True)
    body = "".join(l for l in lines if not l.startswith("check "))
    corrupt = body.replace("b2 ", "b2 1", 1) + [l for l in lines if l.startswith("check ")][0]
    wrong_body = body.replace("dims 7 16 4", "dims 7 16 3", 1)
    wrong_dims = wrong_body + "check %016x\n" % fnv1a(wrong_body)
    assert fnv1a(body) == int([l for l in lines if l.startswith("check ")][0].split()[1], 16)
    g = stc.gen_workload(1)
    gold = goldens["workloads"]["1"]
    want = (gold["balanced"], gold["unbalanced"])
    ctx.prepare(g)
    ctx.count_prepared(dispatch="fixed")
    fixed_table = ctx.last_table()
    for name, text in (("missing", None), ("corrupt", corrupt), ("dims", wrong_dims)):
        say("fallback", name)
        with stc.Counter(0, model_dir=_model_dir(tmp_path, name, text)) as c2:
            assert c2.model_status == stc.STC_E_MODEL, name
            assert "fixed rule" in c2.last_error()
            assert c2.count(g, dispatch="learned") == want, name
            assert c2.stats()["dispatch_used"] == stc.DISPATCH["fixed"], name
            assert (c2.last_table() == fixed_table).all(), name
            assert c2.count(g, dispatch="merge") == want, name
    # reloading: a bad directory drops the model, the shipped one restores it
    assert ctx.load_models(_model_dir(tmp_path, "reload", corrupt)) == stc.STC_E_MODEL
    assert ctx.count_prepared(dispatch="learned") == want
    assert ctx.stats()["dispatch_used"] == stc.DISPATCH["fixed"]
    assert ctx.load_models() == 0
    ctx.count_prepared(dispatch="learned")
    assert ctx.stats()["dispatch_used"] == stc.DISPATCH["learned"]


def test_error_paths(ctx):
    """Size limits and bad arguments return their codes without touching the
    device arrays (the pointers below are never read)."""
    import ctypes as C
    L = stc.lib()
    out = stc.Counts()
    opts = stc.Opts(0, 1, 0)
    big_n = stc.Edges((1 << 31) - (1 << 20) + 1, 0, None, None, None)
    assert L.stc_count(ctx._h, C.byref(big_n), C.byref(opts), C.byref(out)) == stc.STC_E_TOO_LARGE
    fake = C.c_uint32(0)
    big_m = stc.Edges(10, (1 << 32) - (1 << 20) + 1, C.addressof(fake), C.addressof(fake),
                      C.addressof(fake))
    assert L.stc_count(ctx._h, C.byref(big_m), C.byref(opts), C.byref(out)) == stc.STC_E_TOO_LARGE
    null_arrays = stc.Edges(10, 5, None, None, None)
    assert L.stc_count(ctx._h, C.byref(null_arrays), C.byref(opts), C.byref(out)) == stc.STC_E_ARG
    bad_mode = stc.Opts(17, 1, 0)
    g = stc.Graph(u=[0, 1, 0], v=[1, 2, 2], sign=[1, 1, 1], n=3)
    e = g.edges()
    assert L.stc_count(ctx._h, C.byref(e), C.byref(bad_mode), C.byref(out)) == stc.STC_E_ARG
    batch = (stc.Edges * 2)(big_n, big_n)
    assert L.stc_count_batch(ctx._h, batch, 2, C.byref(opts), out) == stc.STC_E_TOO_LARGE
    assert L.stc_set_tune(ctx._h, b"stc-tune v1\nno_such_key 3\n") == stc.STC_E_ARG
    assert ctx.count(g) == (1, 0)
