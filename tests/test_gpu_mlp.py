"""MLP warm start: GPU FP32 forward / backward kernels vs the FP64 oracle (itself pinned by
finite differences), BF16 MFMA inference vs FP64, training on Newton-generated solves,
and what the trained network gains over the plain Newton start at the same 1e-10
residual."""
import numpy as np
import pytest

DIN, H, DOUT = 12, 128, 8


def params(seed=0):
    rng = np.random.default_rng(seed)
    P = H * DIN + H + H * H + H + DOUT * H + DOUT
    th = np.empty(P)
    o = 0
    for fo, fi in ((H, DIN), (H, 1), (H, H), (H, 1), (DOUT, H), (DOUT, 1)):
        a = np.sqrt(6.0 / (fo + fi)) if fi > 1 else 0.1
        th[o:o + fo * fi] = rng.uniform(-a, a, fo * fi)
        o += fo * fi
    return th


@pytest.mark.gpu
def test_forward_backward_match_fp64(clb, oracle):
    print("case: mlp grad fp32 vs fp64", flush=True)
    th = params(1)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (300, DIN))
    T = rng.uniform(-1, 1, (300, DOUT))
    lat = clb.Lattice(4, 1, 1)
    lat.mlp_config(DIN, H, DOUT)
    lat.mlp_set_params(th.astype(np.float32))
    g, loss = lat.mlp_grad(X, T)
    go, lo = oracle.mlp_grad(DIN, H, DOUT, th.astype(np.float32).astype(np.float64), X.astype(np.float32), T.astype(np.float32))
    assert abs(loss - lo) / lo < 1e-5
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-4
    Yo = oracle.mlp_forward(DIN, H, DOUT, th.astype(np.float32).astype(np.float64), X.astype(np.float32))
    Y32 = lat.mlp_forward(X, bf16=False)
    Ybf = lat.mlp_forward(X, bf16=True)
    lat.close()
    assert np.max(np.abs(Y32 - Yo)) < 1e-5
    # BF16 operands (8-bit mantissa) with FP32 accumulation
    assert np.max(np.abs(Ybf - Yo)) < 3e-2
    assert np.linalg.norm(Ybf - Yo) / np.linalg.norm(Yo) < 1e-2


@pytest.mark.gpu
def test_training_reduces_loss(clb):
    print("case: mlp training", flush=True)
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (8192, DIN)).astype(np.float32)
    W = rng.standard_normal((DIN, DOUT)) / np.sqrt(DIN)
    T = np.tanh(X @ W).astype(np.float32)
    lat = clb.Lattice(4, 1, 1)
    lat.mlp_config(DIN, H, DOUT)
    _, l0 = lat.mlp_grad(X, T)
    l1 = lat.mlp_train(X, T, epochs=30, batch=256, lr=2e-3)
    lat.close()
    assert l1 < 0.05 * l0


@pytest.mark.gpu
def test_warm_start_gain(clb, oracle):
    """Train on plain-Newton solves (log10 totals + 4 jittered logK in, log activities out),
    then solve a fresh batch from the network's guess: same 1e-10 judge, fewer iterations,
    and out-of-box inputs fall back to the reference start."""
    print("case: mlp warm start 8/24", flush=True)
    S, lk = clb.gen_tableau(8, 24, seed=3)
    var = [8, 9, 10, 11]
    rng = np.random.default_rng(5)

    def batch(n, seed):
        r = np.random.default_rng(seed)
        T = 10 ** r.uniform(-6, -2, (n, 8))
        L = lk[var][None, :] + r.uniform(-1, 1, (n, 4))
        return T, L

    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, list(range(8)), tol=1e-10, max_iter=200)
    Ttr, Ltr = batch(200000, 6)
    lat.batch_load(Ttr, var, Ltr)
    lat.batch_solve(clb.WARM_REF)
    _, it_tr, _, cv_tr = lat.batch_fetch()
    assert cv_tr.all()
    lat.mlp_config(12, H, 8)
    loss = lat.mlp_train_from_batch(epochs=30, batch=512, lr=2e-3, seed=7)
    assert loss < 0.01
    Tte, Lte = batch(50000, 9)
    Lte[:10, 0] += 1.5                                        # logK outside the training box
    lat.batch_load(Tte, var, Lte)
    lat.batch_solve(clb.WARM_REF)
    x0, it0, r0, cv0 = lat.batch_fetch()
    lat.batch_solve(clb.WARM_MLP)
    x1, it1, r1, cv1 = lat.batch_fetch()
    lat.close()
    assert cv0.all() and cv1.all()
    assert np.all(r1 < 1e-10)
    assert np.max(np.abs(x1 - x0)) < 1e-7
    print(f"mean iterations: reference start {it0.mean():.2f}, MLP start {it1.mean():.2f}")
    assert it1[10:].mean() < 0.5 * it0[10:].mean()
    assert np.array_equal(it1[:10], it0[:10])                 # fallback = reference start


@pytest.mark.gpu
def test_warm_start_gain_at_full_size(clb, oracle):
    """The warm-start setting at full size: 2 M plain-Newton training solves, 60 epochs, then
    128^3 fresh samples solved from the reference-rule start (the full iteration) and from the
    network's start (the polish), both to the 1e-10 judge, every answer judged again by the
    oracle with its sample's logK. The network's start needs at least 3.5x fewer evaluations
    (deterministic: fixed seeds, a fixed minibatch order and fixed-order gradient sums:
    13.6 -> 3.33, 4.09x) and at least 3x less kernel time, inference included (3.37x to 3.49x
    measured)."""
    print("case: mlp warm start gain, 2 M training solves, 128^3 samples", flush=True)
    S, lk = clb.gen_tableau(8, 24, seed=3)
    var = [8, 9, 10, 11]

    def batch(n, seed):
        r = np.random.default_rng(seed)
        return 10 ** r.uniform(-6, -2, (n, 8)), lk[var][None, :] + r.uniform(-1, 1, (n, 4))

    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, list(range(8)), tol=1e-10, max_iter=200)
    Ttr, Ltr = batch(2000000, 6)
    lat.batch_load(Ttr, var, Ltr)
    lat.batch_solve(clb.WARM_REF)
    lat.mlp_config(12, H, 8)
    loss = lat.mlp_train_from_batch(epochs=60, batch=512, lr=2e-3, seed=7)
    print(f"training loss {loss:.3e}", flush=True)
    Tte, Lte = batch(128 ** 3, 9)
    lat.batch_load(Tte, var, Lte)
    out = {}
    for name, mode in (("plain", clb.WARM_REF), ("mlp", clb.WARM_MLP)):
        lat.batch_solve(mode)                     # warm-up pass
        lat.sync()
        lat.reset_perf()
        lat.set_timing(True)
        lat.batch_solve(mode)
        lat.sync()
        p = lat.perf()
        lat.set_timing(False)
        x, it, r, cv = lat.batch_fetch()
        assert cv.all() and np.all(r < 1e-10)
        f = np.linalg.norm(oracle.pcf_residual(S, lk, x, Tte, var_species=var, logk_var=Lte), axis=1)
        assert f.max() < 1e-10, f"{name}: worst sample {int(f.argmax())}: {f.max():.3e}"
        out[name] = (p.spec_ms + p.mlp_ms, it.mean(), x)
    lat.close()
    assert np.max(np.abs(out["mlp"][2] - out["plain"][2])) < 1e-7
    eval_ratio = out["plain"][1] / out["mlp"][1]
    time_ratio = out["plain"][0] / out["mlp"][0]
    print(f"evaluations {out['plain'][1]:.2f} -> {out['mlp'][1]:.2f} ({eval_ratio:.2f}x); "
          f"kernel time {out['plain'][0]:.2f} -> {out['mlp'][0]:.2f} ms ({time_ratio:.2f}x)", flush=True)
    assert eval_ratio >= 3.5
    assert time_ratio >= 3.0


@pytest.mark.gpu
def test_warm_start_small_tableau_without_logk_inputs(clb):
    """The fused warm start with padded shapes: 4 components and no varying logK, so the
    network is 4 -> 128 -> 128 -> 4 and most input slots and output rows of the MFMA tiles
    are padding. Same solutions, fewer evaluations, and a sample count that leaves a
    partial 32-sample tile."""
    print("case: mlp warm start 4/8, no logK inputs", flush=True)
    S, lk = clb.gen_tableau(4, 8, seed=11)
    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, list(range(4)), tol=1e-10, max_iter=200)
    rng = np.random.default_rng(12)
    lat.batch_load(10 ** rng.uniform(-6, -2, (100000, 4)))
    lat.batch_solve(clb.WARM_REF)
    lat.mlp_config(4, H, 4)
    loss = lat.mlp_train_from_batch(epochs=20, batch=512, lr=2e-3, seed=7)
    assert loss < 0.01
    T = 10 ** np.random.default_rng(13).uniform(-6, -2, (20011, 4))
    lat.batch_load(T)
    lat.batch_solve(clb.WARM_REF)
    x0, it0, r0, cv0 = lat.batch_fetch()
    lat.batch_solve(clb.WARM_MLP)
    x1, it1, r1, cv1 = lat.batch_fetch()
    lat.close()
    assert cv0.all() and cv1.all() and np.all(r1 < 1e-10)
    assert np.max(np.abs(x1 - x0)) < 1e-6
    print(f"mean iterations: reference start {it0.mean():.2f}, MLP start {it1.mean():.2f}", flush=True)
    assert it1.mean() < it0.mean()
