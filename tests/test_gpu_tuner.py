"""Learned tuner on the GPU: the bf16 MFMA forward / backward / Adam kernels
against the CPU float64 reference (oracle/mlp_ref.py, which rounds to bf16 at
the same points), the sub-block statistics against numpy, training that
lowers the loss, and the shipped model's choices against the exhaustive
trial (regret) and against the untuned configuration (gain)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


def _problem(n, seed):
    import mlp_ref
    rng = np.random.default_rng(seed)
    P = np.concatenate([rng.normal(0, 0.35, size=16 * 64), rng.normal(0, 0.1, 64),
                        rng.normal(0, 0.18, size=64 * 64), rng.normal(0, 0.1, 64),
                        rng.normal(0, 0.18, size=64 * mlp_ref.OUT), rng.normal(0, 0.1, mlp_ref.OUT)]).astype(np.float32)
    assert P.size == mlp_ref.N_PARAMS
    X = rng.normal(0, 1, size=(n, 16)).astype(np.float32)
    T = rng.dirichlet(np.full(mlp_ref.OUT, 0.3), size=n).astype(np.float32)
    return P, X, T


@pytest.mark.parametrize("n", [16, 100, 257])
def test_mlp_forward_matches_cpu(stz, ctx, torch_dev, n):
    import mlp_ref
    from stz_amd.tuner import Mlp
    torch, dev = torch_dev
    P, X, _ = _problem(n, 1 + n)
    m = Mlp(ctx, params=P)
    y = m.forward(X).cpu().numpy().astype(np.float64)
    want, _ = mlp_ref.forward(P, X, bf=True)
    # identical bf16 operands; fp32 accumulation on the device vs float64 here
    assert np.max(np.abs(y - want)) <= 1e-4 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("n", [64, 100, 513])
def test_mlp_gradient_matches_cpu_float64(stz, ctx, torch_dev, n):
    import mlp_ref
    from stz_amd.tuner import Mlp
    P, X, T = _problem(n, 7 + n)
    m = Mlp(ctx, params=P)
    G, loss = m.grad(X, T)
    g = G.cpu().numpy().astype(np.float64)
    want_loss, want = mlp_ref.loss_grad(P, X, T, bf=True)
    assert abs(loss - want_loss) <= 1e-4 * abs(want_loss)
    # a bf16 delta can round the other way when the fp32 and float64 logits
    # differ in their last bits: normwise 1e-2, per parameter block 2e-2
    assert np.linalg.norm(g - want) <= 1e-2 * np.linalg.norm(want)
    o = 0
    for k in (16 * 64, 64, 64 * 64, 64, 64 * mlp_ref.OUT, mlp_ref.OUT):
        blk = slice(o, o + k)
        assert np.linalg.norm(g[blk] - want[blk]) <= 2e-2 * np.linalg.norm(want[blk]) + 1e-7, (o, k)
        o += k
    # and the float64 network's own gradient (no rounding) is close too
    _, exact = mlp_ref.loss_grad(P, X, T, bf=False)
    assert np.linalg.norm(g - exact) <= 5e-2 * np.linalg.norm(exact)


def test_adam_matches_cpu(stz, ctx, torch_dev):
    import mlp_ref
    from stz_amd.tuner import Mlp
    P, X, T = _problem(128, 3)
    m = Mlp(ctx, params=P)
    Mw = np.zeros(P.size)
    Vw = np.zeros(P.size)
    Pw = P.astype(np.float64)
    for step in range(1, 4):
        G, _ = m.grad(X, T)
        g = G.cpu().numpy().astype(np.float64)
        from stz_amd._lib import check, lib
        from stz_amd.api import _ptr
        check(lib().stz_mlp_adam(ctx.h, _ptr(m.P), _ptr(m.G), _ptr(m.M), _ptr(m.V), step, 1e-2, 1e-3), "adam")
        Pw, Mw, Vw = mlp_ref.adam(Pw, g, Mw, Vw, step, 1e-2, 1e-3)
        got = m.P.cpu().numpy().astype(np.float64)
        assert np.max(np.abs(got - Pw)) <= 1e-5, step
        Pw = got  # continue from the device state (its gradient is the next input)
        assert float(m.G.abs().max().item()) == 0.0   # the step clears the gradient


def test_training_lowers_loss(stz, ctx, torch_dev):
    """A learnable task: the target configuration is a fixed function of the features."""
    from stz_amd.tuner import Mlp
    from stz_amd._lib import MLP_OUTPUTS as K
    rng = np.random.default_rng(5)
    X = rng.normal(size=(2048, 16)).astype(np.float32)
    W = rng.normal(size=(16, K))
    lab = np.argmax(X @ W, axis=1)
    T = np.full((2048, K), 0.02 / (K - 1), np.float32)
    T[np.arange(2048), lab] = 0.98
    m = Mlp(ctx, seed=1)
    first = np.mean([m.step(X[b], T[b], lr=3e-3) for b in np.split(rng.permutation(2048)[:512], 2)])
    losses = []
    for it in range(600):
        b = rng.choice(2048, 256, replace=False)
        losses.append(m.step(X[b], T[b], lr=3e-3))
    assert np.mean(losses[-50:]) < 0.5 * first
    acc = np.mean(np.argmax(m.forward(X).cpu().numpy(), axis=1) == lab)
    assert acc > 0.8


def _block_sums_np(f, eb):
    nz, ny, nx = f.shape
    out = []
    for cz in range(0, nz, 64):
        for cy in range(0, ny, 64):
            for cx in range(0, nx, 64):
                zs = np.arange(cz, min(cz + 64, nz), 2)
                ys = np.arange(cy, min(cy + 64, ny), 2)
                xs = np.arange(cx, min(cx + 64, nx), 2)
                Z, Y, Xx = np.meshgrid(zs, ys, xs, indexing="ij")
                v_all = f[Z, Y, Xx]
                ok = (Z >= 12) & (Z + 12 < nz) & (Y >= 12) & (Y + 12 < ny) & (Xx >= 12) & (Xx + 12 < nx)
                Z, Y, Xx = Z[ok], Y[ok], Xx[ok]
                v = f[Z, Y, Xx]
                S = np.zeros(18)
                S[0], S[1], S[2], S[3] = v.size, v.sum(), (v * v).sum(), (v ** 4).sum()
                S[4], S[5] = v_all.min(), v_all.max()
                ml1 = np.zeros(v.size)
                mc4 = np.zeros(v.size)
                d1 = np.zeros(v.size)
                for a, (dz, dy, dx) in enumerate([(0, 0, 1), (0, 1, 0), (1, 0, 0)]):
                    g = lambda k: f[Z + k * dz, Y + k * dy, Xx + k * dx]
                    l4 = np.abs(v - 0.5 * (g(-4) + g(4)))
                    c4 = np.abs(v - (-0.0625 * g(-12) + 0.5625 * g(-4) + 0.5625 * g(4) - 0.0625 * g(12)))
                    l1 = np.abs(v - 0.5 * (g(-1) + g(1)))
                    S[6 + a], S[9 + a], S[12 + a] = l4.sum(), c4.sum(), l1.sum()
                    ml1 = np.maximum(ml1, l1)
                    mc4 = np.maximum(mc4, c4)
                    d1 += np.abs(g(1) - v)
                S[15], S[16], S[17] = (ml1 <= eb).sum(), (mc4 <= eb).sum(), (d1 / 3).sum()
                out.append(S)
    return out


def _features_np(S, eb):
    lg = lambda v: np.log2(v + 1 / 1024) / 8
    n = S[0] if S[0] > 0 else 1.0
    mean = S[1] / n
    var = max(S[2] / n - mean * mean, 0)
    f = [lg(np.sqrt(var) / eb), lg((S[5] - S[4]) / eb)]
    f += [lg(S[6 + a] / n / eb) for a in range(3)] + [lg(S[9 + a] / n / eb) for a in range(3)]
    f += [lg(S[12 + a] / n / eb) for a in range(3)]
    f += [S[15] / n, S[16] / n, lg(S[17] / n / eb)]
    m2, m4 = S[2] / n, S[3] / n
    f += [np.log2(m4 / (m2 * m2 + 1e-300) + 1e-30) / 8]
    f += [np.log2((max(S[6:9]) + 1e-300) / (min(S[6:9]) + 1e-300)) / 8]
    f = np.array(f)
    if S[0] <= 0:          # no sampled point: only the value statistics remain
        f[2:] = 0
    return f


def test_features_match_numpy(stz, ctx, torch_dev):
    from stz_amd import fields, tuner
    f = fields.grf((70, 96, 130), -3.0, 9)
    eb = 1e-3 * float(f.max() - f.min())
    got = tuner.features(ctx, f, eb)
    want = np.stack([_features_np(S, eb) for S in _block_sums_np(f, eb)])
    assert got.shape == want.shape == (2 * 2 * 3, 16)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-5)


def test_shipped_model_regret_and_gain(stz, ctx, torch_dev):
    """Held-out generated fields (seeds outside training): the model's choice
    against the exhaustive trial and against the untuned configuration."""
    torch, dev = torch_dev
    from stz_amd import tuner
    if not os.path.exists(tuner.MODEL_PATH):
        pytest.fail("trained tuner model missing: %s" % tuner.MODEL_PATH)
    params, meta = tuner.load_params()
    tuner.set_model(ctx, params)
    rng = np.random.default_rng(777001)
    chosen, best, default = 0, 0, 0
    for _ in range(24):
        f, eb, fam = tuner.make_field(rng, dev)
        torch.cuda.synchronize()
        sizes = tuner.trial_sizes(ctx, f, eb).astype(np.float64)
        k, conf = tuner.choose(ctx, f, eb)
        chosen += sizes[k]
        best += sizes.min()
        default += sizes[tuner.DEFAULT_CONFIG]
    regret = chosen / best - 1.0
    gain = default / chosen - 1.0
    print("tuner held-out: regret %.4f gain vs untuned %.4f" % (regret, gain))
    assert regret <= 0.01          # within 1% of the trial's bytes
    assert gain > 0.0              # better than the untuned path


def test_compress_with_learned_and_trial_choice(stz, ctx, torch_dev, oracle):
    """params.tuner = 1 / 2: the chosen configuration is in the header and the
    payload is the oracle's for that configuration."""
    from stz_amd import fields, tuner
    params, _ = tuner.load_params()
    tuner.set_model(ctx, params)
    f = fields.turbulence(shape=(64, 72, 80), seed=31)
    for mode in (1, 2):
        ds = ctx.compress_device(f, 1e-5, "rel", tuner=mode)
        i = ds.info
        p = ds.download(-1)
        want = oracle.compress(f, 1e-5, oracle.EB_REL, pred=i.pred, interp=i.interp, order=i.order,
                               mult_base=i.mult_base, mult_low=i.mult_low, mult_high=i.mult_high, alpha=i.alpha,
                               beta=i.beta)
        assert np.array_equal(p, want)
        r = ds.decompress()
        assert np.abs(r - f).max() <= 1e-5 * (f.max() - f.min()) * (1 + 1e-12)
        assert i.tuner == mode
        if mode == 2:
            sizes = tuner.trial_sizes(ctx, f, 1e-5)
            assert i.payload_bytes == sizes.min()
        else:
            k, conf = tuner.choose(ctx, f, 1e-5)
            assert i.tune_confidence == np.float32(conf)
            assert tuner.config_index(dict(pred=i.pred, interp=i.interp, order=i.order, mult_base=i.mult_base,
                                           mult_low=i.mult_low, mult_high=i.mult_high, alpha=i.alpha,
                                           beta=i.beta)) == k
        ds.free()
    # below the confidence threshold the learned path falls back to the trial
    tuner.set_model(ctx, params, min_confidence=1.01)
    ds = ctx.compress_device(f, 1e-5, "rel", tuner=1)
    assert ds.payload_bytes == tuner.trial_sizes(ctx, f, 1e-5).min()
    assert ds.info.tuner == 2 and 0 < ds.info.tune_confidence <= 1
    ds.free()
    tuner.set_model(ctx, params)


def _forced_model(k):
    """A model whose every sub-block scores configuration k far above the
    rest: zero last-layer weights, biases +10 at k and -10 elsewhere."""
    from stz_amd import tuner
    from stz_amd._lib import MLP_OUTPUTS, MLP_PARAMS
    p = np.zeros(MLP_PARAMS, np.float32)
    b3 = np.full(MLP_OUTPUTS, -10.0, np.float32)
    b3[k] = 10.0
    p[-MLP_OUTPUTS:] = b3
    return p


@pytest.mark.parametrize("k", [48, 3 * 51 + 49])
def test_learned_choice_of_the_lorenzo_regression_base(stz, ctx, torch_dev, oracle, k):
    """A learned choice of the Lorenzo/regression base (whose base stream
    holds header symbols, so the layout made before the choice is made
    again): single and batch compresses carry the choice in the header,
    equal the oracle's payload for it and decode within the bound."""
    import torch
    from stz_amd import fields, tuner
    params, meta = tuner.load_params()
    prm = tuner.config_params(k)
    assert prm["interp"] == 3
    tuner.set_model(ctx, _forced_model(k))
    try:
        fs = [fields.grf(sh, -3.0, 900 + i, None).astype(np.float32) for i, sh in
              enumerate([(40, 36, 44), (17, 70, 33)])]
        for f in fs:
            print("learned lorenzo/regression", k, f.shape, flush=True)
            ds = ctx.compress_device(f, 1e-3, "rel", tuner=1)
            i = ds.info
            assert i.interp == 3 and i.tuner == 1
            want = oracle.compress(f, 1e-3, oracle.EB_REL, pred=i.pred, interp=i.interp, order=i.order,
                                   mult_base=i.mult_base, mult_low=i.mult_low, mult_high=i.mult_high, alpha=i.alpha,
                                   beta=i.beta)
            assert np.array_equal(ds.download(-1), want)
            r = ds.decompress()
            assert np.abs(r.astype(np.float64) - f).max() <= 1e-3 * (float(f.max()) - float(f.min()))
            ds.free()
        dev = torch.device("cuda:0")
        ts = [torch.from_numpy(f).to(dev) for f in fs]
        bs = ctx.compress_batch(ts, 1e-3, "rel", tuner=1)
        for f, b in zip(fs, bs):
            assert b.info.interp == 3
            want = oracle.compress(f, 1e-3, oracle.EB_REL, pred=prm["pred"], interp=3, order=prm["order"],
                                   mult_base=prm["mult_base"], mult_low=prm["mult_low"], mult_high=prm["mult_high"],
                                   alpha=prm["alpha"], beta=prm["beta"])
            assert np.array_equal(b.download(-1), want)
        outs = [torch.empty_like(t) for t in ts]
        ctx.decompress_batch(bs, outs)
        for f, o in zip(fs, outs):
            assert np.abs(o.cpu().numpy().astype(np.float64) - f).max() <= 1e-3 * (float(f.max()) - float(f.min()))
        for b in bs:
            b.free()
    finally:
        tuner.set_model(ctx, params, float(meta.get("min_confidence", 0.0)))


def _two_way_model(k_rough, k_smooth):
    """A model that picks k_rough for fields whose sub-blocks have a large
    spread in units of the bound (feature 0, log2(std/eb)/8, above 0.5) and
    k_smooth otherwise: one hidden unit relu(f0 - 0.5) passed through, output
    k_rough = 40 x that unit, k_smooth a bias of 10, the rest -10."""
    from stz_amd._lib import MLP_OUTPUTS, MLP_PARAMS
    IN, HID, OUT = 16, 64, MLP_OUTPUTS
    p = np.zeros(MLP_PARAMS, np.float32)
    o_b1 = IN * HID
    o_w2 = o_b1 + HID
    o_b2 = o_w2 + HID * HID
    o_w3 = o_b2 + HID
    o_b3 = o_w3 + HID * OUT
    p[0 * HID + 0] = 1.0          # W1[f0][h0]
    p[o_b1 + 0] = -0.5            # b1[h0]
    p[o_w2 + 0 * HID + 0] = 1.0   # W2[h0][h0]
    p[o_w3 + 0 * OUT + k_rough] = 40.0
    b3 = np.full(OUT, -10.0, np.float32)
    b3[k_smooth] = 10.0
    b3[k_rough] = 0.0
    p[o_b3:] = b3
    assert o_b3 + OUT == MLP_PARAMS
    return p


def test_learned_batch_mixing_staged_and_other_predictors(stz, ctx, torch_dev, oracle):
    """One learned batch whose fields get different predictor sets, the first
    one not staged: each field's payload is the oracle's for its own choice
    (staged fields run their three-stage launches, the others the shared
    persistent launch), and the batch decode is within the bound."""
    import torch
    from stz_amd import fields, tuner
    params, meta = tuner.load_params()
    k_staged = 3 * 51 + 24     # staged predictors, cubic base, order 0, F5 multipliers
    k_other = 48               # F5's predictors, Lorenzo / regression base
    assert tuner.config_params(k_staged)["pred"] == 32 and tuner.config_params(k_other)["interp"] == 3
    tuner.set_model(ctx, _two_way_model(k_staged, k_other))
    try:
        rng = np.random.default_rng(5)
        spiky = []
        for i, sh in enumerate([(72, 80, 96), (70, 66, 130)]):
            f = 0.01 * fields.grf(sh, -3.0, 700 + i, None)
            f.flat[rng.integers(f.size)] = 1e3
            spiky.append(f.astype(np.float32))
        noise = rng.uniform(-1, 1, size=(40, 48, 56)).astype(np.float32)
        fs = [spiky[0], noise, spiky[1]]
        want_k = [k_other, k_staged, k_other]
        dev = torch.device("cuda:0")
        ts = [torch.from_numpy(f).to(dev) for f in fs]
        bs = ctx.compress_batch(ts, 1e-3, "rel", tuner=1)
        for f, b, k in zip(fs, bs, want_k):
            i = b.info
            print("mixed learned batch", f.shape, i.pred, i.interp, flush=True)
            prm = tuner.config_params(k)
            assert (i.pred, i.interp, i.order) == (prm["pred"], prm["interp"], prm["order"])
            want = oracle.compress(f, 1e-3, oracle.EB_REL, pred=prm["pred"], interp=prm["interp"], order=prm["order"],
                                   mult_base=prm["mult_base"], mult_low=prm["mult_low"], mult_high=prm["mult_high"],
                                   alpha=prm["alpha"], beta=prm["beta"])
            assert np.array_equal(b.download(-1), want)
        outs = [torch.empty_like(t) for t in ts]
        ctx.decompress_batch(bs, outs)
        for f, o in zip(fs, outs):
            assert np.abs(o.cpu().numpy().astype(np.float64) - f).max() <= 1e-3 * (float(f.max()) - float(f.min()))
        for b in bs:
            b.free()
    finally:
        tuner.set_model(ctx, params, float(meta.get("min_confidence", 0.0)))


def test_learned_batch_fallback_is_the_trial(stz, ctx, torch_dev, oracle):
    """Below the confidence threshold every field of a learned batch falls
    back to the exhaustive trial (run as one batch per configuration): each
    payload equals the single-field trial's, and the header records the
    trial (tuner 2) with the model's confidence."""
    import torch
    from stz_amd import fields, tuner
    params, meta = tuner.load_params()
    tuner.set_model(ctx, params, min_confidence=1.01)
    try:
        dev = torch.device("cuda:0")
        fs = [torch.from_numpy(fields.grf(sh, -3.0, 600 + i, None).astype(np.float32)).to(dev)
              for i, sh in enumerate([(24, 20, 28), (12, 33, 17)])]
        bs = ctx.compress_batch(fs, 1e-3, "rel", tuner=1)
        for t, b in zip(fs, bs):
            single = ctx.compress_device(t, 1e-3, "rel", tuner=2)
            assert np.array_equal(b.download(-1), single.download(-1))
            assert b.info.tuner == 2 and 0 < b.info.tune_confidence <= 1
            single.free()
        for b in bs:
            b.free()
    finally:
        tuner.set_model(ctx, params, float(meta.get("min_confidence", 0.0)))
