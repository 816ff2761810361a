"""Many fields in one call: streams and decoded fields bit-identical to one
call per field, the captured batch decode replayed several times with eager
work in between (every replay checked), the autotuned launch geometry
reported, and the argument checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(33, 47, 61), (64, 64, 64), (5, 90, 17), (40, 128, 72), (7, 7, 7), (36, 52, 100), (3, 4, 130)]


def _fields(dt, dev, seed=0):
    import torch
    from stz_amd import fields
    out = []
    for i, sh in enumerate(SHAPES):
        g = fields.grf(sh, -2.0 - (i % 3), seed + i)
        out.append(torch.from_numpy(g.astype(dt)).to(dev))
    torch.cuda.synchronize()
    return out


def _bits(t):
    a = t.cpu().numpy() if hasattr(t, "cpu") else t
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_batch_bit_identical_to_single(stz, ctx, dt):
    import torch
    dev = torch.device("cuda:0")
    fs = _fields(dt, dev, seed=10)
    eb = 1e-3 if dt == np.float32 else 1e-6
    batch = ctx.compress_batch(fs, eb, "rel")
    singles = [ctx.compress_device(f, eb, "rel") for f in fs]
    for b, s, sh in zip(batch, singles, SHAPES):
        pb, ps = b.download(-1), s.download(-1)
        assert pb.size == ps.size and np.array_equal(pb, ps), sh
    outs = [torch.empty_like(f) for f in fs]
    want = [s.decompress() for s in singles]
    for rep in range(3):
        for o in outs:
            o.fill_(-7.0)
        ctx.decompress_batch(batch, outs)
        for o, w, sh in zip(outs, want, SHAPES):
            assert np.array_equal(_bits(o), _bits(w)), (rep, sh)
        # eager work on the same context between replays
        extra = singles[1].decompress()
        assert np.array_equal(_bits(extra), _bits(want[1]))
    for d in batch + singles:
        d.free()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_batch_decode_with_staged_streams(stz, ctx, dt):
    """A batch decode mixing streams under the staged predictors (their
    levels as stage launches inside the batch plan) with plain ones: eager,
    captured and replayed decodes all equal one call per stream, bit for bit,
    with eager work in between."""
    import torch
    dev = torch.device("cuda:0")
    fs = _fields(dt, dev, seed=40)
    eb = 1e-3 if dt == np.float32 else 1e-6
    st = ctx.compress_batch(fs[:4], eb, "rel", pred=32)
    pl = ctx.compress_batch(fs[4:], eb, "rel")
    mixed = [st[0], pl[0], st[1], st[2], pl[1], st[3], pl[2]]
    want = [d.decompress() for d in mixed]
    outs = [torch.empty(w.shape, dtype=torch.float32 if dt == np.float32 else torch.float64, device=dev)
            for w in want]
    for rep in range(4):
        for o in outs:
            o.fill_(-7.0)
        ctx.decompress_batch(mixed, outs)
        for k, (o, w) in enumerate(zip(outs, want)):
            assert np.array_equal(_bits(o), _bits(w)), (rep, k)
        extra = st[1].decompress()
        assert np.array_equal(_bits(extra), _bits(want[2]))
    for d in st + pl:
        d.free()


def test_batch_autotune_and_new_plan(stz, ctx):
    import torch
    dev = torch.device("cuda:0")
    fs = _fields(np.float32, dev, seed=20)
    batch = ctx.compress_batch(fs, 1e-3, "rel", chunk_high=16, chunk_low=8)
    rep = ctx.autotune()
    assert any(k.startswith("f32/high_encode") for k in rep) and any(k.startswith("f32/low_encode") for k in rep)
    for v in rep.values():
        assert v["waves"] in (4, 8) and v["grid"] > 0 and len(v["trials"]) == 6
    outs = [torch.empty_like(f) for f in fs]
    ctx.decompress_batch(batch, outs)
    # a different set of outputs captures a new plan
    outs2 = [torch.empty_like(f) for f in fs]
    ctx.decompress_batch(batch[::-1], outs2[::-1])
    for a, b, f in zip(outs, outs2, fs):
        assert np.array_equal(_bits(a), _bits(b))
        assert float((a.double() - f.double()).abs().max()) <= 1e-3 * float(f.max() - f.min()) * (1 + 1e-9)
    assert any(k.startswith("f32/high_decode") for k in ctx.autotune())
    for d in batch:
        d.free()


def test_batch_argument_checks(stz, ctx):
    import torch
    dev = torch.device("cuda:0")
    a = torch.rand((8, 8, 8), dtype=torch.float32, device=dev)
    b = torch.rand((8, 8, 8), dtype=torch.float64, device=dev)
    with pytest.raises(TypeError):
        ctx.compress_batch([a, b], 1e-3)
    d32 = ctx.compress_device(a, 1e-3)
    d64 = ctx.compress_device(b, 1e-3)
    with pytest.raises(stz.StzError):
        ctx.decompress_batch([d32, d64], [torch.empty_like(a), torch.empty_like(b)])
    with pytest.raises(stz.StzError):   # outputs must be device memory
        ctx.decompress_batch([d32], [np.empty((8, 8, 8), np.float32)])
    assert ctx.compress_batch([], 1e-3) == []
    d32.free()
    d64.free()


def test_batch_learned_and_trial_choices(stz, ctx):
    """tuner = 1 in a batch picks what single calls pick (same model, no
    fallback); tuner = 2 keeps each field's smallest payload over the 204
    configurations."""
    import torch
    from stz_amd import tuner
    dev = torch.device("cuda:0")
    fs = _fields(np.float32, dev, seed=40)[:4]
    params, _ = tuner.load_params()
    tuner.set_model(ctx, params, min_confidence=0.0)
    batch = ctx.compress_batch(fs, 1e-3, "rel", tuner=1)
    for b, f in zip(batch, fs):
        single = ctx.compress_device(f, 1e-3, "rel", tuner=1)
        assert np.array_equal(b.download(-1), single.download(-1))
        single.free()
    trial = ctx.compress_batch(fs, 1e-3, "rel", tuner=2)
    sizes = [tuner.trial_sizes(ctx, f, 1e-3) for f in fs]
    for t, sz in zip(trial, sizes):
        assert t.payload_bytes == sz.min() and t.info.tuner == 2
    # a threshold between the fields' confidences: the less confident fields
    # fall back to their own trial, the others keep the model's choice
    confs = [b.info.tune_confidence for b in batch]
    assert all(b.info.tuner == 1 for b in batch) and all(0 < c <= 1 for c in confs)
    thr = float(np.median(confs))
    tuner.set_model(ctx, params, min_confidence=thr)
    mixed = ctx.compress_batch(fs, 1e-3, "rel", tuner=1)
    for m, b, sz, c in zip(mixed, batch, sizes, confs):
        if c < thr:
            assert m.info.tuner == 2 and m.payload_bytes == sz.min(), (c, thr)
        else:
            assert m.info.tuner == 1 and np.array_equal(m.download(-1), b.download(-1)), (c, thr)
        assert m.info.tune_confidence == c
    assert any(c < thr for c in confs) or len(set(confs)) == 1
    tuner.set_model(ctx, params, min_confidence=0.0)
    for d in batch + trial + mixed:
        d.free()
