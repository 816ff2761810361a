"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Bar: bit-exact codes, payload bytes and reconstructions (integer/byte work
and the FP predictors are reproduced operation for operation), max pointwise
error <= the requested ABS / REL bound (strictly; under F5's REF meaning the
level bound plus the T rounding of pred+dec, at most 2 ulp of the largest
magnitude), ROI / slice / level == the oracle's, which itself equals a crop
of the full decode.
"""
import sys

import numpy as np
import pytest

from payload_util import first_difference, max_error, nan_sign_cleared, nan_sign_cleared_values, requested_bound

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _field(kind, shape, dtype, seed):
    from stz_amd import fields
    if kind == "sin":
        return fields.sinusoids(shape=shape, seed=seed, dtype=dtype)
    if kind == "lognormal":
        return fields.lognormal(shape=shape, seed=seed, dtype=dtype)
    if kind == "grf":
        return fields.grf(shape, -3.0, seed).astype(dtype)
    if kind == "noise":      # white noise: every stream's codes span thousands of bins
        return np.random.default_rng(seed).standard_normal(shape).astype(dtype)
    raise ValueError(kind)


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _assert_bitexact(got, want, what):
    g, w = _bits(got), _bits(want)
    assert g.shape == w.shape, "%s: shape %s vs %s" % (what, g.shape, w.shape)
    bad = np.nonzero(g.ravel() != w.ravel())[0]
    assert bad.size == 0, "%s: %d of %d values differ, first at %d (%r vs %r)" % (
        what, bad.size, g.size, bad[0], got.ravel()[bad[0]], want.ravel()[bad[0]])


# ------------------------------------------------------------------ stages

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(6, 4, 10), (2, 2, 2), (8, 12, 6)])
def test_split_merge(ctx, oracle, dtype, dims):
    _say("split/merge", dtype.__name__, dims)
    rng = np.random.default_rng(11)
    full = rng.normal(size=dims).astype(dtype)
    subs = ctx.split(full)
    want = oracle.split(full)
    for k in range(8):
        _assert_bitexact(subs[k], want[k], "sub %d" % k)
    back = ctx.merge(subs)
    _assert_bitexact(back, full, "merge")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(9, 7, 11), (4, 4, 4), (3, 5, 6), (2, 3, 1), (17, 12, 20)])
def test_residual_predict_reconstruct(ctx, oracle, dtype, dims):
    rng = np.random.default_rng(5)
    ref = np.cumsum(rng.normal(size=dims), axis=2).astype(dtype)
    sub = (ref + rng.normal(scale=0.1, size=dims)).astype(dtype)
    ref[0] = 0  # flat zero plane: exercises FTZ in the inverse-variance weights
    for b in range(7):
        _say("residual block", b, dtype.__name__, dims)
        _assert_bitexact(ctx.residual_predict(b, sub, ref), oracle.preprocess(b, sub, ref), "predict b%d" % b)
        _assert_bitexact(ctx.residual_reconstruct(b, sub, ref), oracle.depreprocess(b, sub, ref), "recon b%d" % b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_quantize_dequantize(ctx, oracle, dtype):
    rng = np.random.default_rng(9)
    d = np.concatenate([rng.normal(size=7000) * rng.choice([1e-3, 1, 100, 1e5], size=7000),
                        [np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-300]]).astype(dtype)
    for eb in (1e-3, 0.37, 1e-7):
        _say("quantize", dtype.__name__, eb)
        codes, un = ctx.quantize(d, eb)
        wc, wu, _ = oracle.quantize(d, eb)
        np.testing.assert_array_equal(codes, wc)
        _assert_bitexact(un, wu, "unpred")
        out = ctx.dequantize(codes, un, eb, dtype)
        _assert_bitexact(out, oracle.dequantize(wc, wu, eb, dtype), "dequant")


# ------------------------------------------------------------------ whole path

CASES = [
    ("sin", (64, 64, 64), np.float32, 1e-3, "rel"),
    ("sin", (40, 40, 40), np.float64, 1e-3, "rel"),
    ("sin", (37, 35, 33), np.float32, 1e-3, "rel"),     # padding + partial chunks
    ("lognormal", (64, 96, 32), np.float32, 1e-2, "rel"),
    ("lognormal", (64, 64, 64), np.float32, 1e-4, "rel"),
    ("grf", (48, 40, 56), np.float64, 1e-6, "rel"),
    ("grf", (16, 16, 16), np.float32, 0.05, "abs"),
    ("sin", (8, 8, 8), np.float32, 1e-3, "ref"),
    ("sin", (5, 3, 7), np.float64, 1e-3, "rel"),        # dims below 4
    # used symbols wider than the encoder's 2048-entry LDS window (the window
    # placed around the centre, the rest read from global memory)
    ("noise", (40, 48, 56), np.float32, 1e-5, "rel"),
    ("noise", (24, 32, 40), np.float64, 1e-7, "rel"),
]


def _compress_both(ctx, oracle, f, eb, mode, **kw):
    ds = ctx.compress_device(f, eb, mode, **kw)
    gpu = ds.download(zstd_level=-1)
    want = oracle.compress(f, eb, {"abs": 0, "rel": 1, "ref": 2}[mode], **kw)
    return ds, gpu, want


@pytest.mark.parametrize("case", CASES, ids=[("%s-%s-%s-%g-%s" % (c[0], "x".join(map(str, c[1])), c[2].__name__, c[3], c[4])) for c in CASES])
def test_payload_and_decode_bitexact(ctx, oracle, case):
    kind, shape, dtype, eb, mode = case
    _say("whole path", case)
    f = _field(kind, shape, dtype, seed=1001)
    ds, gpu, want = _compress_both(ctx, oracle, f, eb, mode)
    diff = first_difference(gpu, want)
    assert diff is None, diff
    rec = ds.decompress()
    ref = oracle.decompress(want)
    _assert_bitexact(rec, ref, "decompress")
    err = max_error(rec, f)
    if mode in ("abs", "rel"):
        # the requested bound holds strictly (the level unit leaves room for
        # the reconstruction's rounding)
        E = requested_bound(f, eb, mode)
        assert err <= E, (err, E)
    else:
        # REF: F5's level bounds, plus the T rounding of pred + dec
        h = oracle.header(want)
        bound = max(h["eb_base"], h["eb_low"], h["eb_high"])
        slack = 2 * np.spacing(np.abs(f).max().astype(dtype)).astype(np.float64)
        assert err <= bound + slack, (err, bound)
    ds.free()


def test_every_tuner_configuration_bitexact(stz, ctx, oracle):
    """All 204 configurations the learned tuner chooses among (predictor set x
    base codec -- interpolator x axis order, or Lorenzo -- x level-bound
    scaling): payload bytes equal the oracle's, f32 and f64, and a spread of
    them decode bit-exactly."""
    from stz_amd import tuner
    f32 = _field("grf", (36, 40, 44), np.float32, seed=17)
    f64 = _field("sin", (20, 24, 28), np.float64, seed=5)
    from stz_amd._lib import TUNE_CONFIGS
    assert TUNE_CONFIGS == 204
    for k in range(TUNE_CONFIGS):
        prm = tuner.config_params(k)
        for f, eb in ((f32, 1e-3), (f64, 1e-5)):
            _say("tuner configuration", (k, f.dtype.name))
            ds, gpu, want = _compress_both(ctx, oracle, f, eb, "rel", **prm)
            diff = first_difference(gpu, want)
            assert diff is None, (k, f.dtype.name, diff)
            if k in (0, 13, 30, 47, 48, 61, 100, 150, 152):
                _assert_bitexact(ds.decompress(), oracle.decompress(want), "decompress config %d" % k)
            ds.free()


@pytest.mark.parametrize("params", [dict(interp=0, order=3), dict(interp=1, order=5, alpha=1.5, beta=4.0),
                                    dict(chunk_high=8, chunk_low=4), dict(mult_base=1.0, mult_low=2.5, mult_high=6.25)])
def test_params_bitexact(ctx, oracle, params):
    _say("params", params)
    f = _field("sin", (48, 32, 40), np.float32, seed=7)
    ds, gpu, want = _compress_both(ctx, oracle, f, 1e-3, "rel", **params)
    diff = first_difference(gpu, want)
    assert diff is None, diff
    _assert_bitexact(ds.decompress(), oracle.decompress(want), "decompress")


def test_container_roundtrip(ctx, oracle):
    _say("container")
    f = _field("sin", (32, 48, 40), np.float32, seed=3)
    s = ctx.compress(f, 1e-3, "rel")
    rec = ctx.decompress(s)
    want = oracle.compress(f, 1e-3, 1)
    _assert_bitexact(rec, oracle.decompress(want), "container decode")
    info = __import__("stz_amd").stream_info(s)
    assert (info.nx, info.ny, info.nz) == (40, 48, 32)


ROI_QUERIES = [(3, 5, 7, 20, 17, 30), (0, 0, 0, 63, 63, 63), (10, 10, 10, 10, 40, 50), (62, 0, 1, 63, 63, 2),
               (40, 2, 9, 5, 30, 0), (-5, -5, -5, 100, 3, 3), (17, 9, 33, 17, 9, 33)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_roi_slice_level(ctx, oracle, dtype):
    f = _field("lognormal", (64, 56, 64), dtype, seed=21)
    ds, gpu, want = _compress_both(ctx, oracle, f, 1e-3, "rel")
    assert first_difference(gpu, want) is None
    full = oracle.decompress(want)
    for q in ROI_QUERIES:
        _say("roi", dtype.__name__, q)
        got = ds.decompress_roi(*q)
        ref = oracle.decompress_roi(want, *q)
        _assert_bitexact(got, ref, "roi %s" % (q,))
        x0, x1 = sorted((max(0, min(q[0], 63)), max(0, min(q[3], 63))))
        y0, y1 = sorted((max(0, min(q[1], 55)), max(0, min(q[4], 55))))
        z0, z1 = sorted((max(0, min(q[2], 63)), max(0, min(q[5], 63))))
        _assert_bitexact(got, full[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1], "roi crop %s" % (q,))
    s = ds.download(3)
    for axis, idx in [(0, 0), (1, 55), (2, 31), (0, 17)]:
        _say("slice", axis, idx)
        got = ctx.decompress_slice(s, axis, idx)
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(idx, idx + 1)
        _assert_bitexact(got, full[tuple(sl)], "slice %d %d" % (axis, idx))
    for lv in range(4):
        _say("level", lv)
        _assert_bitexact(ds.decompress_level(lv), oracle.decompress_level(want, lv), "level %d" % lv)


def _degenerate(kind, dtype):
    rng = np.random.default_rng(23)
    shape = (20, 24, 28)
    if kind == "constant":
        return np.full(shape, 3.5, dtype)
    if kind == "zeros":
        return np.zeros(shape, dtype)
    if kind == "point":
        return np.array([[[2.0]]], dtype)
    g = rng.standard_normal(shape).astype(dtype)
    if kind == "nan-first":          # a NaN where a sequential range scan starts
        g[0, 0, 0] = np.nan
    elif kind == "nan-inf":
        g[3, 4, 5] = np.nan
        g[10, 1, 2] = np.inf
        g[19, 23, 27] = -np.inf
    return g


# (field, eb, mode): a zero value range under a relative bound (bound 0: every
# value stored verbatim), one symbol per stream, a 1-point field, a bound so
# large that every code is the centre, and non-finite inputs (skipped by the
# value range; unpredictable, and their neighbours' predictions with them)
DEGENERATE = [("constant", 1e-3, "rel"), ("zeros", 1e-3, "abs"), ("point", 1e-3, "rel"), ("normal", 1e30, "abs"),
              ("nan-first", 1e-3, "rel"), ("nan-inf", 1e-3, "abs")]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", DEGENERATE, ids=[c[0] + "-" + c[2] for c in DEGENERATE])
def test_degenerate_fields_bitexact(ctx, oracle, case, dtype):
    kind, eb, mode = case
    _say("degenerate field", kind, np.dtype(dtype).name)
    f = _degenerate(kind, dtype)
    ds, gpu, want = _compress_both(ctx, oracle, f, eb, mode)
    got, ref = ds.decompress(), oracle.decompress(want)
    if kind.startswith("nan"):
        # NaN side values and decoded NaNs may differ in the sign bit only
        # (x86 keeps a NaN subtrahend's sign, the GPU flips it): equal with
        # that bit cleared, NaNs at the same points, every other bit equal
        gpu, want = nan_sign_cleared(gpu, dtype), nan_sign_cleared(want, dtype)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        got, ref = nan_sign_cleared_values(got), nan_sign_cleared_values(ref)
    diff = first_difference(gpu, want)
    assert diff is None, diff
    _assert_bitexact(got, ref, "decompress " + kind)
    ds.free()


def test_c1_config_bitexact(ctx, oracle):
    """128^3 fp32 sinusoid field (seed 1001), rel eb 1e-3, cubic, full path."""
    from stz_amd import fields
    _say("C1 128^3")
    f = fields.sinusoids(128, seed=1001)
    ds, gpu, want = _compress_both(ctx, oracle, f, 1e-3, "rel")
    diff = first_difference(gpu, want)
    assert diff is None, diff
    _assert_bitexact(ds.decompress(), oracle.decompress(want), "C1 decompress")


def test_device_pointers(ctx, oracle):
    import torch
    _say("device pointers")
    f = _field("sin", (32, 32, 32), np.float32, seed=2)
    t = torch.from_numpy(f).cuda()
    ds = ctx.compress_device(t, 1e-3, "rel")
    out = torch.empty_like(t)
    ds.decompress(out)
    torch.cuda.synchronize()
    want = oracle.compress(f, 1e-3, 1)
    _assert_bitexact(out.cpu().numpy(), oracle.decompress(want), "device decode")
    assert first_difference(ds.download(-1), want) is None


def test_corrupt_stream_rejected(ctx, oracle):
    import stz_amd
    _say("corrupt stream")
    f = _field("sin", (16, 16, 16), np.float32, seed=4)
    s = ctx.compress(f, 1e-3, "rel")
    hdr = 280  # container header; the raw file-header frame follows it
    bad = s.copy()
    bad[hdr] ^= 0xFF  # payload magic
    for b in (bad, s[:100], s[: len(s) - 9]):
        with pytest.raises(stz_amd.StzError):
            ctx.decompress(b)
    # a payload whose chunk table lies about the bitstream is refused before any decode
    raw = ctx.compress_device(f, 1e-3, "rel").download(-1)
    from payload_util import stream_sections
    base, offs = stream_sections(raw, 9)
    raw2 = raw.copy()
    raw2[base + offs[1] + 8] ^= 0x01  # chunk_bit[1]
    with pytest.raises(stz_amd.StzError):
        ctx.decompress(raw2)


def test_flipped_bitstream_bit_rejected(ctx):
    """A flipped bit inside a stream's bitstream (every table intact).  In
    one-symbol streams (a constant field: code "0" only) every flip makes an
    invalid code and is refused.  Otherwise a flip that changes the bits a
    sub-chunk uses is refused; one that turns a code into another of the same
    length cannot be seen without a checksum.  In low-entropy streams (about
    1-1.5 bits per symbol, short codes of many lengths) a flip nearly always
    desynchronises the codes: 65 of these 70 flips on this field (the count is
    deterministic); none may fault."""
    import stz_amd
    from payload_util import stream_sections
    _say("flipped bitstream bit")

    def flips(raw, streams, fracs):
        caught = total = 0
        for s_ in streams:
            base, offs = stream_sections(raw, s_)
            lo, hi = base + offs[7], base + offs[8]     # bitstream (16 bytes of padding at its end)
            assert hi - lo > 64, (s_, hi - lo)
            for frac in fracs:
                bad = raw.copy()
                bad[lo + int((hi - lo - 16) * frac)] ^= 0x10
                total += 1
                try:
                    ctx.decompress(bad)
                except stz_amd.StzError:
                    caught += 1
        return caught, total

    const = np.full((64, 64, 64), 3.5, np.float32)
    raw = ctx.compress_device(const, 1e-3, "abs").download(-1)
    caught, total = flips(raw, (9, 12, 14), (0.1, 0.5, 0.8))
    assert caught == total, (caught, total)
    f = _field("lognormal", (64, 64, 64), np.float32, seed=3)
    raw = ctx.compress_device(f, 1e-2, "rel").download(-1)
    caught, total = flips(raw, range(1, 15), (0.1, 0.3, 0.5, 0.7, 0.9))
    assert total == 70 and caught >= 60, (caught, total)

# predictor options (stz_params.pred): F5's three build options each and
# together; stz2.cpp's plane (8) and cube (16) predictors; everything
PRED_OPTS = [1, 2, 4, 7, 8, 16, 24, 31]


@pytest.mark.parametrize("pred", PRED_OPTS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_predictor_options_bitexact(ctx, oracle, dtype, pred):
    """Predictor options through the whole path: payload bytes equal the
    oracle's (whose predictors equal F5 built with its PRED_* switches and
    stz2.cpp's preprocess_block), the option lands in the header, decode is
    bit-exact."""
    _say("predictor options", pred, np.dtype(dtype).name)
    f = _field("lognormal" if dtype == np.float32 else "grf", (40, 44, 52), dtype, seed=31)
    eb = 1e-3 if dtype == np.float32 else 1e-6
    ds, gpu, want = _compress_both(ctx, oracle, f, eb, "rel", pred=pred)
    diff = first_difference(gpu, want)
    assert diff is None, diff
    assert ds.info.pred == pred
    rec = ds.decompress()
    _assert_bitexact(rec, oracle.decompress(want), "decompress pred %d" % pred)
    assert max_error(rec, f) <= requested_bound(f, eb, "rel")
    # the option changes the payload (it is not silently ignored)
    plain = oracle.compress(f, eb, 1)
    assert first_difference(plain, want) is not None
    ds.free()


def test_predictor_options_random_access_and_batch(ctx, oracle):
    """Random access (box, slice, levels) and a batch of uneven fields under
    the options: equal to the oracle / to one call per field."""
    import torch
    f = _field("lognormal", (48, 40, 56), np.float32, seed=8)
    ds, gpu, want = _compress_both(ctx, oracle, f, 1e-3, "rel", pred=31, chunk_high=8, chunk_low=4)
    assert first_difference(gpu, want) is None
    full = oracle.decompress(want)
    for q in [(3, 5, 7, 50, 30, 40), (10, 10, 10, 10, 39, 47), (0, 0, 0, 55, 39, 47)]:
        _assert_bitexact(ds.decompress_roi(*q), oracle.decompress_roi(want, *q), "roi %s" % (q,))
    for lv in range(4):
        _assert_bitexact(ds.decompress_level(lv), oracle.decompress_level(want, lv), "level %d" % lv)
    ds.prepare(keep_base=True)
    reqs = [("box", 1, 2, 3, 40, 30, 20), ("slice", 2, 17), ("level", 2)]
    for r, got in zip(reqs, ds.decompress_batch(reqs)):
        if r[0] == "box":
            want_r = full[3:21, 2:31, 1:41]
        elif r[0] == "slice":
            want_r = full[17:18]
        else:
            want_r = full[::2, ::2, ::2]
        _assert_bitexact(got, want_r, "batch request %s" % (r,))
    ds.free()
    dev = torch.device("cuda:0")
    fs = [torch.from_numpy(_field("grf", sh, np.float32, seed=40 + i)).to(dev)
          for i, sh in enumerate([(33, 47, 61), (16, 20, 24), (40, 36, 52)])]
    batch = ctx.compress_batch(fs, 1e-3, "rel", pred=30)
    outs = [torch.empty_like(t) for t in fs]
    ctx.decompress_batch(batch, outs)
    for t, b, o in zip(fs, batch, outs):
        s = ctx.compress_device(t, 1e-3, "rel", pred=30)
        pb, ps = b.download(-1), s.download(-1)
        assert pb.size == ps.size and np.array_equal(pb, ps)
        assert first_difference(pb, oracle.compress(t.cpu().numpy(), 1e-3, 1, pred=30)) is None
        _assert_bitexact(o.cpu().numpy(), s.decompress(), "batch decode")
        s.free()
    for b in batch:
        b.free()


# dependency.cpp's staged inverse-variance predictors (stz_params.pred = 32):
# blocks 2/4/5 from the reconstructed 0/1/3, 6 from the reconstructed 2/4/5
STAGED_CASES = [("lognormal", (40, 44, 52), np.float32, 1e-3, "rel"), ("grf", (36, 30, 42), np.float64, 1e-6, "rel"),
                ("noise", (21, 18, 27), np.float32, 1e-5, "rel"), ("sin", (9, 70, 66), np.float32, 1e-2, "abs"),
                ("grf", (4, 4, 4), np.float64, 1e-3, "rel")]


@pytest.mark.parametrize("case", STAGED_CASES, ids=[("%s-%s-%s-%g-%s" % (c[0], "x".join(map(str, c[1])),
                                                                          c[2].__name__, c[3], c[4]))
                                                     for c in STAGED_CASES])
def test_staged_predictor_bitexact(ctx, oracle, case):
    """The staged predictors through the whole path: payload bytes equal the
    oracle's (whose staged levels equal dependency.cpp's own functions), the
    option lands in the header, decode is bit-exact, the bound holds; ROI,
    slice and levels equal the oracle's."""
    kind, shape, dtype, eb, mode = case
    _say("staged predictors", case)
    f = _field(kind, shape, dtype, seed=61)
    ds, gpu, want = _compress_both(ctx, oracle, f, eb, mode, pred=32)
    diff = first_difference(gpu, want)
    assert diff is None, diff
    assert ds.info.pred == 32
    rec = ds.decompress()
    _assert_bitexact(rec, oracle.decompress(want), "decompress staged")
    assert max_error(rec, f) <= requested_bound(f, eb, mode)
    nz, ny, nx = shape
    for q in [(1, 2, 0, nx - 2, ny - 3, nz - 1), (nx // 2, 0, nz // 3, nx // 2, ny - 1, nz // 3 + 2)]:
        _assert_bitexact(ds.decompress_roi(*q), oracle.decompress_roi(want, *q), "roi %s" % (q,))
    for lv in range(4):
        _assert_bitexact(ds.decompress_level(lv), oracle.decompress_level(want, lv), "level %d" % lv)
    ds.free()


def test_staged_predictor_requests_and_batch(ctx, oracle):
    """Batched random-access requests and a batch of uneven fields under the
    staged predictors: equal to the crops of the full decode and to one call
    per field; the option does not combine with another bit."""
    import torch
    import stz_amd
    f = _field("lognormal", (48, 40, 56), np.float32, seed=9)
    ds, gpu, want = _compress_both(ctx, oracle, f, 1e-3, "rel", pred=32, chunk_high=8, chunk_low=4)
    assert first_difference(gpu, want) is None
    full = oracle.decompress(want)
    ds.prepare(keep_base=True)
    reqs = [("box", 1, 2, 3, 40, 30, 20), ("slice", 2, 17), ("level", 2), ("level", 0), ("level", 3)]
    for r, got in zip(reqs, ds.decompress_batch(reqs)):
        if r[0] == "box":
            want_r = full[3:21, 2:31, 1:41]
        elif r[0] == "slice":
            want_r = full[17:18]
        elif r[1] == 3:
            want_r = full
        else:
            want_r = oracle.decompress_level(want, r[1])
        _assert_bitexact(got, want_r, "batch request %s" % (r,))
    ds.free()
    dev = torch.device("cuda:0")
    fs = [torch.from_numpy(_field("grf", sh, np.float32, seed=50 + i)).to(dev)
          for i, sh in enumerate([(33, 47, 61), (16, 20, 24)])]
    batch = ctx.compress_batch(fs, 1e-3, "rel", pred=32)
    outs = [torch.empty_like(t) for t in fs]
    ctx.decompress_batch(batch, outs)
    for t, b, o in zip(fs, batch, outs):
        pb = b.download(-1)
        assert first_difference(pb, oracle.compress(t.cpu().numpy(), 1e-3, 1, pred=32)) is None
        _assert_bitexact(o.cpu().numpy(), b.decompress(), "batch decode")
    for b in batch:
        b.free()
    with pytest.raises(stz_amd.StzError):
        ctx.compress_device(fs[1], 1e-3, "rel", pred=32 | 1)


def test_repeated_compress_of_one_field(ctx, oracle):
    """Compressing the same field again and again on one context (its
    buffers come back from the pool, its layout from the cache): every
    payload equals the oracle's, including after a different field and a
    profiled call in between."""
    import torch
    dev = torch.device("cuda:0")
    f = _field("lognormal", (64, 56, 72), np.float32, seed=5)
    g = _field("grf", (40, 44, 36), np.float32, seed=6)
    want_f = oracle.compress(f, 1e-3, oracle.EB_REL)
    want_g = oracle.compress(g, 1e-3, oracle.EB_REL)
    tf, tg = torch.from_numpy(f).to(dev), torch.from_numpy(g).to(dev)
    for k, (t, want) in enumerate([(tf, want_f)] * 4 + [(tg, want_g)] + [(tf, want_f)] * 3):
        _say("repeated compress", k)
        if k == 6:
            ctx.profile(True)
        ds = ctx.compress_device(t, 1e-3, "rel")
        if k == 6:
            ctx.profile(False)
        assert first_difference(ds.download(-1), want) is None, k
        ds.free()


# ------------------------------------------------------------------ Lorenzo base
# interp = 2: the base grid in independent 16^3 Lorenzo tiles (SZ3's
# ALGO_INTERP_LORENZO alternative); interp = 3: Lorenzo or a fitted plane per
# 8^3 sub-block of those tiles (SZ3's Lorenzo + regression, sz.cpp:63-76).
# Base grids from 1 to 2x2x3 tiles, with partial tiles (and empty sub-blocks)
# on every axis; a 64x32x16 base (one 32768-symbol chunk of points) whose
# tiles' header symbols start a second base-stream chunk.
LZ_CASES = [("lognormal", (64, 56, 64), np.float32, 1e-3, "rel"), ("grf", (48, 40, 56), np.float64, 1e-6, "rel"),
            ("sin", (9, 130, 70), np.float32, 1e-2, "abs"), ("noise", (20, 24, 28), np.float64, 1e-5, "rel"),
            ("grf", (4, 4, 4), np.float32, 1e-3, "rel"), ("grf", (128, 96, 160), np.float32, 1e-3, "rel"),
            ("grf", (64, 128, 256), np.float32, 1e-3, "rel")]


@pytest.mark.parametrize("interp", [2, 3])
@pytest.mark.parametrize("case", LZ_CASES, ids=[("%s-%s-%s-%g-%s" % (c[0], "x".join(map(str, c[1])), c[2].__name__,
                                                                      c[3], c[4])) for c in LZ_CASES])
def test_lorenzo_base_bitexact(ctx, oracle, case, interp):
    kind, shape, dtype, eb, mode = case
    _say("lorenzo base", interp, case)
    f = _field(kind, shape, dtype, seed=77)
    ds, gpu, want = _compress_both(ctx, oracle, f, eb, mode, interp=interp)
    diff = first_difference(gpu, want)
    assert diff is None, diff
    assert ds.info.interp == interp and oracle.header(want)["interp"] == interp
    rec = ds.decompress()
    _assert_bitexact(rec, oracle.decompress(want), "decompress")
    assert max_error(rec, f) <= requested_bound(f, eb, mode)
    ds.free()


@pytest.mark.parametrize("interp", [2, 3])
def test_lorenzo_base_degenerate_fields(ctx, oracle, interp):
    """Constant, zero, one-point, huge-bound and non-finite fields with the
    Lorenzo bases: payload and decode equal the oracle's."""
    for kind, eb, mode in DEGENERATE:
        _say("lorenzo degenerate", interp, kind)
        f = _degenerate(kind, np.float32)
        ds, gpu, want = _compress_both(ctx, oracle, f, eb, mode, interp=interp)
        got, ref = ds.decompress(), oracle.decompress(want)
        if kind.startswith("nan"):
            gpu, want = nan_sign_cleared(gpu, np.float32), nan_sign_cleared(want, np.float32)
            assert np.array_equal(np.isnan(got), np.isnan(ref))
            got, ref = nan_sign_cleared_values(got), nan_sign_cleared_values(ref)
        diff = first_difference(gpu, want)
        assert diff is None, (kind, diff)
        _assert_bitexact(got, ref, "decompress " + kind)
        ds.free()


@pytest.mark.parametrize("interp", [2, 3])
def test_lorenzo_base_random_access_and_batch(ctx, oracle, interp):
    """Box, slice and progressive-level decodes of a Lorenzo-based stream
    equal the oracle's; a batch decode mixing Lorenzo and interpolation bases
    equals one call per stream; a batch compress equals single compresses."""
    import torch
    _say("lorenzo random access and batch", interp)
    f = _field("lognormal", (64, 56, 64), np.float32, seed=21)
    ds, gpu, want = _compress_both(ctx, oracle, f, 1e-3, "rel", interp=interp, pred=24, chunk_high=8, chunk_low=4)
    assert first_difference(gpu, want) is None
    full = oracle.decompress(want)
    for q in [(3, 5, 7, 50, 30, 40), (10, 10, 10, 10, 55, 63), (0, 0, 0, 63, 55, 63)]:
        _assert_bitexact(ds.decompress_roi(*q), oracle.decompress_roi(want, *q), "roi %s" % (q,))
    for lv in range(4):
        _assert_bitexact(ds.decompress_level(lv), oracle.decompress_level(want, lv), "level %d" % lv)
    s = ds.download(3)
    got = ctx.decompress_slice(s, 1, 17)
    _assert_bitexact(got, full[:, 17:18, :], "slice")
    ds.prepare(keep_base=True)
    reqs = [("box", 1, 2, 3, 40, 30, 20), ("slice", 2, 17), ("level", 0)]
    for r, g in zip(reqs, ds.decompress_batch(reqs)):
        w = full[3:21, 2:31, 1:41] if r[0] == "box" else full[17:18] if r[0] == "slice" else full[::8, ::8, ::8]
        _assert_bitexact(g, w, "batch request %s" % (r,))
    ds.free()
    dev = torch.device("cuda:0")
    fs = [torch.from_numpy(_field("grf", sh, np.float32, seed=50 + i)).to(dev)
          for i, sh in enumerate([(33, 47, 61), (16, 20, 24), (70, 36, 90)])]
    lz = ctx.compress_batch(fs, 1e-3, "rel", interp=interp)
    it = ctx.compress_batch(fs, 1e-3, "rel", interp=1)
    mixed = [lz[0], it[1], lz[2], it[0]]
    outs = [torch.empty_like(fs[i]) for i in (0, 1, 2, 0)]
    ctx.decompress_batch(mixed, outs)
    for b, o in zip(mixed, outs):
        _assert_bitexact(o.cpu().numpy(), b.decompress(), "mixed batch decode")
    for t, b in zip(fs, lz):
        single = ctx.compress_device(t, 1e-3, "rel", interp=interp)
        pb = b.download(-1)
        assert np.array_equal(pb, single.download(-1))
        assert first_difference(pb, oracle.compress(t.cpu().numpy(), 1e-3, 1, interp=interp)) is None
        single.free()
    for b in lz + it:
        b.free()
