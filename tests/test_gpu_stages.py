"""GPU parity of the per-stream stage entry points against the oracle:

* stz_residual_encode / stz_residual_decode (one block's predict + quantize
  and its inverse) against the oracle's preprocess, quantizer, dequantizer
  and depreprocess;
* stz_huff_encode_chunks: the codes of a high (or low) block of a field,
  encoded as one standalone stream, equal that block's stream inside the
  oracle's payload byte for byte (all but the two header words that name
  the stream's kind and block);
* stz_huff_decode_chunks: every chunk, and row-major chunk lists, decode to
  the oracle's dequantized values at the listed chunks' points only, with
  the reference's symbol and zero counts (F5:1841-1911).
"""
import numpy as np
import pytest

from payload_util import stream_table

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _eq(got, want, what):
    g, w = _bits(got).ravel(), _bits(want).ravel()
    assert g.shape == w.shape, what
    bad = np.flatnonzero(g != w)
    assert bad.size == 0, "%s: %d differ, first at %d" % (what, bad.size, bad[0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(9, 7, 11), (4, 4, 4), (2, 3, 1), (17, 12, 20)])
def test_residual_encode_decode(ctx, oracle, dtype, dims):
    rng = np.random.default_rng(31)
    ref = np.cumsum(rng.normal(size=dims), axis=2).astype(dtype)
    sub = (ref + rng.normal(scale=0.1, size=dims)).astype(dtype)
    sub.ravel()[::37] = 1e6          # unpredictable residuals
    for b in range(7):
        for eb in (1e-3, 0.05):
            codes, un, rec = ctx.residual_encode(b, sub, ref, eb)
            d = oracle.preprocess(b, sub, ref)
            wc, wu, _ = oracle.quantize(d, eb)
            np.testing.assert_array_equal(codes.ravel(), wc, err_msg="codes b%d" % b)
            _eq(un, wu, "unpred b%d" % b)
            dec = oracle.dequantize(wc, wu, eb, dtype).reshape(dims)
            _eq(rec, oracle.depreprocess(b, dec, ref), "rec b%d" % b)
            out = ctx.residual_decode(b, codes, un, ref, eb)
            _eq(out, oracle.depreprocess(b, dec, ref), "decode b%d" % b)


def _level_case(oracle, shape, dtype, seed, **params):
    """A field (dims multiples of 4), its oracle payload, and per level the
    targets and reconstructed predictor grid of every block."""
    from stz_amd import fields
    f = fields.sinusoids(shape=shape, seed=seed, dtype=dtype)
    f.ravel()[::53] += 50.0          # a few unpredictable values
    want = oracle.compress(f, 1e-3, oracle.EB_REL, **params)
    h = oracle.header(want)
    subs = oracle.split(f)
    mid = oracle.decompress_level(want, 2)         # reconstructed sub-block 0
    base = oracle.decompress_level(want, 1)        # reconstructed base level
    lows = oracle.split(np.ascontiguousarray(subs[0]))
    return f, want, h, subs, mid, base, lows


def _stream_bytes(payload, s):
    so, sb = stream_table(payload)
    return payload[so[s]:so[s] + sb[s]]


def _same_but_kind(got, want):
    """Equal bytes except StreamHeader.kind and .block (bytes 4..11)."""
    assert got.size == want.size, (got.size, want.size)
    diff = np.flatnonzero(got != want)
    assert np.all((diff >= 4) & (diff < 12)), diff[:8]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_huff_encode_chunks_equals_payload_streams(ctx, oracle, dtype):
    f, want, h, subs, mid, base, lows = _level_case(oracle, (48, 40, 56), dtype, 3, chunk_high=8, chunk_low=4)
    for b in range(7):
        # high level: block b of the full field, predicted from the mid grid
        codes, un, _ = ctx.residual_encode(b, subs[b + 1], mid, h["eb_high"], with_rec=False)
        s = ctx.huff_encode_chunks(codes, un, h["eb_high"], 8, dtype)
        _same_but_kind(s, _stream_bytes(want, 8 + b))
        # low level: block b of sub-block 0, predicted from the base
        codes, un, _ = ctx.residual_encode(b, lows[b + 1], base, h["eb_low"], with_rec=False)
        s = ctx.huff_encode_chunks(codes, un, h["eb_low"], 4, dtype)
        _same_but_kind(s, _stream_bytes(want, 1 + b))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims,chunk", [((40, 24, 36), 8), ((33, 17, 9), 16), ((5, 3, 7), 4), ((64, 64, 64), 32)])
def test_huff_decode_chunks(ctx, oracle, dtype, dims, chunk):
    rng = np.random.default_rng(sum(dims) + chunk)
    dz, dy, dx = dims
    d = rng.normal(scale=rng.choice([0.01, 1.0, 30.0]), size=dims).astype(dtype)
    d.ravel()[::41] = np.inf        # unpredictable
    eb = 0.01
    wc, wu, _ = oracle.quantize(d, eb)
    codes = wc.reshape(dims)
    vals = oracle.dequantize(wc, wu, eb, dtype).reshape(dims)
    s = ctx.huff_encode_chunks(codes, wu, eb, chunk, dtype)
    ncx, ncy, ncz = [(v + chunk - 1) // chunk for v in (dx, dy, dz)]
    nch = ncx * ncy * ncz
    # every chunk (SZ_decompress_separated)
    out, co, ns, nz = ctx.huff_decode_chunks(s, (dx, dy, dz), dtype, codes_out=np.zeros(dims, np.uint16))
    _eq(out, vals, "all chunks")
    np.testing.assert_array_equal(co, codes)
    assert ns == d.size and nz == int((wc == 0).sum())
    # row-major chunk lists: only their points are written
    for ids in ([0], [nch - 1], list(range(0, nch, 3)), [nch - 1, 0], []):
        sentinel = np.full(dims, -7.25, dtype)
        out, _, ns, nz = ctx.huff_decode_chunks(s, (dx, dy, dz), dtype, chunk_ids=ids, out=sentinel)
        mask = np.zeros(dims, bool)
        for c in ids:
            cx, cy, cz = c % ncx, (c // ncx) % ncy, c // (ncx * ncy)
            mask[cz * chunk:(cz + 1) * chunk, cy * chunk:(cy + 1) * chunk, cx * chunk:(cx + 1) * chunk] = True
        _eq(out[mask], vals[mask], "chunks %s" % ids[:4])
        assert np.all(out[~mask] == np.asarray(-7.25, dtype))
        assert ns == int(mask.sum()) and nz == int((codes[mask] == 0).sum())


def test_stage_argument_errors(stz, ctx):
    d = np.zeros((4, 4, 4), np.float32)
    codes = np.full((4, 4, 4), 32768, np.uint16)
    with pytest.raises(stz.StzError):
        ctx.residual_encode(7, d, d, 1e-3)
    with pytest.raises(stz.StzError):     # n_unpred differs from the zero codes
        ctx.huff_encode_chunks(codes, np.zeros(3, np.float32), 1e-3, 4, np.float32)
    s = ctx.huff_encode_chunks(codes, np.zeros(0, np.float32), 1e-3, 4, np.float32)
    with pytest.raises(stz.StzError):
        ctx.huff_decode_chunks(s, (4, 4, 4), np.float32, chunk_ids=[1])
    bad = s.copy()
    bad[4] = 2                            # not a block stream
    with pytest.raises(stz.StzError):
        ctx.huff_decode_chunks(bad, (4, 4, 4), np.float32)
