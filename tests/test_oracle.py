"""CPU checks that pin the oracle (oracle/stz_oracle.c via oracle/pyoracle.py).

1. Golden vectors written by the reference's own functions (tests/golden/
   ref_vectors.npz, made by tests/golden/make_golden.py from
   oracle/_ref/libstz_ref.so): residual predictors for all 7 blocks
   (F5:141-882, F5:1145-1191), the quantizer (F5:1204-1253), the parity split
   (F5:107-122) and the ROI planner (F5:1602-1830).
2. The live reference build, when oracle/_ref is present, on more shapes.
3. What the mathematics fixes, for the parts with no reference counterpart
   here (the base codec and the Huffman code, whose reference lives in the
   absent SZ3-ncp): polynomial reproduction, an independent level-by-level
   evaluator of the base codec, Huffman optimality against a textbook heap
   construction and brute force, Kraft equality, error bounds, ROI == crop.
"""
import ctypes as C
import heapq
import itertools
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD_PATH = os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libstz_ref.so")
N_SHAPES = 5


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD_PATH)


@pytest.fixture(scope="module")
def pred_gold():
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_pred_vectors.npz"))


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


# ---------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("suf", ["f32", "f64"])
def test_golden_residual_predictors(oracle, gold, suf):
    for si in range(N_SHAPES):
        sub = gold["res_%s_%d_sub" % (suf, si)]
        ref = gold["res_%s_%d_ref" % (suf, si)]
        de = gold["res_%s_%d_de" % (suf, si)]
        for b in range(7):
            got = oracle.preprocess(b, sub, ref)
            assert np.array_equal(_bits(got), _bits(gold["res_%s_%d_pre%d" % (suf, si, b)])), (suf, si, b)
            got = oracle.depreprocess(b, de, ref)
            assert np.array_equal(_bits(got), _bits(gold["res_%s_%d_depre%d" % (suf, si, b)])), (suf, si, b)


@pytest.mark.parametrize("suf", ["f32", "f64"])
def test_golden_quantizer(oracle, gold, suf):
    d = gold["q_%s_in" % suf]
    for ei in range(3):
        eb = float(gold["q_%s_%d_eb" % (suf, ei)])
        codes, unpred, _ = oracle.quantize(d, eb)
        assert np.array_equal(codes.astype(np.int32), gold["q_%s_%d_codes" % (suf, ei)]), (suf, eb)
        assert np.array_equal(_bits(unpred), _bits(gold["q_%s_%d_unpred" % (suf, ei)])), (suf, eb)


def test_golden_split_merge(oracle, gold):
    for si in range(2):
        full = gold["split_%d_full" % si]
        subs = oracle.split(full)
        assert np.array_equal(np.stack(subs), gold["split_%d_subs" % si])
        assert np.array_equal(oracle.merge(subs), full)


def test_golden_planner(oracle, gold):
    pin = gold["plan_in"]
    hoff = loff = 0
    hall, lall = gold["plan_high_chunks"], gold["plan_low_chunks"]
    for k in range(pin.shape[0]):
        dims = tuple(int(v) for v in pin[k, :3])
        ch, cl = int(pin[k, 3]), int(pin[k, 4])
        q = [int(v) for v in pin[k, 5:11]]
        qa, is_slice, high, low = oracle.plan_query(dims, q, ch, cl, 4096)
        gq = gold["plan_query"][k]
        assert qa == tuple(int(v) for v in gq[:6]) and is_slice == bool(gq[6]), k
        for plans, table, allc, off in ((high, gold["plan_high"][k], hall, hoff), (low, gold["plan_low"][k], lall, loff)):
            for b in range(7):
                row = table[b]
                assert bool(row[0]) == (b in plans), (k, b)
                if row[0]:
                    assert plans[b][0] == tuple(int(v) for v in row[1:7]), (k, b)
                    n = int(row[7])
                    assert np.array_equal(plans[b][1], allc[off:off + n]), (k, b)
                    off += n
            if plans is high:
                hoff = off
            else:
                loff = off


# ---------------------------------------------------------------- live reference build
@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_PATH):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = C.CDLL(REF_PATH)
    R.ref_plan_query.argtypes = [C.c_int] * 11 + [C.c_void_p] * 4 + [C.c_void_p, C.c_int64]
    return R


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_reference_predictors_live(oracle, ref, dt):
    suf = "f32" if dt == np.float32 else "f64"
    rng = np.random.default_rng(77)
    for shape in [(17, 9, 33), (5, 31, 8), (20, 20, 20), (1, 3, 4)]:
        dz, dy, dx = shape
        r = np.cumsum(rng.normal(size=shape), axis=1).astype(dt)
        s = (r + rng.normal(scale=0.3, size=shape)).astype(dt)
        out = np.empty_like(s)
        for b in range(7):
            getattr(ref, "ref_preprocess_" + suf)(b, _p(s), _p(r), _p(out), dx, dy, dz)
            assert np.array_equal(_bits(oracle.preprocess(b, s, r)), _bits(out)), (shape, b)
            getattr(ref, "ref_depreprocess_" + suf)(b, _p(s), _p(r), _p(out), dx, dy, dz)
            assert np.array_equal(_bits(oracle.depreprocess(b, s, r)), _bits(out)), (shape, b)


# F5 built with its predictor switches (oracle/build_ref.sh): name -> Params.pred bits
PRED_VARIANTS = {"wpow2": 1, "plane": 2, "cube": 4, "all": 7}


def _pred_inputs(rng, shape, dt):
    """Rough references with a zero plane and a scaled row: every boundary
    rule, the flush-to-zero mode and wide variance ranges in the weights."""
    r = np.cumsum(rng.normal(size=shape), axis=1)
    r[:, :, 2::5] *= 40.0
    s = r + rng.normal(scale=0.3, size=shape)
    r = r.astype(dt)
    s = s.astype(dt)
    r[0] = 0
    return s, r


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("name", sorted(PRED_VARIANTS))
def test_reference_predictor_options_live(oracle, dt, name):
    """The oracle's PRED_WPOW2 / PRED_PLANE_AXIS / PRED_CUBE_AXIS predictors
    against F5 compiled with those switches, bit for bit, all 7 blocks (the
    switches change blocks 2, 4, 5 and 6 only)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libstz_ref_%s.so" % name)
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = C.CDLL(path)
    suf = "f32" if dt == np.float32 else "f64"
    bits = PRED_VARIANTS[name]
    rng = np.random.default_rng(91 + bits)
    for shape in [(17, 9, 33), (6, 7, 8), (20, 20, 20), (1, 3, 4), (4, 4, 4)]:
        dz, dy, dx = shape
        s, r = _pred_inputs(rng, shape, dt)
        out = np.empty_like(s)
        for b in range(7):
            getattr(R, "ref_preprocess_" + suf)(b, _p(s), _p(r), _p(out), dx, dy, dz)
            got = oracle.preprocess(b, s, r, pred=bits)
            assert np.array_equal(_bits(got), _bits(out)), (name, shape, b)
            getattr(R, "ref_depreprocess_" + suf)(b, _p(s), _p(r), _p(out), dx, dy, dz)
            assert np.array_equal(_bits(oracle.depreprocess(b, s, r, pred=bits)), _bits(out)), (name, shape, b)
        # the option changes something: blocks 2/4/5 (plane, wpow2) or 6 (cube, wpow2)
        if shape == (20, 20, 20):
            changed = [b for b in range(7) if not np.array_equal(_bits(oracle.preprocess(b, s, r, pred=bits)),
                                                                  _bits(oracle.preprocess(b, s, r)))]
            want = {1: [2, 4, 5, 6], 2: [2, 4, 5], 4: [6], 7: [2, 4, 5, 6]}[bits]
            assert changed == want, (name, changed)


def test_reference_stz2_predictors_live(oracle):
    """The oracle's stz2-style predictors (option bits 8: 12-point truncated
    bicubic for blocks 2/4/5; 16: 16 points for block 6) against stz2.cpp's
    own preprocess_block (stz2.cpp:82-306, built by oracle/build_ref.sh), bit
    for bit in f32 on blocks 2, 4, 5 and 6 (its blocks 0/1/3 use other
    weights, which the option leaves at F5's)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libstz_ref_stz2.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = C.CDLL(path)
    rng = np.random.default_rng(5)
    for shape in [(17, 9, 33), (6, 7, 8), (20, 20, 20), (4, 4, 4)]:
        dz, dy, dx = shape
        s, r = _pred_inputs(rng, shape, np.float32)
        out = np.empty_like(s)
        for b in (2, 4, 5, 6):
            R.ref2_preprocess_f32(b, _p(s), _p(r), _p(out), dx, dy, dz)
            got = oracle.preprocess(b, s, r, pred=8 | 16)
            assert np.array_equal(_bits(got), _bits(out)), (shape, b)
        # option 8 alone leaves block 6 dynamic, option 16 alone blocks 2/4/5
        assert np.array_equal(_bits(oracle.preprocess(6, s, r, pred=8)), _bits(oracle.preprocess(6, s, r)))
        assert np.array_equal(_bits(oracle.preprocess(2, s, r, pred=16)), _bits(oracle.preprocess(2, s, r)))


def test_stz2_predictors_reproduce_polynomials(oracle):
    """Closed-form checks of the two stz2 predictors on exactly representable
    data: both are weighted sums with weights summing to 1 (the bicubic's
    0.31640625 x 4 - 0.033203125 x 8, the 16-point's 0.140625 x 8 -
    0.015625 x 8), so constants are reproduced; and the values they give at
    a point equal the closed forms for a quadratic in each axis."""
    for dt in (np.float32, np.float64):
        c = np.full((8, 8, 8), 3.0, dt)
        for b in (2, 4, 5, 6):
            d = oracle.preprocess(b, c, c, pred=8 | 16)
            assert np.all(d == 0), (dt, b)
        z, y, x = np.meshgrid(np.arange(8.0), np.arange(8.0), np.arange(8.0), indexing="ij")
        ref = (x * x + 2 * y * y + 3 * z * z).astype(dt)
        sub = np.zeros_like(ref)
        i = (3, 3, 3)          # interior point (z, y, x)
        # bicubic12 in the (x, y) plane at target offset (+1/2, +1/2):
        # core points (0|1, 0|1), edges one step outside on one axis
        want2 = 0.31640625 * (ref[3, 3, 3] + ref[3, 3, 4] + ref[3, 4, 3] + ref[3, 4, 4]) - 0.033203125 * (
            ref[3, 3, 2] + ref[3, 3, 5] + ref[3, 4, 2] + ref[3, 4, 5] + ref[3, 2, 3] + ref[3, 2, 4] + ref[3, 5, 3] +
            ref[3, 5, 4])
        got2 = -oracle.preprocess(2, sub, ref, pred=8)[i]
        assert abs(got2 - want2) <= 1e-5 * abs(want2), (dt, got2, want2)
        want6 = 0.140625 * (ref[3, 3, 3] + ref[4, 4, 4] + ref[4, 4, 3] + ref[4, 3, 4] + ref[3, 4, 4] + ref[3, 3, 4] +
                             ref[3, 4, 3] + ref[4, 3, 3]) - 0.015625 * (
            ref[2, 2, 2] + ref[5, 5, 5] + ref[5, 2, 5] + ref[5, 5, 2] + ref[2, 5, 5] + ref[5, 2, 2] + ref[2, 5, 2] +
            ref[2, 2, 5])
        got6 = -oracle.preprocess(6, sub, ref, pred=16)[i]
        assert abs(got6 - want6) <= 1e-5 * abs(want6), (dt, got6, want6)


def _parity_blocks(g):
    """The 7 parity blocks of a level's 2x grid, in block order (block b at
    the parity of sub-block b + 1: x = bit 0, y = bit 1, z = bit 2)."""
    return np.stack([g[(k >> 2)::2, ((k >> 1) & 1)::2, (k & 1)::2] for k in range(1, 8)])


def test_golden_staged_level(oracle, pred_gold):
    """The oracle's staged level decode (option bit 32) against dependency.cpp's
    own depreprocess_block / depreprocess_block_staged composed in its order
    (dependency.cpp:594-636; vectors by tests/golden/make_golden.py from
    oracle/_ref/libstz_ref_dep.so), bit for bit: every boundary rule, the
    inverse-variance blends and which reconstructed block feeds which."""
    n = 0
    for key in pred_gold.files:
        if not (key.startswith("staged-") and key.endswith("-ref")):
            continue
        stem = key[:-4]
        ref, res = pred_gold[stem + "-ref"], pred_gold[stem + "-res"]
        g = oracle.staged_level(ref, residuals=res)
        assert np.array_equal(_bits(_parity_blocks(g)), _bits(pred_gold[stem + "-recs"])), stem
        assert np.array_equal(_bits(g[::2, ::2, ::2]), _bits(ref)), stem
        n += 1
    assert n >= 5


def test_reference_staged_level_live(oracle):
    """The same against the live dependency.cpp build on more shapes, rough
    residuals (denormal-range and huge values included)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libstz_ref_dep.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = C.CDLL(path)
    rng = np.random.default_rng(11)
    for shape in [(3, 3, 3), (8, 5, 2), (13, 17, 9), (20, 16, 24)]:
        dz, dy, dx = shape
        _, ref = _pred_inputs(rng, shape, np.float32)
        res = rng.normal(size=(7,) + shape).astype(np.float32)
        res[:, ::3] *= 1e-39
        res[:, :, 1::3] *= 1e20
        rec = [np.empty(shape, np.float32) for _ in range(7)]
        for b in (0, 1, 3):
            R.ref3_depreprocess_f32(b, _p(np.ascontiguousarray(res[b])), _p(ref), _p(rec[b]), dx, dy, dz)
        rb = (C.c_void_p * 7)(*[r.ctypes.data for r in rec])
        for b in (2, 4, 5, 6):
            R.ref3_depreprocess_staged_f32(b, _p(np.ascontiguousarray(res[b])), rb, _p(rec[b]), dx, dy, dz)
        g = oracle.staged_level(ref, residuals=res)
        assert np.array_equal(_bits(_parity_blocks(g)), _bits(np.stack(rec))), shape


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_staged_level_closed_loop(oracle, dt):
    """Encoding a level with the staged predictors and decoding its decoded
    residuals rebuild the same grid bit for bit (each block is predicted from
    reconstructions, not originals), and every coded point is within the
    bound.  The f64 form is the same template as the reference-pinned f32."""
    rng = np.random.default_rng(12)
    shape = (9, 11, 14)
    _, ref = _pred_inputs(rng, shape, dt)
    tgt = (ref[None] + rng.normal(scale=0.7, size=(7,) + shape)).astype(dt)
    tgt[3, 2, 2, 2] = np.nan
    eb = 0.05
    g, q, dv = oracle.staged_level(ref, eb=eb, targets=tgt)
    g2 = oracle.staged_level(ref, residuals=dv)
    assert np.array_equal(_bits(g), _bits(g2))
    rec = _parity_blocks(g)
    coded = q != 0
    assert coded.mean() > 0.9
    assert np.all(np.abs(rec[coded].astype(np.float64) - tgt[coded]) <= eb * (1 + 1e-6) + 1e-12)
    assert q[3, 2, 2, 2] == 0


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_staged_payload_roundtrip(oracle, dt):
    """A whole payload with the staged predictors (option bit 32): the
    requested bound holds, ROI / slice outputs equal the crop of the full
    decode, the 1/2 level is the decoded mid grid, and the option cannot be
    combined with another bit."""
    rng = np.random.default_rng(13)
    z, y, x = np.meshgrid(np.arange(22), np.arange(30), np.arange(37), indexing="ij")
    f = (np.sin(x / 5.0) * np.cos(y / 7.0) + 0.3 * np.sin(z / 3.0) + rng.normal(scale=1e-3, size=x.shape)).astype(dt)
    eb = 1e-3
    p = oracle.compress(f, eb, oracle.EB_REL, pred=oracle.PRED_STAGED)
    assert oracle.header(p)["pred"] == oracle.PRED_STAGED
    d = oracle.decompress(p)
    assert np.max(np.abs(d.astype(np.float64) - f)) <= eb * (float(f.max()) - float(f.min()))
    plain = oracle.compress(f, eb, oracle.EB_REL)
    assert plain.size != p.size or not np.array_equal(plain, p)
    roi = oracle.decompress_roi(p, 3, 5, 2, 30, 21, 17)
    assert np.array_equal(_bits(roi), _bits(d[2:18, 5:22, 3:31]))
    mid = oracle.decompress_level(p, 2)
    assert mid.shape == (11, 15, 19)
    with pytest.raises(Exception):
        oracle.compress(f, eb, oracle.EB_REL, pred=oracle.PRED_STAGED | oracle.PRED_WPOW2)


def test_golden_predictor_options(oracle, pred_gold):
    """The same against the committed vectors of those reference builds."""
    for key in pred_gold.files:
        if not key.endswith("_out"):
            continue
        name, dtn, si, b = key[:-4].split("-")
        s = pred_gold["%s-%s-%s-sub" % (name, dtn, si)]
        r = pred_gold["%s-%s-%s-ref" % (name, dtn, si)]
        bits = 8 | 16 if name == "stz2" else PRED_VARIANTS[name]
        got = oracle.preprocess(int(b), s, r, pred=bits)
        assert np.array_equal(_bits(got), _bits(pred_gold[key])), key


def test_reference_planner_live(oracle, ref):
    rng = np.random.default_rng(3)
    cap = 4096
    for it in range(400):
        ch, cl = [(32, 16), (8, 4), (4, 4), (16, 8)][it % 4]
        lcm = 4 * cl * 2 * ch // math.gcd(4 * cl, 2 * ch)
        dims = [lcm * int(rng.integers(1, 4)) for _ in range(3)]
        q = [int(rng.integers(-3, d + 3)) for d in dims] * 1
        q += [int(rng.integers(-3, d + 3)) for d in dims]
        if it % 5 == 0:
            q[3] = q[0]
        qa, sl, high, low = oracle.plan_query(tuple(dims), q, ch, cl, cap)
        query = np.zeros(7, np.int32)
        hp = np.zeros((7, 8), np.int32)
        lp = np.zeros((7, 8), np.int32)
        hc = np.zeros((7, cap), np.int64)
        lc = np.zeros((7, cap), np.int64)
        ref.ref_plan_query(*dims, ch, cl, *q, _p(query), _p(hp), _p(hc), _p(lp), _p(lc), cap)
        assert qa == tuple(int(v) for v in query[:6]) and sl == bool(query[6])
        for plans, P, Cc in ((high, hp, hc), (low, lp, lc)):
            assert set(plans) == {b for b in range(7) if P[b, 0]}
            for b in plans:
                assert plans[b][0] == tuple(int(v) for v in P[b, 1:7])
                assert np.array_equal(plans[b][1], Cc[b, :P[b, 7]])


# ---------------------------------------------------------------- quantizer properties
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_quantizer_properties(oracle, dt):
    rng = np.random.default_rng(11)
    eb = 1e-2
    r = 32768
    d = np.concatenate([rng.normal(size=20000) * 10, [0.0, np.nan, np.inf, -np.inf, (2 * r - 1) * eb * 1.01]]).astype(dt)
    codes, unpred, dec = oracle.quantize(d, eb)
    ok = codes != 0
    assert codes[ok].min() >= 1 and codes[ok].max() <= 2 * r - 1
    assert np.all(np.abs(dec[ok].astype(np.float64) - d[ok].astype(np.float64)) <= eb)
    assert codes[-5] == r                                   # 0 -> the centre code
    assert np.all(codes[-4:] == 0)                          # NaN, +-Inf, out of range -> unpredictable
    assert np.array_equal(_bits(unpred), _bits(d[~ok]))    # stored raw, in element order
    back = oracle.dequantize(codes, unpred, eb, dt)
    # equal as values; the encoder's dec keeps the residual's sign at code r (-0.0)
    assert np.array_equal(back[ok], dec[ok])
    q = codes[ok].astype(np.int64)                           # dequantized value is the closed form 2(q-r)eb
    assert np.array_equal(back[ok], (2.0 * (q - r) * eb).astype(dt))


# ---------------------------------------------------------------- Huffman code lengths
def _textbook_huffman_cost(freq):
    """Total bits of an optimal (unrestricted) prefix code, by the classic
    two-smallest merge: the sum of all internal node weights."""
    h = [int(f) for f in freq if f > 0]
    if len(h) == 1:
        return h[0]
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a = heapq.heappop(h)
        b = heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


def _cost(freq, lens):
    return int(np.sum(freq.astype(np.uint64) * lens.astype(np.uint64)))


def _kraft(lens, lmax=24):
    nz = lens[lens > 0].astype(np.int64)
    return int(np.sum(np.int64(1) << (lmax - nz)))


def test_huffman_optimal_and_complete(oracle):
    rng = np.random.default_rng(5)
    for n in [2, 3, 17, 300, 4096]:
        f = rng.integers(1, 1000, size=n).astype(np.uint64)
        f[rng.random(n) < 0.2] = 0
        if np.count_nonzero(f) < 2:
            f[:2] = 1
        L = oracle.huff_lengths(f)
        assert np.all((L == 0) == (f == 0))
        assert _cost(f, L) == _textbook_huffman_cost(f)
        assert _kraft(L) == 1 << 24


def test_huffman_single_symbol(oracle):
    f = np.zeros(100, np.uint64)
    f[42] = 9
    L = oracle.huff_lengths(f)
    assert L[42] == 1 and np.count_nonzero(L) == 1


def test_huffman_length_limit(oracle):
    # Fibonacci weights make the unrestricted tree 39 deep
    fib = [1, 1]
    while len(fib) < 40:
        fib.append(fib[-1] + fib[-2])
    f = np.array(fib, np.uint64)
    L = oracle.huff_lengths(f, 24)
    assert L.max() <= 24
    assert _kraft(L) <= 1 << 24
    assert _cost(f, L) >= _textbook_huffman_cost(f)


def _brute_force_best(f, lmax):
    best = None
    for lens in itertools.product(range(1, lmax + 1), repeat=len(f)):
        if sum(2.0 ** -l for l in lens) <= 1.0:
            c = sum(int(a) * b for a, b in zip(f, lens))
            best = c if best is None or c < best else best
    return best


def test_huffman_brute_force_optimal(oracle):
    """With no binding limit the lengths reach the minimum over every prefix code."""
    rng = np.random.default_rng(9)
    for n in [2, 3, 5, 6, 7]:
        for _ in range(4):
            f = rng.integers(1, 60, size=n).astype(np.uint64)
            L = oracle.huff_lengths(f, 24)
            assert _cost(f, L) == _brute_force_best(f, n - 1), (f, L)


def _annex_k3(bits, lmax):
    """ITU-T T.81 Annex K.3 (Figure K.3, Adjust_BITS) on per-length counts:
    while a code longer than lmax exists, take two symbols of the longest
    length i, give one prefix to the next shorter length and put the pair
    under a leaf taken from the longest length j < i - 1 that has one."""
    bits = list(bits)
    for i in range(len(bits) - 1, lmax, -1):
        while bits[i] > 0:
            j = i - 2
            while bits[j] == 0:
                j -= 1
            bits[i] -= 2
            bits[i - 1] += 1
            bits[j + 1] += 2
            bits[j] -= 1
    return bits


def test_huffman_length_limit_annex_k3(oracle):
    """A binding limit: the per-length counts are Annex K.3's adjustment of the
    unlimited code's, Kraft equality still holds, and lengths never grow as
    frequency rises (new lengths go out in old-length order)."""
    rng = np.random.default_rng(21)
    cases = []
    fib = [1, 1]
    while len(fib) < 40:
        fib.append(fib[-1] + fib[-2])
    cases.append((np.array(fib, np.uint64), 24))
    for n, lmax in [(6, 3), (8, 4), (12, 5), (300, 9)]:
        f = (rng.pareto(0.7, size=n) * 10 + 1).astype(np.uint64)
        cases.append((f, lmax))
    for f, lmax in cases:
        free = oracle.huff_lengths(f, 64)
        if free.max() <= lmax:
            continue
        L = oracle.huff_lengths(f, lmax)
        want = _annex_k3(np.bincount(free, minlength=int(free.max()) + 1), lmax)
        got = np.bincount(L, minlength=len(want))
        assert list(got[1:]) == list(want[1:len(got)]) and not any(want[len(got):])
        assert L.max() <= lmax
        assert sum(2.0 ** -int(l) for l in L) == 1.0
        assert _monotone_by_old_length(free, L)
        if len(f) <= 8:
            assert _cost(f, L) >= _brute_force_best(f, lmax)


def _monotone_by_old_length(free, L):
    o = np.lexsort((np.arange(len(free)), free))
    return bool(np.all(np.diff(L[o].astype(int)) >= 0))


# ---------------------------------------------------------------- base codec
ORDERS = [(2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2), (0, 2, 1), (0, 1, 2)]  # axis order, 0 = x


def _base_codec_levelwise(oracle, v, eb, interp, order):
    """Dyadic multilevel interpolation codec, written level by level with
    numpy slices: at level l (stride s = 2^(l-1)) each axis in turn predicts
    the points at odd multiples of s along it from neighbours at +-s and +-3s
    (cubic (-1, 9, 9, -1)/16 inside, 3-point or 2-point rules near the ends,
    or linear), quantizes the residual with the oracle's quantizer and
    writes the reconstruction back before the next axis.  Codes are emitted
    pass by pass, x fastest.  Also returns, per code, the rule that predicted
    it: 0 origin, 1 copy of the left neighbour, 2 two-point mean, 3 linear
    extrapolation, 4 four-point cubic, 5/6 three-point end rules."""
    T = v.dtype.type
    w = v.copy()
    shape = w.shape                      # (z, y, x)
    dims = [shape[2], shape[1], shape[0]]
    L = 0
    while (1 << L) < max(dims):
        L += 1
    codes_all, side_all, rules_all = [], [], []

    def q(vals, ebl):
        c, _, dec = oracle.quantize(np.ascontiguousarray(vals).ravel(), ebl)
        return c, dec

    c, dec = q(np.array([w[0, 0, 0]], dtype=v.dtype), eb)
    codes_all.append(c)
    rules_all.append(np.zeros(1, np.int8))
    if c[0]:
        w[0, 0, 0] = dec[0]
    else:
        side_all.append(np.array([v[0, 0, 0]]))
    for lev in range(L, 0, -1):
        s = 1 << (lev - 1)
        ordv = ORDERS[order]
        for pi, a in enumerate(ordv):
            idx = []
            for ax in range(3):
                rank = ordv.index(ax)
                if rank == pi:
                    idx.append(np.arange(s, dims[ax], 2 * s))
                elif rank < pi:
                    idx.append(np.arange(0, dims[ax], s))
                else:
                    idx.append(np.arange(0, dims[ax], 2 * s))
            ix, iy, iz = idx
            if min(len(ix), len(iy), len(iz)) == 0:
                continue
            ax_np = 2 - a                  # numpy axis of x/y/z
            dim = dims[a]

            def take(pos):
                sel = [iz, iy, ix]
                sel[ax_np] = np.array([pos])
                return w[np.ix_(*sel)]

            pred = np.empty((len(iz), len(iy), len(ix)), dtype=v.dtype)
            rule = np.empty(pred.shape, np.int8)
            for k, pos in enumerate(idx[a]):
                b = take(pos - s)
                r1, l3, r3 = pos + s < dim, pos >= 3 * s, pos + 3 * s < dim
                if interp == 1:
                    if r1:
                        cc = take(pos + s)
                        if l3 and r3:
                            p, rl = T(-0.0625) * take(pos - 3 * s) + T(0.5625) * b + T(0.5625) * cc - T(0.0625) * take(pos + 3 * s), 4
                        elif l3:
                            p, rl = T(-0.125) * take(pos - 3 * s) + T(0.75) * b + T(0.375) * cc, 5
                        elif r3:
                            p, rl = T(0.375) * b + T(0.75) * cc - T(0.125) * take(pos + 3 * s), 6
                        else:
                            p, rl = T(0.5) * (b + cc), 2
                    elif l3:
                        p, rl = T(1.5) * b - T(0.5) * take(pos - 3 * s), 3
                    else:
                        p, rl = b, 1
                else:
                    if r1:
                        p, rl = T(0.5) * (b + take(pos + s)), 2
                    elif l3:
                        p, rl = T(1.5) * b - T(0.5) * take(pos - 3 * s), 3
                    else:
                        p, rl = b, 1
                sl = [slice(None)] * 3
                sl[ax_np] = slice(k, k + 1)
                pred[tuple(sl)] = p
                rule[tuple(sl)] = rl
            orig = w[np.ix_(iz, iy, ix)]
            c, dec = q(orig - pred, eb)
            c = c.reshape(pred.shape)
            rec = np.where(c != 0, pred + dec.reshape(pred.shape), orig)
            side_all.append(orig[c == 0])
            codes_all.append(c.ravel())
            rules_all.append(rule.ravel())
            w[np.ix_(iz, iy, ix)] = rec
    side = np.concatenate(side_all) if side_all else np.zeros(0, v.dtype)
    return w, np.concatenate(codes_all), side, np.concatenate(rules_all)


@pytest.mark.parametrize("shape,dt,interp,order", [
    ((9, 7, 13), np.float64, 1, 0), ((8, 8, 8), np.float32, 1, 3), ((5, 12, 6), np.float32, 0, 5),
    ((17, 3, 10), np.float64, 0, 1), ((1, 6, 9), np.float32, 1, 2), ((16, 16, 16), np.float32, 1, 4),
])
def test_base_codec_levelwise(oracle, shape, dt, interp, order):
    rng = np.random.default_rng(sum(shape))
    v = (np.cumsum(rng.normal(size=shape), axis=0) + 3 * np.sin(np.arange(shape[2]) / 2.0)).astype(dt)
    eb = 0.05
    want_rec, want_codes, want_side, _ = _base_codec_levelwise(oracle, v, eb, interp, order)
    rec, codes, side = oracle.interp_compress(v, eb, interp, order)
    assert np.array_equal(codes, want_codes)
    assert np.array_equal(_bits(side), _bits(want_side))
    assert np.array_equal(_bits(rec), _bits(want_rec))
    assert np.abs(rec.astype(np.float64) - v).max() <= eb


@pytest.mark.parametrize("interp,order", [(0, 0), (1, 0), (1, 5), (0, 3)])
def test_base_codec_reproduces_linear(oracle, interp, order):
    """Every rule but the left-neighbour copy is exact on linear data."""
    z, y, x = np.meshgrid(np.arange(11), np.arange(13), np.arange(19), indexing="ij")
    v = (x + 2 * y - 3 * z).astype(np.float32)
    rec, codes, side = oracle.interp_compress(v, 0.25, interp, order)
    _, want, _, rules = _base_codec_levelwise(oracle, v, 0.25, interp, order)
    assert np.array_equal(codes, want)
    assert np.all(codes[rules != 1] == 32768)
    assert np.count_nonzero(rules == 1) > 0 and side.size == 0
    assert np.array_equal(rec, v)


def test_base_codec_cubic_reproduces_cubics(oracle):
    """On a cubic along x (constant along y, z) the four-point rule is exact."""
    x = np.arange(37, dtype=np.float64)
    v = np.broadcast_to(((x - 20) ** 3 / 512.0)[None, None, :], (5, 6, 37)).copy()
    _, codes, _ = oracle.interp_compress(v, 1e-9, 1, 0)
    _, want, _, rules = _base_codec_levelwise(oracle, v, 1e-9, 1, 0)
    assert np.array_equal(codes, want)
    assert np.count_nonzero(rules == 4) > 100
    assert np.all(codes[rules == 4] == 32768)
    assert np.any(codes[rules == 5] != 32768) or np.any(codes[rules == 6] != 32768)


# ---------------------------------------------------------------- Lorenzo base codec
def _tile_local(shape):
    """Per point (stream order) its coordinates inside its 16^3 tile, the
    tile's origin, and the point's grid coordinates."""
    bz, by, bx = shape
    T = 16
    out = []
    for tz in range(0, bz, T):
        for ty in range(0, by, T):
            for tx in range(0, bx, T):
                dz, dy, dx = min(T, bz - tz), min(T, by - ty), min(T, bx - tx)
                for z in range(dz):
                    for y in range(dy):
                        for x in range(dx):
                            out.append((x, y, z, tx + x, ty + y, tz + z))
    return np.array(out, dtype=np.int64)


@pytest.mark.parametrize("shape", [(21, 18, 35), (16, 16, 16), (5, 33, 17), (1, 1, 40)])
def test_lorenzo_base_closed_form_on_linear_field(oracle, shape):
    """f = 5 + 2x - 3y + z (integers) with eb = 0.5, where every residual is
    an integer and quantizes exactly: in a tile's interior and on its faces
    the 3D / 2D Lorenzo predictors are exact (code = radius); on an edge the
    prediction is the previous point on that axis (residual = the slope);
    a tile's corner is predicted by zero (residual = the value).  Pins the
    stencil, the 16^3 tiles, the zero outside a tile and the stream order."""
    bz, by, bx = shape
    z, y, x = np.meshgrid(np.arange(bz), np.arange(by), np.arange(bx), indexing="ij")
    v = (5 + 2 * x - 3 * y + z).astype(np.float64)
    rec, codes, side = oracle.interp_compress(v, 0.5, oracle.BASE_LORENZO)
    P = _tile_local(shape)
    lx, ly, lz = P[:, 0], P[:, 1], P[:, 2]
    val = 5 + 2 * P[:, 3] - 3 * P[:, 4] + P[:, 5]
    zeros = (lx == 0).astype(int) + (ly == 0) + (lz == 0)
    slope = np.where(lx > 0, 2, np.where(ly > 0, -3, 1))
    want = np.where(zeros == 3, val, np.where(zeros == 2, slope, 0)) + 32768
    assert np.array_equal(codes.astype(np.int64), want)
    assert side.size == 0 and np.array_equal(rec, v)


def test_lorenzo_base_exact_on_sums_of_two_axis_functions(oracle):
    """Inside a tile the residual of the 3D Lorenzo predictor is the mixed
    third difference: zero on g(x,y) + h(y,z) + k(x,z) for any g, h, k, and
    exactly 1 per point on xyz (integer data, eb = 0.5 so quantization is
    exact).  A wrong sign or a missing term of the seven-point stencil
    leaves a residual on the first field."""
    rng = np.random.default_rng(3)
    n = 20
    g, h, k = (rng.integers(-50, 50, size=(n, n)) for _ in range(3))
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = (g[x, y] + h[y, z] + k[x, z]).astype(np.float64)
    _, codes, _ = oracle.interp_compress(v, 0.5, oracle.BASE_LORENZO)
    P = _tile_local(v.shape)
    inner = (P[:, 0] > 0) & (P[:, 1] > 0) & (P[:, 2] > 0)
    assert inner.sum() > 3000 and np.all(codes[inner] == 32768)
    w = (x * y * z).astype(np.float64)
    _, codes, _ = oracle.interp_compress(w, 0.5, oracle.BASE_LORENZO)
    assert np.all(codes[inner] == 32769)


@pytest.mark.parametrize("shape,dt,eb", [((21, 18, 35), np.float32, 1e-2), ((40, 9, 33), np.float64, 1e-4),
                                         ((3, 4, 5), np.float32, 0.3)])
def test_lorenzo_base_round_trip_and_bound(oracle, shape, dt, eb):
    """decode(encode) is the encoder's reconstruction bit for bit, within eb;
    values beyond the code range are unpredictable and kept exactly."""
    rng = np.random.default_rng(sum(shape))
    v = np.cumsum(rng.normal(size=shape), axis=2).astype(dt)
    v.flat[7] = 1e6
    rec, codes, side = oracle.interp_compress(v, eb, oracle.BASE_LORENZO)
    assert side.size >= 1 and np.count_nonzero(codes == 0) == side.size
    got = oracle.interp_decompress(v.shape, dt, eb, codes, side, oracle.BASE_LORENZO)
    assert np.array_equal(_bits(got), _bits(rec))
    assert rec.flat[7] == v.flat[7]
    assert np.abs(rec.astype(np.float64) - v).max() <= eb


def test_lorenzo_base_whole_path(oracle):
    """interp = 2 through the whole oracle path: the header names it, the
    decode is within the bound, and the payload differs from interpolation's."""
    f = _field((40, 36, 44), np.float32, 9)
    p = oracle.compress(f, 1e-3, oracle.EB_REL, interp=2)
    assert oracle.header(p)["interp"] == 2
    r = oracle.decompress(p)
    from payload_util import max_error, requested_bound
    E = requested_bound(f, 1e-3, "rel")
    assert max_error(r, f) <= E
    assert not np.array_equal(p, oracle.compress(f, 1e-3, oracle.EB_REL, interp=1))
    with pytest.raises(Exception):
        oracle.compress(f, 1e-3, oracle.EB_REL, interp=4)


# ------------------------------------------------- Lorenzo / regression base codec
def _lzreg_layout(shape):
    """Stream positions of the Lorenzo/regression base: per tile (raster
    order) its 32 header symbols (4 per 8^3 sub-block, sub-block index
    x>=8 | (y>=8)<<1 | (z>=8)<<2), then its points; returns (header rows:
    tile origin, sub-block, extents, first position), and per point its
    position, tile-local and grid coordinates."""
    bz, by, bx = shape
    T = 16
    hdr, pts = [], []
    pos = 0
    for tz in range(0, bz, T):
        for ty in range(0, by, T):
            for tx in range(0, bx, T):
                dz, dy, dx = min(T, bz - tz), min(T, by - ty), min(T, bx - tx)
                for b in range(8):
                    sx, sy, sz = (b & 1) * 8, ((b >> 1) & 1) * 8, (b >> 2) * 8
                    hdr.append((tx, ty, tz, b, min(8, dx - sx), min(8, dy - sy), min(8, dz - sz), pos + 4 * b))
                pos += 32
                for z in range(dz):
                    for y in range(dy):
                        for x in range(dx):
                            pts.append((pos, x, y, z, tx + x, ty + y, tz + z))
                            pos += 1
    return np.array(hdr, dtype=np.int64), np.array(pts, dtype=np.int64)


@pytest.mark.parametrize("shape", [(21, 18, 35), (16, 16, 16), (5, 33, 17), (1, 1, 40)])
def test_lzreg_base_planes_on_linear_field(oracle, shape):
    """On f = 5 + 2.5x - 3.25y + 1.5z (exact in binary) every non-empty
    sub-block takes its plane: the header carries code 0 and the slopes and
    the sub-block origin's value (an axis of extent 1 gets slope 0, its
    d absorbing nothing), every point's residual is zero (code = radius), and
    the unpredictables are exactly those coefficients in stream order.  Pins
    the least-squares fit (a dropped centring term or a wrong sign moves d
    or a slope), the header layout, the tiles and sub-blocks, and the decode."""
    bz, by, bx = shape
    z, y, x = np.meshgrid(np.arange(bz), np.arange(by), np.arange(bx), indexing="ij")
    v = (5 + 2.5 * x - 3.25 * y + 1.5 * z).astype(np.float32)
    rec, codes, side = oracle.interp_compress(v, 1e-2, oracle.BASE_LZREG)
    H, P = _lzreg_layout(shape)
    assert codes.size == oracle.base_symbols(shape, oracle.BASE_LZREG) == len(P) + 4 * len(H)
    want_side = []
    for tx, ty, tz, b, nx, ny, nz, p0 in H:
        if nx <= 0 or ny <= 0 or nz <= 0:
            assert np.all(codes[p0:p0 + 4] == 32768)
            continue
        assert np.all(codes[p0:p0 + 4] == 0)
        ox, oy, oz = tx + (b & 1) * 8, ty + ((b >> 1) & 1) * 8, tz + (b >> 2) * 8
        want_side += [2.5 if nx > 1 else 0.0, -3.25 if ny > 1 else 0.0, 1.5 if nz > 1 else 0.0,
                      5 + 2.5 * ox - 3.25 * oy + 1.5 * oz]
    assert np.all(codes[P[:, 0]] == 32768)
    assert np.array_equal(side, np.array(want_side, np.float32))
    assert np.array_equal(rec, v)
    back = oracle.interp_decompress(v.shape, np.float32, 1e-2, codes, side, oracle.BASE_LZREG, lo=v.min(), hi=v.max())
    assert np.array_equal(back, v)


def test_lzreg_base_clamps_planes_to_the_range(oracle):
    """A plane prediction above hi is hi: with hi set below the field's top,
    the points whose exact plane value exceeds hi get residual v - hi
    (quantized exactly at eb = 0.5 on integer data), the others zero."""
    n = 16
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = (x + 2 * y + 3 * z).astype(np.float64)
    hi = 40.0
    rec, codes, side = oracle.interp_compress(v, 0.5, oracle.BASE_LZREG, lo=0.0, hi=hi)
    H, P = _lzreg_layout(v.shape)
    val = P[:, 4] + 2 * P[:, 5] + 3 * P[:, 6]
    reg = {int(b) for (_, _, _, b, _, _, _, p0) in H if codes[p0] == 0}
    sb = (P[:, 1] >= 8) | ((P[:, 2] >= 8) << 1) | ((P[:, 3] >= 8) << 2)
    on = np.isin(sb, sorted(reg))
    assert 0 in reg and on.sum() > 0
    q = codes[P[:, 0]].astype(np.int64) - 32768
    assert np.array_equal(q[on], np.where(val[on] > hi, val[on] - hi, 0))
    assert np.array_equal(rec, v)


def test_lzreg_base_keeps_lorenzo_where_it_is_exact(oracle):
    """On g(x,y) + h(y,z) + k(x,z) (random integers) the Lorenzo residual is
    zero wherever all seven neighbours lie in the tile, so a sub-block away
    from the tile's low faces (index 7) keeps Lorenzo (no plane fits such a
    field better than zero) and its points code zero residuals."""
    rng = np.random.default_rng(3)
    n = 32
    g, h, k = (rng.integers(-50, 50, size=(n, n)) for _ in range(3))
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = (g[x, y] + h[y, z] + k[x, z]).astype(np.float64)
    _, codes, _ = oracle.interp_compress(v, 0.5, oracle.BASE_LZREG)
    H, P = _lzreg_layout(v.shape)
    seven = H[H[:, 3] == 7]
    assert len(seven) == 8 and np.all(codes[seven[:, 7]] == 32768)
    inner = (P[:, 1] >= 8) & (P[:, 2] >= 8) & (P[:, 3] >= 8)
    assert inner.sum() == 8 * 512 and np.all(codes[P[inner, 0]] == 32768)


def test_lzreg_base_prefers_planes_on_noise(oracle):
    """Independent noise: a plane's error is about the noise, Lorenzo's
    prediction sums seven noisy neighbours (residual about sqrt(8) times the
    noise), so nearly every full sub-block takes its plane and the codes'
    entropy drops below the pure Lorenzo base's."""
    rng = np.random.default_rng(11)
    v = rng.normal(size=(32, 32, 32)).astype(np.float32)
    _, c2, _ = oracle.interp_compress(v, 1e-2, oracle.BASE_LORENZO)
    _, c3, s3 = oracle.interp_compress(v, 1e-2, oracle.BASE_LZREG)
    H, _ = _lzreg_layout(v.shape)
    assert np.mean(c3[H[:, 7]] == 0) > 0.9

    def bits(c):
        p = np.bincount(c, minlength=65536).astype(float)
        p = p[p > 0] / c.size
        return -(p * np.log2(p)).sum() * c.size

    assert bits(c3) < 0.9 * bits(c2)


@pytest.mark.parametrize("shape,dt,eb", [((21, 18, 35), np.float32, 1e-2), ((40, 9, 33), np.float64, 1e-4),
                                         ((3, 4, 5), np.float32, 0.3)])
def test_lzreg_base_round_trip_and_bound(oracle, shape, dt, eb):
    """decode(encode) is the encoder's reconstruction bit for bit, within eb."""
    rng = np.random.default_rng(sum(shape) + 1)
    v = (np.cumsum(rng.normal(size=shape), axis=2) + rng.normal(scale=3, size=shape)).astype(dt)
    v.flat[7] = 1e6
    rec, codes, side = oracle.interp_compress(v, eb, oracle.BASE_LZREG)
    got = oracle.interp_decompress(v.shape, dt, eb, codes, side, oracle.BASE_LZREG, lo=v.min(), hi=v.max())
    assert np.array_equal(_bits(got), _bits(rec))
    assert rec.flat[7] == v.flat[7]
    assert np.abs(rec.astype(np.float64) - v).max() <= eb


def test_lzreg_base_whole_path(oracle):
    """interp = 3 through the whole oracle path: the header names it, the
    base stream holds the header symbols, the decode is within the bound."""
    f = _field((40, 36, 44), np.float32, 9)
    p = oracle.compress(f, 1e-3, oracle.EB_REL, interp=3)
    assert oracle.header(p)["interp"] == 3
    r = oracle.decompress(p)
    from payload_util import max_error, requested_bound
    assert max_error(r, f) <= requested_bound(f, 1e-3, "rel")
    assert not np.array_equal(p, oracle.compress(f, 1e-3, oracle.EB_REL, interp=2))


# ---------------------------------------------------------------- whole oracle path
def _field(shape, dt, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(8):
        k = rng.integers(1, 4, size=3)
        f += rng.uniform(0.1, 1) * np.sin(2 * np.pi * (k[0] * x / shape[2] + k[1] * y / shape[1] + k[2] * z / shape[0])
                                          + rng.uniform(0, 6.28))
    return (f + rng.normal(0, 1e-3, size=shape)).astype(dt)


@pytest.mark.parametrize("shape,dt", [((32, 32, 32), np.float32), ((24, 40, 16), np.float64),
                                      ((13, 22, 9), np.float32), ((4, 4, 4), np.float32)])
def test_oracle_round_trip(oracle, shape, dt):
    f = _field(shape, dt, 1)
    p = oracle.compress(f, 1e-3, oracle.EB_REL)
    assert np.array_equal(p, oracle.compress(f, 1e-3, oracle.EB_REL))          # deterministic
    h = oracle.header(p)
    assert (h["nz"], h["ny"], h["nx"]) == shape
    r = oracle.decompress(p)
    E = 1e-3 * (float(f.max()) - float(f.min()))
    assert np.abs(r.astype(np.float64) - f.astype(np.float64)).max() <= E * (1 + 1e-9)
    assert h["eb_high"] <= E * (1 + 1e-12)
    # ROI and slices are crops of the full decode (F5:2609-2640)
    nz, ny, nx = shape
    for (x0, y0, z0, x1, y1, z1) in [(1, 2, 3, nx - 2, ny - 1, nz - 3), (0, 0, 0, nx - 1, ny - 1, nz - 1),
                                     (nx - 1, 0, 2, 0, ny - 1, 2), (0, ny // 2, 0, nx - 1, ny // 2, nz - 1)]:
        roi = oracle.decompress_roi(p, x0, y0, z0, x1, y1, z1)
        want = r[min(z0, z1):max(z0, z1) + 1, min(y0, y1):max(y0, y1) + 1, min(x0, x1):max(x0, x1) + 1]
        assert np.array_equal(_bits(roi), _bits(want))


def test_oracle_levels_are_subsamples(oracle):
    f = _field((32, 24, 40), np.float32, 2)
    p = oracle.compress(f, 1e-3, oracle.EB_REL)
    r = oracle.decompress(p)
    assert np.array_equal(oracle.decompress_level(p, 2), r[::2, ::2, ::2])
    assert np.array_equal(oracle.decompress_level(p, 1), r[::4, ::4, ::4])
    assert np.array_equal(oracle.decompress_level(p, 3), r)


def test_oracle_abs_mode_and_params(oracle):
    f = _field((16, 16, 32), np.float64, 3)
    p = oracle.compress(f, 1e-4, oracle.EB_ABS, interp=0, order=4, mult_low=2.0, chunk_high=8, chunk_low=4)
    r = oracle.decompress(p)
    assert np.abs(r - f).max() <= 1e-4 * (1 + 1e-12)


# ---------------------------------------------------------------- tuner MLP reference
def test_mlp_reference_gradient_by_finite_differences():
    """The float64 reference of the tuner's MLP, whose bf16-rounded variant
    checks the device kernels: its analytic gradient against central
    differences of its own loss."""
    import mlp_ref
    rng = np.random.default_rng(0)
    P = rng.normal(0, 0.3, size=mlp_ref.N_PARAMS)
    X = rng.normal(size=(24, 16))
    T = rng.dirichlet(np.ones(mlp_ref.OUT), size=24)
    _, g = mlp_ref.loss_grad(P, X, T, bf=False)
    eps = 1e-6
    blocks = np.cumsum([0, 16 * 64, 64, 64 * 64, 64, 64 * mlp_ref.OUT, mlp_ref.OUT])
    for lo, hi in zip(blocks[:-1], blocks[1:]):     # a few parameters of every block
        for i in rng.choice(np.arange(lo, hi), 6, replace=False):
            Pp, Pm = P.copy(), P.copy()
            Pp[i] += eps
            Pm[i] -= eps
            fd = (mlp_ref.loss_grad(Pp, X, T, False)[0] - mlp_ref.loss_grad(Pm, X, T, False)[0]) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_reference():
    import mlp_ref
    one = 1.0
    ulp = 2.0 ** -7
    cases = {one: one, one + ulp / 2: one, one + 3 * ulp / 2: one + 2 * ulp, one + ulp / 2 + 2 ** -20: one + ulp,
             -2.5: -2.5, 3.0e-39: 3.0e-39 - (3.0e-39 % 2 ** -133) if False else None}
    for x, want in cases.items():
        if want is None:
            continue
        assert mlp_ref.bf16(np.float32(x)) == want, x
    v = np.float32(np.pi)
    assert mlp_ref.bf16(v) == 3.140625          # 0x4049 (pi rounds down)


def test_mlp_reference_adam_matches_torch_adamw():
    """The reference Adam step (decoupled weight decay) against
    torch.optim.AdamW in float64 over several steps with changing gradients."""
    import torch

    import mlp_ref
    rng = np.random.default_rng(3)
    n = 257
    P = rng.normal(size=n)
    M = np.zeros(n)
    V = np.zeros(n)
    lr, wd = 3e-3, 1e-2
    t = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.AdamW([t], lr=lr, betas=(0.9, 0.999), eps=1e-8, weight_decay=wd)
    for step in range(1, 8):
        G = rng.normal(size=n) * (10.0 ** rng.integers(-3, 2))
        P, M, V = mlp_ref.adam(P, G, M, V, step, lr, wd)
        opt.zero_grad()
        t.grad = torch.tensor(G, dtype=torch.float64)
        opt.step()
        np.testing.assert_allclose(P, t.detach().numpy(), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dt,ebs", [(np.float32, (1e-2, 1e-3, 1e-4, 1e-5)), (np.float64, (1e-4, 1e-7, 1e-10))])
def test_oracle_requested_bound_is_strict(oracle, dt, ebs):
    """White noise (every residual a large fraction of the bound) and an
    offset field (|values| far above the range): the decoded error never
    exceeds the requested ABS / REL bound, without any rounding slack."""
    from payload_util import max_error, requested_bound
    rng = np.random.default_rng(17)
    noise = rng.standard_normal((24, 28, 32))
    for f in (noise.astype(dt), (1000.0 + noise).astype(dt)):
        for eb in ebs:
            for mode, code in (("rel", oracle.EB_REL), ("abs", oracle.EB_ABS)):
                e = eb if mode == "rel" else eb * float(f.max() - f.min())
                p = oracle.compress(f, e, code)
                err = max_error(oracle.decompress(p), f)
                assert err <= requested_bound(f, e, mode), (dt.__name__, eb, mode, err)
