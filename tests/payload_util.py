"""Payload inspection helpers for the parity tests (test infrastructure)."""
import numpy as np

SECTIONS = ["codebook", "chunk_bit", "chunk_unpred", "chunk_sub", "sub_bits", "sub_unpred", "unpred", "bits"]


def stream_table(payload):
    h = np.frombuffer(payload[:512].tobytes(), dtype=np.uint8)
    u64 = h.view(np.uint64)
    so = u64[28:43]  # FileHeader.stream_off, bytes 224..344
    sb = u64[43:58]  # FileHeader.stream_bytes, bytes 344..464
    return [int(v) for v in so], [int(v) for v in sb]


def stream_sections(payload, s):
    so, sb = stream_table(payload)
    base = so[s]
    sh = np.frombuffer(payload[base:base + 160].tobytes(), dtype=np.uint8)
    u64 = sh.view(np.uint64)
    offs = [int(v) for v in u64[11:19]] + [int(u64[19])]
    return base, offs


def first_difference(a, b):
    """Describe where two payloads first differ (stream and section)."""
    a = np.asarray(a, dtype=np.uint8)
    b = np.asarray(b, dtype=np.uint8)
    if a.size != b.size:
        msg = "sizes differ: %d vs %d" % (a.size, b.size)
    else:
        msg = ""
    n = min(a.size, b.size)
    d = np.nonzero(a[:n] != b[:n])[0]
    if d.size == 0:
        return msg or None
    i = int(d[0])
    if i < 512:
        return msg + " first diff at %d (file header)" % i
    so, sb = stream_table(a)
    for s in range(15):
        if so[s] <= i < so[s] + sb[s]:
            base, offs = stream_sections(a, s)
            rel = i - base
            sec = "stream header"
            for k in range(8):
                if offs[k] <= rel < offs[k + 1]:
                    sec = SECTIONS[k]
            return msg + " first diff at %d: stream %d section %s (+%d), %d bytes differ" % (
                i, s, sec, rel, d.size)
    return msg + " first diff at %d" % i


def _clear_nan_sign(words, bits):
    """Sign bit cleared in place for every NaN bit pattern of a uint32/uint64 view."""
    exp = np.uint64((1 << (bits - 1)) - 1) if bits == 64 else np.uint32((1 << 31) - 1)
    inf = np.uint64(0x7FF0000000000000) if bits == 64 else np.uint32(0x7F800000)
    nan = (words & exp) > inf
    words[nan] &= exp


def nan_sign_cleared(payload, dtype):
    """A copy of a payload whose unpredictable-value sections hold every NaN
    with its sign bit cleared.  A subtraction whose subtrahend is NaN keeps
    the NaN's sign on x86 and flips it on the GPU (its subtract is an add with
    a negated operand), so fields holding NaNs differ in that bit only."""
    out = np.array(payload, dtype=np.uint8, copy=True)
    size = np.dtype(dtype).itemsize
    so, sb = stream_table(out)
    for s in range(15):
        if sb[s] == 0:
            continue
        base, offs = stream_sections(out, s)
        lo, hi = base + offs[6], base + offs[7]
        n = (hi - lo) // size
        if n == 0:
            continue
        assert lo % size == 0, "unaligned unpredictable section"
        view = out[lo:lo + n * size].view(np.uint64 if size == 8 else np.uint32)
        _clear_nan_sign(view, 8 * size)
    return out


def nan_sign_cleared_values(a):
    """A copy of a float array with every NaN's sign bit cleared."""
    a = np.array(a, copy=True)
    w = a.view(np.uint64 if a.dtype == np.float64 else np.uint32)
    _clear_nan_sign(w.reshape(-1), 8 * a.dtype.itemsize)
    return a


def requested_bound(f, eb, mode):
    """The maximum pointwise error an ABS / REL compress guarantees: eb, or
    eb times the value range of the finite values (in double)."""
    if mode == "abs":
        return float(eb)
    g = np.asarray(f, dtype=np.float64)
    g = g[np.isfinite(g)]
    return float(eb) * (float(g.max()) - float(g.min())) if g.size else 0.0


def max_error(rec, f):
    """max |rec - f| over the finite points of f (non-finite points must match)."""
    a = np.asarray(rec, dtype=np.float64)
    b = np.asarray(f, dtype=np.float64)
    fin = np.isfinite(b)
    assert np.array_equal(np.isnan(a[~fin]), np.isnan(b[~fin]))
    return float(np.max(np.abs(a[fin] - b[fin]))) if fin.any() else 0.0
