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
