"""CPU checks that a corrupt or crafted stream is refused on the host, before
any kernel could read through it (stz_stream_check runs the same container and
payload validation a decode runs before its first launch).  Payloads come from
the oracle; containers are assembled here with the system libzstd."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

from payload_util import stream_table

STZ_E_STREAM = -3
N_STREAMS = 15


def _payload(oracle):
    rng = np.random.default_rng(21)
    f = rng.normal(size=(20, 24, 36)).astype(np.float32)
    f[3, 4, 5] = np.nan  # at least one unpredictable value
    return oracle.compress(f, 1e-3, oracle.EB_REL)


def _rc(stz, buf):
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    return stz.lib().stz_stream_check(buf.ctypes.data_as(C.c_void_p), buf.size)


def test_oracle_payload_passes(stz, oracle):
    assert _rc(stz, _payload(oracle)) == 0


def test_every_stream_header_offset_checked(stz, oracle):
    p = _payload(oracle)
    so, _ = stream_table(p)
    # StreamHeader u64 words 11..19: off_codebook, off_chunk_bit, off_chunk_unpred,
    # off_chunk_sub, off_sub_bits, off_sub_unpred, off_unpred, off_bits, stream_bytes
    for s in range(N_STREAMS):
        for word in range(11, 20):
            for delta in (8, -8, 1 << 62):
                bad = p.copy()
                v = bad[so[s] + 8 * word: so[s] + 8 * word + 8].view(np.uint64)
                v[0] = np.uint64((int(v[0]) + delta) % (1 << 64))
                assert _rc(stz, bad) == STZ_E_STREAM, (s, word, delta)


def test_stream_table_offsets_checked(stz, oracle):
    p = _payload(oracle)
    for s in range(N_STREAMS):
        for word, delta in ((28 + s, 8), (28 + s, 1 << 63), (43 + s, 8), (43 + s, (1 << 64) - 8)):
            bad = p.copy()
            v = bad[8 * word: 8 * word + 8].view(np.uint64)
            v[0] = np.uint64((int(v[0]) + delta) % (1 << 64))
            assert _rc(stz, bad) == STZ_E_STREAM, (s, word, delta)
    # counts that would wrap the section offsets: n_unpred (u64 word 7) and
    # total_bits (word 8) of a stream header
    so, _ = stream_table(p)
    for word in (7, 8):
        bad = p.copy()
        bad[so[3] + 8 * word: so[3] + 8 * word + 8].view(np.uint64)[0] = np.uint64(1 << 62)
        assert _rc(stz, bad) == STZ_E_STREAM, word


# ---------------------------------------------------------------- containers

def _zstd():
    name = ctypes.util.find_library("zstd") or "libzstd.so.1"
    try:
        z = C.CDLL(name)
    except OSError:
        pytest.skip("libzstd not loadable")
    z.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    z.ZSTD_compress.restype = C.c_size_t
    z.ZSTD_compressBound.argtypes = [C.c_size_t]
    z.ZSTD_compressBound.restype = C.c_size_t
    z.ZSTD_isError.argtypes = [C.c_size_t]
    return z


def _frame(z, data):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    cap = z.ZSTD_compressBound(data.size)
    out = np.empty(cap, np.uint8)
    r = z.ZSTD_compress(out.ctypes.data_as(C.c_void_p), cap, data.ctypes.data_as(C.c_void_p), data.size, 3)
    assert not z.ZSTD_isError(r)
    return out[:r].copy()


def _container(z, p, raw_override=None):
    """STZC container: u32 magic, version, n_frames, level; u64 payload bytes;
    u64 raw[16]; u64 comp[16]; the 512-byte file header stored raw, then one
    zstd frame per stream."""
    so, sb = stream_table(p)
    frames = [p[:512].copy()] + [_frame(z, p[so[s]:so[s] + sb[s]]) for s in range(N_STREAMS)]
    raw = [512] + [sb[s] for s in range(N_STREAMS)]
    if raw_override:
        for k, v in raw_override.items():
            raw[k] = v
    hdr = np.zeros(280, np.uint8)
    hdr[:16].view(np.uint32)[:] = [0x435A5453, 1, N_STREAMS + 1, 3]
    u64 = hdr[16:].view(np.uint64)
    u64[0] = p.size
    u64[1:17] = np.array(raw, dtype=np.uint64)
    u64[17:33] = [f.size for f in frames]
    return np.concatenate([hdr] + frames)


def test_container_passes_and_wrapped_sizes_refused(stz, oracle):
    z = _zstd()
    p = _payload(oracle)
    assert _rc(stz, _container(z, p)) == 0
    _, sb = stream_table(p)
    # two frames whose raw sizes wrap the 64-bit sum back to the payload size
    big = 1 << 63
    wrapped = _container(z, p, {1: sb[0] + big, 2: sb[1] + big})
    assert _rc(stz, wrapped) == STZ_E_STREAM
    # a raw size that disagrees with the frame's own content size
    shifted = _container(z, p, {1: sb[0] + 8, 2: sb[1] - 8})
    assert _rc(stz, shifted) == STZ_E_STREAM
    # truncated container
    good = _container(z, p)
    assert _rc(stz, good[:-5]) == STZ_E_STREAM
    # a compressed size pointing past the end
    bad = good.copy()
    bad[16 + 8 * 17 + 8 * 3: 16 + 8 * 17 + 8 * 4].view(np.uint64)[0] = np.uint64(1 << 40)
    assert _rc(stz, bad) == STZ_E_STREAM


@pytest.mark.parametrize("shape", [(64, 128, 256), (37, 50, 70)])
def test_every_base_codec_layout_matches_the_oracle(stz, oracle, shape):
    """The library's host layout accepts the oracle's payload for every base
    codec, including the Lorenzo/regression base whose tiles' header symbols
    lengthen the base stream (on 64x128x256 they open a second 32768-symbol
    base chunk: 64x32x16 points plus 8 tiles x 32 symbols)."""
    rng = np.random.default_rng(5)
    f = np.cumsum(rng.normal(size=shape), axis=2).astype(np.float32)
    for interp in (0, 1, 2, 3):
        p = oracle.compress(f, 1e-3, oracle.EB_REL, interp=interp)
        assert oracle.header(p)["interp"] == interp
        assert _rc(stz, p) == 0, interp
