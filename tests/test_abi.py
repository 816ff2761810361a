"""CPU checks of the C ABI library (sc25-stz2_amd/lib/libstz_hip.so): it loads,
exports every entry point include/stz.h declares, and its host-side logic
(argument checks, stream-header parsing, Huffman code lengths, the ROI
planner) agrees with the oracle.  No kernel runs here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stz.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(stz_\w+)\s*\(", text, flags=re.M))


def test_header_and_exports_agree(stz):
    from stz_amd import _lib
    declared = _declared()
    assert len(declared) >= 30
    assert declared == set(_lib.EXPORTS)
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing


def test_strerror_and_defaults(stz):
    L = stz.lib()
    for code in [0, -1, -2, -3, -4, -5, -6, -7, -8]:
        assert L.stz_strerror(code)
    p = stz.default_params()
    # F5:2127-2129 level multipliers, F5:34-35 chunk dims
    assert (p.mult_base, p.mult_low, p.mult_high) == (0.7, 2.7, 8.2)
    assert (p.chunk_high, p.chunk_low) == (32, 16)


def test_null_arguments_rejected(stz):
    L = stz.lib()
    info = stz._lib.Info()
    assert L.stz_query_info(None, 0, C.byref(info)) == -1
    assert L.stz_compress(None, 0, None, 1, 1, 1, 1e-3, 1, None, None, None) == -1
    assert L.stz_decompress(None, None, 0, None) == -1
    assert L.stz_set_grid_share(None, 0.5) == -1


def test_no_device_is_an_error_not_a_fallback(stz):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    assert stz.lib().stz_create(C.byref(ctx), 0) == -7   # STZ_E_NODEV
    with pytest.raises(stz.StzError):
        stz.Context(0)


def test_query_info_reads_oracle_payload(stz, oracle):
    rng = np.random.default_rng(4)
    f = rng.normal(size=(12, 20, 28)).astype(np.float64)
    p = oracle.compress(f, 1e-4, oracle.EB_ABS, interp=0, order=2)
    info = stz.stream_info(p)
    h = oracle.header(p)
    assert (info.nx, info.ny, info.nz) == (28, 20, 12)
    assert (info.px, info.py, info.pz) == (h["px"], h["py"], h["pz"])
    assert info.dtype == 1 and info.interp == 0 and info.order == 2
    assert info.eb_high == h["eb_high"] and info.payload_bytes == p.size
    bad = p.copy()
    bad[:4] = 0
    with pytest.raises(stz.StzError):
        stz.stream_info(bad)


def test_huff_lengths_match_oracle(stz, oracle):
    from stz_amd import api
    rng = np.random.default_rng(8)
    for n, lmax in [(65536, 24), (1000, 24), (40, 12), (2, 24)]:
        f = (rng.pareto(0.8, size=n) * 5).astype(np.uint64)
        f[:2] += 1
        assert np.array_equal(api.huff_lengths(f, lmax), oracle.huff_lengths(f, lmax))


def test_plan_query_matches_golden(stz):
    from stz_amd import api
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
    pin = g["plan_in"]
    # the golden chunk ids of case k, block b follow those of every earlier
    # (case, block) in the concatenated arrays
    offs = {"high": 0, "low": 0}
    for k in range(pin.shape[0]):
        dims = tuple(int(v) for v in pin[k, :3])
        q = [int(v) for v in pin[k, 5:11]]
        qa, sl, high, low = api.plan_query(dims, q, int(pin[k, 3]), int(pin[k, 4]), 4096)
        gq = g["plan_query"][k]
        assert qa == tuple(int(v) for v in gq[:6]) and sl == bool(gq[6]), k
        for name, plans, table in (("high", high, g["plan_high"][k]), ("low", low, g["plan_low"][k])):
            allc = g["plan_%s_chunks" % name]
            assert set(plans) == {b for b in range(7) if table[b, 0]}, (k, name)
            for b in range(7):
                if not table[b, 0]:
                    continue
                assert plans[b][0] == tuple(int(v) for v in table[b, 1:7]), (k, name, b)
                n = int(table[b, 7])
                assert np.array_equal(plans[b][1], allc[offs[name]:offs[name] + n]), (k, name, b)
                offs[name] += n
    assert offs["high"] == g["plan_high_chunks"].size and offs["low"] == g["plan_low_chunks"].size


def _cli():
    return os.path.join(ROOT, "sc25-stz2_amd", "lib", "stz_cli")


def test_cli_argument_errors(tmp_path):
    import subprocess
    if not os.path.exists(_cli()):
        pytest.skip("stz_cli not built")
    r = subprocess.run([_cli()], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stderr
    f = tmp_path / "t.f32"
    np.zeros((4, 5, 6), np.float32).tofile(f)
    r = subprocess.run([_cli(), "1e-3", str(f), "6", "5", "5", "-f"], capture_output=True, text=True)
    assert r.returncode == 1 and "size mismatch" in r.stderr
    r = subprocess.run([_cli(), "1e-3", str(f), "6", "5", "4", "-f", "1", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "exactly six integers" in r.stderr
    r = subprocess.run([_cli(), "1e-3", str(f), "6", "5", "4", "-f", "--trial"], capture_output=True, text=True)
    assert r.returncode == 1 and "--trial needs" in r.stderr
