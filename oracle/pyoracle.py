"""ctypes wrapper of the CPU oracle (oracle/_build/libstz_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product path
(sc25-stz2_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# STZ_ORACLE_LIB: another build of the oracle (the sanitized one, `make -C oracle asan`)
_LIB_PATH = os.environ.get("STZ_ORACLE_LIB") or os.path.join(_HERE, "_build", "libstz_oracle.so")
_lib = None

F32, F64 = 0, 1
EB_ABS, EB_REL, EB_REF = 0, 1, 2


class Params(C.Structure):
    _fields_ = [
        ("mult_base", C.c_double), ("mult_low", C.c_double), ("mult_high", C.c_double),
        ("chunk_high", C.c_int), ("chunk_low", C.c_int),
        ("interp", C.c_int), ("order", C.c_int),
        ("alpha", C.c_double), ("beta", C.c_double),
        ("pred", C.c_int),
    ]

# predictor options as bits of Params.pred: F5's PRED_WPOW2, PRED_PLANE_AXIS,
# PRED_CUBE_AXIS; stz2.cpp's 12-point bicubic (blocks 2/4/5) and 16-point (block 6);
# dependency.cpp's staged inverse-variance predictors (alone)
PRED_WPOW2, PRED_PLANE_AXIS, PRED_CUBE_AXIS, PRED_BICUBIC12, PRED_DIAG16, PRED_STAGED = 1, 2, 4, 8, 16, 32
# base codec (Params.interp): interpolation linear / cubic, or Lorenzo over
# independent LZ_TILE^3 tiles
BASE_LINEAR, BASE_CUBIC, BASE_LORENZO, BASE_LZREG, LZ_TILE, LZ_HDR = 0, 1, 2, 3, 16, 32


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, sz, dbl, i32 = C.c_void_p, C.c_size_t, C.c_double, C.c_int
        L.ora_compress.restype = C.c_longlong
        L.ora_compress.argtypes = [i32, vp, sz, sz, sz, dbl, i32, C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.ora_decompress.argtypes = [vp, sz, vp]
        L.ora_decompress_roi.argtypes = [vp, sz, i32, i32, i32, i32, i32, i32, vp, C.POINTER(sz)]
        L.ora_decompress_level.argtypes = [vp, sz, i32, vp, C.POINTER(sz)]
        L.ora_free.argtypes = [vp]
        L.ora_default_params.argtypes = [C.POINTER(Params)]
        for t in ("f32", "f64"):
            getattr(L, "ora_split_" + t).argtypes = [vp, C.POINTER(vp), i32, i32, i32]
            getattr(L, "ora_merge_" + t).argtypes = [C.POINTER(vp), vp, i32, i32, i32]
            getattr(L, "ora_preprocess_" + t).argtypes = [i32, vp, vp, vp, i32, i32, i32]
            getattr(L, "ora_depreprocess_" + t).argtypes = [i32, vp, vp, vp, i32, i32, i32]
            getattr(L, "ora_preprocess_pv_" + t).argtypes = [i32, i32, vp, vp, vp, i32, i32, i32]
            getattr(L, "ora_depreprocess_pv_" + t).argtypes = [i32, i32, vp, vp, vp, i32, i32, i32]
            getattr(L, "ora_quantize_" + t).restype = sz
            getattr(L, "ora_quantize_" + t).argtypes = [vp, sz, dbl, i32, vp, vp, vp]
            getattr(L, "ora_dequantize_" + t).argtypes = [vp, sz, dbl, i32, vp, vp]
            getattr(L, "ora_interp_compress_" + t).restype = sz
            getattr(L, "ora_interp_compress_" + t).argtypes = [vp, i32, i32, i32, dbl, i32, i32, dbl, dbl, dbl, dbl,
                                                               vp, vp]
            getattr(L, "ora_interp_decompress_" + t).argtypes = [vp, i32, i32, i32, dbl, i32, i32, dbl, dbl, dbl, dbl,
                                                                 vp, vp]
            getattr(L, "ora_staged_level_" + t).argtypes = [vp, i32, i32, i32, dbl, vp, vp, vp, vp, vp]
        L.ora_huff_lengths.argtypes = [vp, i32, i32, vp]
        L.ora_plan_query.argtypes = [i32] * 11 + [vp, vp, vp, vp, vp, C.c_int64]
        L.ora_payload_info.argtypes = [vp, sz, vp, vp, vp]
        _lib = L
    return _lib


def default_params(**kw):
    p = Params()
    lib().ora_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _tsuf(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def compress(field, eb, mode=EB_REL, **params):
    """field: numpy (nz, ny, nx) float32/float64 -> payload bytes (numpy uint8)."""
    field = np.ascontiguousarray(field)
    nz, ny, nx = field.shape
    dt = F32 if field.dtype == np.float32 else F64
    p = default_params(**params)
    out = C.c_void_p()
    n = lib().ora_compress(dt, _ptr(field), nx, ny, nz, float(eb), mode, C.byref(p), C.byref(out))
    if n <= 0:
        raise RuntimeError("ora_compress failed (%d)" % n)
    buf = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint8)), shape=(n,)).copy()
    lib().ora_free(out)
    return buf


def header(payload):
    h = np.frombuffer(payload[:512].tobytes(), dtype=np.uint8)
    u32 = h.view(np.uint32)
    u64 = h.view(np.uint64)
    f64 = h.view(np.float64)
    return {
        "dtype": int(u32[2]), "eb_mode": int(u32[3]),
        "nx": int(u64[2]), "ny": int(u64[3]), "nz": int(u64[4]),
        "px": int(u64[5]), "py": int(u64[6]), "pz": int(u64[7]),
        "eb_user": float(f64[8]), "eb_unit": float(f64[9]),
        "eb_base": float(f64[13]), "eb_low": float(f64[14]), "eb_high": float(f64[15]),
        "vmin": float(f64[16]), "vmax": float(f64[17]),
        "total_bytes": int(u64[24]),
        "interp": int(u32[40]), "order": int(u32[41]), "pred": int(u64[25]),
    }


def decompress(payload):
    h = header(payload)
    dt = np.float32 if h["dtype"] == F32 else np.float64
    out = np.empty((h["nz"], h["ny"], h["nx"]), dtype=dt)
    rc = lib().ora_decompress(_ptr(payload), payload.size, _ptr(out))
    if rc != 0:
        raise RuntimeError("ora_decompress failed (%d)" % rc)
    return out


def decompress_roi(payload, x0, y0, z0, x1, y1, z1):
    h = header(payload)
    dt = np.float32 if h["dtype"] == F32 else np.float64
    cap = (abs(x1 - x0) + 1) * (abs(y1 - y0) + 1) * (abs(z1 - z0) + 1)
    out = np.empty(cap, dtype=dt)
    od = (C.c_size_t * 3)()
    rc = lib().ora_decompress_roi(_ptr(payload), payload.size, x0, y0, z0, x1, y1, z1, _ptr(out), od)
    if rc != 0:
        raise RuntimeError("ora_decompress_roi failed (%d)" % rc)
    return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])


def decompress_level(payload, level):
    h = header(payload)
    dt = np.float32 if h["dtype"] == F32 else np.float64
    out = np.empty(h["nx"] * h["ny"] * h["nz"], dtype=dt)
    od = (C.c_size_t * 3)()
    rc = lib().ora_decompress_level(_ptr(payload), payload.size, level, _ptr(out), od)
    if rc != 0:
        raise RuntimeError("ora_decompress_level failed (%d)" % rc)
    return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])


def payload_info(payload):
    ebl = np.zeros(3)
    sb = np.zeros(15, dtype=np.int64)
    nu = np.zeros(15, dtype=np.int64)
    rc = lib().ora_payload_info(_ptr(payload), payload.size, _ptr(ebl), _ptr(sb), _ptr(nu))
    if rc != 0:
        raise RuntimeError("bad payload")
    return ebl, sb, nu


def split(full):
    full = np.ascontiguousarray(full)
    nz, ny, nx = full.shape
    subs = [np.empty((nz // 2, ny // 2, nx // 2), dtype=full.dtype) for _ in range(8)]
    arr = (C.c_void_p * 8)(*[s.ctypes.data for s in subs])
    getattr(lib(), "ora_split_" + _tsuf(full.dtype))(_ptr(full), arr, nx, ny, nz)
    return subs


def merge(subs):
    sz, sy, sx = subs[0].shape
    out = np.empty((2 * sz, 2 * sy, 2 * sx), dtype=subs[0].dtype)
    subs = [np.ascontiguousarray(s) for s in subs]
    arr = (C.c_void_p * 8)(*[s.ctypes.data for s in subs])
    getattr(lib(), "ora_merge_" + _tsuf(out.dtype))(arr, _ptr(out), sx, sy, sz)
    return out


def preprocess(block, sub, ref, pred=0):
    sub = np.ascontiguousarray(sub)
    ref = np.ascontiguousarray(ref)
    dz, dy, dx = sub.shape
    out = np.empty_like(sub)
    getattr(lib(), "ora_preprocess_pv_" + _tsuf(sub.dtype))(block, pred, _ptr(sub), _ptr(ref), _ptr(out), dx, dy, dz)
    return out


def depreprocess(block, de, ref, pred=0):
    de = np.ascontiguousarray(de)
    ref = np.ascontiguousarray(ref)
    dz, dy, dx = de.shape
    out = np.empty_like(de)
    getattr(lib(), "ora_depreprocess_pv_" + _tsuf(de.dtype))(block, pred, _ptr(de), _ptr(ref), _ptr(out), dx, dy, dz)
    return out


def staged_level(ref, eb=0.0, targets=None, residuals=None):
    """One residual level with dependency.cpp's staged predictors.  ref: the
    coarser grid (dz, dy, dx); targets (encode) or residuals (decode): the 7
    blocks, shape (7, dz, dy, dx).  Returns the level's 2x grid and, when
    encoding, the codes and decoded residuals (grid order, per block)."""
    ref = np.ascontiguousarray(ref)
    dz, dy, dx = ref.shape
    g = np.empty((2 * dz, 2 * dy, 2 * dx), dtype=ref.dtype)
    f = getattr(lib(), "ora_staged_level_" + _tsuf(ref.dtype))
    if targets is not None:
        t = np.ascontiguousarray(targets, dtype=ref.dtype)
        q = np.empty(t.shape, dtype=np.uint16)
        dv = np.empty_like(t)
        f(_ptr(ref), dx, dy, dz, float(eb), _ptr(g), _ptr(t), _ptr(q), _ptr(dv), None)
        return g, q, dv
    r = np.ascontiguousarray(residuals, dtype=ref.dtype)
    f(_ptr(ref), dx, dy, dz, float(eb), _ptr(g), None, None, None, _ptr(r))
    return g


def quantize(d, eb, radius=32768):
    d = np.ascontiguousarray(d).ravel()
    codes = np.empty(d.size, dtype=np.uint16)
    unpred = np.empty(d.size, dtype=d.dtype)
    dec = np.empty(d.size, dtype=d.dtype)
    nu = getattr(lib(), "ora_quantize_" + _tsuf(d.dtype))(_ptr(d), d.size, float(eb), radius, _ptr(codes), _ptr(unpred), _ptr(dec))
    return codes, unpred[:nu].copy(), dec


def dequantize(codes, unpred, eb, dtype, radius=32768):
    codes = np.ascontiguousarray(codes, dtype=np.uint16)
    unpred = np.ascontiguousarray(unpred, dtype=dtype)
    out = np.empty(codes.size, dtype=dtype)
    getattr(lib(), "ora_dequantize_" + _tsuf(dtype))(_ptr(codes), codes.size, float(eb), radius, _ptr(unpred), _ptr(out))
    return out


def huff_lengths(freq, lmax=24):
    freq = np.ascontiguousarray(freq, dtype=np.uint64)
    out = np.zeros(freq.size, dtype=np.uint8)
    lib().ora_huff_lengths(_ptr(freq), freq.size, lmax, _ptr(out))
    return out


def base_symbols(shape, interp):
    """Symbols of the base stream of a grid: its points, plus LZ_HDR header
    symbols per 16^3 tile for the Lorenzo/regression base."""
    n = int(np.prod(shape))
    if interp == BASE_LZREG:
        n += LZ_HDR * int(np.prod([(d + LZ_TILE - 1) // LZ_TILE for d in shape]))
    return n


def interp_compress(v, eb, interp=1, order=0, alpha=1.0, beta=1.0, lo=None, hi=None):
    """The base codec on one grid: (reconstruction, codes, unpredictables);
    lo/hi clamp the Lorenzo/regression base's plane predictions (default:
    the grid's range)."""
    v = np.ascontiguousarray(v).copy()
    bz, by, bx = v.shape
    lo = float(v.min()) if lo is None else float(lo)
    hi = float(v.max()) if hi is None else float(hi)
    n = base_symbols(v.shape, interp)
    codes = np.empty(n, dtype=np.uint16)
    unpred = np.empty(n, dtype=v.dtype)
    nu = getattr(lib(), "ora_interp_compress_" + _tsuf(v.dtype))(_ptr(v), bx, by, bz, float(eb), interp, order, alpha,
                                                                 beta, lo, hi, _ptr(codes), _ptr(unpred))
    return v, codes, unpred[:nu].copy()


def interp_decompress(shape, dtype, eb, codes, unpred, interp=1, order=0, alpha=1.0, beta=1.0, lo=0.0, hi=0.0):
    """The base codec's decode: the reconstruction from codes and
    unpredictables (lo/hi: the clamp range the encode used)."""
    bz, by, bx = shape
    v = np.zeros(shape, dtype=dtype)
    codes = np.ascontiguousarray(codes, np.uint16)
    assert codes.size == base_symbols(shape, interp)
    unpred = np.ascontiguousarray(unpred, dtype)
    getattr(lib(), "ora_interp_decompress_" + _tsuf(np.dtype(dtype)))(_ptr(v), bx, by, bz, float(eb), interp, order,
                                                                        alpha, beta, float(lo), float(hi), _ptr(codes),
                                                                        _ptr(unpred))
    return v


def plan_query(full_dims, q, chunk_high=32, chunk_low=16, cap=1 << 16):
    fx, fy, fz = full_dims
    query = np.zeros(7, dtype=np.int32)
    hp = np.zeros((7, 8), dtype=np.int32)
    lp = np.zeros((7, 8), dtype=np.int32)
    hc = np.zeros((7, cap), dtype=np.int64)
    lc = np.zeros((7, cap), dtype=np.int64)
    rc = lib().ora_plan_query(fx, fy, fz, chunk_high, chunk_low, *q, _ptr(query), _ptr(hp), _ptr(hc), _ptr(lp), _ptr(lc), cap)
    if rc != 0:
        raise RuntimeError("plan failed %d" % rc)
    high = {b: (tuple(hp[b, 1:7]), hc[b, : hp[b, 7]].copy()) for b in range(7) if hp[b, 0]}
    low = {b: (tuple(lp[b, 1:7]), lc[b, : lp[b, 7]].copy()) for b in range(7) if lp[b, 0]}
    return tuple(query[:6]), bool(query[6]), high, low
