"""Python host interface over the C ABI (include/stz.h).

Mirrors the reference's codec entry points (F5 = src/zero_all_line_dynamic_
random_access_optfast3_optcr_lb_fastdec5.cpp): compress = the compress half
of run_typed (F5:2086-2314), decompress = F5:2339-2451, roi / slice =
the --auto-query path (F5:2460-2605), level = progressive 1/8..full.  Arrays
are numpy (host memory) or torch tensors (host or HIP device memory); x is the
fastest axis, so a numpy array of shape (nz, ny, nx) maps to dims (nx, ny, nz).
"""
import ctypes as C

import numpy as np

from ._lib import REQ_BOX, REQ_LEVEL, REQ_SLICE, Info, Params, Request, StzError, check, lib

F32, F64 = 0, 1
EB_MODES = {"abs": 0, "rel": 1, "ref": 2}


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _ptr(a):
    if _is_torch(a):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr())
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def _dtype_id(a):
    name = str(a.dtype)
    if name.endswith("float32"):
        return F32
    if name.endswith("float64"):
        return F64
    raise TypeError("field must be float32 or float64, got %s" % name)


def _np_dtype(dt):
    return np.float32 if dt == F32 else np.float64


def default_params(**kw):
    p = Params()
    lib().stz_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _params(params, kw):
    if params is None:
        params = default_params()
    for k, v in kw.items():
        setattr(params, k, v)
    return params


def _empty_like_kind(like, shape, dt):
    if _is_torch(like):
        import torch
        return torch.empty(shape, dtype=torch.float32 if dt == F32 else torch.float64, device=like.device)
    return np.empty(shape, dtype=_np_dtype(dt))


def stream_info(stream):
    buf = np.frombuffer(stream, dtype=np.uint8) if not isinstance(stream, np.ndarray) else stream
    info = Info()
    check(lib().stz_query_info(_ptr(buf), buf.size, C.byref(info)), "stz_query_info")
    return info


def stream_check(stream):
    """Host-only validation of a container or payload (stz_stream_check); raises StzError."""
    buf = np.frombuffer(stream, dtype=np.uint8) if not isinstance(stream, np.ndarray) else stream
    check(lib().stz_stream_check(_ptr(buf), buf.size), "stz_stream_check")


class DStream:
    """A compressed stream resident in HBM (stz_dstream)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = C.c_void_p(handle)

    def free(self):
        if self.h:
            lib().stz_dstream_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def info(self):
        i = Info()
        check(lib().stz_dstream_info(self.h, C.byref(i)), "stz_dstream_info")
        return i

    @property
    def payload_bytes(self):
        return self.info.payload_bytes

    def payload_ptr(self):
        p = C.c_void_p()
        n = C.c_size_t()
        check(lib().stz_dstream_payload(self.h, C.byref(p), C.byref(n)), "stz_dstream_payload")
        return p.value, n.value

    def download(self, zstd_level=3):
        """Host copy: a zstd container (zstd_level >= 0) or the bare payload (< 0)."""
        p = C.c_void_p()
        n = C.c_size_t()
        check(lib().stz_dstream_download(self.ctx.h, self.h, zstd_level, C.byref(p), C.byref(n)),
              "stz_dstream_download")
        out = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value,)).copy()
        lib().stz_free(p)
        return out

    def decompress(self, out=None, like=None, ctx=None):
        i = self.info
        c = ctx if ctx is not None else self.ctx
        if out is None:
            out = _empty_like_kind(like if like is not None else np.empty(0), (i.nz, i.ny, i.nx), i.dtype)
        check(lib().stz_dstream_decompress(c.h, self.h, _ptr(out)), "stz_dstream_decompress")
        return out

    def decompress_roi(self, x0, y0, z0, x1, y1, z1, like=None):
        i = self.info
        cap = (abs(x1 - x0) + 1) * (abs(y1 - y0) + 1) * (abs(z1 - z0) + 1)
        out = _empty_like_kind(like if like is not None else np.empty(0), (cap,), i.dtype)
        od = (C.c_size_t * 3)()
        check(lib().stz_dstream_decompress_roi(self.ctx.h, self.h, x0, y0, z0, x1, y1, z1, _ptr(out), od),
              "stz_dstream_decompress_roi")
        return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])

    def decompress_level(self, level, like=None):
        i = self.info
        out = _empty_like_kind(like if like is not None else np.empty(0), (i.nx * i.ny * i.nz,), i.dtype)
        od = (C.c_size_t * 3)()
        check(lib().stz_dstream_decompress_level(self.ctx.h, self.h, level, _ptr(out), od),
              "stz_dstream_decompress_level")
        return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])


    # -- random access in batches
    def prepare(self, keep_base=True, ctx=None):
        """Build the decode tables and (keep_base) keep the decoded base level on
        the device, so random-access requests skip it."""
        c = ctx if ctx is not None else self.ctx
        check(lib().stz_dstream_prepare(c.h, self.h, 1 if keep_base else 0), "stz_dstream_prepare")

    def request_capacity(self, req):
        """Elements a request's output needs at most."""
        i = self.info
        kind = req[0]
        if kind == "box":
            x0, y0, z0, x1, y1, z1 = req[1:7]
            return (abs(x1 - x0) + 1) * (abs(y1 - y0) + 1) * (abs(z1 - z0) + 1)
        if kind == "slice":
            d = [i.nx, i.ny, i.nz]
            d[req[1]] = 1
            return d[0] * d[1] * d[2]
        if kind == "level":
            sh = 3 - req[1]
            return (((i.nx + (1 << sh) - 1) >> sh) * ((i.ny + (1 << sh) - 1) >> sh) *
                    ((i.nz + (1 << sh) - 1) >> sh))
        raise ValueError("request kind %r" % (kind,))

    def decompress_batch(self, requests, outs=None, like=None, ctx=None):
        """Serve several random-access requests in one decode.  requests:
        ("box", x0, y0, z0, x1, y1, z1) | ("slice", axis, index) | ("level", level).
        Returns one (z, y, x) array per request."""
        c = ctx if ctx is not None else self.ctx
        n = len(requests)
        R = (Request * n)()
        kinds = {"box": REQ_BOX, "slice": REQ_SLICE, "level": REQ_LEVEL}
        for k, r in enumerate(requests):
            R[k].kind = kinds[r[0]]
            for j, v in enumerate(r[1:]):
                R[k].a[j] = int(v)
        i = self.info
        if outs is None:
            base = like if like is not None else np.empty(0)
            outs = [_empty_like_kind(base, (self.request_capacity(r),), i.dtype) for r in requests]
        ptrs = (C.c_void_p * n)(*[_ptr(o) for o in outs])
        od = (C.c_size_t * (3 * n))()
        check(lib().stz_dstream_decompress_batch(c.h, self.h, n, R, ptrs, od), "stz_dstream_decompress_batch")
        res = []
        for k in range(n):
            a, b, z = od[3 * k], od[3 * k + 1], od[3 * k + 2]
            res.append(outs[k][: a * b * z].reshape(z, b, a))
        return res


class Context:
    """One HIP stream + memory pool on one device (stz_ctx).  own_queue: the
    context's streams get hardware queues of their own (STZ_CTX_OWN_QUEUE),
    so its work overlaps other streams' whatever GPU_MAX_HW_QUEUES is."""

    def __init__(self, device=0, own_queue=False):
        h = C.c_void_p()
        check(lib().stz_create_ex(C.byref(h), device, 1 if own_queue else 0), "stz_create_ex")
        self.h = h

    def close(self):
        if self.h:
            lib().stz_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stream / profiling
    def set_stream(self, hip_stream):
        check(lib().stz_set_stream(self.h, C.c_void_p(hip_stream)), "stz_set_stream")

    def stream(self):
        return lib().stz_get_stream(self.h)

    def set_grid_share(self, share):
        """Persistent full-level decode kernels of this context take at most
        this fraction of the resident workgroup slots (serving)."""
        check(lib().stz_set_grid_share(self.h, float(share)), "stz_set_grid_share")

    def profile(self, enable=True):
        check(lib().stz_profile(self.h, 1 if enable else 0), "stz_profile")

    def profile_report(self):
        import json
        buf = C.create_string_buffer(1 << 16)
        check(lib().stz_profile_report(self.h, buf, len(buf)), "stz_profile_report")
        return json.loads(buf.value.decode())

    # -- whole path
    def compress(self, field, eb, mode="rel", params=None, **kw):
        nz, ny, nx = field.shape
        p = _params(params, kw)
        s = C.c_void_p()
        n = C.c_size_t()
        check(lib().stz_compress(self.h, _dtype_id(field), _ptr(field), nx, ny, nz, float(eb), EB_MODES[mode],
                                 C.byref(p), C.byref(s), C.byref(n)), "stz_compress")
        out = np.ctypeslib.as_array(C.cast(s, C.POINTER(C.c_uint8)), shape=(n.value,)).copy()
        lib().stz_free(s)
        return out

    def compress_device(self, field, eb, mode="rel", params=None, **kw):
        nz, ny, nx = field.shape
        p = _params(params, kw)
        h = C.c_void_p()
        check(lib().stz_compress_device(self.h, _dtype_id(field), _ptr(field), nx, ny, nz, float(eb),
                                        EB_MODES[mode], C.byref(p), C.byref(h)), "stz_compress_device")
        return DStream(self, h.value)

    # -- many fields
    def compress_batch(self, fields, eb, mode="rel", params=None, **kw):
        """Uneven fields of one dtype in one call (bit-identical to one call each)."""
        n = len(fields)
        if n == 0:
            return []
        dt = _dtype_id(fields[0])
        ptrs = (C.c_void_p * n)(*[_ptr(f) for f in fields])
        dims = (C.c_size_t * (3 * n))()
        for i, f in enumerate(fields):
            if _dtype_id(f) != dt:
                raise TypeError("one dtype per batch")
            nz, ny, nx = f.shape
            dims[3 * i], dims[3 * i + 1], dims[3 * i + 2] = nx, ny, nz
        p = _params(params, kw)
        out = (C.c_void_p * n)()
        check(lib().stz_compress_batch(self.h, dt, n, ptrs, dims, float(eb), EB_MODES[mode], C.byref(p), out),
              "stz_compress_batch")
        return [DStream(self, out[i]) for i in range(n)]

    def decompress_batch(self, dstreams, outs):
        """Decode several device-resident streams into device tensors (one
        captured launch sequence, replayed while the streams and outputs stay)."""
        n = len(dstreams)
        hs = (C.c_void_p * n)(*[d.h.value for d in dstreams])
        ps = (C.c_void_p * n)(*[_ptr(o) for o in outs])
        check(lib().stz_decompress_batch(self.h, n, hs, ps), "stz_decompress_batch")
        return outs

    def autotune(self, enable=None):
        """Launch geometries chosen so far (and switch the timing on/off)."""
        import json
        buf = C.create_string_buffer(1 << 16)
        check(lib().stz_autotune(self.h, -1 if enable is None else (1 if enable else 0), buf, len(buf)),
              "stz_autotune")
        return json.loads(buf.value.decode())

    def upload(self, stream):
        buf = np.ascontiguousarray(np.frombuffer(stream, dtype=np.uint8))
        h = C.c_void_p()
        check(lib().stz_dstream_upload(self.h, _ptr(buf), buf.size, C.byref(h)), "stz_dstream_upload")
        return DStream(self, h.value)

    def decompress(self, stream, out=None, like=None):
        buf = np.ascontiguousarray(np.frombuffer(stream, dtype=np.uint8))
        i = stream_info(buf)
        if out is None:
            out = _empty_like_kind(like if like is not None else np.empty(0), (i.nz, i.ny, i.nx), i.dtype)
        check(lib().stz_decompress(self.h, _ptr(buf), buf.size, _ptr(out)), "stz_decompress")
        return out

    def decompress_roi(self, stream, x0, y0, z0, x1, y1, z1):
        buf = np.ascontiguousarray(np.frombuffer(stream, dtype=np.uint8))
        i = stream_info(buf)
        cap = (abs(x1 - x0) + 1) * (abs(y1 - y0) + 1) * (abs(z1 - z0) + 1)
        out = np.empty(cap, dtype=_np_dtype(i.dtype))
        od = (C.c_size_t * 3)()
        check(lib().stz_decompress_roi(self.h, _ptr(buf), buf.size, x0, y0, z0, x1, y1, z1, _ptr(out), od),
              "stz_decompress_roi")
        return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])

    def decompress_slice(self, stream, axis, index):
        buf = np.ascontiguousarray(np.frombuffer(stream, dtype=np.uint8))
        i = stream_info(buf)
        cap = max(i.nx * i.ny, i.ny * i.nz, i.nx * i.nz)
        out = np.empty(cap, dtype=_np_dtype(i.dtype))
        od = (C.c_size_t * 3)()
        check(lib().stz_decompress_slice(self.h, _ptr(buf), buf.size, axis, index, _ptr(out), od),
              "stz_decompress_slice")
        return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])

    def decompress_level(self, stream, level):
        buf = np.ascontiguousarray(np.frombuffer(stream, dtype=np.uint8))
        i = stream_info(buf)
        out = np.empty(i.nx * i.ny * i.nz, dtype=_np_dtype(i.dtype))
        od = (C.c_size_t * 3)()
        check(lib().stz_decompress_level(self.h, _ptr(buf), buf.size, level, _ptr(out), od),
              "stz_decompress_level")
        return out[: od[0] * od[1] * od[2]].reshape(od[2], od[1], od[0])

    # -- stage entry points
    def split(self, full):
        nz, ny, nx = full.shape
        subs = [np.empty((nz // 2, ny // 2, nx // 2), dtype=full.dtype) for _ in range(8)]
        arr = (C.c_void_p * 8)(*[s.ctypes.data for s in subs])
        check(lib().stz_split(self.h, _dtype_id(full), _ptr(full), nx, ny, nz, arr), "stz_split")
        return subs

    def merge(self, subs):
        sz, sy, sx = subs[0].shape
        out = np.empty((2 * sz, 2 * sy, 2 * sx), dtype=subs[0].dtype)
        subs = [np.ascontiguousarray(s) for s in subs]
        arr = (C.c_void_p * 8)(*[s.ctypes.data for s in subs])
        check(lib().stz_merge(self.h, _dtype_id(out), arr, sx, sy, sz, _ptr(out)), "stz_merge")
        return out

    def residual_predict(self, block, sub, ref):
        dz, dy, dx = sub.shape
        out = np.empty_like(sub)
        check(lib().stz_residual_predict(self.h, _dtype_id(sub), block, _ptr(sub), _ptr(ref), dx, dy, dz,
                                         _ptr(out)), "stz_residual_predict")
        return out

    def residual_reconstruct(self, block, de, ref):
        dz, dy, dx = de.shape
        out = np.empty_like(de)
        check(lib().stz_residual_reconstruct(self.h, _dtype_id(de), block, _ptr(de), _ptr(ref), dx, dy, dz,
                                             _ptr(out)), "stz_residual_reconstruct")
        return out

    def quantize(self, data, eb):
        d = np.ascontiguousarray(data).ravel()
        codes = np.empty(d.size, dtype=np.uint16)
        un = np.empty(d.size, dtype=d.dtype)
        n = C.c_size_t()
        check(lib().stz_quantize(self.h, _dtype_id(d), _ptr(d), d.size, float(eb), _ptr(codes), _ptr(un),
                                 C.byref(n)), "stz_quantize")
        return codes, un[: n.value].copy()

    def dequantize(self, codes, unpred, eb, dtype):
        codes = np.ascontiguousarray(codes, dtype=np.uint16)
        unpred = np.ascontiguousarray(unpred, dtype=dtype)
        out = np.empty(codes.size, dtype=dtype)
        dt = F32 if np.dtype(dtype) == np.float32 else F64
        up = _ptr(unpred) if unpred.size else C.c_void_p()
        check(lib().stz_dequantize(self.h, dt, _ptr(codes), codes.size, float(eb), up, _ptr(out)),
              "stz_dequantize")
        return out

    def residual_encode(self, block, sub, ref, eb, with_rec=True):
        """Fused predict + quantize of one block: (codes (z,y,x) uint16,
        unpredictable residuals in element order, reconstruction or None)."""
        sub = np.ascontiguousarray(sub)
        ref = np.ascontiguousarray(ref, dtype=sub.dtype)
        dz, dy, dx = sub.shape
        codes = np.empty(sub.shape, dtype=np.uint16)
        un = np.empty(sub.size, dtype=sub.dtype)
        rec = np.empty_like(sub) if with_rec else None
        n = C.c_size_t()
        check(lib().stz_residual_encode(self.h, _dtype_id(sub), block, _ptr(sub), _ptr(ref), dx, dy, dz, float(eb),
                                        _ptr(codes), _ptr(un), C.byref(n),
                                        _ptr(rec) if rec is not None else C.c_void_p()), "stz_residual_encode")
        return codes, un[: n.value].copy(), rec

    def residual_decode(self, block, codes, unpred, ref, eb):
        codes = np.ascontiguousarray(codes, dtype=np.uint16)
        ref = np.ascontiguousarray(ref)
        unpred = np.ascontiguousarray(unpred, dtype=ref.dtype)
        dz, dy, dx = ref.shape
        out = np.empty_like(ref)
        up = _ptr(unpred) if unpred.size else C.c_void_p()
        check(lib().stz_residual_decode(self.h, _dtype_id(ref), block, _ptr(codes), up, _ptr(ref), dx, dy, dz,
                                        float(eb), _ptr(out)), "stz_residual_decode")
        return out

    def huff_encode_chunks(self, codes, unpred, eb, chunk, dtype):
        """One block's codes (z,y,x) + unpredictables -> a chunk-marked stream (uint8 array)."""
        codes = np.ascontiguousarray(codes, dtype=np.uint16)
        unpred = np.ascontiguousarray(unpred, dtype=dtype)
        dz, dy, dx = codes.shape
        dt = F32 if np.dtype(dtype) == np.float32 else F64
        s = C.c_void_p()
        n = C.c_size_t()
        up = _ptr(unpred) if unpred.size else C.c_void_p()
        check(lib().stz_huff_encode_chunks(self.h, dt, _ptr(codes), up, unpred.size, dx, dy, dz, int(chunk),
                                           float(eb), C.byref(s), C.byref(n)), "stz_huff_encode_chunks")
        out = np.ctypeslib.as_array(C.cast(s, C.POINTER(C.c_uint8)), shape=(n.value,)).copy()
        lib().stz_free(s)
        return out

    def huff_decode_chunks(self, stream, dims, dtype, chunk_ids=None, out=None, codes_out=None):
        """Decode the listed chunks (row-major ids; None: all) of a block stream
        into out (z,y,x), leaving other points as they were.  Returns
        (out, codes_out, decoded_symbols, unpred_zeros)."""
        stream = np.ascontiguousarray(stream, dtype=np.uint8)
        dx, dy, dz = dims
        if out is None:
            out = np.zeros((dz, dy, dx), dtype=dtype)
        ids = None if chunk_ids is None else np.ascontiguousarray(chunk_ids, dtype=np.int64)
        ns, nz = C.c_size_t(), C.c_size_t()
        check(lib().stz_huff_decode_chunks(self.h, _ptr(stream), stream.size,
                                           _ptr(ids) if ids is not None else C.c_void_p(),
                                           -1 if ids is None else ids.size, _ptr(out),
                                           _ptr(codes_out) if codes_out is not None else C.c_void_p(),
                                           C.byref(ns), C.byref(nz)), "stz_huff_decode_chunks")
        return out, codes_out, ns.value, nz.value


def huff_lengths(freq, lmax=24):
    freq = np.ascontiguousarray(freq, dtype=np.uint64)
    out = np.zeros(freq.size, dtype=np.uint8)
    check(lib().stz_huff_lengths(_ptr(freq), freq.size, lmax, _ptr(out)), "stz_huff_lengths")
    return out


def plan_query(full_dims, q, chunk_high=32, chunk_low=16, cap=1 << 16):
    fx, fy, fz = full_dims
    query = np.zeros(7, dtype=np.int32)
    hp = np.zeros((7, 8), dtype=np.int32)
    lp = np.zeros((7, 8), dtype=np.int32)
    hc = np.zeros((7, cap), dtype=np.int64)
    lc = np.zeros((7, cap), dtype=np.int64)
    check(lib().stz_plan_query(fx, fy, fz, chunk_high, chunk_low, *q, _ptr(query), _ptr(hp), _ptr(hc), _ptr(lp),
                               _ptr(lc), cap), "stz_plan_query")
    high = {b: (tuple(int(v) for v in hp[b, 1:7]), hc[b, : hp[b, 7]].copy()) for b in range(7) if hp[b, 0]}
    low = {b: (tuple(int(v) for v in lp[b, 1:7]), lc[b, : lp[b, 7]].copy()) for b in range(7) if lp[b, 0]}
    return tuple(int(v) for v in query[:6]), bool(query[6]), high, low
