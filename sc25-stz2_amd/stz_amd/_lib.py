"""ctypes loader for libstz_hip.so (the C ABI declared in include/stz.h).

Fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# STZ_HIP_LIB: another build of the same library (the host-sanitized one, for instance)
LIB_PATH = os.environ.get("STZ_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libstz_hip.so")
_lib = None

STZ_OK = 0
ERRORS = {-1: "STZ_E_ARG", -2: "STZ_E_DIMS", -3: "STZ_E_STREAM", -4: "STZ_E_OOM", -5: "STZ_E_HIP",
          -6: "STZ_E_ZSTD", -7: "STZ_E_NODEV", -8: "STZ_E_SCRATCH"}

# every entry point include/stz.h declares
EXPORTS = [
    "stz_create", "stz_create_ex", "stz_destroy", "stz_set_stream", "stz_get_stream", "stz_set_grid_share", "stz_profile",
    "stz_profile_report",
    "stz_default_params", "stz_strerror", "stz_free",
    "stz_compress", "stz_decompress", "stz_decompress_roi", "stz_decompress_slice", "stz_decompress_level",
    "stz_query_info", "stz_stream_check",
    "stz_compress_device", "stz_dstream_upload", "stz_dstream_download", "stz_dstream_info",
    "stz_dstream_decompress", "stz_dstream_decompress_roi", "stz_dstream_decompress_level",
    "stz_dstream_payload", "stz_dstream_free", "stz_dstream_prepare", "stz_dstream_decompress_batch",
    "stz_split", "stz_merge", "stz_residual_predict", "stz_residual_reconstruct", "stz_quantize",
    "stz_dequantize", "stz_huff_lengths", "stz_plan_query",
    "stz_residual_encode", "stz_residual_decode", "stz_huff_encode_chunks", "stz_huff_decode_chunks",
    "stz_residual_encode_noalloc",
    "stz_tune_config", "stz_tune_features", "stz_tune_set_model", "stz_tune", "stz_tune_trial",
    "stz_mlp_forward", "stz_mlp_grad", "stz_mlp_adam",
    "stz_compress_batch", "stz_decompress_batch", "stz_autotune",
]
TUNE_CONFIGS, TUNE_FEATURES, MLP_PARAMS, MLP_OUTPUTS = 204, 16, 18768, 208


class StzError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("%s failed: %s (%d)" % (what, ERRORS.get(code, "?"), code))


class Params(C.Structure):
    _fields_ = [
        ("mult_base", C.c_double), ("mult_low", C.c_double), ("mult_high", C.c_double),
        ("chunk_high", C.c_int), ("chunk_low", C.c_int),
        ("interp", C.c_int), ("order", C.c_int),
        ("alpha", C.c_double), ("beta", C.c_double),
        ("zstd_level", C.c_int), ("tuner", C.c_int), ("pred", C.c_int),
    ]


REQ_BOX, REQ_SLICE, REQ_LEVEL = 0, 1, 2


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_int * 6)]


class Info(C.Structure):
    _fields_ = [
        ("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
        ("px", C.c_size_t), ("py", C.c_size_t), ("pz", C.c_size_t),
        ("dtype", C.c_int),
        ("eb_base", C.c_double), ("eb_low", C.c_double), ("eb_high", C.c_double),
        ("payload_bytes", C.c_size_t),
        ("interp", C.c_int), ("order", C.c_int),
        ("mult_base", C.c_double), ("mult_low", C.c_double), ("mult_high", C.c_double),
        ("alpha", C.c_double), ("beta", C.c_double),
        ("pred", C.c_int), ("tuner", C.c_int), ("tune_confidence", C.c_float),
    ]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libstz_hip.so not built (%s): run python sc25-stz2_amd/build.py" % LIB_PATH)
    # torch ships its own libamdhip64 (same SONAME as /opt/rocm's).  Loading
    # torch first makes this library bind to that one runtime, so device
    # pointers from torch tensors are valid here; the other order would put
    # two HIP runtimes in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, sz, i32, dbl = C.c_void_p, C.c_size_t, C.c_int, C.c_double
    P = C.POINTER
    L.stz_create.argtypes = [P(vp), i32]
    L.stz_create_ex.argtypes = [P(vp), i32, C.c_uint]
    L.stz_destroy.argtypes = [vp]
    L.stz_destroy.restype = None
    L.stz_set_stream.argtypes = [vp, vp]
    L.stz_get_stream.argtypes = [vp]
    L.stz_get_stream.restype = vp
    L.stz_set_grid_share.argtypes = [vp, C.c_float]
    L.stz_profile.argtypes = [vp, i32]
    L.stz_profile_report.argtypes = [vp, C.c_char_p, sz]
    L.stz_default_params.argtypes = [P(Params)]
    L.stz_default_params.restype = None
    L.stz_strerror.argtypes = [i32]
    L.stz_strerror.restype = C.c_char_p
    L.stz_free.argtypes = [vp]
    L.stz_free.restype = None
    L.stz_compress.argtypes = [vp, i32, vp, sz, sz, sz, dbl, i32, P(Params), P(vp), P(sz)]
    L.stz_decompress.argtypes = [vp, vp, sz, vp]
    L.stz_decompress_roi.argtypes = [vp, vp, sz, i32, i32, i32, i32, i32, i32, vp, P(sz)]
    L.stz_decompress_slice.argtypes = [vp, vp, sz, i32, i32, vp, P(sz)]
    L.stz_decompress_level.argtypes = [vp, vp, sz, i32, vp, P(sz)]
    L.stz_query_info.argtypes = [vp, sz, P(Info)]
    L.stz_stream_check.argtypes = [vp, sz]
    L.stz_compress_device.argtypes = [vp, i32, vp, sz, sz, sz, dbl, i32, P(Params), P(vp)]
    L.stz_dstream_upload.argtypes = [vp, vp, sz, P(vp)]
    L.stz_dstream_download.argtypes = [vp, vp, i32, P(vp), P(sz)]
    L.stz_dstream_info.argtypes = [vp, P(Info)]
    L.stz_dstream_decompress.argtypes = [vp, vp, vp]
    L.stz_dstream_decompress_roi.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, vp, P(sz)]
    L.stz_dstream_decompress_level.argtypes = [vp, vp, i32, vp, P(sz)]
    L.stz_dstream_payload.argtypes = [vp, P(vp), P(sz)]
    L.stz_dstream_prepare.argtypes = [vp, vp, i32]
    L.stz_dstream_decompress_batch.argtypes = [vp, vp, i32, P(Request), P(vp), P(sz)]
    L.stz_dstream_free.argtypes = [vp]
    L.stz_dstream_free.restype = None
    L.stz_split.argtypes = [vp, i32, vp, sz, sz, sz, P(vp)]
    L.stz_merge.argtypes = [vp, i32, P(vp), sz, sz, sz, vp]
    L.stz_residual_predict.argtypes = [vp, i32, i32, vp, vp, sz, sz, sz, vp]
    L.stz_residual_reconstruct.argtypes = [vp, i32, i32, vp, vp, sz, sz, sz, vp]
    L.stz_quantize.argtypes = [vp, i32, vp, sz, dbl, vp, vp, P(sz)]
    L.stz_dequantize.argtypes = [vp, i32, vp, sz, dbl, vp, vp]
    L.stz_huff_lengths.argtypes = [vp, i32, i32, vp]
    L.stz_residual_encode.argtypes = [vp, i32, i32, vp, vp, sz, sz, sz, dbl, vp, vp, P(sz), vp]
    L.stz_residual_decode.argtypes = [vp, i32, i32, vp, vp, vp, sz, sz, sz, dbl, vp]
    L.stz_huff_encode_chunks.argtypes = [vp, i32, vp, vp, sz, sz, sz, sz, i32, dbl, P(vp), P(sz)]
    L.stz_residual_encode_noalloc.argtypes = [vp, i32, i32, vp, vp, sz, sz, sz, dbl, vp, vp, vp, vp]
    L.stz_huff_decode_chunks.argtypes = [vp, vp, sz, vp, C.c_int64, vp, vp, P(sz), P(sz)]
    L.stz_plan_query.argtypes = [i32] * 11 + [vp, vp, vp, vp, vp, C.c_int64]
    L.stz_compress_batch.argtypes = [vp, i32, i32, P(vp), P(sz), dbl, i32, P(Params), P(vp)]
    L.stz_decompress_batch.argtypes = [vp, i32, P(vp), P(vp)]
    L.stz_autotune.argtypes = [vp, i32, C.c_char_p, sz]
    f32 = C.c_float
    L.stz_tune_config.argtypes = [i32, P(Params)]
    L.stz_tune_features.argtypes = [vp, i32, vp, sz, sz, sz, dbl, vp, sz, P(sz)]
    L.stz_tune_set_model.argtypes = [vp, vp, sz, f32]
    L.stz_tune.argtypes = [vp, i32, vp, sz, sz, sz, dbl, i32, P(Params), P(f32), P(i32)]
    L.stz_tune_trial.argtypes = [vp, i32, vp, sz, sz, sz, dbl, i32, P(Params), vp, P(i32)]
    L.stz_mlp_forward.argtypes = [vp, vp, vp, i32, vp]
    L.stz_mlp_grad.argtypes = [vp, vp, vp, vp, i32, vp, vp]
    L.stz_mlp_adam.argtypes = [vp, vp, vp, vp, vp, i32, f32, f32]
    _lib = L
    return L


def check(rc, what):
    if rc != STZ_OK:
        raise StzError(rc, what)
