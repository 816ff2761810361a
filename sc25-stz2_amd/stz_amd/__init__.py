"""STZ compress/decompress hot path on MI355X (gfx950): Python host interface.

The compute path is libstz_hip.so (hand-written HIP kernels behind the C ABI
in include/stz.h); this package only binds it.
"""
from ._lib import EXPORTS, LIB_PATH, Info, Params, StzError, lib  # noqa: F401
from .api import (F32, F64, Context, DStream, default_params, huff_lengths, plan_query,  # noqa: F401
                  stream_info)
