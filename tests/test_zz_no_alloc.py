"""The allocation-free residual encode (stz_residual_encode_noalloc) through
the C ABI: device buffers only, the residual scratch a per-GPU static array.
Its codes, unpredictable list and reconstruction equal stz_residual_encode's
bit for bit on consecutive calls (the second on a smaller block, so the
scratch still holds the first call's larger contents), f32 and f64; sizes
past the scratch and host buffers are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STZ_E_ARG, STZ_E_SCRATCH = -1, -8


def _p(t):
    return C.c_void_p(t.data_ptr())


def _noalloc(stz, ctx, block, sub, ref, eb, with_rec=True):
    import torch
    dz, dy, dx = sub.shape
    codes = torch.empty(sub.shape, dtype=torch.int16, device=sub.device)
    unpred = torch.empty(sub.numel(), dtype=sub.dtype, device=sub.device)
    cnt = torch.zeros(1, dtype=torch.int64, device=sub.device)
    rec = torch.empty_like(sub) if with_rec else None
    torch.cuda.synchronize()
    dt = 0 if sub.dtype == torch.float32 else 1
    rc = stz.lib().stz_residual_encode_noalloc(ctx.h, dt, block, _p(sub), _p(ref), dx, dy, dz, float(eb), _p(codes),
                                              _p(unpred), _p(cnt), _p(rec) if rec is not None else None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    n = int(cnt.item())
    return (codes.cpu().numpy().view(np.uint16), unpred[:n].cpu().numpy(),
            rec.cpu().numpy() if rec is not None else None)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _case(dims, dtype, seed):
    rng = np.random.default_rng(seed)
    ref = np.cumsum(rng.normal(size=dims), axis=2).astype(dtype)
    sub = (ref + rng.normal(scale=0.2, size=dims)).astype(dtype)
    sub.ravel()[::29] = 1e7              # unpredictable residuals
    return sub, ref


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_residual_encode(stz, ctx, dtype):
    import torch
    dev = torch.device("cuda:0")
    # two calls in a row, the second smaller; then a third with no reconstruction
    for dims, block, eb, with_rec in (((40, 48, 56), 6, 1e-2, True), ((9, 7, 11), 2, 1e-3, True),
                                      ((5, 6, 3), 0, 0.5, False)):
        sub, ref = _case(dims, dtype, sum(dims) + block)
        codes, un, rec = _noalloc(stz, ctx, block, torch.from_numpy(sub).to(dev), torch.from_numpy(ref).to(dev), eb,
                                  with_rec)
        wc, wu, wr = ctx.residual_encode(block, sub, ref, eb)
        assert np.array_equal(codes, wc), (dims, block)
        assert un.size == wu.size and np.array_equal(_bits(un), _bits(wu)), (dims, block)
        assert un.size > 0
        if with_rec:
            assert np.array_equal(_bits(rec), _bits(wr)), (dims, block)


def test_noalloc_refuses_before_launch(stz, ctx):
    import torch
    dev = torch.device("cuda:0")
    L = stz.lib()
    t = torch.zeros(64, dtype=torch.float32, device=dev)
    c16 = torch.zeros(64, dtype=torch.int16, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    cap = 32 << 20
    # more points than the scratch holds (the size check comes first)
    assert L.stz_residual_encode_noalloc(ctx.h, 0, 1, _p(t), _p(t), cap // 4 + 1, 1, 1, 1e-3, _p(c16), _p(t),
                                         _p(cnt), None) == STZ_E_SCRATCH
    assert L.stz_residual_encode_noalloc(ctx.h, 1, 1, _p(t), _p(t), 2048, 2048, 2, 1e-3, _p(c16), _p(t), _p(cnt),
                                         None) == STZ_E_SCRATCH
    assert L.stz_residual_encode_noalloc(ctx.h, 0, 1, _p(t), _p(t), 1 << 40, 1 << 40, 1 << 40, 1e-3, _p(c16),
                                         _p(t), _p(cnt), None) == STZ_E_SCRATCH
    # a host buffer is refused (nothing would stage it)
    h = np.zeros(64, np.float32)
    assert L.stz_residual_encode_noalloc(ctx.h, 0, 1, h.ctypes.data_as(C.c_void_p), _p(t), 4, 4, 4, 1e-3, _p(c16),
                                         _p(t), _p(cnt), None) == STZ_E_ARG
    assert L.stz_residual_encode_noalloc(ctx.h, 0, 7, _p(t), _p(t), 4, 4, 4, 1e-3, _p(c16), _p(t), _p(cnt),
                                         None) == STZ_E_ARG
