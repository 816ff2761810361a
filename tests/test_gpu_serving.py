"""Random access in batches and the continuous-batching server: every result
is bit-identical alone and in a mix, and equals the crop of the full decode
(the reference's own invariant, F5:2609-2640)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _requests(dims, rng, n_box=12):
    nx, ny, nz = dims
    reqs = []
    for _ in range(n_box):
        a = [int(rng.integers(-4, d + 4)) for d in dims]
        b = [int(rng.integers(-4, d + 4)) for d in dims]
        for k, d in enumerate(dims):      # a box wholly past the far end is an error
            if min(a[k], b[k]) >= d:
                a[k] = d - 1
        reqs.append(("box", a[0], a[1], a[2], b[0], b[1], b[2]))
    reqs += [("box", 0, 0, 0, nx - 1, ny - 1, nz - 1), ("box", 5, 6, 7, 5, 30, 40), ("box", 3, 3, 3, 3, 3, 3)]
    reqs += [("slice", 0, nx // 3), ("slice", 1, ny - 1), ("slice", 2, 0), ("slice", 2, nz // 2)]
    reqs += [("level", 0), ("level", 1), ("level", 2), ("level", 3)]
    return reqs


def _expected(req, full):
    nz, ny, nx = full.shape
    if req[0] == "box":
        x0, y0, z0, x1, y1, z1 = req[1:]

        def ax(a, b, d):      # make_query / normalize_axis, F5:1622-1644
            if a == b:
                v = min(max(a, 0), d - 1)
                return slice(v, v + 1)
            lo, hi = max(0, min(a, b)), min(d - 1, max(a, b))
            return slice(lo, max(hi, lo) + 1)
        return full[ax(z0, z1, nz), ax(y0, y1, ny), ax(x0, x1, nx)]
    if req[0] == "slice":
        sl = [slice(None)] * 3
        sl[2 - req[1]] = slice(req[2], req[2] + 1)
        return full[tuple(sl)]
    step = {0: 8, 1: 4, 2: 2, 3: 1}[req[1]]
    return full[::step, ::step, ::step]


@pytest.mark.parametrize("shape,dt,keep_base", [((60, 72, 88), np.float32, True), ((64, 64, 64), np.float32, False),
                                                ((37, 50, 45), np.float64, True)])
def test_batch_matches_alone_and_full(stz, ctx, shape, dt, keep_base):
    from stz_amd import fields
    f = fields.sinusoids(shape=shape, seed=7, dtype=dt)
    ds = ctx.compress_device(f, 1e-3, "rel", chunk_high=8, chunk_low=4)
    full = ds.decompress()
    ds.prepare(keep_base=keep_base)
    rng = np.random.default_rng(sum(shape))
    reqs = _requests(shape[::-1], rng)
    together = ds.decompress_batch(reqs)
    for r, got in zip(reqs, together):
        alone = ds.decompress_batch([r])[0]
        want = _expected(r, full)
        assert got.shape == want.shape, r
        assert np.array_equal(_bits(got), _bits(want)), r
        assert np.array_equal(_bits(alone), _bits(got)), r
    ds.free()


def test_roi_entry_points_use_same_path(stz, ctx):
    from stz_amd import fields
    f = fields.sinusoids(shape=(48, 40, 56), seed=3)
    ds = ctx.compress_device(f, 1e-3, "rel")
    full = ds.decompress()
    roi = ds.decompress_roi(50, 2, 30, 10, 33, 5)
    assert np.array_equal(_bits(roi), _bits(full[5:31, 2:34, 10:51]))
    for lv, st in [(0, 8), (1, 4), (2, 2)]:
        assert np.array_equal(_bits(ds.decompress_level(lv)), _bits(full[::st, ::st, ::st]))
    ds.free()


def test_box_past_the_field_is_an_error(stz, ctx):
    from stz_amd import fields
    f = fields.sinusoids(shape=(16, 16, 16), seed=4)
    ds = ctx.compress_device(f, 1e-3, "rel")
    with pytest.raises(stz.StzError):
        ds.decompress_batch([("box", 0, 0, 17, 15, 15, 19)])
    with pytest.raises(stz.StzError):
        ds.decompress_batch([("slice", 2, 16)])
    assert ds.decompress_batch([("box", 0, 0, -5, 15, 15, -2)])[0].shape == (1, 16, 16)
    ds.free()


def test_grid_share_leaves_results_unchanged(stz, ctx):
    # 256^3: the persistent level grids exceed a quarter of the resident slots
    import torch
    from stz_amd import fields
    dev = torch.device("cuda:0")
    g = fields.grf_device((256, 256, 256), -3.0, 21, dev, torch.float32)
    ds = ctx.compress_device(g, 1e-3, "rel")
    want = ds.decompress()
    want2 = ds.decompress_batch([("level", 2)])[0]
    c2 = stz.Context(0)
    try:
        c2.set_grid_share(0.25)
        got = ds.decompress(ctx=c2)
        got2 = ds.decompress_batch([("level", 2)], ctx=c2)[0]
        with pytest.raises(stz.StzError):
            c2.set_grid_share(0.0)
        with pytest.raises(stz.StzError):
            c2.set_grid_share(1.5)
    finally:
        c2.close()
    assert np.array_equal(_bits(got), _bits(want))
    assert np.array_equal(_bits(got2), _bits(want2))
    ds.free()


def test_server_results_identical_alone_and_in_mix(stz, ctx):
    import torch
    from stz_amd import fields, serve
    dev = torch.device("cuda:0")
    f = fields.grf((96, 112, 128), slope=-3.0, seed=11).astype(np.float32)
    ds = ctx.compress_device(torch.from_numpy(f).to(dev), 1e-3, "rel")
    full = ds.decompress()
    reqs, arr = serve.poisson_trace((128, 112, 96), 240, rate=20000.0, seed=5, box_edge=(8, 64))
    srv = serve.RoiServer(ds, workers=3, max_batch=16, mem_budget=1 << 30, keep_results=True)
    try:
        recs, t0 = srv.run_trace(reqs, arr)
    finally:
        srv.close()
    sizes = [b[1] for b in srv.batches]
    assert max(sizes) > 1, "no batching happened"
    spans = sorted((b[2], b[3], b[0]) for b in srv.batches)
    overlap = any(s2 < e1 and w1 != w2 for (s1, e1, w1), (s2, e2, w2) in zip(spans, spans[1:]))
    assert overlap, "no two batches were in flight together"
    for p in recs:
        got = p.out.cpu().numpy()
        alone = ds.decompress_batch([p.req])[0]
        want = _expected(p.req, full)
        assert np.array_equal(_bits(got), _bits(alone)), p.req
        assert np.array_equal(_bits(got), _bits(want)), p.req
    stats, gbs, wall = serve.summarize(recs, t0, 4)
    assert stats["all"]["n"] == len(reqs) and gbs > 0
    ds.free()


def test_server_coalesces_identical_requests(stz, ctx):
    """Identical requests queued together are served by one batch, decoded
    once, and each gets the request's result."""
    from stz_amd import serve
    from stz_amd import fields
    f = fields.sinusoids(shape=(40, 48, 56), seed=4)
    ds = ctx.compress_device(f, 1e-3, "rel")
    full = ds.decompress()
    reqs = [("level", 3)] * 3 + [("box", 2, 3, 4, 20, 30, 25)] * 4 + [("level", 1)]
    srv = serve.RoiServer(ds, workers=1, max_batch=16, mem_budget=1 << 30, keep_results=True)
    try:
        recs, _ = srv.run_trace(reqs, [0.0] * len(reqs))
    finally:
        srv.close()
    for p in recs:
        assert np.array_equal(_bits(p.out.cpu().numpy()), _bits(_expected(p.req, full))), p.req
    # one worker, all queued at once: the duplicates ride in shared batches
    assert len(srv.batches) < len(reqs) and max(b[1] for b in srv.batches) > 1
    ds.free()


def test_server_tiny_memory_budget_serializes(stz, ctx):
    import torch
    from stz_amd import fields, serve
    f = fields.sinusoids(shape=(40, 48, 56), seed=2)
    ds = ctx.compress_device(f, 1e-3, "rel")
    full = ds.decompress()
    reqs = [("box", 1, 2, 3, 30, 20, 10), ("slice", 1, 7), ("level", 2), ("box", 0, 0, 0, 55, 47, 39)]
    srv = serve.RoiServer(ds, workers=2, max_batch=8, mem_budget=1, keep_results=True)
    try:
        recs, _ = srv.run_trace(reqs, [0.0] * len(reqs))
    finally:
        srv.close()
    assert all(b[1] == 1 for b in srv.batches)
    for p in recs:
        assert np.array_equal(_bits(p.out.cpu().numpy()), _bits(_expected(p.req, full)))
    ds.free()
