"""Writes tests/golden/ref_vectors.npz: input/output vectors of the reference's
own functions, computed by oracle/_ref/libstz_ref.so (those functions compiled
from the reference source where it lies, by oracle/build_ref.sh).

Covered (F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp):
  preprocess_block_zslab / depreprocess_block_zslab   F5:1145-1191 (predictors F5:141-882)
  quantize_range_nodec                                F5:1204-1253
  slice_full_data / merge_sub_blocks_to_full          F5:68-122
  make_query + build_residual_plan + build_low_residual_plan   F5:1602-1830

and tests/golden/ref_pred_vectors.npz: preprocess_block_zslab of F5 built
with its predictor switches PRED_WPOW2 (F5:177-203), PRED_PLANE_AXIS
(F5:579-607, F5:627-633) and PRED_CUBE_AXIS (F5:609-622, F5:675-686), one
build each and one with all three (oracle/_ref/libstz_ref_<name>.so); and
preprocess_block of stz2.cpp (stz2.cpp:82-306, its 12-point truncated
bicubic and 16-point predictors; oracle/_ref/libstz_ref_stz2.so).

Run in a container that holds the reference sources (after build_ref.sh):
    python tests/golden/make_golden.py
The vectors are data only; nothing of the reference's source is stored.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libstz_ref.so")
OUT = os.path.join(HERE, "ref_vectors.npz")

SHAPES = [(9, 7, 11), (4, 4, 4), (3, 5, 6), (6, 10, 12), (2, 2, 2)]   # (dz, dy, dx)
PLAN_CAP = 4096


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def residual_inputs(rng, shape, dt):
    """A rough field and a nearby target; a zero plane, denormals and large
    values put every boundary rule and the flush-to-zero mode to work."""
    ref = np.cumsum(rng.normal(size=shape), axis=2)
    sub = ref + rng.normal(scale=0.1, size=shape)
    ref = ref.astype(dt)
    sub = sub.astype(dt)
    ref[0] = 0
    tiny = np.float32(1e-39) if dt == np.float32 else np.float64(1e-310)
    sub[0] = tiny
    if shape[1] > 2:
        ref[:, 1, :] *= dt(1e6)
    return sub, ref


def quantize_inputs(rng, dt):
    d = rng.normal(size=3000) * rng.choice([1e-4, 1e-2, 1.0, 70.0, 1e5], size=3000)
    special = [np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-300, 1e-39, 65535.0, -65535.0]
    return np.concatenate([d, special]).astype(dt)


def plan_cases(rng, n):
    cases = []
    for it in range(n):
        ch, cl = [(32, 16), (8, 4), (4, 4), (16, 8)][it % 4]
        lcm = 4 * cl * 2 * ch // math.gcd(4 * cl, 2 * ch)
        dims = [lcm * int(rng.integers(1, 4)) for _ in range(3)]
        q = [int(rng.integers(-3, d + 3)) for d in dims] + [int(rng.integers(-3, d + 3)) for d in dims]
        if it % 5 == 0:
            q[5] = q[2]      # a z slice
        cases.append((dims, ch, cl, q))
    return cases


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/libstz_ref.so is missing: run oracle/build_ref.sh first")
    R = C.CDLL(REF)
    R.ref_plan_query.argtypes = [C.c_int] * 11 + [C.c_void_p] * 4 + [C.c_void_p, C.c_int64]
    rng = np.random.default_rng(20251)
    out = {}
    for dt, suf in [(np.float32, "f32"), (np.float64, "f64")]:
        for si, shape in enumerate(SHAPES):
            dz, dy, dx = shape
            sub, ref = residual_inputs(rng, shape, dt)
            de = (rng.normal(scale=0.05, size=shape)).astype(dt)
            out["res_%s_%d_sub" % (suf, si)] = sub
            out["res_%s_%d_ref" % (suf, si)] = ref
            out["res_%s_%d_de" % (suf, si)] = de
            for b in range(7):
                r = np.empty_like(sub)
                getattr(R, "ref_preprocess_" + suf)(b, _p(sub), _p(ref), _p(r), dx, dy, dz)
                out["res_%s_%d_pre%d" % (suf, si, b)] = r.copy()
                getattr(R, "ref_depreprocess_" + suf)(b, _p(de), _p(ref), _p(r), dx, dy, dz)
                out["res_%s_%d_depre%d" % (suf, si, b)] = r.copy()
        d = quantize_inputs(rng, dt)
        for ei, eb in enumerate([1e-3, 0.37, 1e-7]):
            codes = np.empty(d.size, np.int32)
            un = np.empty(d.size, dt)
            nu = C.c_size_t()
            getattr(R, "ref_quantize_" + suf)(_p(d), C.c_size_t(d.size), C.c_double(eb), 32768, _p(codes),
                                             _p(un), C.byref(nu))
            out["q_%s_%d_codes" % (suf, ei)] = codes
            out["q_%s_%d_unpred" % (suf, ei)] = un[:nu.value].copy()
            out["q_%s_%d_eb" % (suf, ei)] = np.float64(eb)
        out["q_%s_in" % suf] = d
    # split / merge
    for si, (nz, ny, nx) in enumerate([(8, 6, 10), (4, 4, 4)]):
        full = rng.normal(size=(nz, ny, nx)).astype(np.float32)
        subs = [np.empty((nz // 2, ny // 2, nx // 2), np.float32) for _ in range(8)]
        arr = (C.c_void_p * 8)(*[_p(s) for s in subs])
        R.ref_split_f32(_p(full), arr, nx, ny, nz)
        out["split_%d_full" % si] = full
        out["split_%d_subs" % si] = np.stack(subs)
    # planner
    cases = plan_cases(rng, 96)
    pin, pquery, phigh, plow, phc, plc = [], [], [], [], [], []
    for dims, ch, cl, q in cases:
        query = np.zeros(7, np.int32)
        hp = np.zeros((7, 8), np.int32)
        lp = np.zeros((7, 8), np.int32)
        hc = np.zeros((7, PLAN_CAP), np.int64)
        lc = np.zeros((7, PLAN_CAP), np.int64)
        rc = R.ref_plan_query(*dims, ch, cl, *q, _p(query), _p(hp), _p(hc), _p(lp), _p(lc), PLAN_CAP)
        assert rc == 0
        pin.append(dims + [ch, cl] + q)
        pquery.append(query)
        phigh.append(hp)
        plow.append(lp)
        phc.append(np.concatenate([hc[b, :hp[b, 7]] for b in range(7)]))
        plc.append(np.concatenate([lc[b, :lp[b, 7]] for b in range(7)]))
    out["plan_in"] = np.array(pin, np.int32)
    out["plan_query"] = np.stack(pquery)
    out["plan_high"] = np.stack(phigh)
    out["plan_low"] = np.stack(plow)
    out["plan_high_chunks"] = np.concatenate(phc).astype(np.int32)
    out["plan_low_chunks"] = np.concatenate(plc).astype(np.int32)
    np.savez_compressed(OUT, **out)
    print("wrote %s (%d arrays, %d bytes)" % (OUT, len(out), os.path.getsize(OUT)))


PRED_OUT = os.path.join(HERE, "ref_pred_vectors.npz")
PRED_SHAPES = [(6, 7, 8), (9, 7, 11), (12, 10, 14)]   # (dz, dy, dx)


def pred_inputs(rng, shape, dt):
    """Rough references with a zero plane and scaled columns (wide variance
    ranges in the inverse-variance weights)."""
    r = np.cumsum(rng.normal(size=shape), axis=1)
    r[:, :, 2::5] *= 40.0
    s = r + rng.normal(scale=0.3, size=shape)
    r = r.astype(dt)
    s = s.astype(dt)
    r[0] = 0
    return s, r


def main_pred():
    out = {}
    for name in ("wpow2", "plane", "cube", "all"):
        R = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libstz_ref_%s.so" % name))
        rng = np.random.default_rng(2024)
        for dt, dtn in ((np.float32, "f32"), (np.float64, "f64")):
            for si, shape in enumerate(PRED_SHAPES):
                dz, dy, dx = shape
                sub, ref = pred_inputs(rng, shape, dt)
                out["%s-%s-%d-sub" % (name, dtn, si)] = sub
                out["%s-%s-%d-ref" % (name, dtn, si)] = ref
                for b in (2, 4, 5, 6):
                    d = np.empty_like(sub)
                    getattr(R, "ref_preprocess_" + dtn)(b, _p(sub), _p(ref), _p(d), dx, dy, dz)
                    out["%s-%s-%d-%d_out" % (name, dtn, si, b)] = d
    # stz2.cpp's preprocess_block (f32): its plane and cube predictors
    R = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libstz_ref_stz2.so"))
    rng = np.random.default_rng(2025)
    for si, shape in enumerate(PRED_SHAPES):
        dz, dy, dx = shape
        sub, ref = pred_inputs(rng, shape, np.float32)
        out["stz2-f32-%d-sub" % si] = sub
        out["stz2-f32-%d-ref" % si] = ref
        for b in (2, 4, 5, 6):
            d = np.empty_like(sub)
            R.ref2_preprocess_f32(b, _p(sub), _p(ref), _p(d), dx, dy, dz)
            out["stz2-f32-%d-%d_out" % (si, b)] = d
    # dependency.cpp's staged level decode (f32), composed in the order of
    # dependency.cpp:594-636: blocks 0, 1, 3 by depreprocess_block from the
    # coarser grid, then 2, 4, 5 and last 6 by depreprocess_block_staged
    # from the reconstructed blocks
    R = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libstz_ref_dep.so"))
    rng = np.random.default_rng(2026)
    for si, shape in enumerate(PRED_SHAPES + [(4, 4, 4), (5, 3, 9)]):
        dz, dy, dx = shape
        _, ref = pred_inputs(rng, shape, np.float32)
        res = rng.normal(scale=0.5, size=(7,) + shape).astype(np.float32)
        res[:, :, :, 1::4] *= 30.0
        rec = [np.empty(shape, np.float32) for _ in range(7)]
        for b in (0, 1, 3):
            R.ref3_depreprocess_f32(b, _p(res[b]), _p(ref), _p(rec[b]), dx, dy, dz)
        rb = (C.c_void_p * 7)(*[r.ctypes.data for r in rec])
        for b in (2, 4, 5, 6):
            R.ref3_depreprocess_staged_f32(b, _p(np.ascontiguousarray(res[b])), rb, _p(rec[b]), dx, dy, dz)
        out["staged-f32-%d-ref" % si] = ref
        out["staged-f32-%d-res" % si] = res
        out["staged-f32-%d-recs" % si] = np.stack(rec)
    np.savez_compressed(PRED_OUT, **out)
    print("wrote %s (%d arrays, %d bytes)" % (PRED_OUT, len(out), os.path.getsize(PRED_OUT)))


if __name__ == "__main__":
    main()
    main_pred()
