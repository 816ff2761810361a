"""Benchmark: STZ compress + decompress throughput on one MI355X.

Workload (the metric's configuration): 512^3 fp32 log-normal density field,
rho = exp(1.5 g), g a unit-variance Gaussian random field with P(k) ~ k^-3
(seed 1002), value-range relative error bound 1e-3, cubic base interpolation,
F5 level multipliers 0.7 / 2.7 / 8.2, marker chunks 32^3 / 16^3.

A step = one compress (field resident in HBM -> device payload) plus one
decompress (device payload -> field in HBM).  value = 2 * N * 4 bytes / (step
time), GB/s.  Host zstd of the payload is timed separately (zstd_ms), as is
the D2H copy it needs.  The dominant kernel's roofline figure comes from HIP
events around each launch on the library's stream, in extra profiled steps
after the timed ones.  cpu_baseline: the CPU oracle (C/OpenMP restatement of
the reference) on the same field, compress + decompress, about 10 s of work.

Also reported: "learned" (the same steps with the learned tuner choosing each
compress's configuration), "tuner_c3" (learned tuner vs exhaustive trial on a
256x384x384 fp64 turbulence field at rel eb 1e-6) and "serving" (random
access on a resident 1024^3 stream under a Poisson request load; roi_p99_ms).

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--n 512] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sc25-stz2_amd"))

# GPU_MAX_HW_QUEUES is left as the caller set it (HIP's default: 4); the
# serving workers' contexts take hardware queues of their own
# (STZ_CTX_OWN_QUEUE) and need no more.  Recorded in the result line.
HW_QUEUES = os.environ.get("GPU_MAX_HW_QUEUES", "4 (HIP default, unset)")

METRIC = "GB/s comp+decomp at rel eb 1e-3 (learned vs plain, same GPU); ROI p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_bound(t, eb):
    """The requested value-range-relative bound of a (finite) field tensor,
    in double: the maximum pointwise error the library guarantees."""
    return float(eb) * (float(t.max().item()) - float(t.min().item()))


def kernel_bytes(name, n, s, payload):
    """Algorithmic HBM bytes of one launch of the per-point kernels at field size n
    (values), value size s, device payload bytes (minimum each kernel must move)."""
    n8 = n / 8.0
    if name == "high_encode":      # read 7/8 targets + 1/8 mid reference, write 7/8 codes (2 B)
        return n * (7 / 8.0) * s + n8 * s + n * (7 / 8.0) * 2
    if name == "high_decode":      # read 7/8 codes + 1/8 mid, write the full field
        return n * (7 / 8.0) * 2 + n8 * s + n * s
    if name == "huff_decode":      # read the payload, write every code (2 B)
        return payload + n * 2
    if name == "huff_pack":        # read every code, write the payload
        return n * 2 + payload
    if name == "huff_len":         # read every code
        return n * 2
    if name == "low_encode":
        return n / 64.0 * 7 * s + n / 64.0 * s + n / 64.0 * 7 * 2 + n8 * s
    if name == "low_decode":
        return n / 64.0 * 7 * 2 + n / 64.0 * s + n8 * s
    return None


TRAFFIC_FILE = "profiles/hbm_traffic_512cubed.json"


def measured_traffic(name, n, s):
    """Memory bytes per launch of kernel `name` measured with rocprofv3 on the
    512^3 fp32 workload (read requests by size + WRITE_SIZE), or None for
    another workload or a kernel the profile does not cover."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if n != 512 or s != 4 or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d["kernels"].get(name)
    if k is None:
        return None
    return {"bytes_per_launch": int(k["bytes_per_launch"]), "method": d.get("method", "")}


def host_cpu():
    """The host CPU the baseline runs on: model name, sockets, logical CPUs
    visible to this process, and the OpenMP thread binding in effect."""
    model, sockets = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model is None:
                    model = v.strip()
                elif k == "physical id":
                    sockets.add(v.strip())
    except OSError:
        pass
    try:
        visible = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        visible = os.cpu_count()
    return {"cpu_model": model, "sockets": len(sockets) or None, "visible_cpus": visible,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "unset"),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS", "unset")}


def copy_peak(torch, dev, mib=1024, reps=10):
    """Device-to-device copy rate (read + write bytes per second) of a
    `mib` MiB buffer: the practical HBM ceiling beside the 8 TB/s peak."""
    a = torch.empty(mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return round(2.0 * (mib << 20) / (ms * 1e-3) / 1e9, 1)


def psnr_db(ref, rec):
    """PSNR of rec against ref over the value range (the reference's Global
    PSNR line, F5:3003-3016): 20 log10(range) - 10 log10(MSE)."""
    import math
    d = (rec.double() - ref.double())
    mse = float((d * d).mean())
    rng = float(ref.max() - ref.min())
    if mse == 0.0:
        return float("inf")
    return round(20.0 * math.log10(rng) - 10.0 * math.log10(mse), 3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--eb", type=float, default=1e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--calib", action="store_true",
                    help="run a 1 GiB device-to-device copy first (known-size reference for HBM counters)")
    ap.add_argument("--no-serve", action="store_true", help="skip the random-access serving run")
    ap.add_argument("--no-tuner", action="store_true", help="skip the learned-tuner runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the many-fields run")
    ap.add_argument("--batch-fields", type=int, default=96)
    ap.add_argument("--serve-n", type=int, default=1024)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=[2000.0, 4000.0],
                    help="further arrival rates, each a trace of max(1000, rate) requests")
    ap.add_argument("--serve-workers", type=int, default=4,
                    help="worker contexts of the request server (one HIP stream each, two for the full-decode worker)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the compress/decompress path runs on a single MI355X (how it would "
              "shard over several GPUs is a design note, not a measured run). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import stz_amd
    from stz_amd import fields

    dev = torch.device("cuda:0")
    if args.calib:
        a = torch.empty(256 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB written")
        del a, b
    n = args.n
    t0 = time.time()
    f_np = fields.lognormal(n=n, seed=1002, dtype=np.float32, device=dev)
    field = torch.from_numpy(f_np).to(dev)
    torch.cuda.synchronize()
    log("field %d^3 generated in %.1fs" % (n, time.time() - t0))
    N = field.numel()
    s = 4

    ctx = stz_amd.Context(0)
    stream = torch.cuda.Stream(device=dev)
    ctx.set_stream(stream.cuda_stream)
    out = torch.empty_like(field)

    def step():
        ds = ctx.compress_device(field, args.eb, "rel")
        ds.decompress(out)
        return ds

    for _ in range(args.warmup):
        step().free()
    torch.cuda.synchronize()
    # timed region: K steps, each compress + decompress, HIP events on the library stream
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    tc = td = 0.0
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    last = None
    for _ in range(args.steps):
        if last is not None:
            last.free()
        a = time.perf_counter()
        last = ctx.compress_device(field, args.eb, "rel")
        b = time.perf_counter()
        last.decompress(out)
        c = time.perf_counter()
        tc += b - a
        td += c - b
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    dev_ms = ev0.elapsed_time(ev1)
    ms_per_step = dev_ms / args.steps
    value = 2.0 * N * s / (ms_per_step * 1e-3) / 1e9

    # correctness of the measured work
    info = last.info
    err = float((out.double() - field.double()).abs().max())
    eb_high = info.eb_high
    payload = last.payload_bytes
    eb_req = rel_bound(field, args.eb)
    ok = err <= eb_req
    psnr = psnr_db(field, out)
    log("step %.3f ms (device), wall %.3f ms; compress %.3f ms, decompress %.3f ms (host clock); max err %.3g "
        "eb %.3g ok=%s psnr %.2f dB" % (ms_per_step, wall * 1e3 / args.steps, tc * 1e3 / args.steps,
                                        td * 1e3 / args.steps, err, eb_high, ok, psnr))

    # host zstd of the payload (timed separately, D2H included)
    z0 = time.perf_counter()
    host_stream = last.download(3)
    zstd_ms = (time.perf_counter() - z0) * 1e3
    cr_payload = N * s / payload
    cr_zstd = N * s / host_stream.size

    # per-kernel profile (HIP events around every launch on the library stream)
    prof = {}
    if args.profile_steps > 0:
        ctx.profile(True)
        for _ in range(args.profile_steps):
            step().free()
        prof = ctx.profile_report()
        ctx.profile(False)
    per_step = {k: (v[0] / args.profile_steps, v[1] / args.profile_steps) for k, v in prof.items()}
    roof = None
    kb = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1][1])
        dom_name, (dom_launches, dom_total_ms) = dom
        dom_avg_ms = dom_total_ms / dom_launches
        kb = kernel_bytes(dom_name, N, s, payload)
    copy_gbps = copy_peak(torch, dev)
    log("device copy: %.1f GB/s (read + write)" % copy_gbps)
    if kb is not None:
        ach = kb / (dom_avg_ms * 1e-3) / 1e9
        roof = {"kernel": dom_name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(dom_avg_ms, 4), "algorithmic_bytes": int(kb),
                "measured_copy_GBps": copy_gbps, "frac_of_copy": round(ach / copy_gbps, 4)}
        tr = measured_traffic(dom_name, args.n, s)
        if tr is not None:
            roof["traffic"] = tr["bytes_per_launch"]
            roof["traffic_over_algorithmic"] = round(tr["bytes_per_launch"] / kb, 3)
            roof["traffic_source"] = TRAFFIC_FILE + " (rocprofv3: " + tr["method"] + ")"
    step_alg_bytes = 2.0 * (N * s + payload)
    whole = {"algorithmic_bytes_per_step": int(step_alg_bytes),
             "achieved_GBps": round(step_alg_bytes / (ms_per_step * 1e-3) / 1e9, 1),
             "frac_of_hbm_peak": round(step_alg_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "frac_of_copy": round(step_alg_bytes / (ms_per_step * 1e-3) / 1e9 / copy_gbps, 4)}

    cpu = None
    if not args.no_cpu:
        # the oracle (C/OpenMP restatement of the reference's CPU path) on the
        # whole field, repeated until about 10 s of CPU work; median rate
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rates, tcs, tds = [], [], []
        t_start = time.perf_counter()
        while len(rates) < 3 or (time.perf_counter() - t_start < 10.0 and len(rates) < 20):
            c0 = time.perf_counter()
            p = pyoracle.compress(f_np, args.eb, pyoracle.EB_REL)
            c1 = time.perf_counter()
            pyoracle.decompress(p)
            c2 = time.perf_counter()
            rates.append(2.0 * f_np.nbytes / (c2 - c0) / 1e9)
            tcs.append(c1 - c0)
            tds.append(c2 - c1)
        med = float(np.median(rates))
        cpu = {"value": round(med, 4), "unit": "GB/s", "cores": threads, "kind": "port", **host_cpu(),
               "sample": "the whole %d^3 field, compress + decompress (pre-zstd payload) by the C/OpenMP oracle, "
                         "%d repetitions in %.1f s, median; compress %.2f s, decompress %.2f s (medians)"
                         % (n, len(rates), time.perf_counter() - t_start, float(np.median(tcs)),
                            float(np.median(tds)))}
        log("cpu baseline", cpu)

    learned = tuner_c3 = None
    if not args.no_tuner:
        learned, tuner_c3 = tuner_bench(args, ctx, stream, field, out, N, s, payload, dev)

    batch = None
    if not args.no_batch:
        batch = batch_bench(args, dev)

    serving = None
    if not args.no_serve:
        serving = serve_bench(args, dev)

    rec = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%d^3 fp32 log-normal GRF (k^-3, sigma 1.5, seed 1002), rel eb %g, plain tuning, "
                               "compress+decompress device path (zstd excluded)" % (n, args.eb),
                   "field": "%dx%dx%d" % (n, n, n), "eb_rel": args.eb, "tuner": "plain",
                   "quantizer_arith": "f64"},
        "compress_ms_host_clock": round(tc * 1e3 / args.steps, 3),
        "decompress_ms_host_clock": round(td * 1e3 / args.steps, 3),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 3),
        "payload_bytes": int(payload), "cr_payload": round(cr_payload, 3), "cr_zstd": round(cr_zstd, 3),
        "zstd_ms": round(zstd_ms, 2), "max_err": err, "eb_abs": eb_req, "eb_high_level": eb_high,
        "error_bound_ok": bool(ok), "psnr_db": psnr,
        "roofline": roof, "whole_step": whole, "gpu_max_hw_queues": HW_QUEUES,
        "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(per_step.items(), key=lambda kv: -kv[1][1])},
        "cpu_baseline": cpu,
        "batch": batch,
        "learned": learned,
        "tuner_c3": tuner_c3,
        "roi_p99_ms": serving["latency"]["box"]["p99_ms"] if serving else None,
        "serving": serving,
    }
    print(json.dumps(rec), flush=True)
    last.free()
    ctx.close()


def tuner_bench(args, ctx, stream, field, out, N, s, payload_plain, dev):
    """(1) The same comp+decomp steps with the learned tuner choosing the
    configuration of every compress (its statistics + MLP run inside the
    timed step).  (2) The tuner workload: 256x384x384 fp64 turbulence-like
    field (GRF k^-11/3 + 8 tanh fronts of width 2, seed 1003), rel eb 1e-6:
    learned choice vs the exhaustive trial over all 204 configurations --
    tuning time and payload bytes -- and vs the untuned configuration.  The
    model runs with its calibrated confidence threshold, below which the
    learned path falls back to the trial (reported as "fallback")."""
    import numpy as np
    import torch
    from stz_amd import fields, tuner
    meta = tuner.load_model(ctx)
    min_conf = float(meta.get("min_confidence", 0.0))

    def step():
        ds = ctx.compress_device(field, args.eb, "rel", tuner=1)
        ds.decompress(out)
        return ds

    for _ in range(args.warmup):
        step().free()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    last = None
    for _ in range(args.steps):
        if last is not None:
            last.free()
        last = step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    info = last.info
    err = float((out.double() - field.double()).abs().max())
    learned = {"value": round(2.0 * N * s / (ms * 1e-3) / 1e9, 2), "unit": "GB/s", "ms_per_step": round(ms, 4),
               "payload_bytes": int(last.payload_bytes), "cr_payload": round(N * s / last.payload_bytes, 3),
               "cr_vs_plain": round(payload_plain / last.payload_bytes, 4),
               "choice": {"pred": info.pred, "interp": info.interp, "order": info.order,
                          "mult": [info.mult_base, info.mult_low, info.mult_high], "alpha": info.alpha,
                          "beta": info.beta},
               "confidence": round(float(info.tune_confidence), 4), "min_confidence": round(min_conf, 4),
               "fallback": bool(info.tuner == 2),
               "max_err": err, "eb_abs": rel_bound(field, args.eb),
               "error_bound_ok": bool(err <= rel_bound(field, args.eb))}
    last.free()
    log("learned", json.dumps(learned))

    t0 = time.time()
    f3 = fields.turbulence(shape=(256, 384, 384), seed=1003, device=dev)
    f3t = torch.from_numpy(f3).to(dev)
    torch.cuda.synchronize()
    log("tuner field 256x384x384 fp64 generated in %.1fs" % (time.time() - t0))
    tuner.choose(ctx, f3t, 1e-6)   # warm the pool
    t0 = time.perf_counter()
    k, conf = tuner.choose(ctx, f3t, 1e-6)
    t_learned = time.perf_counter() - t0
    t0 = time.perf_counter()
    sizes = tuner.trial_sizes(ctx, f3t, 1e-6).astype(np.float64)
    t_trial = time.perf_counter() - t0
    nbytes = f3.nbytes
    c3 = {"workload": "256x384x384 fp64 turbulence-like field (GRF k^-11/3 + 8 tanh fronts, seed 1003), rel eb 1e-6",
          "configs": tuner.TUNE_CONFIGS, "tuning_ms_learned": round(t_learned * 1e3, 3), "tuning_ms_trial": round(t_trial * 1e3, 1),
          "tuning_speedup": round(t_trial / t_learned, 1), "choice": int(k), "trial_best": int(sizes.argmin()),
          "confidence": round(conf, 4), "min_confidence": round(min_conf, 4), "fallback": bool(conf < min_conf),
          "cr_learned": round(nbytes / sizes[k], 4),
          "cr_trial": round(nbytes / sizes.min(), 4), "cr_untuned": round(nbytes / sizes[tuner.DEFAULT_CONFIG], 4),
          "cr_learned_vs_trial": round(sizes.min() / sizes[k], 5),
          "gain_vs_untuned": round(sizes[tuner.DEFAULT_CONFIG] / sizes[k] - 1.0, 4),
          "model": {"holdout_mean_regret": meta.get("holdout_eval", {}).get("mean_regret"),
                    "holdout_gain_vs_untuned": meta.get("holdout_eval", {}).get("mean_gain_vs_default"),
                    "holdout_fallback_rate": meta.get("calibration", {}).get("fallback_rate")}}
    log("tuner_c3", json.dumps(c3))
    del f3t
    return learned, c3


def batch_bench(args, dev):
    """Many uneven fields: 96 fp32 GRFs, each dim drawn from 32..384, spectral
    slopes in [-4, -2] (seeds 1004 + i), rel eb 1e-3, compressed in one batch
    call and decoded in one (graph-replayed) batch call; GB/s over the total
    bytes, next to one call per field on the same fields."""
    import numpy as np
    import torch
    import stz_amd
    from stz_amd import fields
    t0 = time.time()
    fs = []
    for i in range(args.batch_fields):
        # host FFTs: uneven lengths would each need a device FFT plan built at run time
        rng = np.random.default_rng(1004 + i)
        shape = tuple(int(v) for v in rng.integers(32, 385, size=3))
        slope = float(rng.uniform(-4.0, -2.0))
        fs.append(torch.from_numpy(fields.grf(shape, slope, 1004 + i).astype(np.float32)).to(dev))
        if (i + 1) % 16 == 0:
            log("batch: %d fields generated, %.1fs" % (i + 1, time.time() - t0))
    torch.cuda.synchronize()
    pts = sum(f.numel() for f in fs)
    log("batch: %d fields, %.3g points, generated in %.1fs" % (len(fs), pts, time.time() - t0))
    ctx = stz_amd.Context(0)
    outs = [torch.empty_like(f) for f in fs]

    def batch_step():
        ds = ctx.compress_batch(fs, 1e-3, "rel")
        ctx.decompress_batch(ds, outs)
        return ds

    t0 = time.perf_counter()
    first = batch_step()            # autotunes the level launches
    t_first = time.perf_counter() - t0
    for d in first:
        d.free()
    steps = max(2, min(args.steps, 5))
    # keep the streams of the last step: the decode graph is replayed while the
    # same streams and outputs come back
    ds = ctx.compress_batch(fs, 1e-3, "rel")
    ctx.decompress_batch(ds, outs)
    torch.cuda.synchronize()
    tc = td = 0.0
    for _ in range(steps):
        for d in ds:
            d.free()
        a = time.perf_counter()
        ds = ctx.compress_batch(fs, 1e-3, "rel")
        b = time.perf_counter()
        ctx.decompress_batch(ds, outs)
        c_ = time.perf_counter()
        tc += b - a
        td += c_ - b
    # the decode graph replays only while the streams stay: time that on its own
    # (a new set of streams decodes eagerly; the second call captures the graph)
    ctx.decompress_batch(ds, outs)
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(steps):
        ctx.decompress_batch(ds, outs)
    t_replay = (time.perf_counter() - a) / steps
    payload = sum(d.payload_bytes for d in ds)
    err_ok = all(float((o.double() - f.double()).abs().max()) <= rel_bound(f, 1e-3) for o, f in zip(outs, fs))
    # one call per field, same fields
    torch.cuda.synchronize()
    a = time.perf_counter()
    for f, o in zip(fs, outs):
        d = ctx.compress_device(f, 1e-3, "rel")
        d.decompress(o)
        d.free()
    t_single = time.perf_counter() - a
    nbytes = pts * 4
    ms = (tc + td) / steps * 1e3
    # learned tuner on the same batch (one statistics pass + one MLP launch for
    # all fields inside the compress call), and the exhaustive trial once
    from stz_amd import tuner
    tuner.load_model(ctx)
    ld = ctx.compress_batch(fs, 1e-3, "rel", tuner=1)
    for d in ld:
        d.free()
    torch.cuda.synchronize()
    a = time.perf_counter()
    ld = ctx.compress_batch(fs, 1e-3, "rel", tuner=1)
    t_lc = time.perf_counter() - a
    louts = [torch.empty_like(f) for f in fs]
    ctx.decompress_batch(ld, louts)
    a = time.perf_counter()
    ctx.decompress_batch(ld, louts)
    t_ld = time.perf_counter() - a
    lpay = sum(d.payload_bytes for d in ld)
    n_fallback = sum(int(d.info.tuner == 2) for d in ld)
    l_ok = all(float((o.double() - f.double()).abs().max()) <= rel_bound(f, 1e-3) for o, f in zip(louts, fs))
    for d in ld:
        d.free()
    del louts
    a = time.perf_counter()
    td_ = ctx.compress_batch(fs, 1e-3, "rel", tuner=2)
    t_trial = time.perf_counter() - a
    tpay = sum(d.payload_bytes for d in td_)
    for d in td_:
        d.free()
    learned = {"compress_ms": round(t_lc * 1e3, 3), "decompress_ms_graph": round(t_ld * 1e3, 3),
               "value": round(2.0 * nbytes / (t_lc + t_ld) / 1e9, 2), "unit": "GB/s",
               "payload_bytes": int(lpay), "cr_payload": round(nbytes / lpay, 3),
               "cr_vs_plain": round(payload / lpay, 4), "error_bound_ok": bool(l_ok),
               "trial_compress_ms": round(t_trial * 1e3, 1), "trial_payload_bytes": int(tpay),
               "cr_learned_vs_trial": round(tpay / lpay, 5), "fallback_fields": n_fallback}
    out = {"workload": "%d fp32 GRF fields, dims U{32..384}, slopes U[-4,-2] (seeds 1004+i), rel eb 1e-3" % len(fs),
           "points": int(pts), "bytes": int(nbytes), "payload_bytes": int(payload),
           "cr_payload": round(nbytes / payload, 3),
           "ms_per_step": round(ms, 3), "value": round(2.0 * nbytes / (ms * 1e-3) / 1e9, 2), "unit": "GB/s",
           "compress_ms": round(tc / steps * 1e3, 3), "decompress_ms": round(td / steps * 1e3, 3),
           "decompress_graph_replay_ms": round(t_replay * 1e3, 3),
           "one_call_per_field_ms": round(t_single * 1e3, 3),
           "speedup_vs_one_call_per_field": round(t_single * 1e3 / ms, 3),
           "first_step_ms_incl_autotune_and_capture": round(t_first * 1e3, 1),
           "error_bound_ok": bool(err_ok), "learned": learned, "autotune": ctx.autotune()}
    log("batch", json.dumps(out))
    for d in ds:
        d.free()
    ctx.close()
    del fs, outs
    torch.cuda.empty_cache()
    return out


def serve_bench(args, dev):
    """Random-access serving: an N^3 fp32 GRF (k^-3, seed 1005) compressed once
    at rel eb 1e-3 and kept in HBM; a Poisson stream of box / slice / level
    requests replayed in real time through the continuous-batching server."""
    import numpy as np
    import torch
    import stz_amd
    from stz_amd import fields, serve
    n = args.serve_n
    t0 = time.time()
    g = fields.grf_device((n, n, n), -3.0, 1005, dev, torch.float32)
    torch.cuda.synchronize()
    log("serving field %d^3 generated in %.1fs" % (n, time.time() - t0))
    ctx = stz_amd.Context(0)
    ds = ctx.compress_device(g, 1e-3, "rel")
    info = ds.info
    del g
    torch.cuda.empty_cache()
    ctx.close()  # its streams would take hardware queues the workers need; the stream keeps its memory
    def run(rate, n_req):
        reqs, arr = serve.poisson_trace((n, n, n), n_req, rate, seed=1005)
        srv = serve.RoiServer(ds, workers=args.serve_workers, max_batch=32, mem_budget=32 << 30)
        log("serving %d requests at %.0f/s on %d workers" % (len(reqs), rate, args.serve_workers))
        try:
            recs, ts = srv.run_trace(reqs, arr)
        finally:
            srv.close()
        stats, gbs, wall = serve.summarize(recs, ts, 4)
        return reqs, stats, gbs, wall, [b[1] for b in srv.batches]

    reqs, stats, gbs, wall, sizes = run(args.serve_rate, args.serve_requests)
    # rate sweep: shorter traces at higher arrival rates, stopping at the
    # first failure
    sweep = []
    for rate in args.serve_sweep:
        try:
            _, st, g_, w_, sz = run(rate, max(1000, int(rate)))
        except Exception as e:  # recorded, and the sweep ends
            sweep.append({"rate_per_s": rate, "error": repr(e)})
            break
        sweep.append({"rate_per_s": rate, "requests": max(1000, int(rate)), "box_p99_ms": st["box"]["p99_ms"],
                      "all_p99_ms": st["all"]["p99_ms"], "level3_p99_ms": st.get("level3", {}).get("p99_ms"),
                      "mean_batch": round(float(np.mean(sz)), 2) if sz else 0, "max_batch_seen": max(sz) if sz else 0,
                      "delivered_GBps": round(g_, 2)})
        log("sweep", json.dumps(sweep[-1]))
    out = {"workload": "%d^3 fp32 GRF (k^-3, seed 1005), rel eb 1e-3, compressed once and resident in HBM "
                       "(payload %d bytes, CR %.1f)" % (n, info.payload_bytes, n ** 3 * 4 / info.payload_bytes),
           "requests": len(reqs), "rate_per_s": args.serve_rate,
           "mix": {"box (edge U{16..256}, uniform in-bounds origin)": 0.70, "axis slice": 0.18,
                   "level 1/8": 0.03, "level 1/4": 0.03, "level 1/2": 0.03, "level full": 0.03},
           "workers": args.serve_workers, "max_batch": 32, "batches": len(sizes),
           "mean_batch": round(float(np.mean(sizes)), 2) if sizes else 0, "max_batch_seen": max(sizes) if sizes else 0,
           "wall_s": round(wall, 3), "delivered_GBps": round(gbs, 2), "latency": stats,
           "gpu_max_hw_queues": HW_QUEUES, "rate_sweep": sweep}
    log("serving", json.dumps(out))
    ds.free()
    ctx.close()
    return out


if __name__ == "__main__":
    main()
