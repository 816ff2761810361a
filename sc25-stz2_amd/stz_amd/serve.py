"""Continuous-batching random-access server over HBM-resident streams.

Streams stay compressed in HBM (stz_dstream, decode tables and the decoded
base level prepared once).  Requests -- 3D boxes, axis slices and
progressive levels -- queue on arrival.  Each worker thread owns a context
(its own HIP stream and memory pool).  When a worker frees up it admits the
queued requests of its class (worker 0: full-resolution decodes; the
others: boxes, slices and the 1/8, 1/4, 1/2 levels) that fit its slot limit and the server's
memory budget and serves them with ONE stz_dstream_decompress_batch call: one Huffman decode of
the union of their chunks, one launch for all their low regions and one for
all their high boxes; identical requests in a batch are decoded once and
share the result.  Requests that arrive while batches are on the GPU wait
only for the next free worker; batches of different workers overlap on the
device (the library call releases the GIL).  Latency is arrival to
completion of the batch that served the request.

The workers' batches overlap only when their streams sit on different
hardware queues: HIP spreads ordinary streams over GPU_MAX_HW_QUEUES queues
(4 by default, shared with torch's streams), and streams that share one run
in order, so a box batch would wait for a full decode queued before it.  The
workers' contexts are therefore created with STZ_CTX_OWN_QUEUE: their streams
carry a CU mask over every CU, which gives each a hardware queue of its own
at any GPU_MAX_HW_QUEUES.  The full-decode worker's persistent level kernels
take only `full_share` of the resident workgroup slots (stz_set_grid_share),
so a box batch finds free CUs while a full decode runs.
"""
import collections
import threading
import time

import numpy as np

__all__ = ["RoiServer", "poisson_trace", "summarize"]


class _Pending:
    __slots__ = ("idx", "req", "t_arrival", "nbytes", "t_done", "out", "batch", "large")

    def __init__(self, idx, req, t_arrival, nbytes):
        self.idx = idx
        self.req = req
        self.t_arrival = t_arrival
        self.nbytes = nbytes
        self.t_done = None
        self.out = None
        self.batch = None
        # a full-resolution level decodes every stream: tens of ms at 1024^3
        self.large = req[0] == "level" and req[1] == 3


class RoiServer:
    def __init__(self, ds, workers=3, max_batch=32, mem_budget=16 << 30, device=0, keep_base=True,
                 keep_results=False, full_share=0.5):
        import torch
        from .api import Context
        self.torch = torch
        self.ds = ds
        self.dev = torch.device("cuda", device)
        self.info = ds.info
        self.tsize = 4 if self.info.dtype == 0 else 8
        self.tdtype = torch.float32 if self.info.dtype == 0 else torch.float64
        self.max_batch = max_batch
        self.mem_budget = mem_budget
        self.keep_results = keep_results
        self.ctxs = [Context(device, own_queue=True) for _ in range(workers)]
        if workers > 1:
            # the full-resolution decodes' persistent kernels leave part of
            # the GPU to the other workers' batches
            self.ctxs[0].set_grid_share(full_share)
        ds.prepare(keep_base=keep_base, ctx=self.ctxs[0])
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.inflight_bytes = 0
        self.closed = False
        self.batches = []          # (worker, n requests, t_start, t_end)
        self.errors = []
        self.threads = [threading.Thread(target=self._worker, args=(w,), daemon=True) for w in range(workers)]
        for t in self.threads:
            t.start()

    # ------------------------------------------------------------------ admission
    def _cost(self, req):
        return self.ds.request_capacity(req) * self.tsize

    def submit(self, idx, req, t_arrival=None):
        p = _Pending(idx, req, time.perf_counter() if t_arrival is None else t_arrival, self._cost(req))
        with self.cv:
            self.queue.append(p)
            self.cv.notify()
        return p

    def _take_batch(self, w):
        """Admission for worker w.  With more than one worker, worker 0 takes
        the full-resolution decodes (tens of ms at 1024^3) and the others
        everything else, so a box never queues behind a full decode; worker 0
        falls back to the other requests when no full decode waits.  FIFO within a class, up to max_batch requests
        and the memory budget (a request larger than the budget goes alone
        once nothing else is in flight)."""
        if len(self.ctxs) == 1:
            picks = list(self.queue)
        elif w == 0:
            picks = [p for p in self.queue if p.large] or list(self.queue)
        else:
            picks = [p for p in self.queue if not p.large]
        batch = []
        for p in picks:
            if len(batch) >= self.max_batch:
                break
            if self.inflight_bytes + p.nbytes > self.mem_budget and (batch or self.inflight_bytes > 0):
                break
            self.inflight_bytes += p.nbytes
            batch.append(p)
        for p in batch:
            self.queue.remove(p)
        return batch

    def _worker(self, w):
        ctx = self.ctxs[w]
        torch = self.torch
        while True:
            with self.cv:
                batch = []
                while not batch:
                    if self.closed and not self.queue:
                        return
                    batch = self._take_batch(w)
                    if not batch:
                        self.cv.wait(0.01)
            t0 = time.perf_counter()
            try:
                # identical requests in one batch (the same level, box or
                # slice) are decoded once and share the result
                uniq = list(dict.fromkeys(tuple(p.req) for p in batch))
                outs = [torch.empty(self.ds.request_capacity(r), dtype=self.tdtype, device=self.dev) for r in uniq]
                res = dict(zip(uniq, self.ds.decompress_batch(uniq, outs=outs, ctx=ctx)))
                t1 = time.perf_counter()
                for p in batch:
                    p.t_done = t1
                    p.batch = len(self.batches)
                    p.out = res[tuple(p.req)] if self.keep_results else None
            except Exception as e:  # reported by close()
                t1 = time.perf_counter()
                self.errors.append(repr(e))
            with self.cv:
                self.batches.append((w, len(batch), t0, t1))
                for p in batch:
                    self.inflight_bytes -= p.nbytes
                self.cv.notify_all()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        for c in self.ctxs:
            c.close()
        if self.errors:
            raise RuntimeError("serving errors: %s" % self.errors[:3])

    # ------------------------------------------------------------------ trace replay
    def run_trace(self, requests, arrivals):
        """Replay requests at their arrival offsets (seconds from now); returns
        the pending records once every request has completed."""
        t_start = time.perf_counter()
        recs = []
        for i, (r, a) in enumerate(zip(requests, arrivals)):
            target = t_start + a
            while True:
                now = time.perf_counter()
                if now >= target:
                    break
                time.sleep(min(target - now, 0.0005) if target - now > 0.0002 else 0)
            recs.append(self.submit(i, r, target))
        with self.cv:
            while any(p.t_done is None for p in recs) and not self.errors:
                self.cv.wait(0.05)
        return recs, t_start


def poisson_trace(dims, n, rate, seed=1005, mix=None, box_edge=(16, 256)):
    """Poisson arrivals (rate per second) of a request mix over a field of
    dims (nx, ny, nz).  mix: fractions of box, slice and the four progressive
    levels (1/8, 1/4, 1/2, full)."""
    mix = mix or {"box": 0.70, "slice": 0.18, "level0": 0.03, "level1": 0.03, "level2": 0.03, "level3": 0.03}
    rng = np.random.default_rng(seed)
    kinds = list(mix)
    p = np.array([mix[k] for k in kinds], dtype=np.float64)
    p /= p.sum()
    arr = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    nx, ny, nz = dims
    for k in rng.choice(len(kinds), size=n, p=p):
        kind = kinds[k]
        if kind == "box":
            e = [int(rng.integers(box_edge[0], min(box_edge[1], d) + 1)) for d in (nx, ny, nz)]
            o = [int(rng.integers(0, d - ed + 1)) for d, ed in zip((nx, ny, nz), e)]
            reqs.append(("box", o[0], o[1], o[2], o[0] + e[0] - 1, o[1] + e[1] - 1, o[2] + e[2] - 1))
        elif kind == "slice":
            ax = int(rng.integers(0, 3))
            reqs.append(("slice", ax, int(rng.integers(0, (nx, ny, nz)[ax]))))
        else:
            reqs.append(("level", int(kind[-1])))
    return reqs, arr


def summarize(recs, t_start, tsize):
    """p50/p99 latency (ms) per request kind and overall; delivered GB/s."""
    lat = collections.defaultdict(list)
    delivered = 0
    t_end = t_start
    for p in recs:
        k = p.req[0] if p.req[0] != "level" else "level%d" % p.req[1]
        lat[k].append((p.t_done - p.t_arrival) * 1e3)
        lat["all"].append((p.t_done - p.t_arrival) * 1e3)
        delivered += p.nbytes
        t_end = max(t_end, p.t_done)
    out = {}
    for k, v in sorted(lat.items()):
        v = np.array(v)
        out[k] = {"n": int(v.size), "p50_ms": round(float(np.percentile(v, 50)), 3),
                  "p99_ms": round(float(np.percentile(v, 99)), 3), "max_ms": round(float(v.max()), 3)}
    wall = t_end - t_start
    return out, delivered / wall / 1e9, wall
