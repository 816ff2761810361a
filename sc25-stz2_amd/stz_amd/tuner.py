"""Learned tuner: a 16 -> 64 -> 64 -> 208 MLP over per-sub-block field
statistics that picks the residual predictor set, base codec (interpolator
and axis order, or Lorenzo / regression) and level-eb scaling (204 configurations, the
first 204 outputs) in place of the exhaustive trial.

Labels come from the exhaustive trial itself (the payload size of every
configuration on generated fields); the target of a sub-block is the soft
distribution exp(-(log size_k - log size_best) / tau) of its field, so
near-ties share probability.  Training runs on the GPU through the library's
bf16 MFMA kernels (forward + backward in stz_mlp_grad, Adam in stz_mlp_adam);
the host only shuffles batches.  A field's choice maximises the mean log
probability over its sub-blocks.

    python -m stz_amd.tuner train --out sc25-stz2_amd/stz_amd/tuner_model.npz
"""
import argparse
import ctypes as C
import json
import os
import time

import numpy as np

from ._lib import MLP_OUTPUTS, MLP_PARAMS, TUNE_CONFIGS, TUNE_FEATURES, Params, check, lib
from .api import EB_MODES, _dtype_id, _ptr

DEFAULT_CONFIG = 24          # cubic, order 0, F5 multipliers: the untuned path
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuner_model.npz")
_IN, _HID, _OUT = 16, 64, MLP_OUTPUTS   # outputs past TUNE_CONFIGS are never chosen


def config_params(k):
    p = Params()
    lib().stz_default_params(C.byref(p))
    check(lib().stz_tune_config(int(k), C.byref(p)), "stz_tune_config")
    return {"pred": p.pred, "interp": p.interp, "order": p.order, "mult_base": p.mult_base, "mult_low": p.mult_low,
            "mult_high": p.mult_high, "alpha": p.alpha, "beta": p.beta}


def config_index(prm):
    """The configuration whose parameters are prm (a dict like config_params'), or -1."""
    for k in range(TUNE_CONFIGS):
        if config_params(k) == {key: prm[key] for key in config_params(0)}:
            return k
    return -1


def features(ctx, field, eb_abs):
    """16 statistics per 64^3 sub-block -> (n_blocks, 16) float32."""
    nz, ny, nx = field.shape
    cap = ((nx + 63) // 64) * ((ny + 63) // 64) * ((nz + 63) // 64)
    out = np.zeros((cap, TUNE_FEATURES), np.float32)
    nb = C.c_size_t()
    check(lib().stz_tune_features(ctx.h, _dtype_id(field), _ptr(field), nx, ny, nz, float(eb_abs), _ptr(out), cap,
                                  C.byref(nb)), "stz_tune_features")
    return out[: nb.value]


def informative(X):
    """Rows of blocks with sampled points (thin edge blocks have features 2.. = 0)."""
    m = np.any(X[:, 2:] != 0, axis=1)
    return m if m.any() else np.ones(X.shape[0], bool)


def trial_sizes(ctx, field, eb, mode="rel"):
    """Payload bytes of every configuration (the exhaustive trial)."""
    nz, ny, nx = field.shape
    sizes = np.zeros(TUNE_CONFIGS, np.uint64)
    best = C.c_int()
    check(lib().stz_tune_trial(ctx.h, _dtype_id(field), _ptr(field), nx, ny, nz, float(eb), EB_MODES[mode], None,
                               _ptr(sizes), C.byref(best)), "stz_tune_trial")
    return sizes


def soft_target(sizes, tau=0.01):
    """Target distribution over the MLP's outputs: configurations by payload
    size, zero on the outputs past them."""
    lg = np.log(sizes.astype(np.float64))
    w = np.exp(-(lg - lg.min()) / tau)
    t = np.zeros(_OUT, np.float32)
    t[: sizes.size] = w / w.sum()
    return t


class Mlp:
    """fp32 master parameters + Adam state on the device (torch tensors);
    every product runs in the library's bf16 MFMA kernels."""

    def __init__(self, ctx, seed=0, device="cuda:0", params=None):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.dev = torch.device(device)
        if params is None:
            params = self.init_params(seed)
        self.P = torch.tensor(np.asarray(params, np.float32), device=self.dev)
        self.G = torch.zeros(MLP_PARAMS, dtype=torch.float32, device=self.dev)
        self.M = torch.zeros_like(self.G)
        self.V = torch.zeros_like(self.G)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self.step_no = 0

    @staticmethod
    def init_params(seed):
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, fan_out in ((_IN, _HID), (_HID, _HID), (_HID, _OUT)):
            lim = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, size=(fan_in, fan_out)).ravel())
            parts.append(np.zeros(fan_out))
        p = np.concatenate(parts).astype(np.float32)
        assert p.size == MLP_PARAMS
        return p

    def params(self):
        return self.P.cpu().numpy()

    def forward(self, X):
        torch = self.torch
        X = torch.as_tensor(X, dtype=torch.float32, device=self.dev).contiguous()
        Y = torch.empty((X.shape[0], _OUT), dtype=torch.float32, device=self.dev)
        torch.cuda.synchronize(self.dev)
        check(lib().stz_mlp_forward(self.ctx.h, _ptr(self.P), _ptr(X), X.shape[0], _ptr(Y)), "stz_mlp_forward")
        return Y

    def grad(self, X, T):
        torch = self.torch
        X = torch.as_tensor(X, dtype=torch.float32, device=self.dev).contiguous()
        T = torch.as_tensor(T, dtype=torch.float32, device=self.dev).contiguous()
        torch.cuda.synchronize(self.dev)
        check(lib().stz_mlp_grad(self.ctx.h, _ptr(self.P), _ptr(X), _ptr(T), X.shape[0], _ptr(self.G),
                                 _ptr(self.loss)), "stz_mlp_grad")
        return self.G, float(self.loss.item())

    def step(self, X, T, lr=3e-3, wd=1e-4):
        _, loss = self.grad(X, T)
        self.step_no += 1
        check(lib().stz_mlp_adam(self.ctx.h, _ptr(self.P), _ptr(self.G), _ptr(self.M), _ptr(self.V), self.step_no,
                                 float(lr), float(wd)), "stz_mlp_adam")
        return loss

    def save(self, path, meta=None):
        np.savez(path, params=self.params(), meta=json.dumps(meta or {}))


def load_params(path=MODEL_PATH):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), json.loads(str(z["meta"]))


def set_model(ctx, params, min_confidence=0.0):
    p = np.ascontiguousarray(params, np.float32)
    check(lib().stz_tune_set_model(ctx.h, _ptr(p), p.size, float(min_confidence)), "stz_tune_set_model")


def load_model(ctx, path=MODEL_PATH):
    """Set the trained model on ctx with its calibrated confidence threshold
    (below it the learned path falls back to the exhaustive trial); returns
    the model's metadata."""
    params, meta = load_params(path)
    set_model(ctx, params, float(meta.get("min_confidence", 0.0)))
    return meta


def choose(ctx, field, eb, mode="rel"):
    """The model's configuration for a field (model set with set_model)."""
    nz, ny, nx = field.shape
    p = Params()
    lib().stz_default_params(C.byref(p))
    conf = C.c_float()
    k = C.c_int()
    check(lib().stz_tune(ctx.h, _dtype_id(field), _ptr(field), nx, ny, nz, float(eb), EB_MODES[mode], C.byref(p),
                         C.byref(conf), C.byref(k)), "stz_tune")
    return k.value, conf.value


# ---------------------------------------------------------------- training data
def make_field(rng, device):
    """One generated field of a random family and shape, with a rel bound."""
    import torch
    from . import fields
    shape = tuple(int(v) for v in rng.integers(48, 161, size=3))
    fam = rng.choice(["grf", "aniso", "lognormal", "sines", "fronts", "rough", "blocks"])
    seed = int(rng.integers(1 << 30))
    if fam in ("grf", "rough", "lognormal", "fronts"):
        g = fields.grf(shape, float(rng.uniform(-5.0, -1.5)), seed, device)
        if fam == "rough":
            g = g + rng.normal(0, float(rng.uniform(0.01, 0.2)), size=shape)
        if fam == "lognormal":
            g = np.exp(float(rng.uniform(0.5, 2.0)) * g)
        if fam == "fronts":
            z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
            for _ in range(int(rng.integers(2, 9))):
                nrm = rng.normal(size=3)
                nrm /= np.linalg.norm(nrm)
                d = nrm @ np.array([rng.uniform(0, shape[2]), rng.uniform(0, shape[1]), rng.uniform(0, shape[0])])
                g += np.tanh((nrm[0] * x + nrm[1] * y + nrm[2] * z - d) / float(rng.uniform(1, 4)))
        f = g
    elif fam == "aniso":
        # correlation lengths differ per axis: the axis order matters
        w = rng.standard_normal(size=shape)
        F = np.fft.rfftn(w)
        nz, ny, nx = shape
        a = rng.uniform(0.25, 4.0, size=3)
        kz = np.fft.fftfreq(nz)[:, None, None] * nz * a[2]
        ky = np.fft.fftfreq(ny)[None, :, None] * ny * a[1]
        kx = np.fft.rfftfreq(nx)[None, None, :] * nx * a[0]
        k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
        s = float(rng.uniform(-5.0, -2.0))
        with np.errstate(divide="ignore"):
            F *= np.where(k > 0, k ** (s / 2.0), 0.0)
        f = np.fft.irfftn(F, s=shape, axes=(0, 1, 2))
        f = (f - f.mean()) / f.std()
    elif fam == "blocks":
        # piecewise-constant cells of one random size (AMR-like patches) over
        # a weak smooth field: where the Lorenzo / regression base beats interpolation
        b = int(rng.integers(3, 17))
        vals = rng.normal(size=[(n + b - 1) // b for n in shape])
        f = np.kron(vals, np.ones((b, b, b)))[: shape[0], : shape[1], : shape[2]]
        f = f + float(rng.uniform(0.0, 0.1)) * fields.grf(shape, -3.0, seed, device)
    else:
        f = fields.sinusoids(shape=shape, seed=seed, dtype=np.float64)
    dt = np.float64 if rng.random() < 0.3 else np.float32
    eb = float(10.0 ** rng.uniform(-6.0 if dt == np.float64 else -5.0, -2.0))
    return torch.from_numpy(np.ascontiguousarray(f, dtype=dt)).to(device), eb, str(fam)


def build_dataset(ctx, n_fields, seed, device="cuda:0", log=None, keep_from=None):
    """Features and trial sizes of n_fields generated fields; the fields from
    index keep_from on are kept (device tensors, with their bounds) for the
    library-side evaluation."""
    import torch
    rng = np.random.default_rng(seed)
    X, fid, sizes, meta, kept = [], [], [], [], {}
    t0 = time.time()
    for i in range(n_fields):
        f, eb, fam = make_field(rng, device)
        torch.cuda.synchronize()
        rngv = float((f.max() - f.min()).item())
        Xi = features(ctx, f, eb * rngv)
        X.append(Xi[informative(Xi)])
        fid.append(np.full(X[-1].shape[0], i, np.int32))
        sizes.append(trial_sizes(ctx, f, eb, "rel"))
        meta.append({"family": fam, "shape": list(f.shape), "dtype": str(f.dtype), "eb": eb})
        if keep_from is not None and i >= keep_from:
            kept[i] = (f, eb)
        if log and (i + 1) % 50 == 0:
            log("dataset: %d/%d fields, %.1fs" % (i + 1, n_fields, time.time() - t0))
    return {"X": np.concatenate(X), "fid": np.concatenate(fid), "sizes": np.stack(sizes), "meta": meta,
            "kept": kept}


def train(model, data, steps=3000, batch=256, lr=3e-3, tau=0.01, seed=0, log=None, fields_idx=None):
    rng = np.random.default_rng(seed)
    T_field = np.stack([soft_target(s, tau) for s in data["sizes"]])
    rows = np.arange(data["X"].shape[0]) if fields_idx is None else np.nonzero(np.isin(data["fid"], fields_idx))[0]
    X = data["X"][rows]
    T = T_field[data["fid"][rows]]
    losses = []
    for it in range(steps):
        b = rng.choice(X.shape[0], size=min(batch, X.shape[0]), replace=False)
        cur_lr = lr * 0.5 * (1 + np.cos(np.pi * it / steps))   # cosine decay
        losses.append(model.step(X[b], T[b], lr=cur_lr))
        if log and (it + 1) % 500 == 0:
            log("train step %d loss %.4f" % (it + 1, float(np.mean(losses[-100:]))))
    return losses


def field_choice(model, X_field):
    """argmax of the mean log-softmax over a field's sub-blocks; confidence =
    mean probability of the chosen configuration."""
    X_field = X_field[informative(X_field)]
    Y = model.forward(X_field).cpu().numpy().astype(np.float64)[:, :TUNE_CONFIGS]
    Y -= Y.max(axis=1, keepdims=True)
    lp = Y - np.log(np.exp(Y).sum(axis=1, keepdims=True))
    k = int(np.argmax(lp.mean(axis=0)))
    return k, float(np.exp(lp[:, k]).mean())


def evaluate(model, data, fields_idx):
    """Regret of the model's choice against the exhaustive trial, and its
    gain over the untuned configuration, in payload bytes."""
    regrets, gains, confs = [], [], []
    for i in fields_idx:
        k, conf = field_choice(model, data["X"][data["fid"] == i])
        s = data["sizes"][i].astype(np.float64)
        regrets.append(s[k] / s.min() - 1.0)
        gains.append(s[DEFAULT_CONFIG] / s[k] - 1.0)
        confs.append(conf)
    r, g = np.array(regrets), np.array(gains)
    return {"fields": len(fields_idx), "mean_regret": float(r.mean()), "p90_regret": float(np.percentile(r, 90)),
            "max_regret": float(r.max()), "mean_gain_vs_default": float(g.mean()),
            "total_bytes_ratio_vs_trial": float(sum(data["sizes"][i][field_choice(model, data["X"][data["fid"] == i])[0]]
                                                    for i in fields_idx) / sum(data["sizes"][i].min() for i in fields_idx)),
            "mean_confidence": float(np.mean(confs))}


def library_choices(ctx, data):
    """The library's own choice and confidence (stz_tune: block-weighted vote)
    on every kept field, with its regret against the trial."""
    out = []
    for i, (f, eb) in sorted(data["kept"].items()):
        k, conf = choose(ctx, f, eb)
        s = data["sizes"][i].astype(np.float64)
        out.append((i, k, conf, float(s[k] / s.min() - 1.0)))
    return out


def calibrate_min_confidence(choices, max_regret=0.01, floor_percentile=2.0):
    """The confidence threshold below which the learned path falls back to
    the trial: above the confidence of every held-out field whose regret
    exceeds max_regret, and no lower than the floor_percentile-th percentile
    of the held-out confidences (a field less certain than almost every
    held-out one is treated as unlike the training set)."""
    conf = np.array([c for _, _, c, _ in choices])
    reg = np.array([r for _, _, _, r in choices])
    t = float(np.percentile(conf, floor_percentile))
    bad = conf[reg > max_regret]
    if bad.size:
        t = max(t, float(np.nextafter(np.float32(bad.max()), np.float32(np.inf))))
    fb = conf < t
    reg_fb = np.where(fb, 0.0, reg)
    return t, {"fields": int(conf.size), "fallback_rate": float(fb.mean()), "max_regret_learned": float(reg.max()),
               "fields_over_max_regret": int((reg > max_regret).sum()), "mean_regret_with_fallback": float(reg_fb.mean()),
               "max_regret_with_fallback": float(reg_fb.max()), "max_regret_target": max_regret,
               "confidence_min": float(conf.min()), "confidence_median": float(np.median(conf))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--fields", type=int, default=480)
    ap.add_argument("--holdout", type=int, default=80)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--data-out", default=None)
    args = ap.parse_args()
    import torch
    from .api import Context

    def log(*a):
        print(*a, flush=True)

    ctx = Context(0)
    n_train = args.fields - args.holdout
    data = build_dataset(ctx, args.fields, args.seed, log=log, keep_from=n_train)
    if args.data_out:
        np.savez_compressed(args.data_out, X=data["X"], fid=data["fid"], sizes=data["sizes"],
                            meta=json.dumps(data["meta"]))
    model = Mlp(ctx, seed=args.seed)
    t0 = time.time()
    train(model, data, steps=args.steps, fields_idx=np.arange(n_train), log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    ev_tr = evaluate(model, data, range(n_train))
    ev_ho = evaluate(model, data, range(n_train, args.fields))
    log("train", json.dumps(ev_tr))
    log("holdout", json.dumps(ev_ho))
    set_model(ctx, model.params(), 0.0)
    choices = library_choices(ctx, data)
    min_conf, cal = calibrate_min_confidence(choices)
    log("calibration", min_conf, json.dumps(cal))
    model.save(args.out, {"fields": args.fields, "holdout": args.holdout, "steps": args.steps, "seed": args.seed,
                          "configs": TUNE_CONFIGS, "train_seconds": t_train, "train_eval": ev_tr,
                          "holdout_eval": ev_ho, "min_confidence": min_conf, "calibration": cal})
    log("saved", args.out)
    ctx.close()


if __name__ == "__main__":
    main()
