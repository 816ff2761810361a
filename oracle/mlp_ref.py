"""CPU float64 reference of the learned tuner's MLP (16 -> 64 -> 64 -> 208,
ReLU, soft-target cross entropy) and of one Adam step.

TEST INFRASTRUCTURE ONLY: the checker of the library's bf16 MFMA kernels
(stz_mlp_forward / stz_mlp_grad / stz_mlp_adam).  With bf16=True the values
are rounded to bfloat16 (round to nearest even, from the format's
definition) at the points where the device kernels round them: inputs,
weights, hidden activations and the backward deltas; every product and sum
is float64 here (fp32 accumulation on the device).  With bf16=False it is
the plain float64 network, whose gradient the tests pin by finite
differences.

Parameter layout (fp32 vector of 18768): W1[16][64], b1[64], W2[64][64],
b2[64], W3[64][208], b3[208]; W[k][n] maps input k to output n.
"""
import numpy as np

IN, HID, OUT = 16, 64, 208
N_PARAMS = IN * HID + HID + HID * HID + HID + HID * OUT + OUT


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) and return float64."""
    u = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, np.float64)
    o = 0
    out = []
    for shape in ((IN, HID), (HID,), (HID, HID), (HID,), (HID, OUT), (OUT,)):
        n = int(np.prod(shape))
        out.append(P[o:o + n].reshape(shape))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])


def forward(P, X, bf=True):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    x = r(X)
    W1r, W2r, W3r = (r(W1), r(W2), r(W3)) if bf else (W1, W2, W3)
    h1 = r(np.maximum(x @ W1r + b1, 0.0))
    h2 = r(np.maximum(h1 @ W2r + b2, 0.0))
    y = h2 @ W3r + b3
    return y, (x, h1, h2, W1r, W2r, W3r)


def loss_grad(P, X, T, bf=True):
    """Mean soft-target cross entropy over the rows and its gradient."""
    n = X.shape[0]
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, (x, h1, h2, W1r, W2r, W3r) = forward(P, X, bf)
    m = y.max(axis=1, keepdims=True)
    lse = m + np.log(np.exp(y - m).sum(axis=1, keepdims=True))
    p = np.exp(y - lse)
    T = np.asarray(T, np.float64)
    loss = float(-(T * (y - lse)).sum() / n)
    dy = r((p - T) / n)
    gW3 = h2.T @ dy
    gb3 = dy.sum(axis=0)
    dh2 = r(dy @ W3r.T) * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(axis=0)
    dh1 = r(dh2 @ W2r.T) * (h1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(axis=0)
    return loss, pack([gW1, gb1, gW2, gb2, gW3, gb3])


def adam(P, G, M, V, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    M = b1 * M + (1 - b1) * G
    V = b2 * V + (1 - b2) * G * G
    mh = M / (1 - b1 ** step)
    vh = V / (1 - b2 ** step)
    return P - lr * (mh / (np.sqrt(vh) + eps) + wd * P), M, V
