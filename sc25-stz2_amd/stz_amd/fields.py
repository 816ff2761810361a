"""Seeded synthetic fields for the benchmark workloads (C-order (z, y, x)).

All randomness comes from numpy PCG64 with the stated seed; spectra are
applied in float64 with numpy's FFT, or torch's when a device is given
(faster at 512^3 and above; the white noise is still drawn by numpy).
"""
import os

import numpy as np


def sinusoids(n=128, n_waves=32, seed=1001, dtype=np.float32, shape=None):
    """Sum of random 3D sinusoids (wavenumbers 1-8, amplitudes U(0.1,1),
    random phases) plus N(0, 1e-3) noise."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape if shape is not None else (n, n, n)
    z = np.arange(nz, dtype=np.float64)[:, None, None]
    y = np.arange(ny, dtype=np.float64)[None, :, None]
    x = np.arange(nx, dtype=np.float64)[None, None, :]
    f = np.zeros((nz, ny, nx), dtype=np.float64)
    for _ in range(n_waves):
        k = rng.integers(1, 9, size=3)
        a = rng.uniform(0.1, 1.0)
        ph = rng.uniform(0.0, 2 * np.pi)
        f += a * np.sin(2 * np.pi * (k[0] * x / nx + k[1] * y / ny + k[2] * z / nz) + ph)
    f += rng.normal(0.0, 1e-3, size=f.shape)
    return f.astype(dtype)


def _kmag(shape):
    nz, ny, nx = shape
    kz = np.fft.fftfreq(nz) * nz
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.rfftfreq(nx) * nx
    return np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)


def grf(shape, slope, seed, device=None):
    """Unit-variance Gaussian random field with P(k) ~ k^slope, P(0) = 0 (float64)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(size=shape)
    if device is not None:
        import torch
        t = torch.from_numpy(w).to(device)
        del w
        F = torch.fft.rfftn(t)
        del t
        nz, ny, nx = shape
        kz = torch.fft.fftfreq(nz, device=device, dtype=torch.float64) * nz
        ky = torch.fft.fftfreq(ny, device=device, dtype=torch.float64) * ny
        kx = torch.fft.rfftfreq(nx, device=device, dtype=torch.float64) * nx
        k = torch.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
        amp = torch.where(k > 0, k.clamp(min=1e-30) ** (slope / 2.0), torch.zeros_like(k))
        F *= amp
        del k, amp
        g = torch.fft.irfftn(F, s=shape)
        del F
        g = (g - g.mean()) / g.std()
        return g.cpu().numpy()
    try:  # multithreaded pocketfft when scipy is there (any length, no plan compilation)
        import scipy.fft as sfft
        workers = min(16, os.cpu_count() or 1)
        F = sfft.rfftn(w, workers=workers)
    except ImportError:
        sfft = None
        F = np.fft.rfftn(w)
    del w
    k = _kmag(shape)
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** (slope / 2.0), 0.0)
    F *= amp
    g = sfft.irfftn(F, s=shape, workers=workers) if sfft is not None else np.fft.irfftn(F, s=shape)
    g -= g.mean()
    g /= g.std()
    return g


def grf_device(shape, slope, seed, device, dtype=None):
    """grf() computed on the GPU and returned as a device tensor (torch dtype,
    float64 by default).  The white noise is the same numpy PCG64 stream,
    drawn plane by plane so no full host copy is held."""
    import torch
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    t = torch.empty(shape, dtype=torch.float64, device=device)
    for z in range(nz):
        t[z].copy_(torch.from_numpy(rng.standard_normal(size=(ny, nx))))
    F = torch.fft.rfftn(t)
    del t
    kz = torch.fft.fftfreq(nz, device=device, dtype=torch.float64) * nz
    ky = torch.fft.fftfreq(ny, device=device, dtype=torch.float64) * ny
    kx = torch.fft.rfftfreq(nx, device=device, dtype=torch.float64) * nx
    for z in range(nz):   # k^(slope/2) amplitude, plane by plane to bound memory
        k = torch.sqrt(kz[z] ** 2 + ky[:, None] ** 2 + kx[None, :] ** 2)
        F[z] *= torch.where(k > 0, k.clamp(min=1e-30) ** (slope / 2.0), torch.zeros_like(k))
    g = torch.fft.irfftn(F, s=shape)
    del F
    g = (g - g.mean()) / g.std()
    return g if dtype is None else g.to(dtype)


def lognormal(n=512, sigma=1.5, slope=-3.0, seed=1002, dtype=np.float32, device=None, shape=None):
    """exp(sigma * g), g a unit GRF with P(k) ~ k^-3 (density-like field)."""
    shape = shape if shape is not None else (n, n, n)
    g = grf(shape, slope, seed, device)
    return np.exp(sigma * g).astype(dtype)


def turbulence(shape=(256, 384, 384), seed=1003, n_fronts=8, width=2.0, dtype=np.float64, device=None):
    """Unit GRF with P(k) ~ k^(-11/3) plus tanh fronts of the given width at random planes."""
    g = grf(shape, -11.0 / 3.0, seed, device)
    rng = np.random.default_rng(seed + 7919)
    nz, ny, nx = shape
    z = np.arange(nz, dtype=np.float64)[:, None, None]
    y = np.arange(ny, dtype=np.float64)[None, :, None]
    x = np.arange(nx, dtype=np.float64)[None, None, :]
    for _ in range(n_fronts):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        c = np.array([rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz)])
        d = nrm @ c
        g += np.tanh((nrm[0] * x + nrm[1] * y + nrm[2] * z - d) / width)
    return g.astype(dtype)


def uneven_batch(count=96, lo=32, hi=384, seed0=1004, dtype=np.float32, device=None, max_points=None):
    """Fields of uneven non-cubic sizes (each dim ~ U{lo..hi}), GRF spectra with slopes U[-4,-2]."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        shape = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        if max_points is not None:
            while shape[0] * shape[1] * shape[2] > max_points:
                shape = tuple(max(lo, v // 2) for v in shape)
        slope = float(rng.uniform(-4.0, -2.0))
        out.append(grf(shape, slope, seed0 + i, device).astype(dtype))
    return out
