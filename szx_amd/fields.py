"""Seeded synthetic scientific fields for the codec's workloads.

* ``sinusoid_field`` -- a sum of random plane waves with integer wave vectors
  (amplitude |k|^-slope, uniform phase), Gaussian noise relative to the field's
  std, and axis-aligned boxes set to a constant (0.0 or the field mean) so that
  constant blocks occur.  100x500x500 fp32 with 32 waves, k in [1,16], slope 2,
  noise 1e-3, one zero box of 25% is the CPU workload; 512^3 fp32 with 64
  waves, k in [1,32], slope 2.5, noise 1e-4, two mean boxes of 30% in total is
  the single-GPU workload.
* ``grf_field`` -- a 2-D Gaussian random field with P(k) ~ k^-3 (FFT of seeded
  complex white noise) times exp(3u) for a second such field u, with 10% of
  the entries set to 1e-300.

``grf_field_torch`` builds the same kind of field in device memory.

Every generator takes a seed and is deterministic.  The numpy versions serve
tests and the CPU baseline; ``sinusoid_field_torch`` builds the same kind of
field directly in device memory for the large single-GPU workloads (CPU and GPU
generators need not agree bit for bit: parity checks feed the same bytes to
both codecs).
"""
from __future__ import annotations

import numpy as np


def _waves(rng, ndim, n_waves, kmax, slope):
    k = rng.integers(1, kmax + 1, size=(n_waves, ndim))
    amp = np.linalg.norm(k, axis=1) ** (-float(slope))
    phase = rng.uniform(0.0, 2.0 * np.pi, size=n_waves)
    return k, amp, phase


def _boxes(rng, shape, n_boxes, total_frac):
    """Random axis-aligned boxes, each covering total_frac/n_boxes of the volume."""
    out = []
    ndim = len(shape)
    side = (total_frac / n_boxes) ** (1.0 / ndim)
    for _ in range(n_boxes):
        sl = []
        for d in shape:
            ln = max(1, int(round(side * d)))
            st = int(rng.integers(0, d - ln + 1))
            sl.append(slice(st, st + ln))
        out.append(tuple(sl))
    return out


def sinusoid_field(shape, seed, n_waves=32, kmax=16, slope=2.0, noise=1e-3, n_boxes=1,
                   box_frac=0.25, box_value="zero", dtype=np.float32) -> np.ndarray:
    shape = tuple(int(s) for s in shape)
    rng = np.random.default_rng(seed)
    ndim = len(shape)
    k, amp, phase = _waves(rng, ndim, n_waves, kmax, slope)
    grids = [np.arange(s, dtype=np.float64) / s for s in shape]
    f = np.zeros(shape, np.float64)
    # sin(A + B) = sin A cos B + cos A sin B with B along the fastest axis:
    # two multiply-adds per value instead of one sine.
    for w in range(n_waves):
        a = np.full(shape[:-1] + (1,), phase[w])
        for d in range(ndim - 1):
            sh = [1] * ndim
            sh[d] = shape[d]
            a = a + 2.0 * np.pi * k[w, d] * grids[d].reshape(sh)
        b = 2.0 * np.pi * k[w, ndim - 1] * grids[ndim - 1]
        f += amp[w] * (np.sin(a) * np.cos(b) + np.cos(a) * np.sin(b))
    std = f.std()
    if noise > 0:
        f += rng.normal(0.0, noise * std, size=shape)
    fill = 0.0 if box_value == "zero" else float(f.mean())
    for box in _boxes(rng, shape, n_boxes, box_frac):
        f[box] = fill
    return f.astype(dtype)


def sinusoid_field_torch(shape, seed, n_waves=64, kmax=32, slope=2.5, noise=1e-4, n_boxes=2,
                         box_frac=0.30, box_value="mean", dtype="float32", device="cuda"):
    """Device-resident version of sinusoid_field (same construction, torch).

    sin(A_w + B_w) with A_w over the leading axes and B_w along the fastest
    one is S_A C_B + C_A S_B, so the whole sum over waves is two
    (leading cells x waves) @ (waves x fastest axis) products in float64."""
    import torch
    shape = tuple(int(s) for s in shape)
    rng = np.random.default_rng(seed)
    ndim = len(shape)
    k, amp, phase = _waves(rng, ndim, n_waves, kmax, slope)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    f64 = torch.float64
    kt = torch.tensor(k, dtype=f64, device=device)
    lead = shape[:-1]
    A = torch.tensor(phase, dtype=f64, device=device).reshape(n_waves, *([1] * len(lead)))
    for d, s in enumerate(lead):
        g = torch.arange(s, dtype=f64, device=device) / s
        sh = [1] + [1] * len(lead)
        sh[1 + d] = s
        A = A + 2.0 * np.pi * kt[:, d].reshape(n_waves, *([1] * len(lead))) * g.reshape(sh)
    A = A.reshape(n_waves, -1)
    z = torch.arange(shape[-1], dtype=f64, device=device) / shape[-1]
    B = 2.0 * np.pi * kt[:, ndim - 1:ndim] * z.reshape(1, -1)
    at = torch.tensor(amp, dtype=f64, device=device).reshape(-1, 1)
    f = (torch.sin(A).t().contiguous() @ (torch.cos(B) * at).contiguous()
         + torch.cos(A).t().contiguous() @ (torch.sin(B) * at).contiguous()).reshape(shape)
    del A, B
    std = float(f.std())
    if noise > 0:
        f += torch.randn(shape, dtype=f64, device=device, generator=gen) * (noise * std)
    fill = 0.0 if box_value == "zero" else float(f.mean())
    for box in _boxes(rng, shape, n_boxes, box_frac):
        f[box] = fill
    out = f.to(torch.float32 if dtype == "float32" else torch.float64)
    del f
    return out


def grf_field(shape, seed, slope=3.0, dyn_exp=3.0, tiny_frac=0.10, tiny=1e-300,
              dtype=np.float64) -> np.ndarray:
    """2-D Gaussian random field with P(k) ~ k^-slope, times exp(dyn_exp*u)."""
    rng = np.random.default_rng(seed)

    def one():
        noise = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        ky = np.fft.fftfreq(shape[0])[:, None]
        kx = np.fft.fftfreq(shape[1])[None, :]
        kk = np.sqrt(kx * kx + ky * ky)
        kk[0, 0] = np.inf
        g = np.fft.ifft2(noise * kk ** (-slope / 2.0)).real
        return (g - g.mean()) / g.std()

    g = one()
    u = 0.5 * one()
    f = g * np.exp(dyn_exp * u)
    mask = rng.random(shape) < tiny_frac
    f[mask] = tiny
    return f.astype(dtype)


def grf_field_torch(shape, seed, slope=3.0, dyn_exp=3.0, tiny_frac=0.10, tiny=1e-300,
                    dtype="float64", device="cuda"):
    """Device-resident version of grf_field (same construction, torch FFTs)."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    f64 = torch.float64
    ky = torch.fft.fftfreq(shape[0], dtype=f64, device=device).reshape(-1, 1)
    kx = torch.fft.fftfreq(shape[1], dtype=f64, device=device).reshape(1, -1)
    kk = torch.sqrt(kx * kx + ky * ky)
    kk[0, 0] = float("inf")
    amp = kk ** (-slope / 2.0)
    del kk

    def one():
        noise = torch.complex(torch.randn(shape, dtype=f64, device=device, generator=gen),
                              torch.randn(shape, dtype=f64, device=device, generator=gen))
        g = torch.fft.ifft2(noise * amp).real
        del noise
        return (g - g.mean()) / g.std()

    g = one()
    u = 0.5 * one()
    f = g * torch.exp(dyn_exp * u)
    del g, u
    mask = torch.rand(shape, dtype=f64, device=device, generator=gen) < tiny_frac
    f[mask] = tiny
    return f.to(torch.float64 if dtype == "float64" else torch.float32)
