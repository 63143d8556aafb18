"""Train the fixed-rate rate model on the GPU with the library's own kernels.

Training data are (chunk features at a bound, achieved log2 bits/value of
that chunk) pairs produced by running the exact codec itself:

* fields: fp32 256^3, each the sum of 32 plane waves with integer wave
  vectors in [1,16], amplitude |k|^-slope with slope ~ U[1.5, 4], Gaussian
  noise with sigma/std log-uniform in [1e-6, 1e-2], and a 25% box of zeros;
  training fields use seeds 20000 + j, held-out validation fields 30000 + j
  (the evaluation fields of the benchmark use 5000 + i);
* bounds: 16 per field, log-spaced over range * [1e-7, 10^-0.5] with a random
  offset per field;
* per bound the codec compresses the whole field (szx_amd_compress_dev) and
  its record index gives every 2^19-value chunk's stream bytes (records,
  constant medians, bitmap bits); the chunk statistics pass
  (szx_amd_rate_stats_dev) and szx_amd_rate_inputs_dev give the features.

The model is trained with szx_amd_mlp_grad_dev / szx_amd_mlp_adam_dev (BF16
matrix-core forward and backward, FP32 accumulation and master weights) on
minibatches, and the FP32 parameters are written to rate_model_fp32.bin.

    python -m szx_amd.rate_train [--fields 400] [--steps 6000] [--out PATH]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import szx_amd as Z  # noqa: E402
from szx_amd.fields import sinusoid_field_torch  # noqa: E402

SHAPE = (256, 256, 256)
N_BOUNDS = 16
BS = 128


def log(*a):
    print(*a, flush=True)


def field(seed, device):
    r = np.random.default_rng(seed)
    slope = r.uniform(1.5, 4.0)
    noise = 10 ** r.uniform(-6, -2)
    x = sinusoid_field_torch(SHAPE, seed=seed, n_waves=32, kmax=16, slope=slope, noise=noise,
                             n_boxes=1, box_frac=0.25, box_value="zero", device=device)
    return x.contiguous().view(-1), r


def chunk_rates(comp, n):
    """Per-chunk bits/value of the stream the compressor just wrote."""
    idx = comp.index.cpu().numpy().astype(np.int64)
    tiles_per_chunk = Z.RATE_CHUNK // 4096
    ntiles = (n + 4095) // 4096
    nch = (n + Z.RATE_CHUNK - 1) // Z.RATE_CHUNK
    out = np.empty(nch)
    for c in range(nch):
        t0, t1 = c * tiles_per_chunk, min((c + 1) * tiles_per_chunk, ntiles)
        vals = min(Z.RATE_CHUNK, n - c * Z.RATE_CHUNK)
        payload = idx[2 * t1] - idx[2 * t0]
        consts = idx[2 * t1 + 1] - idx[2 * t0 + 1]
        blocks = (vals + BS - 1) // BS
        out[c] = 8.0 * (payload + 4 * consts + blocks / 8.0) / vals
    return out


def make_set(seeds, device):
    import torch
    n = int(np.prod(SHAPE))
    comp = Z.DeviceCompressor(n, np.float32, BS, device=device)
    Xs, Ts = [], []
    for i, seed in enumerate(seeds):
        x, r = field(seed, device)
        rng = float(x.max() - x.min())
        u = r.uniform(0, 1)
        rel = 10.0 ** (-7.0 + 6.5 * (np.arange(N_BOUNDS) + u) / N_BOUNDS)
        ebs = np.float32(rng) * rel.astype(np.float32)
        stats = Z.rate_stats(x, BS)
        T = np.empty((N_BOUNDS, (n + Z.RATE_CHUNK - 1) // Z.RATE_CHUNK), np.float32)
        for k, eb in enumerate(ebs):
            comp(x, Z.ABS, absErrBound=float(eb))
            comp.stream_view()   # raises if the call failed
            T[k] = np.log2(chunk_rates(comp, n))
        X = Z.rate_inputs(stats, n, torch.from_numpy(ebs).to(device))
        Xs.append(X)
        Ts.append(torch.from_numpy(T.reshape(-1)).to(device))
        if (i + 1) % 25 == 0:
            log(f"  {i + 1}/{len(seeds)} fields")
        del x
    return torch.cat(Xs), torch.cat(Ts)


def init_params(seed):
    """He-style initial parameters in the library's layout."""
    r = np.random.default_rng(seed)
    h, d = 64, Z.RATE_MODEL_INPUTS
    parts = [r.standard_normal(h * d) * np.sqrt(2.0 / d), np.zeros(h),
             r.standard_normal(h * h) * np.sqrt(2.0 / h), np.zeros(h),
             r.standard_normal(h) * np.sqrt(1.0 / h), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def train(X, T, Xv, Tv, steps, batch, seed, device):
    import torch
    model = Z.RateModel(init_params(seed), device=device)
    m = torch.zeros(Z.RATE_MODEL_PARAMS, device=device)
    v = torch.zeros_like(m)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    lr0, lr1 = 2e-3, 2e-5
    for step in range(1, steps + 1):
        sel = torch.randint(0, X.shape[0], (batch,), device=device, generator=gen)
        g, loss = model.grad(X[sel].contiguous(), T[sel].contiguous())
        lr = lr1 + 0.5 * (lr0 - lr1) * (1 + np.cos(np.pi * step / steps))
        model.adam(g, m, v, float(lr), step)
        if step % 500 == 0 or step == 1:
            with torch.no_grad():
                yv = model.forward(Xv)
                vl = float(((yv - Tv) ** 2).mean())
            log(f"  step {step}: train loss {float(loss.item()):.5f}, held-out loss {vl:.5f}")
    return model


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fields", type=int, default=400)
    ap.add_argument("--val-fields", type=int, default=24)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=Z.RATE_MODEL_PATH)
    args = ap.parse_args()
    import torch
    device = "cuda"
    t0 = time.time()
    log(f"generating {args.fields} training fields x {N_BOUNDS} bounds on the codec")
    X, T = make_set([20000 + j for j in range(args.fields)], device)
    log(f"generating {args.val_fields} held-out fields")
    Xv, Tv = make_set([30000 + j for j in range(args.val_fields)], device)
    log(f"{X.shape[0]} training triples, {Xv.shape[0]} held-out ({time.time() - t0:.1f} s)")
    t1 = time.time()
    model = train(X, T, Xv, Tv, args.steps, args.batch, args.seed, device)
    torch.cuda.synchronize()
    log(f"trained {args.steps} steps in {time.time() - t1:.1f} s")
    with torch.no_grad():
        yv = model.forward(Xv)
        err = (2.0 ** (yv - Tv) - 1.0).abs()
    log(f"held-out chunk rate error: median {float(err.median()):.4f}, "
        f"90th pct {float(err.quantile(0.9)):.4f}")
    p = model.params.cpu().numpy().astype("<f4")
    p.tofile(args.out)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
