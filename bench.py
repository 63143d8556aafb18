#!/usr/bin/env python3
"""Single-GPU throughput of the blockwise fast codec (SZx fastMode 2).

Workload: one fp32 field of 512x512x512 (512 MiB) -- 64 random plane waves
with integer wave vectors in [1,32], amplitude |k|^-2.5, Gaussian noise at
1e-4 of the field's std, two boxes (30% of the volume together) holding the
field mean -- compressed at a REL bound (default 1e-3, blockSize 128) and
decompressed, all device-resident.

A step is one compress of the whole field through the C ABI
(szx_amd_compress_dev, REL mode: range pass + single-pass codec + header
assembly).  value = uncompressed bytes / step time (GB/s).  Also reported:
ABS-mode compress (the bound already known, as the reference's CLI times it),
decompress with the compressor's record index and without it (rebuilt from
the stream), roofline of the dominant kernel from HIP events recorded around
it inside the library, a full-size parity check against the oracle, the
reference's own CPU codec timed on this host (compiled from its sources into
oracle/_ref; the oracle itself if that is absent), and the fixed-rate solve
on 64 fp32 256^3 fields at 1/2/4/8 bits per value: learned proposal vs
codec-driven bisection (codec passes, wall time, rate error), with the
reference's sampled-model bound judged by its real stream beside them; an fp64
8192x8192 Gaussian random field at REL 1e-4/1e-6/1e-8 and blockSize
64/128/256 (throughput and full-size parity per configuration); the 512^3
field at block sizes outside the tiled kernels (default 100, 255, 7: the
per-block segment kernels, throughput and full-size parity); the random-access
container (fastMode 3 and 4), the non-blocked codec (fastMode 1) and the
cuSZx stream writer and reader on the same field; and 1000 mixed
fp32/fp64 fields compressed and decompressed by the batched launch, against
one launch sequence per field.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--rel 1e-3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cmpr/decmpr GB/s + %HBM peak; fixed-rate solve time, learned vs bisection"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_times(steps, *fns):
    """The library's kernel timer over `steps` calls of each fn, in a pass of
    its own: the wall-clock loops run with the timer off, since its events sit
    between the kernels of a call and add a few microseconds each."""
    import torch

    import szx_amd as Z
    torch.cuda.synchronize()
    Z.timing_begin()
    for fn in fns:
        for _ in range(steps):
            fn()
    torch.cuda.synchronize()
    return Z.timing_end()


def cpu_baseline(x_host, eb, sample_values):
    """Time the reference CPU codec on a bounded sample of the same field."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    xs = np.ascontiguousarray(x_host[:sample_values])
    nbytes = xs.nbytes
    out = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def best(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    if O.ref_available():
        kind = "reference"
        t = best(lambda: O.ref_compress(xs, eb, 128))
        s = O.ref_compress(xs, eb, 128)
        td = best(lambda: O.ref_decompress(s, xs.size, np.float32))
        t4 = best(lambda: O.ref_compress_args(xs, mode=0, abs_eb=eb, fast_mode=4))
        out["serial"] = {"compress_gbs": nbytes / t / 1e9, "decompress_gbs": nbytes / td / 1e9,
                         "cores": 1}
        s4 = O.ref_compress_args(xs, mode=0, abs_eb=eb, fast_mode=4)
        td4 = best(lambda: O.ref_decompress(s4, xs.size, np.float32, fast_mode=4))
        out["openmp"] = {"compress_gbs": nbytes / t4 / 1e9, "decompress_gbs": nbytes / td4 / 1e9,
                         "cores": threads}
        # the other modes on the same sample, one core each
        others = {}
        for name, fm in (("fastmode1_serial", 1), ("fastmode3_serial", 3)):
            tc = best(lambda: O.ref_compress_args(xs, mode=0, abs_eb=eb, fast_mode=fm))
            sm = O.ref_compress_args(xs, mode=0, abs_eb=eb, fast_mode=fm)
            tdm = best(lambda: O.ref_decompress(sm, xs.size, np.float32, fast_mode=fm))
            others[name] = {"compress_gbs": nbytes / tc / 1e9, "decompress_gbs": nbytes / tdm / 1e9}
        tb = best(lambda: O.ref_compress(xs, eb, 100))
        others["blocksize100_serial"] = {"compress_gbs": nbytes / tb / 1e9}
        out["other_modes"] = others
    else:
        kind = "port"
        t = best(lambda: O.compress(xs, eb, 128))
        s = O.compress(xs, eb, 128)
        td = best(lambda: O.decompress(s, xs.size, np.float32))
        out["serial"] = {"compress_gbs": nbytes / t / 1e9, "decompress_gbs": nbytes / td / 1e9,
                         "cores": 1}
    return kind, out


def host_cpu_info():
    """Host cores this process may use, all cores, and the CPU model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cfg1_eval(dev, steps, parity=True, cpu=True):
    """The CPU-runnable configuration at full size: one fp32 100x500x500
    field (32 waves, integer k in [1,16], |k|^-2, noise 1e-3 std, a 25% box of
    0.0), blockSize 128, ABS eb = 1e-3 x (max - min).  GPU: compress (ABS and
    REL 1e-3, which yields the same eb) and decompress with the record index,
    with the stream and the decode compared with the oracle's.  CPU: the
    reference compiled from its sources (the oracle if that is absent), serial
    fastMode 2 on one core and OpenMP fastMode 4 on every core, each with the
    eb given (ABS) and with the REL range pass, best of 5 after one warm-up."""
    import numpy as np
    import torch

    import szx_amd as Z
    from szx_amd.fields import sinusoid_field_torch

    shape = (100, 500, 500)
    x = sinusoid_field_torch(shape, seed=1, n_waves=32, kmax=16, slope=2.0, noise=1e-3, n_boxes=1,
                             box_frac=0.25, box_value="zero", device=dev).contiguous().view(-1)
    n = x.numel()
    nbytes = 4 * n
    comp = Z.DeviceCompressor(n, np.float32, 128, device=dev)
    comp(x, Z.REL, relBoundRatio=1e-3)
    view, _, csize, eb = comp.stream_view()
    dec = Z.DeviceDecompressor(n, np.float32, 128, comp.capacity, device=dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_rel = timed(lambda: comp(x, Z.REL, relBoundRatio=1e-3))
    t_abs = timed(lambda: comp(x, Z.ABS, absErrBound=eb))
    view, _, csize, eb2 = comp.stream_view()
    assert eb2 == eb
    t_dec = timed(lambda: dec(view, csize, Z.DECODE_REFERENCE, index=comp.index))
    assert dec.check() == 0
    out = {"workload": "fp32 100x500x500 sinusoid field (32 waves, k<=16, |k|^-2, noise 1e-3 std, "
                       "25% zero box), blockSize 128, ABS eb = 1e-3 x range",
           "eb": eb, "compression_ratio": nbytes / csize,
           "gpu": {"compress_rel_gbs": nbytes / t_rel / 1e9, "compress_abs_gbs": nbytes / t_abs / 1e9,
                   "decompress_gbs": nbytes / t_dec / 1e9,
                   "note": "wall clock per call, inputs resident in HBM"}}
    xh = x.cpu().numpy() if (parity or cpu) else None
    if parity or cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
    if parity:
        s_gpu = view.cpu().numpy().tobytes()
        eb_o = O.rel_eb(xh, 1e-3)
        s_ora = O.compress(xh, eb_o, 128)
        d_ora = O.decompress(s_ora, n, np.float32)
        d_gpu = dec.out.cpu().numpy()
        err = float(np.max(np.abs(d_gpu.astype(np.float64) - xh)))
        out["parity"] = {"eb_identical": bool(np.float32(eb_o) == np.float32(eb)),
                         "stream_identical": s_gpu == s_ora,
                         "decode_identical": d_gpu.tobytes() == d_ora.tobytes(),
                         "max_abs_err": err, "err_within_eb": err <= eb}
        log(f"cfg1 parity {out['parity']}")
    if cpu:
        def best(fn, reps=5):
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return min(ts)

        rows = {}
        if O.ref_available():
            kind = "reference"
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            for name, fm, cores in (("serial_fastmode2", 2, 1), ("openmp_fastmode4", 4, threads)):
                tc_abs = best(lambda: O.ref_compress_args(xh, mode=0, abs_eb=eb, fast_mode=fm))
                tc_rel = best(lambda: O.ref_compress_args(xh, mode=1, rel=1e-3, fast_mode=fm))
                st = O.ref_compress_args(xh, mode=1, rel=1e-3, fast_mode=fm)
                td = best(lambda: O.ref_decompress(st, n, np.float32, fast_mode=fm))
                rows[name] = {"cores": cores, "compress_abs_gbs": nbytes / tc_abs / 1e9,
                              "compress_rel_gbs": nbytes / tc_rel / 1e9,
                              "decompress_gbs": nbytes / td / 1e9,
                              "stream_equals_gpu": (st == s_gpu) if (parity and fm == 2) else None}
        else:
            kind = "port"
            tc = best(lambda: O.compress(xh, eb, 128))
            td = best(lambda: O.decompress(O.compress(xh, eb, 128), n, np.float32))
            rows["serial_fastmode2"] = {"cores": 1, "compress_abs_gbs": nbytes / tc / 1e9,
                                        "decompress_gbs": nbytes / td / 1e9}
        out["cpu"] = {"kind": kind, "host": host_cpu_info(), **rows,
                      "method": "best of 5 after 1 warm-up, in memory; compress_rel includes the "
                                "reference's serial range pass (computeValueRange)"}
        log(f"cfg1 CPU {out['cpu']}")
    del comp, dec, x
    torch.cuda.empty_cache()
    return out


def measured_traffic(shape, rel):
    """HBM bytes per k_compress launch on the default workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a
    plain copy of the input); None for any other workload."""
    if tuple(shape) != (512, 512, 512) or rel != 1e-3:
        return None
    path = os.path.join(ROOT, "profiles", "k_compress_512cubed_hbm_bytes.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch k_compress:"):
                    return int(line.split(":")[1])
    except OSError:
        return None
    return None


def rate_reachability(x, targets, tol, bs=128, depth=40):
    """Which target rates any bound can reach, from the exact stream sizes
    of every octave of bounds between range * 2^-depth and the range
    (k_rate_octave: 512 bounds per octave, and the size each octave tends to
    at its top edge).  Inside an octave the size falls in steps of single
    blocks turning constant; at each octave edge every block's reqLength
    drops one bit and the size jumps.  Per target: 'hit' when a size on the
    ladder is within tolerance, 'gap' when the tolerance window lies wholly
    inside a jump or outside every reachable size (no bound reaches it), and
    'between' otherwise (the window falls between two ladder sizes of one
    octave: reachable unless a single step crosses it)."""
    import math

    import numpy as np
    import szx_amd as Z
    n = x.numel()
    rng = float(x.max()) - float(x.min())
    top = math.floor(math.log2(rng)) if rng > 0 else 0
    pts, spans = [], []
    for e in range(top - depth, top + 2):
        sizes, _, end = Z.rate_octave(x, e, bs, with_end=True)
        bpv = 8.0 * sizes.astype(np.float64) / n
        pts.append(bpv)
        pts.append(np.array([8.0 * end / n]))
        spans.append((8.0 * end / n, float(bpv[0])))
    pts = np.concatenate(pts)
    out = []
    for t in targets:
        lo, hi = t * (1 - tol), t * (1 + tol)
        if np.any((pts >= lo) & (pts <= hi)):
            out.append("hit")
        elif any(a <= hi and b >= lo for a, b in spans):
            out.append("between")
        else:
            out.append("gap")
    return out


def fixed_rate_eval(n_fields, dev, tol=0.02, targets=(1.0, 2.0, 4.0, 8.0)):
    """Fixed-rate solve on fp32 256^3 fields (32 waves, k in [1,16], slope
    U[1.5,4], noise sigma/std log-uniform [1e-6,1e-2], a 25% zero box; seeds
    5000+i): learned proposal vs codec-driven bisection, plus the reference's
    sampled-model FXR bound judged by the real stream.  Wall time per solve
    includes every pass and host synchronisation."""
    import numpy as np
    import torch
    import szx_amd as Z
    from szx_amd.fields import sinusoid_field_torch
    n = 256 ** 3
    model = Z.RateModel(device=dev)
    fr = Z.FixedRateCompressor(n, np.float32, 128, model=model, device=dev)
    fxb = Z.DeviceFixedRateBound(n, np.float32, 128, device=dev)
    comp = Z.DeviceCompressor(n, np.float32, 128, device=dev)
    res = {m: {"passes": [], "ms": [], "err": [], "hit": []} for m in ("learned", "bisection",
                                                                       "reference_model")}

    def gen(i):
        r = np.random.default_rng(5000 + i)
        slope, noise = r.uniform(1.5, 4.0), 10 ** r.uniform(-6, -2)
        return sinusoid_field_torch((256, 256, 256), seed=5000 + i, n_waves=32, kmax=16,
                                    slope=slope, noise=noise, n_boxes=1, box_frac=0.25,
                                    box_value="zero", device=dev).contiguous().view(-1)

    x = gen(0)
    fr(x, 2.0, tol, method=Z.RATE_LEARNED)   # warm-up
    fr(x, 2.0, tol, method=Z.RATE_BISECT)
    reach = []
    for i in range(n_fields):
        x = gen(i)
        torch.cuda.synchronize()
        reach += rate_reachability(x, targets, tol)
        for t in targets:
            for name, meth in (("learned", Z.RATE_LEARNED), ("bisection", Z.RATE_BISECT)):
                t0 = time.perf_counter()
                info = fr(x, t, tol, method=meth)
                ms = (time.perf_counter() - t0) * 1e3
                d = res[name]
                d["passes"].append(info.codecPasses)
                d["ms"].append(ms)
                d["err"].append(abs(info.bitsPerValue - t) / t)
                d["hit"].append(info.hit)
            # the reference's fixed-rate mode: sampled model for CR = 32 / bpv
            t0 = time.perf_counter()
            eb, _ = fxb(x, 32.0 / t, tol)
            comp(x, Z.ABS, absErrBound=eb)
            size = comp.stream_view()[2]
            ms = (time.perf_counter() - t0) * 1e3
            bpv = 8.0 * size / n
            d = res["reference_model"]
            d["passes"].append(1)
            d["ms"].append(ms)
            d["err"].append(abs(bpv - t) / t)
            d["hit"].append(int(abs(bpv - t) <= tol * t))
        if (i + 1) % 16 == 0:
            log(f"  fixed rate: {i + 1}/{n_fields} fields")
    out = {"fields": n_fields, "targets_bpv": list(targets), "tolerance": tol,
           "workload": "fp32 256^3 sinusoid fields (32 waves, k<=16, slope U[1.5,4], noise "
                       "1e-6..1e-2 std, 25% zero box), seeds 5000+i"}
    for name, d in res.items():
        out[name] = {"mean_codec_passes": float(np.mean(d["passes"])),
                     "mean_solve_ms": float(np.mean(d["ms"])),
                     "median_solve_ms": float(np.median(d["ms"])),
                     "hit_rate": float(np.mean(d["hit"])),
                     "mean_rate_err": float(np.mean(d["err"])),
                     "max_passes": int(np.max(d["passes"]))}
    feas = [r == "hit" for r in reach]
    out["reachability"] = {
        "targets": len(reach), "reachable_on_ladder": int(sum(feas)),
        "between_ladder_sizes": int(sum(r == "between" for r in reach)),
        "in_a_gap": int(sum(r == "gap" for r in reach)),
        "note": "exact sizes of every octave of bounds (512 per octave): a target in a gap is "
                "reached by no bound; 'between' targets fall between two ladder sizes of one octave"}
    for name, d in res.items():
        hits = [h for h, f in zip(d["hit"], feas) if f]
        out[name]["hit_rate_on_reachable"] = float(np.mean(hits)) if hits else None
    both = [a == b for a, b in zip(res["learned"]["hit"], res["bisection"]["hit"])]
    out["learned_hits_where_bisection_hits"] = float(np.mean(
        [a for a, b in zip(res["learned"]["hit"], res["bisection"]["hit"]) if b] or [0]))
    out["same_hit_outcome"] = float(np.mean(both))
    return out


def fp64_eval(dev, steps, parity=True):
    """fp64 8192x8192 Gaussian random field (P(k) ~ k^-3, times exp(3u), 10%
    of the entries 1e-300), REL 1e-4 / 1e-6 / 1e-8 at blockSize 64 / 128 /
    256: compress (REL: range pass + codec + assembly) and decompress with the
    record index, per configuration; with parity, the stream and the decode
    are compared with the oracle's at full size."""
    import numpy as np
    import torch

    import szx_amd as Z
    from szx_amd.fields import grf_field_torch

    x = grf_field_torch((8192, 8192), seed=3, device=dev).reshape(-1).contiguous()
    n = x.numel()
    nbytes = n * 8
    xh = x.cpu().numpy() if parity else None
    O = None
    if parity:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
    out = []
    for bs in (64, 128, 256):
        comp = Z.DeviceCompressor(n, np.float64, bs, device=dev)
        dec = Z.DeviceDecompressor(n, np.float64, bs, comp.capacity, device=dev)
        for rel in (1e-4, 1e-6, 1e-8):
            step = lambda: comp(x, Z.REL, relBoundRatio=rel)
            step()
            view, off, csize, eb = comp.stream_view()
            dstep = lambda: dec(view, csize, Z.DECODE_REFERENCE, index=comp.index)
            dstep()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(steps):
                dstep()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            kt = kernel_times(steps, step, dstep)
            assert dec.check() == 0, "decoder rejected the stream"
            kc = kt["compress"][0] / max(1, kt["compress"][1])
            kd = kt["decode"][0] / max(1, kt["decode"][1])
            rec = {"block_size": bs, "rel": rel, "eb": eb, "compression_ratio": nbytes / csize,
                   "compress_gbs": nbytes * steps / (t1 - t0) / 1e9,
                   "decompress_gbs": nbytes * steps / (t2 - t1) / 1e9,
                   "compress_kernel_ms": kc, "decode_kernel_ms": kd,
                   "compress_kernel_pct_hbm": (nbytes + csize) / (kc * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "decode_kernel_pct_hbm": (nbytes + csize) / (kd * 1e-3) / 1e9 / HBM_PEAK_GBS}
            if parity:
                s_gpu = view.cpu().numpy().tobytes()
                eb_o = O.rel_eb(xh, rel)
                rec["stream_identical"] = bool(np.float32(eb_o) == np.float32(eb)) and \
                    s_gpu == O.compress(xh, eb_o, bs)
                d_ora = O.decompress(s_gpu, n, np.float64, ref_quirk=True)
                rec["decode_identical"] = bool(np.array_equal(dec.out.cpu().numpy().view(np.uint64),
                                                              d_ora.view(np.uint64)))
            log(f"fp64 bs {bs} rel {rel:g}: {rec}")
            out.append(rec)
        del comp, dec
    del x
    torch.cuda.empty_cache()
    return out


def generic_eval(x, xh, eb, dev, steps, sizes, parity=True):
    """Block sizes outside the tiled kernels (per-block segment kernels) on
    the same fp32 field at the same ABS bound: compress and decompress (with
    the record index) throughput, kernel-sequence time, and with parity the
    stream and the decode compared with the oracle's at full size."""
    import numpy as np
    import torch

    import szx_amd as Z

    n = x.numel()
    nbytes = n * 4
    O = None
    if parity:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
    out = []
    for bs in sizes:
        comp = Z.DeviceCompressor(n, np.float32, bs, device=dev)
        comp(x, Z.ABS, absErrBound=eb)
        view, off, csize, _ = comp.stream_view()
        dec = Z.DeviceDecompressor(n, np.float32, bs, comp.capacity, device=dev)
        dstep = lambda: dec(view, csize, Z.DECODE_REFERENCE, index=comp.index)
        dstep()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            comp(x, Z.ABS, absErrBound=eb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            dstep()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        kt = kernel_times(steps, lambda: comp(x, Z.ABS, absErrBound=eb), dstep)
        assert dec.check() == 0, "decoder rejected the stream"
        kc = kt["compress"][0] / max(1, kt["compress"][1])
        kd = kt["decode"][0] / max(1, kt["decode"][1])
        rec = {"block_size": bs, "compression_ratio": nbytes / csize,
               "compress_gbs": nbytes * steps / (t1 - t0) / 1e9,
               "decompress_gbs": nbytes * steps / (t2 - t1) / 1e9,
               "compress_kernels_ms": kc, "decode_kernel_ms": kd,
               "compress_kernels_pct_hbm": (nbytes + csize) / (kc * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "decode_kernel_pct_hbm": (nbytes + csize) / (kd * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if parity:
            s_gpu = view.cpu().numpy().tobytes()
            s_ora = O.compress(xh, eb, bs)
            rec["stream_identical"] = s_gpu == s_ora
            d_ora = O.decompress(s_ora, n, np.float32)
            rec["decode_identical"] = bool(np.array_equal(dec.out.cpu().numpy().view(np.uint32),
                                                          d_ora.view(np.uint32)))
        log(f"generic bs {bs}: {rec}")
        out.append(rec)
        del comp, dec, view
    torch.cuda.empty_cache()
    return out


def random_access_eval(x, xh, rel, dev, steps, parity=True):
    """The random-access container (fastMode 3: big-endian medians, fastMode
    4: native medians) on the same fp32 field at the same REL bound: compress
    (whole call), decompress with the record index and from the container's
    size table, and with parity the stream and the decode against the oracle."""
    import numpy as np
    import torch

    import szx_amd as Z

    n = x.numel()
    nbytes = n * 4
    O = None
    if parity:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
    out = []
    for fm in (3, 4):
        comp = Z.DeviceCompressor(n, np.float32, 128, device=dev, fastMode=fm)
        comp(x, Z.REL, relBoundRatio=rel)
        view, off, csize, eb = comp.stream_view()
        dec = Z.DeviceDecompressor(n, np.float32, 128, comp.capacity, device=dev, fastMode=fm)
        dec(view, csize, Z.DECODE_REFERENCE, index=comp.index)
        dec(view, csize, Z.DECODE_REFERENCE)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            comp(x, Z.REL, relBoundRatio=rel)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            dec(view, csize, Z.DECODE_REFERENCE, index=comp.index)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(steps):
            dec(view, csize, Z.DECODE_REFERENCE)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        kt = kernel_times(steps, lambda: comp(x, Z.REL, relBoundRatio=rel),
                          lambda: dec(view, csize, Z.DECODE_REFERENCE, index=comp.index))
        assert dec.check() == 0, "decoder rejected the container"
        kc = kt["compress"][0] / max(1, kt["compress"][1])
        rec = {"fast_mode": fm, "compression_ratio": nbytes / csize,
               "compress_gbs": nbytes * steps / (t1 - t0) / 1e9,
               "compress_kernel_ms": kc,
               "decompress_gbs": nbytes * steps / (t2 - t1) / 1e9,
               "decompress_size_table_gbs": nbytes * steps / (t3 - t2) / 1e9}
        if parity:
            s_gpu = view.cpu().numpy().tobytes()
            s_ora = O.compress_ra(xh, eb, 128, fm)
            rec["stream_identical"] = s_gpu == s_ora
            d_ora = O.decompress_ra(s_ora, n, np.float32, fm)
            rec["decode_identical"] = bool(np.array_equal(dec.out.cpu().numpy().view(np.uint32),
                                                          d_ora.view(np.uint32)))
        log(f"random access fastMode {fm}: {rec}")
        out.append(rec)
        del comp, dec, view
    torch.cuda.empty_cache()
    return out


def non_blocked_eval(x, xh, rel, dev, steps, parity=True):
    """The non-blocked codec (fastMode 1) on the same fp32 field at the same
    REL bound: compress and decompress (whole calls), kernel time, and with
    parity the stream and the decode against the oracle."""
    import numpy as np
    import torch

    import szx_amd as Z

    n = x.numel()
    nbytes = n * 4
    c = Z.DeviceNonBlocked(n, np.float32, device=dev)
    c.compress(x, Z.REL, relBoundRatio=rel)
    view, csize, eb = c.stream_view()
    c.decompress(view, csize)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        c.compress(x, Z.REL, relBoundRatio=rel)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        c.decompress(view, csize)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    kt = kernel_times(steps, lambda: c.compress(x, Z.REL, relBoundRatio=rel), lambda: c.decompress(view, csize))
    assert c.check() == 0, "decoder rejected the stream"
    rec = {"compression_ratio": nbytes / csize, "compress_gbs": nbytes * steps / (t1 - t0) / 1e9,
           "decompress_gbs": nbytes * steps / (t2 - t1) / 1e9,
           "compress_kernels_ms": kt["compress"][0] / max(1, kt["compress"][1]),
           "decode_kernels_ms": kt["decode"][0] / max(1, kt["decode"][1])}
    if parity:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        s_gpu = view.cpu().numpy().tobytes()
        s_ora = O.compress_nb(xh, eb)
        rec["stream_identical"] = s_gpu == s_ora
        d_ora = O.decompress_nb(s_ora, n, np.float32)
        rec["decode_identical"] = bool(np.array_equal(c.dec.cpu().numpy().view(np.uint32), d_ora.view(np.uint32)))
    log(f"non-blocked fastMode 1: {rec}")
    del c, view
    torch.cuda.empty_cache()
    return rec


def cuszx_eval(x, xh, eb, dev, steps, parity=True):
    """The cuSZx stream (the reference's CUDA codec format, blockSize 128) on
    the same fp32 field at the headline's absolute bound: the writer and the
    reader (whole calls and kernel time), and with parity the stream against
    the oracle's restatement of the reference's writer and the decode against
    its decoder."""
    import numpy as np
    import torch

    import szx_amd as Z

    n = x.numel()
    nbytes = n * 4
    w = Z.DeviceCuszxWriter(n, 128, device=dev)
    rd = Z.DeviceCuszxReader(n, device=dev)
    w.compress(x, eb)
    view, csize, bad = w.stream_view()
    assert not bad, "a block needs more than 32 significant bits"
    rd.decompress(view, csize)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        w.compress(x, eb)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        rd.decompress(view, csize)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    kt = kernel_times(steps, lambda: w.compress(x, eb), lambda: rd.decompress(view, csize))
    assert rd.check() == 0, "reader rejected the stream"
    rec = {"compression_ratio": nbytes / csize, "compress_gbs": nbytes * steps / (t1 - t0) / 1e9,
           "decompress_gbs": nbytes * steps / (t2 - t1) / 1e9,
           "compress_kernels_ms": kt["compress"][0] / max(1, kt["compress"][1]),
           "decode_kernels_ms": kt["decode"][0] / max(1, kt["decode"][1])}
    if parity:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        s_gpu = view.cpu().numpy().tobytes()
        s_ora = O.compress_cuszx(xh, eb, 128)
        rec["stream_identical"] = s_gpu == s_ora
        d_ora = O.decompress_cuszx(s_ora, n)
        rec["decode_identical"] = bool(np.array_equal(rd.dec.cpu().numpy().view(np.uint32), d_ora.view(np.uint32)))
    log(f"cuSZx stream: {rec}")
    del w, rd, view
    torch.cuda.empty_cache()
    return rec


def _extents(r, n, nd):
    """Random 1-D/2-D/3-D extents whose product is close to n (at most n)."""
    import numpy as np
    if nd == 1 or n < 8:
        return (n,)
    if nd == 2:
        a = int(min(n, max(1, round(n ** 0.5 * 2 ** r.uniform(-1, 1)))))
        return (a, max(1, n // a))
    a = int(min(n, max(1, round(n ** (1 / 3) * 2 ** r.uniform(-1, 1)))))
    rest = max(1, n // a)
    b = int(min(rest, max(1, round(rest ** 0.5 * 2 ** r.uniform(-0.5, 0.5)))))
    return (a, b, max(1, rest // b))


def batched_eval(n_fields, dev, reps, single_gbs, parity=True):
    """Many fields in one launch sequence per direction, captured in a HIP
    graph, against one launch sequence per field on the same buffers.

    Fields: sizes log-uniform in [64 KiB, 256 MiB], 60% fp32 / 40% fp64,
    random 1-D/2-D/3-D extents, each a 32-wave sinusoid field (k <= 16,
    |k|^-2, noise 1e-3 std, a zero box of 25%) with its own seed, compressed at
    its own REL ratio (log-uniform in [1e-5, 1e-2]), blockSize 128."""
    import numpy as np
    import torch

    import szx_amd as Z
    from szx_amd.fields import sinusoid_field_torch

    r = np.random.default_rng(4)
    fields, total, n32, n64 = [], 0, 0, 0
    t0 = time.perf_counter()
    for i in range(n_fields):
        w = 4 if r.random() < 0.6 else 8
        nb = float(np.exp(r.uniform(np.log(64 * 2 ** 10), np.log(256 * 2 ** 20))))
        shape = _extents(r, max(1, int(nb) // w), int(r.integers(1, 4)))
        rel = float(np.float32(np.exp(r.uniform(np.log(1e-5), np.log(1e-2)))))
        x = sinusoid_field_torch(shape, seed=1000 + i, n_waves=32, kmax=16, slope=2.0, noise=1e-3,
                                 n_boxes=1, box_frac=0.25, box_value="zero",
                                 dtype="float32" if w == 4 else "float64", device=dev).reshape(-1)
        fields.append((x, Z.REL, rel))
        total += x.numel() * w
        n32 += w == 4
        n64 += w == 8
        if (i + 1) % 100 == 0:
            log(f"  generated {i + 1} fields, {total / 2**30:.2f} GiB")
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    log(f"batch: {n_fields} fields ({n32} fp32, {n64} fp64), {total / 2**30:.2f} GiB, "
        f"generated in {gen_s:.1f} s")

    bc = Z.BatchCodec(fields, 128)
    shapes = {}
    for lv, ipt in ((16, 4), (8, 4), (8, 2), (8, 1), (4, 2), (4, 1)):
        bc.set_shape(lv, ipt)
        bc.compress()
        torch.cuda.synchronize()
        t = time.perf_counter()
        bc.compress()
        torch.cuda.synchronize()
        shapes[f"{lv}x{ipt}"] = (time.perf_counter() - t) * 1e3
    lv, ipt, tune_ms = bc.tune(reps=2)
    log(f"tuner: {lv} values per lane, {ipt} index tiles per tile, {tune_ms:.3f} ms")

    def timed(fn, k):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    bc.compress()
    bc.decompress()
    torch.cuda.synchronize()
    Z.timing_begin()
    bc.compress()
    bc.decompress()
    kt = Z.timing_end()
    gc, gd = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    with torch.cuda.graph(gc):
        bc.compress()
    with torch.cuda.graph(gd):
        bc.decompress()
    ms_c = timed(gc.replay, reps)
    res_b = bc.results()
    assert not res_b[:, 3].any(), "a bounded wait between workgroups ran out"
    ms_d = timed(gd.replay, reps)
    st = bc.status().cpu().numpy()
    assert not st.any(), f"decode status {st[st != 0][:4]}"
    comp_bytes = int(res_b[:, 1].sum())

    # a sample of streams, to compare with the per-field launches' output and
    # with the oracle
    sample = r.choice(n_fields, size=min(n_fields, 64), replace=False)
    kept = {int(i): bc.stream(int(i), res_b).clone() for i in sample}
    kept_dec = {int(i): bc.decs[int(i)].cpu().numpy() for i in sample}   # DECODE_REFERENCE

    # every field within its bound (exact decode: for fp64 the reference
    # decoder misreads one byte of some records, which DECODE_REFERENCE keeps)
    bc.decompress(flags=Z.DECODE_EXACT)
    # A constant block decodes to its median rounded to fp32 (the stream's
    # format), so |x - x'| may pass eb by that rounding: about one fp32 ulp of
    # max(|median|, block range).  The same bytes come out of the oracle.
    ebs = np.array(res_b[:, 2], np.uint32).view(np.float32)
    worst, over, within_ulp = 0.0, 0, True
    over_idx = []
    for i, (x, _, _) in enumerate(fields):
        d = bc.decs[i]
        e = (d.double() - x.double()).abs()
        eb = float(ebs[i])
        m = float(e.max())
        worst = max(worst, m / eb if eb > 0 else (0.0 if m == 0 else float("inf")))
        if m > eb:
            over += 1
            over_idx.append(i)
            a = torch.clamp(d.double().abs(), min=2 * eb).float()
            ulp = (torch.nextafter(a, torch.full_like(a, float("inf"))) - a).double()
            within_ulp &= bool((e <= eb + ulp).all())
    log(f"max |x - x'| / eb over the batch: {worst:.7f} ({over} fields above eb, "
        f"all within one fp32 ulp of the median: {within_ulp})")

    # the sampled fields against the oracle (eb, stream bytes, decoded bytes),
    # and every field above eb against the reference's own decoder: the same
    # bytes, so the same elements above eb
    oracle_cmp = None
    if parity:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        log(f"batch parity: {len(sample)} sampled fields and {len(over_idx)} fields above eb vs the oracle")
        cnt = {"eb_identical": 0, "stream_identical": 0, "decode_identical": 0}
        for i in sorted(kept):
            x, _, rel = fields[i]
            xh = x.cpu().numpy()
            eb_o = O.rel_eb(xh, rel)
            s_ora = O.compress(xh, eb_o, 128)
            d_ora = O.decompress(s_ora, xh.size, xh.dtype)
            cnt["eb_identical"] += int(np.float32(eb_o).view(np.uint32) == np.uint32(res_b[i, 2]))
            cnt["stream_identical"] += int(kept[i].cpu().numpy().tobytes() == s_ora)
            cnt["decode_identical"] += int(kept_dec[i].tobytes() == d_ora.tobytes())
        bc.decompress(flags=Z.DECODE_REFERENCE)
        above = {"fields": len(over_idx), "decode_identical_to_reference_decoder": 0,
                 "same_elements_above_eb": 0, "elements_above_eb": 0,
                 "decoder": "oracle/_ref (the reference compiled from its sources)" if O.ref_available()
                 else "oracle"}
        for i in over_idx:
            x = fields[i][0]
            xh = x.cpu().numpy()
            s_i = bc.stream(i, res_b).cpu().numpy().tobytes()
            d_ref = (O.ref_decompress(s_i, xh.size, xh.dtype) if O.ref_available()
                     else O.decompress(s_i, xh.size, xh.dtype))
            d_i = bc.decs[i].cpu().numpy()
            eb = float(ebs[i])
            m_gpu = np.abs(d_i.astype(np.float64) - xh) > eb
            m_ref = np.abs(d_ref.astype(np.float64) - xh) > eb
            above["decode_identical_to_reference_decoder"] += int(d_i.tobytes() == d_ref.tobytes())
            above["same_elements_above_eb"] += int(np.array_equal(m_gpu, m_ref))
            above["elements_above_eb"] += int(m_gpu.sum())
        oracle_cmp = {"sampled_fields": len(kept), **cnt, "fields_above_eb": above}
        oracle_cmp["all_identical"] = (cnt["eb_identical"] == cnt["stream_identical"] == cnt["decode_identical"]
                                       == len(kept)
                                       and above["decode_identical_to_reference_decoder"] == len(over_idx)
                                       and above["same_elements_above_eb"] == len(over_idx))
        log(f"batch parity {oracle_cmp}")

    ms_pc = timed(lambda: bc.compress(mode=Z.BATCH_PER_FIELD), reps)
    res_p = bc.results()
    same = bool((res_p == res_b).all()) and all(torch.equal(bc.stream(i, res_p), s) for i, s in kept.items())
    ms_pd = timed(lambda: bc.decompress(mode=Z.BATCH_PER_FIELD), reps)
    st = bc.status().cpu().numpy()
    assert not st.any(), "per-field decode status"
    gbs = lambda ms: total / (ms * 1e-3) / 1e9
    out = {
        "fields": n_fields, "fp32_fields": n32, "fp64_fields": n64, "input_bytes": total,
        "compressed_bytes": comp_bytes, "compression_ratio": total / comp_bytes,
        "block_size": 128, "rel_range": [1e-5, 1e-2], "graph_captured": True,
        "tuned_shape": {"values_per_lane": lv, "index_tiles_per_tile": ipt, "ms": tune_ms},
        "shape_ms_single_pass": shapes,
        "compress_ms": ms_c, "compress_gbs": gbs(ms_c),
        "decompress_ms": ms_d, "decompress_gbs": gbs(ms_d),
        "compress_kernel_ms": kt["compress"][0], "decode_kernel_ms": kt["decode"][0],
        "per_field_compress_ms": ms_pc, "per_field_compress_gbs": gbs(ms_pc),
        "per_field_decompress_ms": ms_pd, "per_field_decompress_gbs": gbs(ms_pd),
        "speedup_compress": ms_pc / ms_c, "speedup_decompress": ms_pd / ms_d,
        "single_field_compress_gbs": single_gbs,
        "single_over_batched_compress": single_gbs / gbs(ms_c) if single_gbs else None,
        "per_field_streams_identical": same,
        "oracle_parity": oracle_cmp,
        "bound_check": {"max_err_over_eb": worst, "fields_above_eb": over,
                        "all_within_eb_plus_fp32_ulp_of_median": within_ulp,
                        "decode": "exact (fp64: DECODE_EXACT)",
                        "note": "constant blocks decode to the fp32-rounded block median, "
                                "as the reference's streams do"},
        "generation_s": gen_s,
    }
    bc.close()
    del fields, bc, kept
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rel", type=float, default=1e-3)
    ap.add_argument("--shape", default="512,512,512")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-noindex", action="store_true", help="skip the index-free decode")
    ap.add_argument("--fxr-fields", type=int, default=64,
                    help="fixed-rate evaluation fields (0 skips it)")
    ap.add_argument("--no-cfg1", action="store_true",
                    help="skip the fp32 100x500x500 field (GPU, parity and reference CPU rows)")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the fp64 8192^2 field (REL 1e-4/1e-6/1e-8 x blockSize 64/128/256)")
    ap.add_argument("--batch-fields", type=int, default=1000,
                    help="fields of the batched-launch evaluation (0 skips it)")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-nb", action="store_true",
                    help="skip the non-blocked codec (fastMode 1) leg")
    ap.add_argument("--no-cuszx", action="store_true",
                    help="skip the cuSZx stream (reference CUDA codec format) leg")
    ap.add_argument("--no-ra", action="store_true",
                    help="skip the random-access container (fastMode 3/4) leg")
    ap.add_argument("--generic-bs", default="100,255,7",
                    help="block sizes outside the tiled kernels to time on the same field "
                         "('' to skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run 3 plain elementwise copies of the input (y = x * 1, "
                         "N*4 bytes read and written) to calibrate HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the codec has no multi-GPU path "
              "(a field would shard by contiguous block ranges with one gather of "
              "per-shard totals).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import szx_amd as Z
    from szx_amd.fields import sinusoid_field_torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    shape = tuple(int(v) for v in args.shape.split(","))
    log(f"generating field {shape} fp32 on {torch.cuda.get_device_name(0)}")
    x = sinusoid_field_torch(shape, seed=2, n_waves=64, kmax=32, slope=2.5, noise=1e-4,
                             n_boxes=2, box_frac=0.30, box_value="mean", device=dev)
    x = x.contiguous().view(-1)
    n = x.numel()
    nbytes = n * 4
    torch.cuda.synchronize()

    comp = Z.DeviceCompressor(n, np.float32, 128, device=dev)
    comp(x, Z.REL, relBoundRatio=args.rel)
    view, off, csize, eb = comp.stream_view()
    dec = Z.DeviceDecompressor(n, np.float32, 128, comp.capacity, device=dev)
    log(f"stream {csize} bytes (CR {nbytes / csize:.3f}), eb {eb:.6g}")

    def timed(fn, steps):
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps):
            fn()
        t1.record()
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps

    rel_step = lambda: comp(x, Z.REL, relBoundRatio=args.rel)
    abs_step = lambda: comp(x, Z.ABS, absErrBound=eb)
    dec_step = lambda: dec(view, csize, Z.DECODE_REFERENCE, index=comp.index)

    for _ in range(args.warmup):
        rel_step()
        dec_step()
    torch.cuda.synchronize()

    # Wall-clock steps run with the library's kernel timer off: its events
    # sit between the kernels of a call and add a few microseconds each.
    # The kernel durations come from a separate timed pass.
    log(f"timing {args.steps} REL compress steps")
    ms_rel = timed(rel_step, args.steps)
    Z.timing_begin()
    timed(rel_step, args.steps)
    kt = Z.timing_end()
    k_ms = kt["compress"][0] / max(1, kt["compress"][1])
    _, _, csize2, _ = comp.stream_view()
    assert csize2 == csize

    log(f"timing {args.steps} ABS compress steps")
    ms_abs = timed(abs_step, args.steps)

    log(f"timing {args.steps} decompress steps (with record index)")
    ms_dec = timed(dec_step, args.steps)
    Z.timing_begin()
    timed(dec_step, args.steps)
    kd = Z.timing_end()
    kd_ms = kd["decode"][0] / max(1, kd["decode"][1])
    assert dec.check() == 0, "decoder rejected the stream"

    # measured copy bandwidth of this GPU (device-to-device copy of the input:
    # N*W bytes read + N*W written), beside the 8 TB/s peak
    ycopy = torch.empty_like(x)
    ycopy.copy_(x)
    ms_copy = timed(lambda: ycopy.copy_(x), 10)
    copy_gbs = 2 * nbytes / (ms_copy * 1e-3) / 1e9
    # the library's own copy kernel (non-temporal 16-byte loads and
    # stores), timed by the kernel timer: the ceiling the codec kernels share
    Z.copy_dev(x, ycopy)
    Z.timing_begin()
    timed(lambda: Z.copy_dev(x, ycopy), 10)
    kcp = Z.timing_end()
    copy_kernel_ms = kcp["copy"][0] / max(1, kcp["copy"][1])
    copy_kernel_gbs = 2 * nbytes / (copy_kernel_ms * 1e-3) / 1e9
    assert torch.equal(ycopy, x)
    del ycopy
    log(f"device copy {copy_gbs:.0f} GB/s (torch), {copy_kernel_gbs:.0f} GB/s (copy kernel; read + write)")

    if args.calib_copy:
        log("calibration: 3 plain elementwise copies of the input")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        del y

    ms_dec_noidx = None
    ix_ms = None
    if not args.no_noindex:
        log("timing 3 decompresses without the index (index rebuilt from the stream)")
        dec(view, csize, Z.DECODE_REFERENCE)   # warm-up
        ms_dec_noidx = timed(lambda: dec(view, csize, Z.DECODE_REFERENCE), 3)
        Z.timing_begin()
        timed(lambda: dec(view, csize, Z.DECODE_REFERENCE), 3)
        ki = Z.timing_end()
        ix_ms = ki["index"][0] / max(1, ki["index"][1])
        assert dec.check() == 0

    parity = None
    xh = None
    if not args.no_parity or not args.no_cpu:
        xh = x.cpu().numpy()
    if not args.no_parity:
        log("parity check against the oracle at full size")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        def parity_at(rel):
            comp(x, Z.REL, relBoundRatio=rel)
            v, _, cs, e = comp.stream_view()
            s_gpu = v.cpu().numpy().tobytes()
            eb_o = O.rel_eb(xh, rel)
            s_ora = O.compress(xh, eb_o, 128)
            dec(v, cs, Z.DECODE_REFERENCE, index=comp.index)
            d_gpu = dec.out.cpu().numpy()
            d_ora = O.decompress(s_ora, n, np.float32)
            dec(v, cs, Z.DECODE_REFERENCE)
            d_gpu2 = dec.out.cpu().numpy()
            assert dec.check() == 0
            rec = {"eb_identical": bool(np.float32(eb_o) == np.float32(e)),
                   "stream_identical": s_gpu == s_ora,
                   "decode_identical": bool(np.array_equal(d_gpu.view(np.uint32), d_ora.view(np.uint32))),
                   "decode_noindex_identical": bool(np.array_equal(d_gpu2.view(np.uint32),
                                                                   d_ora.view(np.uint32))),
                   "max_abs_err": float(np.max(np.abs(d_gpu.astype(np.float64) - xh))),
                   "eb": e, "compressed_bytes": cs}
            rec["err_within_eb"] = rec["max_abs_err"] <= e
            log(f"parity REL {rel:g}: {rec}")
            return rec

        # the other REL ratios first: the headline's stream stays in comp's
        # buffer for the legs below
        others = {f"rel_{r:g}": parity_at(r) for r in (1e-2, 1e-3, 1e-4) if r != args.rel}
        parity = parity_at(args.rel)
        assert comp.stream_view()[2] == csize
        parity["other_rel"] = others
        parity["all_identical"] = all(
            r["eb_identical"] and r["stream_identical"] and r["decode_identical"]
            and r["decode_noindex_identical"] for r in [parity, *others.values()])

    cpu = None
    if not args.no_cpu:
        sample = min(n, 1 << 25)
        log(f"CPU baseline on {sample} values")
        kind, res = cpu_baseline(xh, eb, sample)
        ref = res.get("openmp", res["serial"])
        cpu = {"value": ref["compress_gbs"], "unit": "GB/s", "cores": ref["cores"],
               "kind": kind,
               "sample": f"first {sample} values ({sample * 4 / 2**20:.0f} MiB) of the same field, "
                         f"ABS eb {eb:.6g}, blockSize 128, best of 3 after 1 warm-up; "
                         + ("reference fastMode-4 OpenMP compress" if "openmp" in res
                            else "oracle serial compress"),
               "host": host_cpu_info(),
               "serial_fastmode2": res["serial"], "openmp_fastmode4": res.get("openmp"),
               "other_modes": res.get("other_modes")}

    host_api = None
    if xh is not None:
        # the drop-in entry points take host arrays: PCIe-inclusive rate
        log("host API (SZx_fast_compress_args / SZx_fast_decompress, host buffers)")
        xs = xh.reshape(shape)
        Z.SZx_fast_compress_args(Z.SZx_WITH_BLOCK_FAST_CMPR, xs, Z.REL, relBoundRatio=args.rel)
        t0 = time.perf_counter()
        hs = Z.SZx_fast_compress_args(Z.SZx_WITH_BLOCK_FAST_CMPR, xs, Z.REL, relBoundRatio=args.rel)
        t1 = time.perf_counter()
        Z.SZx_fast_decompress(Z.SZx_WITH_BLOCK_FAST_CMPR, np.float32, hs, shape)
        t2 = time.perf_counter()
        host_api = {"compress_gbs": nbytes / (t1 - t0) / 1e9,
                    "decompress_gbs": nbytes / (t2 - t1) / 1e9,
                    "note": "one call each, pageable host buffers, H2D + kernels + D2H + malloc"}
        log(f"host API {host_api}")

    cfg1 = None
    if not args.no_cfg1:
        log("fp32 100x500x500 field, ABS 1e-3 x range (GPU, parity, reference CPU codec)")
        cfg1 = cfg1_eval(dev, max(2, args.steps // 4), parity=not args.no_parity, cpu=not args.no_cpu)

    fixed = None
    if args.fxr_fields > 0:
        log(f"fixed-rate evaluation on {args.fxr_fields} fields")
        fixed = fixed_rate_eval(args.fxr_fields, dev)
        log(f"fixed rate {fixed}")

    fp64 = None
    if not args.no_fp64:
        log("fp64 8192^2 field, 9 configurations")
        fp64 = fp64_eval(dev, max(2, args.steps // 4), parity=not args.no_parity)

    ra = None
    if not args.no_ra:
        log("random-access container (fastMode 3 and 4)")
        ra = random_access_eval(x, xh, args.rel, dev, max(2, args.steps // 4),
                                parity=not args.no_parity and xh is not None)

    nb = None
    if not args.no_nb:
        log("non-blocked codec (fastMode 1)")
        nb = non_blocked_eval(x, xh, args.rel, dev, max(2, args.steps // 4),
                              parity=not args.no_parity and xh is not None)

    cz = None
    if not args.no_cuszx:
        log("cuSZx stream (writer and reader)")
        cz = cuszx_eval(x, xh, eb, dev, max(2, args.steps // 4),
                        parity=not args.no_parity and xh is not None)

    generic = None
    gsizes = [int(v) for v in args.generic_bs.split(",") if v.strip()]
    if gsizes:
        log(f"block sizes {gsizes} (per-block segment kernels)")
        generic = generic_eval(x, xh, eb, dev, max(2, args.steps // 4), gsizes,
                               parity=not args.no_parity and xh is not None)

    batched = None
    if args.batch_fields > 0:
        log(f"batched-launch evaluation on {args.batch_fields} fields")
        del view, comp, dec
        torch.cuda.empty_cache()
        batched = batched_eval(args.batch_fields, dev, args.batch_reps,
                               nbytes / (ms_rel * 1e-3) / 1e9, parity=not args.no_parity)
        log(f"batched {batched}")

    alg_c = nbytes + csize
    roof = {"bound": "hbm", "kernel": "k_compress", "achieved": alg_c / (k_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": measured_traffic(shape, args.rel),
            "algorithmic_bytes": alg_c, "kernel_ms": k_ms}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["measured_copy_gbs"] = copy_gbs
    roof["frac_of_copy"] = roof["achieved"] / copy_gbs
    roof["copy_kernel_gbs"] = copy_kernel_gbs
    roof["frac_of_copy_kernel"] = roof["achieved"] / copy_kernel_gbs
    line = {
        "metric": METRIC,
        "value": nbytes / (ms_rel * 1e-3) / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_rel,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{'x'.join(map(str, shape))} fp32 sinusoid field (64 waves, k<=32, "
                               f"|k|^-2.5, noise 1e-4 std, 2 mean boxes 30%), REL {args.rel:g}, "
                               f"blockSize 128, fastMode 2", "values": n, "rel": args.rel,
                   "block_size": 128},
        "roofline": roof,
        "cpu_baseline": cpu,
        "compress_rel_gbs": nbytes / (ms_rel * 1e-3) / 1e9,
        "compress_abs_gbs": nbytes / (ms_abs * 1e-3) / 1e9,
        "compress_pct_hbm": alg_c / (ms_abs * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "decompress_gbs": nbytes / (ms_dec * 1e-3) / 1e9,
        "decompress_pct_hbm": (nbytes + csize) / (ms_dec * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "decode_kernel_ms": kd_ms,
        "decode_kernel_pct_hbm": (nbytes + csize) / (kd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "decode_kernel_frac_of_copy": (nbytes + csize) / (kd_ms * 1e-3) / 1e9 / copy_gbs,
        "decode_kernel_frac_of_copy_kernel": (nbytes + csize) / (kd_ms * 1e-3) / 1e9 / copy_kernel_gbs,
        "copy_kernel_gbs": copy_kernel_gbs,
        "copy_kernel_pct_hbm": copy_kernel_gbs / HBM_PEAK_GBS,
        "decompress_noindex_gbs": (nbytes / (ms_dec_noidx * 1e-3) / 1e9) if ms_dec_noidx else None,
        "index_build_ms": ix_ms,
        "compressed_bytes": csize,
        "compression_ratio": nbytes / csize,
        "eb": eb,
        "parity": parity,
        "fixed_rate": fixed,
        "cfg1_100x500x500": cfg1,
        "fp64_grf_8192sq": fp64,
        "batched": batched,
        "generic_block_sizes": generic,
        "random_access": ra,
        "non_blocked": nb,
        "cuszx": cz,
        "host_api_pcie_inclusive": host_api,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
