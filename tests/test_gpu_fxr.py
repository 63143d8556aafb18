"""GPU parity of the fixed-rate (FXR) entry points and the value range.

The bound the GPU path picks must equal the oracle's restatement of the
reference's search bit for bit (the oracle's search and its whole FXR
compress path are pinned against the compiled reference in
test_oracle.py), and the FXR stream must be the oracle's stream at that bound.
"""
import numpy as np
import pytest

import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _field(r, n, dtype):
    t = np.arange(n)
    x = (np.sin(t * r.uniform(1e-4, 1e-2)) * 10 ** r.uniform(-2, 3)
         + r.standard_normal(n) * 10 ** r.uniform(-7, -3))
    if r.random() < 0.5:   # a constant stretch
        a = int(r.integers(0, n // 2))
        x[a:a + n // 4] = 0.0
    return x.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fxr_bound_device(dtype, device):
    import torch
    r = np.random.default_rng(40 if dtype == np.float32 else 41)
    for trial in range(8):
        n = int(r.integers(20000, 400000))
        x = _field(r, n, dtype)
        fb = Z.DeviceFixedRateBound(n, dtype, 128)
        xt = torch.from_numpy(x).to(device)
        for target, tol in ((3.0, 0.1), (8.0, 0.05), (20.0, 0.02)):
            say("fxr bound", trial, dtype.__name__, n, target, tol)
            eb, ne = fb(xt, target, tol)
            want, ne_want = O.fxr_error_bound(x, target, tol, 128)
            assert np.float32(eb) == np.float32(want), (trial, target, eb, want)
            assert ne == ne_want


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fxr_compress_args(dtype, device):
    r = np.random.default_rng(42)
    for trial in range(6):
        n = int(r.integers(30000, 200000))
        x = _field(r, n, dtype)
        for target in (4.0, 16.0):
            say("fxr compress", trial, dtype.__name__, n, target)
            s = Z.SZx_fast_compress_args(Z.SZx_WITH_BLOCK_FAST_CMPR, x, Z.FXR,
                                         compressionRatio=target, tolerance=0.1)
            eb, _ = O.fxr_error_bound(x, target, 0.1, 128)
            assert s == O.compress(x, eb, 128), (trial, target)


@pytest.mark.parametrize("bs", [128, 64, 100, 256, 7])
def test_estimate_from_init(bs, device):
    r = np.random.default_rng(bs)
    for dtype in (np.float32, np.float64):
        n = int(r.integers(50000, 150000))
        x = _field(r, n, dtype)
        init = float(np.float32(np.ptp(x) * 10 ** r.uniform(-5, -1)))
        say("estimate", bs, dtype.__name__, n, init)
        for target in (5.0, 10.0):
            eb = Z.SZx_estimateErrorBoundbasedonCR(x, target, 0.05, init, bs)
            want, _ = O.fxr_search_from(x, target, 0.05, init, bs)
            assert np.float32(eb) == np.float32(want), (bs, target)


def test_fxr_tiny_inputs(device):
    # fewer values than one sample stride: no sampled block at all
    for n in (1, 100, 1279):
        x = np.linspace(0, 1, n).astype(np.float32)
        say("tiny", n)
        s = Z.SZx_fast_compress_args(Z.SZx_WITH_BLOCK_FAST_CMPR, x, Z.FXR,
                                     compressionRatio=4.0, tolerance=0.1)
        eb, _ = O.fxr_error_bound(x, 4.0, 0.1, 128)
        assert s == O.compress(x, eb, 128)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_value_range(dtype, device):
    r = np.random.default_rng(3)
    for n in (1, 2, 1000, 300001):
        x = (r.standard_normal(n) * 1e3).astype(dtype)
        say("range", dtype.__name__, n)
        rng, rad, med = Z.SZx_computeValueRange(x)
        lo, hi = x.min(), x.max()
        want = hi - lo                                  # in T
        assert rng == float(want)
        want_rad = np.float32(want / dtype(2))          # szx_float.cc:494-495
        want_med = np.float32(lo + dtype(want_rad))
        assert np.float32(rad) == want_rad and np.float32(med) == want_med
    x = np.array([1.0, np.nan, -3.0, 7.0], dtype)      # NaNs after the first are skipped
    assert Z.SZx_computeValueRange(x)[0] == 10.0
    x = np.array([np.nan, 1.0, 2.0], dtype)            # a leading NaN poisons the range
    assert np.isnan(Z.SZx_computeValueRange(x)[0])
