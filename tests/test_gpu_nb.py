"""GPU parity of the non-blocked codec (fastMode 1) against the reference's
own streams (tests/golden/nb*.npz, written by the compiled reference) and the
oracle, through the C ABI."""
import glob
import os

import numpy as np
import pytest

import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu

GOLDEN_NB = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "nb*.npz")))


def say(*a):
    print(*a, flush=True)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def bits_equal(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(
        a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("path", GOLDEN_NB, ids=[os.path.basename(p)[:-4] for p in GOLDEN_NB])
def test_golden_non_blocked_host_api(path, device):
    g = load(path)
    x = g["x"]
    name = os.path.basename(path)
    say("case", name, x.dtype, x.size)
    if "rel" in name:
        s = Z.SZx_fast_compress_args(Z.SZx_NO_BLOCK_FAST_CMPR, x, Z.REL, relBoundRatio=float(g["rel"]))
    else:
        s = Z.SZx_fast_compress_args(Z.SZx_NO_BLOCK_FAST_CMPR, x, Z.ABS, absErrBound=float(g["eb"]))
    assert s == g["stream"].tobytes()
    d = Z.SZx_fast_decompress(Z.SZx_NO_BLOCK_FAST_CMPR, x.dtype, s, x.shape)
    assert bits_equal(d, g["decoded"])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_non_blocked_device_api(dtype, device):
    import torch
    r = np.random.default_rng(400 + (dtype == np.float64))
    for trial in range(8):
        n = int(r.choice([1, 3, 4095, 4097, 65536 + 17, 300001]))
        t = np.arange(n)
        kind = trial % 3
        if kind == 0:
            x = np.sin(t * r.uniform(1e-4, 0.05)) * 10 ** r.uniform(-3, 3) + r.standard_normal(n) * 1e-6
        elif kind == 1:
            x = np.cumsum(r.standard_normal(n)) * 10 ** r.uniform(-6, 2)
        else:
            x = np.where(r.random(n) < 0.7, 0.5, r.standard_normal(n))   # long runs of one value
        x = x.astype(dtype)
        eb = float(np.float32((np.ptp(x) + 1e-30) * 10.0 ** r.uniform(-9, -2)))
        say("nb device", dtype.__name__, n, eb)
        c = Z.DeviceNonBlocked(n, dtype)
        c.compress(torch.from_numpy(x).to(device), Z.ABS, eb)
        view, size, _ = c.stream_view()
        s = view.cpu().numpy().tobytes()
        assert s == O.compress_nb(x, eb)
        c.decompress(view, size, Z.DECODE_REFERENCE)
        assert c.check() == 0
        assert bits_equal(c.dec.cpu().numpy(), O.decompress_nb(s, n, dtype, ref_quirk=True))
        if dtype == np.float64:
            c.decompress(view, size, Z.DECODE_EXACT)
            assert c.check() == 0
            assert bits_equal(c.dec.cpu().numpy(), O.decompress_nb(s, n, dtype, ref_quirk=False))
    # REL on the device
    x = np.sin(np.arange(123457) * 0.002).astype(dtype)
    c = Z.DeviceNonBlocked(x.size, dtype)
    c.compress(torch.from_numpy(x).to(device), Z.REL, 0.0, 1e-5)
    view, size, eb = c.stream_view()
    assert np.float32(eb) == np.float32(O.rel_eb(x, 1e-5))
    assert view.cpu().numpy().tobytes() == O.compress_nb(x, eb)


def test_non_blocked_rejects_damage(device):
    import torch
    x = np.cumsum(np.random.default_rng(5).standard_normal(50000)).astype(np.float32)
    s = bytearray(O.compress_nb(x, 1e-4))
    c = Z.DeviceNonBlocked(x.size, np.float32)
    say("nb truncated")
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    c.decompress(st, len(s) - 10, Z.DECODE_REFERENCE)
    assert c.check() != 0
    say("nb lead-array size")
    s[16] ^= 1
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    c.decompress(st, len(s), Z.DECODE_REFERENCE)
    assert c.check() != 0
    say("nb reqLength 0")
    s[16] ^= 1
    s[4] = 0
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    c.decompress(st, len(s), Z.DECODE_REFERENCE)
    assert c.check() != 0
