"""GPU parity of the random-access container (fastMode 3 and 4) against the
reference's own streams (tests/golden/ra*.npz, written by the compiled
reference) and the oracle, through the C ABI."""
import glob
import os

import numpy as np
import pytest

import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu

GOLDEN_RA = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ra*.npz")))


def say(*a):
    print(*a, flush=True)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def bits_equal(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(
        a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("path", GOLDEN_RA, ids=[os.path.basename(p)[:-4] for p in GOLDEN_RA])
def test_golden_random_access_host_api(path, device):
    g = load(path)
    x, fm = g["x"], int(g["fast_mode"])
    name = os.path.basename(path)
    say("case", name, x.dtype, x.size, fm)
    if "rel" in name:
        s = Z.SZx_fast_compress_args(fm, x, Z.REL, relBoundRatio=float(g["rel"]))
    else:
        s = Z.SZx_fast_compress_args(fm, x, Z.ABS, absErrBound=float(g["eb"]))
    assert s == g["stream"].tobytes()
    d = Z.SZx_fast_decompress(fm, x.dtype, s, x.shape)
    assert bits_equal(d, g["decoded"])


def _field(r, n, dtype):
    t = np.arange(n)
    x = (np.sin(t * r.uniform(1e-4, 0.05)) * 10 ** r.uniform(-3, 3)
         + r.standard_normal(n) * 10 ** r.uniform(-8, -2))
    x[: n // 4] = 0.75
    return x.astype(dtype)


@pytest.mark.parametrize("fm", [3, 4])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_access_device_api(fm, dtype, device):
    import torch
    r = np.random.default_rng(300 + 10 * fm + (dtype == np.float64))
    for bs in (8, 64, 128, 256):
        n = int(r.integers(5000, 120000))
        x = _field(r, n, dtype)
        eb = float(np.float32(np.ptp(x) * 10.0 ** r.uniform(-7, -2)))
        say("ra device", fm, dtype.__name__, bs, n, eb)
        comp = Z.DeviceCompressor(n, dtype, bs, fastMode=fm)
        comp(torch.from_numpy(x).to(device), Z.ABS, eb)
        view, off, size, _ = comp.stream_view()
        s = view.cpu().numpy().tobytes()
        assert s == O.compress_ra(x, eb, bs, fm)
        want = O.decompress_ra(s, n, dtype, fm, ref_quirk=True)
        dec = Z.DeviceDecompressor(n, dtype, bs, size, fastMode=fm)
        dec(view, size, Z.DECODE_REFERENCE, index=comp.index)
        assert dec.check() == 0
        assert bits_equal(dec.out.cpu().numpy(), want)
        dec(view, size, Z.DECODE_REFERENCE)          # index from the size table
        assert dec.check() == 0
        assert bits_equal(dec.out.cpu().numpy(), want)
        if dtype == np.float64:
            dec(view, size, Z.DECODE_EXACT)
            assert dec.check() == 0
            assert bits_equal(dec.out.cpu().numpy(), O.decompress_ra(s, n, dtype, fm, ref_quirk=False))
    # REL through the device API
    x = _field(r, 77777, dtype)
    comp = Z.DeviceCompressor(x.size, dtype, 128, fastMode=fm)
    comp(torch.from_numpy(x).to(device), Z.REL, 0.0, 1e-4)
    view, off, size, eb = comp.stream_view()
    assert np.float32(eb) == np.float32(O.rel_eb(x, 1e-4))
    assert view.cpu().numpy().tobytes() == O.compress_ra(x, eb, 128, fm)


def test_random_access_rejects_damage(device):
    import torch
    x = _field(np.random.default_rng(9), 100000, np.float32)
    s = bytearray(O.compress_ra(x, 1e-5, 128, 3))
    dec = Z.DeviceDecompressor(x.size, np.float32, 128, len(s), fastMode=3)
    say("ra truncated")
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    dec(st[: len(s) - 7], len(s) - 7, Z.DECODE_REFERENCE)
    assert dec.check() != 0
    say("ra size table off by one")
    s[20] ^= 1                                    # first record's size
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    dec(st, len(s), Z.DECODE_REFERENCE)
    assert dec.check() != 0
    say("ra block size field")
    s[20] ^= 1
    s[11] = 64                                    # big-endian u64 blockSize -> 64
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    dec(st, len(s), Z.DECODE_REFERENCE)
    assert dec.check() != 0
    with pytest.raises(Z.SZxError):
        Z.DeviceCompressor(1000, np.float32, 100, fastMode=3)(torch.zeros(1000, device=device), Z.ABS, 1e-3)
