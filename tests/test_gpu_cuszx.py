"""GPU parity of the cuSZx stream reader (the reference's CUDA codec format)
against the oracle's restatement of the reference's decoder, through the C
ABI: streams the oracle writes for every block size the reader takes,
hand-built and randomly built records that reach the decoder's corners, and
the malformed-stream statuses."""
import struct

import numpy as np
import pytest

import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu

BLOCK_SIZES = [32, 64, 96, 128, 160, 192, 224]
STATUS = {-6: -42, -7: -45}   # oracle code -> device status


def say(*a):
    print(*a, flush=True)


def same(a, b):
    """Bit-identical, NaNs matched by position."""
    na, nb_ = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb_):
        return False
    return np.array_equal(a[~na].view(np.uint32), b[~nb_].view(np.uint32))


def field(r, n, kind):
    t = np.arange(n)
    if kind == 0:
        x = np.sin(t * r.uniform(1e-4, 0.05)) * 10 ** r.uniform(-2, 2)
    elif kind == 1:
        x = np.cumsum(r.standard_normal(n)) * 10 ** r.uniform(-3, 1)
    else:
        x = np.where(r.random(n) < 0.6, 0.25, r.standard_normal(n))
    return x.astype(np.float32)


def device_read(stream: bytes, n: int, offset: int = 0):
    import torch
    rd = Z.DeviceCuszxReader(n)
    s = torch.frombuffer(bytearray(stream), dtype=torch.uint8).cuda()
    out = torch.full((n + offset,), float("nan"), dtype=torch.float32, device="cuda")
    rd.decompress(s, len(stream), out=out[offset:])
    torch.cuda.synchronize()
    return rd.check(), out[offset:].cpu().numpy()


@pytest.mark.parametrize("bs", BLOCK_SIZES)
def test_oracle_streams_every_block_size(bs, device):
    r = np.random.default_rng(900 + bs)
    for trial in range(3):
        n = int(bs * r.integers(1, 3000) + r.integers(0, bs))
        x = field(r, n, trial)
        eb = float(10 ** r.uniform(-5, -1)) * float(np.ptp(x) or 1.0)
        try:
            s = O.compress_cuszx(x, eb, bs)
        except ValueError:
            continue   # a block beyond 32 significant bits
        say("case bs", bs, "n", n, "eb", eb, "bytes", len(s))
        want = O.decompress_cuszx(s, n)
        assert same(Z.cuszx_decompress(s, n), want)
        st, got = device_read(s, n)
        assert st == 0 and same(got, want)


def test_known_answer_stream_and_unaligned_output(device):
    from test_oracle_cuszx import KAT_VALUES, kat_stream
    s = kat_stream()
    want = np.array([4.5] * 32 + KAT_VALUES + [7.0, -2.5, 0.125], np.float32)
    say("case known answer")
    assert same(Z.cuszx_decompress(s, 67), want)
    for off in (0, 1, 3):
        st, got = device_read(s, 67, offset=off)
        assert st == 0 and same(got, want)


def test_stream_at_unaligned_device_addresses(device):
    """The reader stages records with aligned word loads from wherever the
    stream starts."""
    import torch
    r = np.random.default_rng(77)
    n = 128 * 500 + 9
    x = field(r, n, 1)
    s = O.compress_cuszx(x, 1e-3 * float(np.ptp(x)), 128)
    want = O.decompress_cuszx(s, n)
    rd = Z.DeviceCuszxReader(n)
    for off in (1, 2, 3, 5):
        say("case stream offset", off)
        buf = torch.zeros(len(s) + 8, dtype=torch.uint8, device="cuda")
        buf[off:off + len(s)] = torch.frombuffer(bytearray(s), dtype=torch.uint8).cuda()
        rd.dec.fill_(float("nan"))
        rd.decompress(buf[off:], len(s))
        torch.cuda.synchronize()
        assert rd.check() == 0 and same(rd.dec.cpu().numpy(), want)


def random_stream(r, bs, nb, rem):
    """A stream whose records are random bytes of random length in [0, 4 bs],
    with reqLength bytes drawn to reach every significant-byte count (0..5) and
    lead codes of every value."""
    state = r.random(nb) < 0.8
    nconst = int((~state).sum())
    bits = bytearray((nb + 7) // 8)
    for i in np.flatnonzero(state):
        bits[i // 8] |= 0x80 >> (i % 8)
    msize = 5 + bs // 4
    counts, recs = [], []
    for _ in range(int(state.sum())):
        leng = int(r.choice([r.integers(0, msize + 1), r.integers(msize, 4 * bs + 1), msize + bs]))
        rec = bytearray(r.integers(0, 256, leng, dtype=np.uint8).tobytes())
        if leng >= 4:
            rec[0:4] = struct.pack("<f", float(r.uniform(-4, 4)))
        if leng >= 5:
            rec[4] = int(r.choice([0, 3, 8, 10, 12, 16, 17, 24, 25, 31, 32, 33, 40]))
        counts.append(leng - msize)
        recs.append(bytes(rec))
    meds = r.uniform(-8, 8, nconst).astype(np.float32).tobytes()
    tail = r.uniform(-8, 8, rem).astype(np.float32).tobytes()
    return (bytes([1, 1, 1, 0, bs]) + struct.pack(">Q", nconst) + bytes(bits) + tail + meds
            + struct.pack(f"<{len(counts)}h", *counts) + b"".join(recs))


@pytest.mark.parametrize("bs", [32, 96, 128, 224])
def test_random_records_reach_the_same_corners(bs, device):
    """Random records: the device reproduces the reference decoder's values,
    or reports the same malformed block."""
    r = np.random.default_rng(1700 + bs)
    seen = set()
    for trial in range(24):
        nb, rem = int(r.integers(1, 40)), int(r.integers(0, bs))
        n = nb * bs + rem
        s = random_stream(r, bs, nb, rem)
        try:
            want, rc = O.decompress_cuszx(s, n), 0
        except ValueError as e:
            want, rc = None, int(str(e).split("rc=")[1].rstrip(")"))
        seen.add(rc)
        st, got = device_read(s, n)
        say("case random bs", bs, "trial", trial, "oracle", rc, "device", st)
        if rc == 0:
            assert st == 0 and same(got, want)
        else:
            assert st == STATUS[rc]
    assert 0 in seen


def test_malformed_headers(device):
    from test_oracle_cuszx import header, kat_stream
    s = kat_stream()
    say("case malformed")
    assert device_read(s[:20], 67)[0] == -41                       # sections past the end
    assert device_read(header(100, 0, []), 400)[0] == -41          # block size
    assert device_read(s[:5] + struct.pack(">Q", 5) + s[13:], 67)[0] == -41   # constants > blocks
    bad = bytearray(s)
    bad[13] = 0xC0                                                   # both blocks non-constant
    assert device_read(bytes(bad), 67)[0] == -43
    assert device_read(s[:-1], 67)[0] == -44                        # record cut short


def test_large_field(device):
    n = (1 << 24) + 77
    r = np.random.default_rng(5)
    x = (np.cumsum(r.standard_normal(n)) * 1e-3).astype(np.float32)
    x[1 << 20:1 << 21] = 3.0
    eb = 1e-3 * float(np.ptp(x))
    s = O.compress_cuszx(x, eb, 128)
    say("case large", n, len(s))
    want = O.decompress_cuszx(s, n)
    st, got = device_read(s, n)
    assert st == 0 and same(got, want)
    # a constant block (radius <= eb, cuszx_float.cu:278) decodes to its
    # float median, which is rounded: half an ulp of the field's magnitude
    slack = float(np.spacing(np.float32(np.abs(x).max())))
    assert np.abs(got.astype(np.float64) - x).max() <= eb + slack


def device_write(x, eb, bs):
    import torch
    w = Z.DeviceCuszxWriter(x.size, bs)
    w.compress(torch.from_numpy(x).cuda(), eb)
    torch.cuda.synchronize()
    view, size, bad = w.stream_view()
    return bad, view.cpu().numpy().tobytes()


@pytest.mark.parametrize("bs", BLOCK_SIZES)
def test_writer_matches_the_oracle_stream(bs, device):
    r = np.random.default_rng(2300 + bs)
    for trial in range(4):
        n = int(r.choice([bs - 1, bs, bs * int(r.integers(2, 4000)) + int(r.integers(0, bs))]))
        x = field(r, n, trial % 3)
        eb = float(10 ** r.uniform(-5, -1)) * float(np.ptp(x) or 1.0)
        try:
            want = O.compress_cuszx(x, eb, bs)
        except ValueError:
            continue
        say("case writer bs", bs, "n", n, "eb", eb)
        assert Z.cuszx_compress(x, eb, bs) == want
        bad, got = device_write(x, eb, bs)
        assert not bad and got == want
        assert same(Z.cuszx_decompress(got, n), O.decompress_cuszx(want, n))


def test_writer_equal_neighbours_and_refusal(device):
    x = np.array([(-1) ** i * (0.5 + 0.01 * i) for i in range(64)], np.float32)
    x[5] = x[4]
    x[6:] = -x[6:]
    say("case writer lead 4")
    assert Z.cuszx_compress(x, 1e-6, 32) == O.compress_cuszx(x, 1e-6, 32)
    # a range of 2^30 bounds needs more than 32 bits: no stream
    y = np.array([0.0, 1.0] * 64, np.float32)
    with pytest.raises(Z.SZxError, match="more than 32"):
        Z.cuszx_compress(y, 1e-9, 128)
    assert device_write(y, 1e-9, 128)[0]
