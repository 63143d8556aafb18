"""CPU tests that pin the oracle (oracle/szx_oracle.c) before anything is
compared against it.

Pins, strongest first:
  * golden vectors written by the compiled reference (tests/golden/*.npz);
  * the compiled reference itself (oracle/_ref), function by function, when
    it is present (it is built from /root/reference in this container);
  * known-answer streams worked out by hand, and invariants of the format
    (all-constant fields, the size formula, the fp32 error bound).
"""
import ctypes as C
import glob
import os
import struct

import numpy as np
import pytest

import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "f*.npz")))   # fastMode 2
GOLDEN_RA = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ra*.npz")))  # fastMode 3/4
GOLDEN_NB = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "nb*.npz")))  # fastMode 1
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built")


def load_golden(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_golden_present():
    assert len(GOLDEN) >= 30


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    g = load_golden(path)
    x, bs, eb = g["x"], int(g["bs"]), float(g["eb"])
    stream = O.compress(x, eb, bs)
    assert stream == g["stream"].tobytes()
    if g["decoded"].size:
        dec = O.decompress(stream, x.size, x.dtype, ref_quirk=True)
        assert np.array_equal(dec.view(np.uint8), g["decoded"].view(np.uint8))


@pytest.mark.parametrize("path", [p for p in GOLDEN if "rel" in os.path.basename(p)],
                         ids=lambda p: os.path.basename(p)[:-4])
def test_rel_bound_matches_golden(path):
    # REL fp32 goldens come from SZx_fast_compress_args(REL): the stream pins
    # the bound the reference derived from the range.
    g = load_golden(path)
    assert np.float32(O.rel_eb(g["x"], float(g["rel"]))) == g["eb"]


@pytest.mark.parametrize("path", GOLDEN_RA, ids=[os.path.basename(p)[:-4] for p in GOLDEN_RA])
def test_oracle_reproduces_golden_random_access(path):
    g = load_golden(path)
    x, eb, fm = g["x"], float(g["eb"]), int(g["fast_mode"])
    stream = O.compress_ra(x, eb, 128, fm)
    assert stream == g["stream"].tobytes()
    dec = O.decompress_ra(stream, x.size, x.dtype, fm, ref_quirk=True)
    assert np.array_equal(dec.view(np.uint8), g["decoded"].view(np.uint8))


def test_random_access_fastmode4_fp64_statistics_differ():
    # a radius of 1 + 2^-31 against eb = 1: non-constant in double (fastMode
    # 3), constant once rounded to float (fastMode 4)
    assert len(GOLDEN_RA) >= 8
    g3 = load_golden([p for p in GOLDEN_RA if "ra3_f64_radius_flip" in p][0])
    g4 = load_golden([p for p in GOLDEN_RA if "ra4_f64_radius_flip" in p][0])
    s3, s4 = g3["stream"].tobytes(), g4["stream"].tobytes()
    assert int.from_bytes(s4[12:20], "big") > int.from_bytes(s3[12:20], "big")
    # and fastMode 3 is fastMode 2's statistics: same records, same medians
    x = g3["x"]
    s2 = O.compress(x, 1.0, 128)
    assert int.from_bytes(s2[5:13], "big") == int.from_bytes(s3[12:20], "big")


@pytest.mark.parametrize("path", GOLDEN_NB, ids=[os.path.basename(p)[:-4] for p in GOLDEN_NB])
def test_oracle_reproduces_golden_non_blocked(path):
    g = load_golden(path)
    x, eb = g["x"], float(g["eb"])
    stream = O.compress_nb(x, eb)
    assert stream == g["stream"].tobytes()
    dec = O.decompress_nb(stream, x.size, x.dtype, ref_quirk=True)
    assert np.array_equal(dec.view(np.uint8), g["decoded"].view(np.uint8))


def test_non_blocked_layout_known_answer():
    # x = [1, 3], eb = 0.5: radius 1, median 2, reqLength 10 + 0 - (-1) = 11
    # (2 bytes, shift 5); s = bits(-1) >> 5 = 0x05FC0000, bits(1) >> 5 =
    # 0x01FC0000, both lead 0; mid bytes 2..3 of each, ascending
    x = np.array([1.0, 3.0], np.float32)
    s = O.compress_nb(x, 0.5)
    want = bytes([1, 1, 1, 0, 11]) + struct.pack(">f", 2.0) + (1).to_bytes(8, "big") \
        + bytes([0x00, 0xFC, 0x05, 0xFC, 0x01])
    assert s == want
    assert np.array_equal(O.decompress_nb(s, 2, np.float32), x)


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_non_blocked_vs_reference_random(dtype):
    r = np.random.default_rng(31 if dtype == np.float32 else 32)
    compared = quirk = 0
    for trial in range(30):
        n = int(r.integers(1, 40000))
        t = np.arange(n)
        k = trial % 3
        if k == 0:
            x = np.sin(t * r.uniform(1e-4, 0.05)) * 10 ** r.uniform(-3, 3) + r.standard_normal(n) * 1e-5
        elif k == 1:
            x = np.cumsum(r.standard_normal(n)) * 10 ** r.uniform(-6, 2)
        else:
            x = r.standard_normal(n) * 10.0 ** r.uniform(-30, 0)
        x = x.astype(dtype)
        if trial % 2 == 0:
            mode, kw = 0, dict(abs_eb=float(np.float32(np.ptp(x) * 10 ** r.uniform(-9, -1)) + 1e-30))
            eb = kw["abs_eb"]
        else:
            mode, kw = 1, dict(rel=10 ** r.uniform(-7, -1))
            eb = O.rel_eb(x, kw["rel"])
        a = O.compress_nb(x, eb)
        if len(a) > x.size * x.itemsize:
            continue   # the reference would overflow its N*W output buffer
        compared += 1
        assert a == O.ref_compress_args(x, mode, fast_mode=1, **kw), (trial, n)
        ref_dec = O.ref_decompress(a, n, dtype, fast_mode=1)
        dec = O.decompress_nb(a, n, dtype, ref_quirk=True)
        assert np.array_equal(dec.view(np.uint8), ref_dec.view(np.uint8)), trial
        if dtype == np.float64:
            quirk += not np.array_equal(dec, O.decompress_nb(a, n, dtype, ref_quirk=False))
    assert compared > 20


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_access_vs_reference_random(dtype):
    r = np.random.default_rng(21 if dtype == np.float32 else 22)
    compared = 0
    for trial in range(24):
        n = int(r.integers(1, 60000))
        t = np.arange(n)
        x = (np.sin(t * r.uniform(1e-4, 0.05)) * 10 ** r.uniform(-3, 3)
             + r.standard_normal(n) * 10 ** r.uniform(-8, -2)).astype(dtype)
        if trial % 3 == 0:
            x[: n // 3] = 0.5
        fm = 3 + trial % 2
        if trial % 4 < 2:
            mode, kw = 0, dict(abs_eb=float(np.float32(np.ptp(x) * 10 ** r.uniform(-7, -2)) + 1e-30))
            eb = kw["abs_eb"]
        else:
            mode, kw = 1, dict(rel=10 ** r.uniform(-6, -2))
            eb = O.rel_eb(x, kw["rel"])
        a = O.compress_ra(x, eb, 128, fm)
        if len(a) > x.size * x.itemsize:
            continue   # the reference would overflow its N*W output buffer
        compared += 1
        assert a == O.ref_compress_args(x, mode, fast_mode=fm, **kw), (trial, n, fm)
        ref_dec = O.ref_decompress(a, n, dtype, fast_mode=fm)
        dec = O.decompress_ra(a, n, dtype, fm, ref_quirk=True)
        assert np.array_equal(dec.view(np.uint8), ref_dec.view(np.uint8)), trial
    assert compared > 16


# ---- function-level pins against the compiled reference --------------------

def _ref_fn(name, restype, argtypes):
    f = getattr(O.ref(), name)
    f.restype = restype
    f.argtypes = argtypes
    return f


@needs_ref
def test_req_length_exhaustive_vs_reference():
    f32 = _ref_fn("_ZN3szx22computeReqLength_floatEdsPiPf", None,
                  [C.c_double, C.c_short, C.POINTER(C.c_int), C.POINTER(C.c_float)])
    f64 = _ref_fn("_ZN3szx23computeReqLength_doubleEfsPiPf", None,
                  [C.c_float, C.c_short, C.POINTER(C.c_int), C.POINTER(C.c_float)])
    rl = C.c_int(0)
    med = C.c_float(1.0)
    # every radius exponent against bounds from subnormal to huge
    ebs = [2.0 ** e for e in range(-149, 128, 3)] + [0.0, 3e-45, 1e-40, 0.7, 1.5e38]
    for e_rad in range(-127, 129):
        if e_rad == -127:
            radius = np.float32(1e-40)          # subnormal: exponent field 0
        elif e_rad == 128:
            radius = np.float32(np.inf)
        else:
            radius = np.float32(2.0 ** e_rad)
        rad_expo = int((np.array([radius], np.float32).view(np.uint32)[0] >> 23) & 0xFF) - 127
        assert rad_expo == e_rad
        for eb in ebs:
            eb = float(np.float32(eb))
            for w, fn in ((4, f32), (8, f64)):
                med.value = 1.0
                fn(eb, rad_expo, C.byref(rl), C.byref(med))
                ln, clamped = O.req_length(w, eb, float(radius))
                assert ln == rl.value, (w, e_rad, eb)
                assert clamped == (med.value == 0.0), (w, e_rad, eb)


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_block_stats_vs_reference(dtype):
    if dtype == np.float32:
        fn = _ref_fn("_ZN3szx30computeStateMedianRadius_floatEPfmfiPhS0_S0_", C.c_size_t,
                     [C.c_void_p, C.c_size_t, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p])
    else:
        fn = _ref_fn("_ZN3szx31computeStateMedianRadius_doubleEPdmfiPhPfS2_", C.c_size_t,
                     [C.c_void_p, C.c_size_t, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p])
    r = np.random.default_rng(7)
    for trial in range(40):
        n = int(r.integers(1, 3000))
        bs = int(r.choice([1, 3, 4, 64, 128, 200]))
        x = (r.standard_normal(n) * 10.0 ** r.uniform(-40, 30)).astype(dtype)
        x[r.random(n) < 0.2] = 0.0
        x[r.random(n) < 0.05] = -0.0
        eb = float(np.float32(10.0 ** r.uniform(-45, 30)))
        nb = -(-n // bs)
        st, med, rad = np.zeros(nb, np.uint8), np.zeros(nb, np.float32), np.zeros(nb, np.float32)
        nc = fn(x.ctypes.data, n, eb, bs, st.ctypes.data, med.ctypes.data, rad.ctypes.data)
        onc, ost, omed, orad = O.block_stats(x, eb, bs)
        assert nc == onc
        assert np.array_equal(st, ost)
        assert np.array_equal(med.view(np.uint32), omed.view(np.uint32))
        assert np.array_equal(rad.view(np.uint32), orad.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_codec_vs_reference_random(dtype):
    r = np.random.default_rng(11 if dtype == np.float32 else 12)
    compared = quirk = 0
    for trial in range(150):
        n = int(r.integers(64, 5000))
        bs = int(r.choice([4, 16, 64, 100, 128, 255]))
        kind = trial % 4
        if kind == 0:
            x = np.sin(np.arange(n) * r.uniform(1e-3, 0.2)) * 10 ** r.uniform(-5, 5)
        elif kind == 1:
            x = np.cumsum(r.standard_normal(n)) * 10 ** r.uniform(-6, 2)
        elif kind == 2:
            x = np.where(r.random(n) < 0.6, 1.25, r.standard_normal(n))
        else:
            x = r.standard_normal(n) * 10.0 ** r.uniform(-300 if dtype == np.float64 else -38, 0)
        x = (x + r.standard_normal(n) * 10 ** r.uniform(-9, -3)).astype(dtype)
        eb = float(np.float32((np.ptp(x) + 1e-30) * 10.0 ** r.uniform(-12, -1)))
        a = O.compress(x, eb, bs)
        if len(a) > x.size * x.itemsize:
            continue   # the reference would overflow its N*W output buffer
        compared += 1
        assert a == O.ref_compress(x, eb, bs), (trial, n, bs, eb)
        ref_dec = O.ref_decompress(a, n, dtype)
        dec = O.decompress(a, n, dtype, ref_quirk=True)
        assert np.array_equal(dec.view(np.uint8), ref_dec.view(np.uint8)), trial
        if dtype == np.float64:
            quirk += not np.array_equal(dec, O.decompress(a, n, dtype, ref_quirk=False))
    assert compared > 100
    if dtype == np.float64:
        assert quirk > 0, "no stream reached the reqBytes=5 decoder branch"


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fxr_search_vs_reference(dtype):
    if dtype == np.float32:
        fn = _ref_fn("_ZN3szx42estimateErrorBoundbasedonCR_buffered_floatEffifimPfS0_S0_", C.c_float,
                     [C.c_float, C.c_float, C.c_int, C.c_float, C.c_int, C.c_size_t,
                      C.c_void_p, C.c_void_p, C.c_void_p])
    else:
        fn = _ref_fn("_ZN3szx43estimateErrorBoundbasedonCR_buffered_doubleEffifimPdPfS1_", C.c_float,
                     [C.c_float, C.c_float, C.c_int, C.c_float, C.c_int, C.c_size_t,
                      C.c_void_p, C.c_void_p, C.c_void_p])
    L = O.lib()
    w = 4 if dtype == np.float32 else 8
    r = np.random.default_rng(21)
    for trial in range(12):
        n = int(r.integers(20000, 60000))
        x = (np.sin(np.arange(n) * r.uniform(1e-4, 1e-2)) * 10 ** r.uniform(-2, 3)
             + r.standard_normal(n) * 10 ** r.uniform(-6, -2)).astype(dtype)
        bs, rate = 128, 10
        ns = n // (bs * rate)
        buf = np.zeros((ns + 1) * bs, dtype)
        med = np.zeros(ns + 1, np.float32)
        rad = np.zeros(ns + 1, np.float32)
        rng_ = L.szxo_fxr_sample(x.ctypes.data, w, n, rate, bs, buf.ctypes.data, med.ctypes.data,
                                 rad.ctypes.data)
        init = float(np.float32(rng_ * 1e-2))
        for target in (3.0, 6.0, 12.0):
            ne = C.c_int(0)
            ours = L.szxo_fxr_search(target, 0.1, rate, init, bs, n, buf.ctypes.data, w,
                                     med.ctypes.data, rad.ctypes.data, C.byref(ne))
            theirs = fn(target, 0.1, rate, init, bs, n, buf.ctypes.data, med.ctypes.data,
                        rad.ctypes.data)
            assert np.float32(ours) == np.float32(theirs), (trial, target)


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fxr_compress_vs_reference(dtype):
    """The whole fixed-rate path (szx.cc:238-243): the oracle's bound and the
    oracle's stream at that bound equal the reference's FXR output.  The
    reference writes into N*W/cr*(1+2 tol) bytes, so it is only called where
    the stream fits."""
    r = np.random.default_rng(33)
    w = np.dtype(dtype).itemsize
    checked = 0
    for trial in range(10):
        n = int(r.integers(30000, 120000))
        x = (np.sin(np.arange(n) * r.uniform(1e-4, 1e-2)) * 10 ** r.uniform(-2, 3)
             + r.standard_normal(n) * 10 ** r.uniform(-7, -3)).astype(dtype)
        for target in (4.0, 8.0, 16.0):
            tol = 0.1
            eb, _ = O.fxr_error_bound(x, target, tol, 128)
            ours = O.compress(x, eb, 128)
            cap = int(w * n / np.float32(target) * (1 + np.float32(tol) * 2))
            if len(ours) + 64 > cap:   # margin for the float sizing
                continue
            theirs = O.ref_compress_args(x, mode=2, fast_mode=2, cr=target, tol=tol)
            assert ours == theirs, (trial, target)
            checked += 1
    assert checked >= 5


# ---- hand-worked known answers and format invariants ------------------------

def _be32(v):
    return struct.pack(">f", v)


def test_known_answer_all_constant():
    # constant field: header, all-zero bitmap, one BE median per block
    x = np.full(300, 2.5, np.float32)
    s = O.compress(x, 0.0, 128)
    want = bytes([1, 1, 1, 0, 128]) + (3).to_bytes(8, "big") + b"\x00" + _be32(2.5) * 3
    assert s == want
    assert np.array_equal(O.decompress(s, 300, np.float32), x)


def test_known_answer_one_block_fp32():
    # x = [0, 1, 2, 3], eb = 0.5, bs = 4:
    #   radius 1.5 (expo 0), median 1.5, eb expo -1 -> reqLength 10+0+1 = 11,
    #   rb = 2, shift = 5.  x - median = -1.5, -0.5, 0.5, 1.5
    #   bits: 0xBFC00000 0xBF000000 0x3F000000 0x3FC00000, >> 5:
    #   0x05FE0000 0x05F80000 0x01F80000 0x01FE0000
    #   lead vs previous (first vs 0): xor 0x05FE0000 -> 0; 0x00060000 -> 1;
    #   0x04000000 -> 0; 0x00060000 -> 1;  lead byte 0b00010001 = 0x11
    #   mid bytes (bytes 2..3 of each, ascending, only below the lead):
    #   e0: FE 05; e1: F8; e2: F8 01; e3: FE
    x = np.array([0, 1, 2, 3], np.float32)
    s = O.compress(x, 0.5, 4)
    rec = bytes([11]) + _be32(1.5) + bytes([0x11]) + bytes([0xFE, 0x05, 0xF8, 0xF8, 0x01, 0xFE])
    want = bytes([1, 1, 1, 0, 4]) + (0).to_bytes(8, "big") + bytes([0x80]) + rec
    assert s == want
    d = O.decompress(s, 4, np.float32)
    assert np.max(np.abs(d - x)) <= 0.5


def test_block_size_256_is_stored_as_zero_byte():
    x = np.linspace(0, 1, 1000, dtype=np.float32)
    s = O.compress(x, 1e-4, 256)
    assert s[4] == 0
    d = O.decompress(s, x.size, np.float32)   # 0 read back as 256
    assert np.max(np.abs(d - x)) <= 1e-4


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_size_formula_and_offsets(dtype):
    r = np.random.default_rng(3)
    x = (np.cumsum(r.standard_normal(10000)) * 1e-2).astype(dtype)
    x[2000:4000] = 0
    s = O.compress(x, 1e-3, 128)
    w = x.itemsize
    nc, st, _, _ = O.block_stats(x, 1e-3, 128)
    offs = O.record_offsets(s, x.size, w)
    nb = st.size
    assert len(offs) == nb - nc
    assert offs[0] == 13 + (nb + 7) // 8 + 4 * nc
    sizes = np.diff(np.append(offs, len(s)))
    arr = np.frombuffer(s, np.uint8)
    blocks = np.nonzero(st)[0]
    for k, o in enumerate(offs):
        nel = min(128, x.size - 128 * int(blocks[k]))
        assert O.lib().szxo_record_size(arr.ctypes.data + int(o), nel, w) == sizes[k]


def test_fp32_error_bound_holds():
    r = np.random.default_rng(5)
    for trial in range(30):
        x = (np.sin(np.arange(4096) * r.uniform(1e-3, 1e-1)) * 10 ** r.uniform(-3, 3)
             + r.standard_normal(4096) * 1e-3).astype(np.float32)
        eb = float(np.float32(np.ptp(x) * 10 ** r.uniform(-6, -1)))
        d = O.decompress(O.compress(x, eb, 128), x.size, np.float32)
        assert np.max(np.abs(d.astype(np.float64) - x.astype(np.float64))) <= eb


def test_fp64_quirk_only_changes_rb5_records():
    # reqBytes = 5 needs reqLength 33..40: radius/eb ~ 2^21..2^28
    r = np.random.default_rng(9)
    x = np.cumsum(r.standard_normal(4096))
    eb = float(np.float32(np.ptp(x) * 1e-8))
    s = O.compress(x, eb, 128)
    a = O.decompress(s, x.size, np.float64, ref_quirk=True)
    b = O.decompress(s, x.size, np.float64, ref_quirk=False)
    assert np.max(np.abs(b - x)) <= eb          # the bytes the encoder wrote decode in bound
    offs = O.record_offsets(s, x.size, 8)
    blocks = np.nonzero(O.block_stats(x, eb, 128)[1])[0]
    rb5 = {int(blocks[k]) for k, o in enumerate(offs) if (s[o] + 7) // 8 == 5}
    assert rb5, "workload did not reach reqBytes 5"
    diff_blocks = set(np.nonzero(a != b)[0] // 128)
    assert diff_blocks and diff_blocks <= rb5


# ---- the rate model's CPU oracle, pinned by finite differences ---------------

def test_mlp_oracle_gradient_matches_finite_differences():
    import mlp_oracle as M
    r = np.random.default_rng(5)
    P = M.init_params(1).astype(np.float64)
    P[M.N_PARAMS - 1] = 0.3
    X = r.standard_normal((40, M.N_IN))
    T = r.standard_normal(40)
    g, _ = M.grad(P, X, T, rounding=False)
    h = 1e-6
    for i in r.choice(M.N_PARAMS, 120, replace=False).tolist() + [M.N_PARAMS - 1, 1536, 5760]:
        e = np.zeros_like(P)
        e[i] = h
        fd = (M.loss(P + e, X, T, rounding=False) - M.loss(P - e, X, T, rounding=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_mlp_oracle_bf16_rounding():
    import mlp_oracle as M
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30], np.float32)
    b = M.bf16(a)
    # bf16 keeps 8 significant bits: 1 + 2^-8 is a tie and rounds to even (1.0)
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0078125
    assert np.all(np.abs(b - a) <= np.abs(a) * 2.0 ** -8)


def test_mlp_oracle_forward_matches_torch():
    """The forward pass against torch's own layers: F.linear for each layer
    and F.gelu(approximate="tanh"), in float64 with rounding off, and with
    rounding on, torch's bfloat16 cast at the same points.  A wrong GELU
    constant, a transposed weight or a swapped layer order fails here."""
    import torch
    import torch.nn.functional as F

    import mlp_oracle as M
    r = np.random.default_rng(11)
    P = M.init_params(3).astype(np.float64)
    P += 0.05 * r.standard_normal(P.size)
    X = r.standard_normal((64, M.N_IN)) * 2.0
    p = {k: torch.from_numpy(v.copy()) for k, v in M.unpack(P).items()}
    xt = torch.from_numpy(X)

    def torch_forward(rnd):
        q = (lambda t: t.to(torch.bfloat16).to(torch.float64)) if rnd else (lambda t: t)
        g1 = F.gelu(F.linear(q(xt), q(p["W1"]), p["b1"]), approximate="tanh")
        g2 = F.gelu(F.linear(q(g1), q(p["W2"]), p["b2"]), approximate="tanh")
        return F.linear(g2, p["w3"].view(1, -1), p["b3"]).view(-1).numpy()

    y, _ = M.forward(P, X, rounding=False)
    np.testing.assert_allclose(y, torch_forward(False), rtol=1e-12, atol=1e-12)
    y, _ = M.forward(P, X, rounding=True)
    np.testing.assert_allclose(y, torch_forward(True), rtol=1e-9, atol=1e-9)
    # bf16 rounding itself: torch's float32 -> bfloat16 cast (round to nearest even)
    a = (r.standard_normal(4096) * 10.0 ** r.uniform(-20, 20, 4096)).astype(np.float32)
    tb = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(a), tb)
