"""The drop-in boundary exercised from C: tests/c_abi/szx_caller, a C program
written against the reference's public API (szx.h:74-98) and linked to
libszx_amd.so, compresses and decompresses the golden inputs; the files it
writes must equal the reference-written golden streams and decodes byte for
byte.  When the caller was also built against the reference's own szx.h
(where its sources were present at build time), that binary runs the same
cases.  Checks beyond the goldens use the oracle (the sampled fixed-rate
model, the random-access container at other block sizes)."""
import glob
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
GOLDEN = os.path.join(HERE, "golden")
CALLERS = [p for p in (os.path.join(HERE, "c_abi", "szx_caller"), os.path.join(HERE, "c_abi", "szx_caller_refhdr"))
           if os.path.exists(p)]


def say(*a):
    print(*a, flush=True)


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def run(exe, *args):
    say("run", os.path.basename(exe), *args[:6])
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, f"{os.path.basename(exe)} {args[0]}: rc={r.returncode} {r.stderr.strip()}"
    return r.stdout


def f9(v):
    return f"{float(np.float32(v)):.9g}"


def dt(x):
    return "f32" if x.dtype == np.float32 else "f64"


@pytest.fixture(params=range(2), ids=["szx_amd.h", "reference szx.h"])
def caller(request):
    if not CALLERS:
        pytest.fail("tests/c_abi/szx_caller is not built (make -C tests/c_abi)")
    if request.param >= len(CALLERS):
        pytest.skip("caller against the reference's szx.h was not built")
    return CALLERS[request.param]


@pytest.mark.parametrize("name", ["f32_sin_rel0.01", "f32_sin_rel0.001", "f32_sin_rel0.0001", "f32_sin_abs1e-3"])
def test_codec_from_c(caller, name, tmp_path, device):
    g = load(name)
    x = g["x"]
    x.tofile(tmp_path / "in.bin")
    mode, bound = ("rel", g["rel"]) if "_rel" in name else ("abs", g["eb"])
    run(caller, "codec", dt(x), mode, f9(bound), x.size, tmp_path / "in.bin", tmp_path / "s.bin", tmp_path / "d.bin")
    assert (tmp_path / "s.bin").read_bytes() == g["stream"].tobytes()
    assert (tmp_path / "d.bin").read_bytes() == g["decoded"].tobytes()


BLOCKED = ["f32_ramp_eb1e-4_bs8", "f32_ramp_eb1e-7_bs100", "f32_ramp_eb1e-10_bs128",
           "f64_grf_rel1e-06_bs64", "f64_grf_rel0.0001_bs256", "f64_walk_eb1e-14"]


@pytest.mark.parametrize("name", BLOCKED)
def test_blocked_and_caller_buffer_from_c(caller, name, tmp_path, device):
    g = load(name)
    x = g["x"]
    bs = int(g["bs"])
    x.tofile(tmp_path / "in.bin")
    run(caller, "blocked", dt(x), f9(g["eb"]), bs, x.size, tmp_path / "in.bin", tmp_path / "s.bin",
        tmp_path / "d.bin")
    assert (tmp_path / "s.bin").read_bytes() == g["stream"].tobytes()
    want = g["decoded"] if g["decoded"].size else O.decompress(g["stream"].tobytes(), x.size, x.dtype)
    assert (tmp_path / "d.bin").read_bytes() == want.tobytes()


@pytest.mark.parametrize("name", ["ra4_f32_sin_rel1e-3", "ra4_f32_sin_abs1e-5", "ra4_f64_grf_rel1e-6",
                                  "ra4_f64_radius_flip"])
def test_random_access_openmp_from_c(caller, name, tmp_path, device):
    g = load(name)
    x = g["x"]
    x.tofile(tmp_path / "in.bin")
    run(caller, "ra", dt(x), f9(g["eb"]), int(g["bs"]), x.size, tmp_path / "in.bin", tmp_path / "s.bin",
        tmp_path / "d.bin")
    assert (tmp_path / "s.bin").read_bytes() == g["stream"].tobytes()
    assert (tmp_path / "d.bin").read_bytes() == g["decoded"].tobytes()


@pytest.mark.parametrize("bs", [4, 32, 64, 256])
def test_random_access_openmp_other_block_sizes(caller, bs, tmp_path, device):
    """blockSize is an argument of this pair: other sizes against the oracle's
    fastMode-4 container (itself pinned to the compiled reference)."""
    for name in ("ra4_f32_sin_rel1e-3", "ra4_f64_grf_rel1e-6"):
        g = load(name)
        x = g["x"]
        x.tofile(tmp_path / "in.bin")
        eb = float(g["eb"])
        run(caller, "ra", dt(x), f9(eb), bs, x.size, tmp_path / "in.bin", tmp_path / "s.bin", tmp_path / "d.bin")
        want = O.compress_ra(x, eb, bs, fast_mode=4)
        assert (tmp_path / "s.bin").read_bytes() == want
        dec = O.decompress_ra(want, x.size, x.dtype, fast_mode=4)
        assert (tmp_path / "d.bin").read_bytes() == dec.tobytes()


@pytest.mark.parametrize("name", ["nb_f32_sin_rel1e-3", "nb_f32_sin_abs1e-5"])
def test_decompress_pred_from_c(caller, name, tmp_path, device):
    """SZx_fast_decompress_pred: the non-blocked decode plus the prediction,
    value by value in float (szxd_float.cc:25-32)."""
    g = load(name)
    n = g["x"].size
    pred = np.random.default_rng(7).standard_normal(n).astype(np.float32)
    (tmp_path / "s.bin").write_bytes(g["stream"].tobytes())
    pred.tofile(tmp_path / "p.bin")
    run(caller, "pred", n, tmp_path / "s.bin", tmp_path / "p.bin", tmp_path / "d.bin")
    want = (g["decoded"] + pred).astype(np.float32)
    assert (tmp_path / "d.bin").read_bytes() == want.tobytes()


@pytest.mark.parametrize("name,rate,bs", [("f32_sin_abs1e-3", 10, 128), ("f32_sin_rel0.001", 3, 64),
                                          ("f64_grf_rel1e-06_bs128", 10, 128), ("f64_walk_eb1e-8", 2, 32)])
def test_radius_buffer_and_buffered_estimate_from_c(caller, name, rate, bs, tmp_path, device):
    """computeRadiusBuffer and estimateCRbasedonErrorBound_buffered against the
    oracle's restatement (szx_oracle_fxr.c, pinned to the compiled reference)."""
    g = load(name)
    x = g["x"]
    eb = float(np.float32(g["eb"]))
    x.tofile(tmp_path / "in.bin")
    out = run(caller, "sample", dt(x), rate, bs, f9(eb), x.size, tmp_path / "in.bin", tmp_path / "o").split()
    got = dict(zip(out[0::2], out[1::2]))
    ns = x.size // (rate * bs)
    buf = np.zeros((ns + 1) * bs, x.dtype)
    med = np.zeros(ns + 1, np.float32)
    rad = np.zeros(ns + 1, np.float32)
    L = O.lib()
    w = x.dtype.itemsize
    rng = L.szxo_fxr_sample(O._ptr(x), w, x.size, rate, bs, O._ptr(buf), O._ptr(med), O._ptr(rad))
    assert np.float32(got["range"]) == np.float32(rng)
    assert (tmp_path / "o.buf").read_bytes() == buf[:ns * bs].tobytes()
    assert (tmp_path / "o.rad").read_bytes() == rad[:ns].tobytes()
    assert (tmp_path / "o.med").read_bytes() == med[:ns].tobytes()
    import ctypes as C
    req, lead = C.c_size_t(0), C.c_size_t(0)
    cr = L.szxo_fxr_estimate_cr(eb, O._ptr(buf), w, O._ptr(med), O._ptr(rad), rate, bs, x.size,
                                C.byref(req), C.byref(lead))
    assert int(got["req"]) == req.value and int(got["lead"]) == lead.value
    assert np.float32(got["cr"]) == np.float32(cr)
