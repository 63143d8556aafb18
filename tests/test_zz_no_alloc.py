"""szx_amd_compress_dev_noalloc (the compress whose workspace is the
library's static device scratch array) against szx_amd_compress_dev (the same
compress with a caller workspace), through the C ABI: result words, stream
bytes and record index must be identical, on consecutive calls where a later
call is smaller than the one before (the scratch keeps the earlier call's
contents), and in both bound modes, both widths, a tiled and a per-block
block size.  Sizes past the scratch array are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import szx_amd as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _field(torch, n, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    t = torch.arange(n, device="cuda", dtype=torch.float64)
    x = torch.sin(t * 1e-3) * 100 + torch.randn(n, device="cuda", dtype=torch.float64, generator=g) * 0.05
    x[n // 3: n // 3 + n // 10] = 0.5   # a constant stretch
    return x.to(dtype).contiguous()


def _compress(torch, x, n, code, mode, eb, rel, bs, with_ws):
    L = Z.lib()
    cap = L.szx_amd_compress_capacity(n, code, bs)
    out = torch.full((cap,), 0xA5, dtype=torch.uint8, device="cuda")
    res = torch.zeros(4, dtype=torch.int64, device="cuda")
    idx = torch.zeros(L.szx_amd_index_words(n, code, bs), dtype=torch.int64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if with_ws:
        wsb = L.szx_amd_compress_workspace(n, code, bs)
        ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        rc = L.szx_amd_compress_dev(x.data_ptr(), n, code, mode, eb, rel, bs, out.data_ptr(), cap, ws.data_ptr(),
                                    wsb, res.data_ptr(), idx.data_ptr(), stream)
    else:
        rc = L.szx_amd_compress_dev_noalloc(x.data_ptr(), n, code, mode, eb, rel, bs, out.data_ptr(), cap,
                                            res.data_ptr(), idx.data_ptr(), stream)
    assert rc == 0, Z.last_error()
    torch.cuda.synchronize()
    off, size, ebbits, err = (int(v) for v in res.cpu().tolist())
    assert err == 0
    return out[off:off + size].cpu().numpy().tobytes(), (off, size, ebbits), idx.cpu().numpy()


# (n, dtype, errBoundMode, bound, blockSize); each later call smaller than
# the one before it on the shared scratch
CALLS = [
    (3_000_001, "f32", Z.REL, 1e-3, 128),
    (1_000_003, "f32", Z.REL, 1e-3, 128),
    (2_000_000, "f64", Z.ABS, 1e-6, 64),
    (77_777, "f64", Z.REL, 1e-4, 256),
    (500_000, "f32", Z.ABS, 1e-2, 100),
    (4_097, "f32", Z.REL, 1e-2, 100),
]


def test_noalloc_matches_the_workspace_compress(device):
    import torch
    for i, (n, dt, mode, bound, bs) in enumerate(CALLS):
        say("case", i, n, dt, "REL" if mode == Z.REL else "ABS", bound, bs)
        tdt = torch.float32 if dt == "f32" else torch.float64
        code = Z.SZx_FLOAT if dt == "f32" else Z.SZx_DOUBLE
        x = _field(torch, n, tdt, seed=100 + i)
        eb = bound if mode == Z.ABS else 0.0
        rel = bound if mode == Z.REL else 0.0
        s_ws, r_ws, i_ws = _compress(torch, x, n, code, mode, eb, rel, bs, True)
        s_na, r_na, i_na = _compress(torch, x, n, code, mode, eb, rel, bs, False)
        assert r_na == r_ws
        assert s_na == s_ws
        assert np.array_equal(i_na, i_ws)


def test_noalloc_refuses_sizes_past_the_scratch(device):
    import torch
    L = Z.lib()
    x = torch.zeros(16, dtype=torch.float32, device="cuda")
    n = 1 << 27
    assert L.szx_amd_compress_workspace(n, Z.SZx_FLOAT, 4) > Z.NOALLOC_SCRATCH_BYTES
    res = torch.zeros(4, dtype=torch.int64, device="cuda")
    rc = L.szx_amd_compress_dev_noalloc(x.data_ptr(), n, Z.SZx_FLOAT, Z.ABS, 1e-3, 0.0, 4, x.data_ptr(),
                                        1 << 40, res.data_ptr(), None,
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == Z.SZx_MERR
    assert "scratch" in Z.last_error()
