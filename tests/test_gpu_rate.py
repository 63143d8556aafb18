"""GPU tests of the fixed-rate machinery on the exact codec: the rate model's
BF16 matrix-core kernels against the CPU oracle (oracle/mlp_oracle.py), the
octave pass against the codec's own stream sizes, the statistics pass
against numpy, and the two solvers."""
import os

import numpy as np
import pytest

import mlp_oracle as M
import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))


def test_mlp_forward_matches_oracle(device):
    import torch
    r = np.random.default_rng(11)
    P = M.init_params(3)
    P[-1] = 0.25
    for n in (1, 15, 16, 1000, 4099):
        X = r.standard_normal((n, M.N_IN)).astype(np.float32)
        say("forward", n)
        model = Z.RateModel(P, device=device)
        y = model.forward(torch.from_numpy(X).to(device)).cpu().numpy()
        want, _ = M.forward(P, X)
        assert np.max(np.abs(y - want)) <= 2e-3 * (np.max(np.abs(want)) + 1e-3)


def test_mlp_grad_matches_oracle(device):
    import torch
    r = np.random.default_rng(12)
    P = M.init_params(4)
    for n in (16, 333, 5000):
        X = r.standard_normal((n, M.N_IN)).astype(np.float32)
        T = r.standard_normal(n).astype(np.float32)
        say("grad", n)
        model = Z.RateModel(P, device=device)
        g, loss = model.grad(torch.from_numpy(X).to(device), torch.from_numpy(T).to(device))
        g = g.cpu().numpy().astype(np.float64)
        want, want_loss = M.grad(P, X, T)
        assert abs(float(loss.item()) - want_loss) <= 1e-3 * want_loss
        o = 0
        for name, shape in M.SIZES:
            k = int(np.prod(shape))
            assert _rel(g[o:o + k], want[o:o + k]) < 5e-3, name
            o += k


def test_mlp_trains(device):
    """Adam on the GPU kernels fits a smooth target."""
    import torch
    r = np.random.default_rng(13)
    X = r.uniform(-1, 1, (8192, M.N_IN)).astype(np.float32)
    T = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3]).astype(np.float32)
    Xd, Td = torch.from_numpy(X).to(device), torch.from_numpy(T).to(device)
    model = Z.RateModel(M.init_params(5), device=device)
    m = torch.zeros(Z.RATE_MODEL_PARAMS, device=device)
    v = torch.zeros_like(m)
    losses = []
    for step in range(1, 301):
        g, loss = model.grad(Xd, Td)
        model.adam(g, m, v, 3e-3, step)
        if step in (1, 300):
            losses.append(float(loss.item()))
    say("loss", losses)
    assert losses[1] < 0.1 * losses[0]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rate_octave_is_exact(dtype, device):
    import torch
    r = np.random.default_rng(14)
    n = 200003
    t = np.arange(n)
    x = (np.sin(t * 0.003) * 50 + r.standard_normal(n) * 1e-2).astype(dtype)
    x[50000:90000] = 1.5
    xt = torch.from_numpy(x).to(device)
    rng = float(np.ptp(x))
    for e in (int(np.floor(np.log2(rng))) - 6, int(np.floor(np.log2(rng))) - 14,
              int(np.floor(np.log2(rng))) - 3):
        sizes, ebs = Z.rate_octave(xt, e, 128)
        say("octave", dtype.__name__, e)
        for k in (0, 1, 100, 255, 256, 511):
            assert sizes[k] == len(O.compress(x, float(ebs[k]), 128)), (e, k)
        assert np.all(np.diff(sizes) <= 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rate_octave_span_is_exact(dtype, device):
    """One pass over an octave pair gives each octave's single-pass sizes,
    and the bounds just outside the pair size like the oracle's streams."""
    import torch
    r = np.random.default_rng(16)
    n = 150001
    t = np.arange(n)
    x = (np.cos(t * 0.002) * 20 + r.standard_normal(n) * 3e-3).astype(dtype)
    x[20000:60000] = -2.25
    xt = torch.from_numpy(x).to(device)
    top = int(np.floor(np.log2(np.ptp(x))))
    for e in (top - 5, top - 12):
        for nocts in (1, 2):
            say("octave span", dtype.__name__, e, nocts)
            sizes, ebs, up, dn = Z.rate_octave_span(xt, e, nocts, 128)
            assert sizes.size == nocts * Z.OCTAVE_BINS
            for j in range(nocts):
                s1, e1 = Z.rate_octave(xt, e - j, 128)
                at = (nocts - 1 - j) * Z.OCTAVE_BINS
                assert np.array_equal(sizes[at:at + Z.OCTAVE_BINS], s1), (e, j)
                assert np.array_equal(ebs[at:at + Z.OCTAVE_BINS], e1), (e, j)
            assert np.all(np.diff(sizes) <= 0)
            for size, eb in (up, dn):
                assert size == len(O.compress(x, float(eb), 128)), (e, nocts, float(eb))
            assert sizes[0] == len(O.compress(x, float(ebs[0]), 128))
            assert sizes[-1] == len(O.compress(x, float(ebs[-1]), 128))


def test_rate_stats_match_numpy(device):
    import torch
    r = np.random.default_rng(15)
    n = Z.RATE_CHUNK + 70001
    x = (np.cumsum(r.standard_normal(n)) * 0.01).astype(np.float32)
    x[1000:5000] = 0.0
    st = Z.rate_stats(torch.from_numpy(x).to(device), 128).cpu().numpy()
    rec = np.dtype([("d", "<f8", 7), ("values", "<u8"), ("blocks", "<u4"), ("zero", "<u4"),
                    ("hist", "<u4", 256)])
    s = st.view(rec)
    assert s.shape[0] == 2
    for c in range(2):
        xc = x[c * Z.RATE_CHUNK:(c + 1) * Z.RATE_CHUNK].astype(np.float64)
        say("stats chunk", c)
        assert s["values"][c] == xc.size
        assert s["d"][c][0] == xc.min() and s["d"][c][1] == xc.max()
        assert np.isclose(s["d"][c][2], xc.sum(), rtol=1e-9, atol=1e-6)
        assert np.isclose(s["d"][c][3], (xc * xc).sum(), rtol=1e-9)
        nb = (xc.size + 127) // 128
        pad = np.concatenate([xc, np.full(nb * 128 - xc.size, np.nan)]).reshape(nb, 128)
        d1 = np.diff(pad, axis=1)
        d2 = np.diff(d1, axis=1)
        assert np.isclose(s["d"][c][4], np.nansum(np.abs(d1)), rtol=1e-9)
        assert np.isclose(s["d"][c][5], np.nansum(d1 * d1), rtol=1e-9)
        assert np.isclose(s["d"][c][6], np.nansum(np.abs(d2)), rtol=1e-9)
        rad = ((np.nanmax(pad, 1) - np.nanmin(pad, 1)) / 2).astype(np.float32)
        assert s["blocks"][c] == nb and s["zero"][c] == int(np.sum(rad == 0))
        hist = np.bincount((rad.view(np.uint32) >> 23) & 0xFF, minlength=256)
        assert np.array_equal(s["hist"][c], hist)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fixed_rate_bisection(dtype, device):
    import torch
    from szx_amd.fields import sinusoid_field
    x = sinusoid_field((64, 96, 100), seed=7, dtype=dtype)
    xt = torch.from_numpy(x.ravel()).to(device)
    fr = Z.FixedRateCompressor(x.size, dtype, 128)
    for target in (2.0, 6.0):
        say("bisect", dtype.__name__, target)
        info = fr(xt, target, 0.02, method=Z.RATE_BISECT)
        s = fr.stream_view().cpu().numpy().tobytes()
        assert s == O.compress(x, info.errorBound, 128)
        assert abs(8.0 * len(s) / x.size - info.bitsPerValue) < 1e-9
        assert info.hit == 1 and abs(info.bitsPerValue - target) <= 0.02 * target


@pytest.mark.skipif(not os.path.exists(Z.RATE_MODEL_PATH), reason="no trained rate model")
def test_fixed_rate_learned_beats_bisection(device):
    """The trained model's proposal needs far fewer codec passes than
    bisection for the same tolerance, and its streams are exact codec
    output."""
    import torch
    from szx_amd.fields import sinusoid_field_torch
    model = Z.RateModel(device=device)
    passes_l, passes_b, hits_l, hits_b, total = 0, 0, 0, 0, 0
    for i in range(6):
        r = np.random.default_rng(5000 + i)
        slope, noise = r.uniform(1.5, 4.0), 10 ** r.uniform(-6, -2)
        x = sinusoid_field_torch((128, 128, 128), seed=5000 + i, n_waves=32, kmax=16, slope=slope,
                                 noise=noise, n_boxes=1, box_frac=0.25, box_value="zero",
                                 device=device).contiguous().view(-1)
        fr = Z.FixedRateCompressor(x.numel(), np.float32, 128, model=model)
        for target in (1.0, 2.0, 4.0, 8.0):
            li = fr(x, target, 0.02, method=Z.RATE_LEARNED)
            sl = fr.stream_view().cpu().numpy().tobytes()
            bi = fr(x, target, 0.02, method=Z.RATE_BISECT)
            say("field", i, "target", target, "learned", li.codecPasses, li.bitsPerValue,
                "bisect", bi.codecPasses, bi.bitsPerValue)
            assert sl == O.compress(x.cpu().numpy(), li.errorBound, 128)
            passes_l += li.codecPasses
            passes_b += bi.codecPasses
            hits_l += li.hit
            hits_b += bi.hit
            total += 1
    say("mean passes learned", passes_l / total, "bisection", passes_b / total,
        "hits", hits_l, hits_b, "of", total)
    assert hits_l >= hits_b - 1
    assert passes_l / total <= 2.5
    assert passes_l < 0.4 * passes_b
