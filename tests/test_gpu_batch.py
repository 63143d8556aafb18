"""GPU parity of the batched launch (szx_amd_batch_*) through the C ABI.

Every field's stream must equal the oracle's stream for that field alone,
byte for byte, whatever the tile shape, the launch mode (one persistent
launch or one launch sequence per field) or the replay of a captured graph;
every decode must equal the oracle decoder's bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(
        a.view(np.uint8), b.view(np.uint8))


def _values(r, n, dtype, kind):
    t = np.arange(n)
    if kind == 0:
        x = np.sin(t * r.uniform(1e-4, 0.05)) * 10 ** r.uniform(-3, 3)
    elif kind == 1:
        x = np.cumsum(r.standard_normal(n)) * 10 ** r.uniform(-4, 2)
    elif kind == 2:   # constant runs (constant blocks) between waves
        x = np.where((t // 3000) % 3 == 0, 0.25, np.sin(t * 0.003))
    else:
        x = r.standard_normal(n) * 10.0 ** r.uniform(-300 if dtype == np.float64 else -30, 0)
    return (x + r.standard_normal(n) * 10 ** r.uniform(-8, -4)).astype(dtype)


SIZES = (1, 5, 127, 128, 2048, 4096, 4097, 8191, 8192, 8193, 65536, 100_003, 1 << 20,
         (1 << 20) + 77, 3_000_001)


def make_fields(seed, count=None):
    """(numpy arrays, modes, bounds, oracle bounds) of a mixed batch."""
    r = np.random.default_rng(seed)
    sizes = list(SIZES if count is None else r.choice(SIZES, count))
    r.shuffle(sizes)
    xs, modes, bounds, ebs = [], [], [], []
    for i, n in enumerate(sizes):
        dtype = np.float32 if r.random() < 0.6 else np.float64
        x = _values(r, int(n), dtype, i % 4)
        if r.random() < 0.6:
            rel = float(np.float32(10 ** r.uniform(-5, -2)))
            modes.append(Z.REL)
            bounds.append(rel)
            ebs.append(O.rel_eb(x, rel))
        else:
            eb = float(np.float32((float(np.ptp(x)) + 1e-30) * 10 ** r.uniform(-6, -2)))
            modes.append(Z.ABS)
            bounds.append(eb)
            ebs.append(eb)
        xs.append(x)
    return xs, modes, bounds, ebs


def check_batch(bc, xs, ebs, want, tag):
    import torch
    torch.cuda.synchronize()
    res = bc.results()
    assert not res[:, 3].any(), f"{tag}: a bounded wait ran out"
    st = bc.status().cpu().numpy()
    for i, x in enumerate(xs):
        s = bc.stream(i, res).cpu().numpy().tobytes()
        assert s == want[i][0], (tag, i, x.dtype, x.size)
        assert np.uint32(res[i, 2]) == np.array([ebs[i]], np.float32).view(np.uint32)[0], (tag, i)
        if bc.decs[i] is not None:
            assert st[i] == 0, (tag, i, st[i])
            assert bits_equal(bc.decs[i].cpu().numpy(), want[i][1]), (tag, i, x.dtype, x.size)


def oracle_of(xs, ebs, bs=128):
    out = []
    for x, eb in zip(xs, ebs):
        s = O.compress(x, eb, bs)
        out.append((s, O.decompress(s, x.size, x.dtype, ref_quirk=True)))
    return out


@pytest.fixture(scope="module")
def batch(device):
    import torch
    xs, modes, bounds, ebs = make_fields(42)
    say("batch fields", [(x.dtype.name, x.size, m) for x, m in zip(xs, modes)])
    ts = [torch.from_numpy(x).to(device) for x in xs]
    bc = Z.BatchCodec(list(zip(ts, modes, bounds)), 128)
    return bc, xs, ebs, oracle_of(xs, ebs)


@pytest.mark.parametrize("shape", [(16, 4), (8, 4), (8, 2), (8, 1), (4, 2), (4, 1)])
def test_batch_shapes_match_oracle(batch, shape):
    bc, xs, ebs, want = batch
    say("shape", shape)
    for d in bc.decs:
        d.zero_()
    bc.set_shape(*shape)
    bc.compress()
    bc.decompress()
    check_batch(bc, xs, ebs, want, f"shape {shape}")


def test_batch_per_field_mode(batch):
    bc, xs, ebs, want = batch
    say("per-field launches")
    for o in bc.outs:
        o.zero_()
    for d in bc.decs:
        d.zero_()
    bc.compress(mode=Z.BATCH_PER_FIELD)
    bc.results()
    bc.decompress(mode=Z.BATCH_PER_FIELD)
    check_batch(bc, xs, ebs, want, "per field")


def test_batch_exact_decode(batch):
    bc, xs, ebs, want = batch
    say("exact decode")
    bc.compress()
    bc.decompress(flags=Z.DECODE_EXACT)
    import torch
    torch.cuda.synchronize()
    res = bc.results()
    for i, x in enumerate(xs):
        s = want[i][0]
        assert bits_equal(bc.decs[i].cpu().numpy(), O.decompress(s, x.size, x.dtype, ref_quirk=False))
        if x.dtype == np.float32:
            err = np.max(np.abs(bc.decs[i].cpu().numpy().astype(np.float64) - x))
            eb = np.array([res[i, 2]], np.uint32).view(np.float32)[0]
            assert err <= eb, (i, err, eb)


def test_batch_tune_then_compress(batch):
    bc, xs, ebs, want = batch
    lv, ipt, ms = bc.tune(reps=2)
    say("tuned", lv, ipt, ms)
    assert (lv, ipt) in {(8, 4), (8, 2), (8, 1), (4, 2), (4, 1)} and ms > 0
    for o in bc.outs:
        o.zero_()
    bc.compress()
    bc.decompress()
    check_batch(bc, xs, ebs, want, "tuned")


def test_batch_graph_replay(device):
    """Capture compress + decompress once, replay three times with eager work
    (clearing every output) between replays, check every replay."""
    import torch
    xs, modes, bounds, ebs = make_fields(7, count=9)
    say("graph fields", [(x.dtype.name, x.size) for x in xs])
    ts = [torch.from_numpy(x).to(device) for x in xs]
    bc = Z.BatchCodec(list(zip(ts, modes, bounds)), 128)
    want = oracle_of(xs, ebs)
    bc.compress()          # eager pass first: code objects loaded, launch shapes known
    bc.decompress()
    check_batch(bc, xs, ebs, want, "eager")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bc.compress()
        bc.decompress()
    for rep in range(3):
        for o in bc.outs:
            o.zero_()
        for d in bc.decs:
            d.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        check_batch(bc, xs, ebs, want, f"replay {rep}")


@pytest.mark.parametrize("bs", [8, 64, 256])
def test_batch_block_sizes(device, bs):
    import torch
    xs, modes, bounds, ebs = make_fields(11 + bs, count=6)
    say("block size", bs, [(x.dtype.name, x.size) for x in xs])
    ts = [torch.from_numpy(x).to(device) for x in xs]
    bc = Z.BatchCodec(list(zip(ts, modes, bounds)), bs)
    bc.compress()
    bc.decompress()
    check_batch(bc, xs, ebs, oracle_of(xs, ebs, bs), f"bs {bs}")


def test_batch_without_decode_buffers(device):
    import torch
    xs, modes, bounds, ebs = make_fields(3, count=4)
    ts = [torch.from_numpy(x).to(device) for x in xs]
    bc = Z.BatchCodec(list(zip(ts, modes, bounds)), 128, decode=False)
    bc.compress()
    check_batch(bc, xs, ebs, oracle_of(xs, ebs), "no decode")
    with pytest.raises(Z.SZxError):
        bc.decompress()


def test_batch_decode_rejects_damaged_index(batch):
    """A field whose record index points past its stream is reported (-20)
    by the batched decode's per-tile setup; the other fields still decode."""
    import torch
    bc, xs, ebs, want = batch
    bc.set_shape(8, 4)
    bc.compress()
    torch.cuda.synchronize()
    victim = next(i for i, x in enumerate(xs) if x.size >= 3 * 4096)
    ix = Z.lib().szx_amd_batch_index_dev(bc.handle, victim)
    base = ix - bc.ws.data_ptr()   # the index lives in the plan's workspace
    assert 0 <= base and base % 16 == 0
    # index pair 1: the record offset (and constants before) of the second tile
    pair = bc.ws[base + 16:base + 32].view(torch.int64)
    saved = pair.clone()
    pair.copy_(torch.tensor([1 << 40, 0], dtype=torch.int64, device=pair.device))
    say("damaged index of field", victim)
    for d in bc.decs:
        d.zero_()
    bc.decompress()
    torch.cuda.synchronize()
    st = bc.status().cpu().numpy()
    assert st[victim] == -20, st[victim]
    for i in range(len(xs)):
        if i != victim:
            assert st[i] == 0, (i, st[i])
            assert bits_equal(bc.decs[i].cpu().numpy(), want[i][1]), i
    pair.copy_(saved)
    bc.decompress()
    torch.cuda.synchronize()
    assert bc.status().cpu().numpy()[victim] == 0
    assert bits_equal(bc.decs[victim].cpu().numpy(), want[victim][1])
