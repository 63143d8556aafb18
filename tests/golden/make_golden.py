"""Write the golden vectors of the fastMode-2 codec with the compiled reference.

Runs the reference's own CPU codec (oracle/_ref/libszx_ref.so, built from the
reference sources by `make -C oracle ref`) on small seeded inputs and stores,
per case, the input, the reference stream and the reference decode as a
compressed .npz (plain arrays, no pickles).  Cases cover both word widths,
ABS and REL bounds, every reqBytes branch the data reaches, constant blocks,
remainder blocks and block sizes 64/128/256 (256 is compressed only: the
reference decoder reads its block-size byte 0 as 0 and divides by it).

    python tests/golden/make_golden.py        # fastMode-2 vectors
    python tests/golden/make_golden.py ra     # random-access containers (ra*.npz)
    python tests/golden/make_golden.py nb     # non-blocked streams (nb*.npz)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from szx_amd.fields import grf_field, sinusoid_field  # noqa: E402


def cases():
    # scaled-down CPU workload: sinusoids + 25% zero box, ABS eb = 1e-3 * range
    x = sinusoid_field((20, 50, 50), seed=1, n_waves=32, kmax=16, slope=2.0, noise=1e-3,
                       n_boxes=1, box_frac=0.25, box_value="zero")
    rng = float(np.float32(np.float32(x.max()) - np.float32(x.min())))
    yield "f32_sin_abs1e-3", x, dict(mode="abs", eb=float(np.float32(rng * np.float32(1e-3))), bs=128)
    # scaled-down single-GPU workload: 64 waves, two mean boxes, REL sweep
    x = sinusoid_field((32, 32, 33), seed=2, n_waves=64, kmax=32, slope=2.5, noise=1e-4,
                       n_boxes=2, box_frac=0.30, box_value="mean")
    for rel in (1e-2, 1e-3, 1e-4):
        yield f"f32_sin_rel{rel:g}", x, dict(mode="rel", rel=rel, bs=128)
    # fp64 GRF with 1e-300 entries, REL sweep x block sizes
    x = grf_field((96, 100), seed=3)
    for rel in (1e-4, 1e-6, 1e-8):
        for bs in (64, 128, 256):
            yield f"f64_grf_rel{rel:g}_bs{bs}", x, dict(mode="rel_blocked", rel=rel, bs=bs)
    # every fp32 reqBytes value and odd block sizes on a smooth ramp + noise
    r = np.random.default_rng(4)
    base = (np.sin(np.arange(3000) * 0.003) * 50 + r.standard_normal(3000) * 1e-3).astype(np.float32)
    for ebexp in (-1, -4, -7, -10):
        for bs in (8, 100, 128):
            eb = float(np.float32(10.0 ** ebexp))
            yield f"f32_ramp_eb1e{ebexp}_bs{bs}", base, dict(mode="abs", eb=eb, bs=bs)
    # fp64: reqBytes 2..8 via a sweep of bounds
    b64 = (np.cumsum(r.standard_normal(2500)) * 1e-2).astype(np.float64)
    for ebexp in (-2, -5, -8, -11, -14, -17):
        eb = float(np.float32(10.0 ** ebexp))
        yield f"f64_walk_eb1e{ebexp}", b64, dict(mode="abs", eb=eb, bs=128)


def ra_cases():
    """Random-access containers (fastMode 3: big-endian medians; fastMode 4:
    native medians, fp64 float-radius statistics) through
    SZx_fast_compress_args (blockSize 128)."""
    x = sinusoid_field((32, 32, 33), seed=2, n_waves=64, kmax=32, slope=2.5, noise=1e-4,
                       n_boxes=2, box_frac=0.30, box_value="mean")
    g = grf_field((96, 100), seed=3)
    for fm in (3, 4):
        yield f"ra{fm}_f32_sin_rel1e-3", x, dict(mode=1, rel=1e-3, fm=fm)
        yield f"ra{fm}_f32_sin_abs1e-5", x, dict(mode=0, eb=1e-5, fm=fm)
        yield f"ra{fm}_f64_grf_rel1e-6", g, dict(mode=1, rel=1e-6, fm=fm)
    # fp64 blocks whose radius exceeds the bound in double but not once rounded
    # to float: constant under fastMode 4 only (szx_double.cc:1120-1135)
    r = np.random.default_rng(6)
    flip = np.repeat(r.uniform(-4, 4, 40), 128)
    flip[::128] += 2.0 + 2.0 ** -30        # radius 1 + 2^-31 against eb 1
    flip[64::128] += r.uniform(0, 1e-3, 40)
    for fm in (3, 4):
        yield f"ra{fm}_f64_radius_flip", flip, dict(mode=0, eb=1.0, fm=fm)


def main_ra():
    manifest = []
    for name, x, prm in ra_cases():
        x = np.ascontiguousarray(x).ravel()
        eb = prm["eb"] if prm["mode"] == 0 else O.rel_eb(x, prm["rel"])
        if len(O.compress_ra(x, eb, 128, prm["fm"])) > x.size * x.itemsize:
            print(f"{name}: skipped (stream would exceed the reference's N*W buffer)")
            continue
        stream = O.ref_compress_args(x, prm["mode"], abs_eb=prm.get("eb", 0.0), rel=prm.get("rel", 0.0),
                                     fast_mode=prm["fm"])
        dec = O.ref_decompress(stream, x.size, x.dtype, fast_mode=prm["fm"])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), x=x,
                            stream=np.frombuffer(stream, np.uint8), decoded=dec,
                            eb=np.float32(eb), bs=np.int32(128), fast_mode=np.int32(prm["fm"]),
                            rel=np.float32(prm.get("rel", 0.0)))
        manifest.append(f"{name}\tn={x.size}\t{x.dtype}\tfastMode={prm['fm']}\teb={eb!r}\tbytes={len(stream)}")
        print(manifest[-1])
    with open(os.path.join(HERE, "MANIFEST_RA.txt"), "w") as f:
        f.write("# produced by make_golden.py ra from the compiled reference codec\n")
        f.write("\n".join(manifest) + "\n")


def nb_cases():
    """Non-blocked streams (fastMode 1) through SZx_fast_compress_args."""
    x = sinusoid_field((32, 32, 33), seed=2, n_waves=64, kmax=32, slope=2.5, noise=1e-4,
                       n_boxes=2, box_frac=0.30, box_value="mean")
    yield "nb_f32_sin_rel1e-3", x, dict(mode=1, rel=1e-3)
    yield "nb_f32_sin_abs1e-5", x, dict(mode=0, eb=1e-5)
    g = grf_field((96, 100), seed=3)
    yield "nb_f64_grf_rel1e-6", g, dict(mode=1, rel=1e-6)
    r = np.random.default_rng(7)
    walk = np.cumsum(r.standard_normal(6000)) * 1e-2
    for ebexp in (-5, -8, -11):
        yield f"nb_f64_walk_eb1e{ebexp}", walk, dict(mode=0, eb=float(np.float32(10.0 ** ebexp)))


def main_nb():
    manifest = []
    for name, x, prm in nb_cases():
        x = np.ascontiguousarray(x).ravel()
        eb = prm["eb"] if prm["mode"] == 0 else O.rel_eb(x, prm["rel"])
        if len(O.compress_nb(x, eb)) > x.size * x.itemsize:
            print(f"{name}: skipped (stream would exceed the reference's N*W buffer)")
            continue
        stream = O.ref_compress_args(x, prm["mode"], abs_eb=prm.get("eb", 0.0), rel=prm.get("rel", 0.0),
                                     fast_mode=1)
        dec = O.ref_decompress(stream, x.size, x.dtype, fast_mode=1)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), x=x,
                            stream=np.frombuffer(stream, np.uint8), decoded=dec,
                            eb=np.float32(eb), fast_mode=np.int32(1),
                            rel=np.float32(prm.get("rel", 0.0)))
        manifest.append(f"{name}\tn={x.size}\t{x.dtype}\tfastMode=1\teb={eb!r}\tbytes={len(stream)}")
        print(manifest[-1])
    with open(os.path.join(HERE, "MANIFEST_NB.txt"), "w") as f:
        f.write("# produced by make_golden.py nb from the compiled reference codec\n")
        f.write("\n".join(manifest) + "\n")


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libszx_ref.so is missing: make -C oracle ref")
    if sys.argv[1:] == ["ra"]:
        main_ra()
        return
    if sys.argv[1:] == ["nb"]:
        main_nb()
        return
    manifest = []
    for name, x, prm in cases():
        x = np.ascontiguousarray(x).ravel()
        bs = prm["bs"]
        if prm["mode"] == "abs":
            eb = prm["eb"]
        elif prm["mode"] == "rel":
            eb = O.rel_eb(x, prm["rel"])
        else:
            eb = O.rel_eb(x, prm["rel"])
        # The reference writes into a buffer of N*W bytes and overflows it
        # when the stream is larger (szx_float.cc:1065-1067): such cases are
        # not run through it.
        if len(O.compress(x, eb, bs)) > x.size * x.itemsize:
            print(f"{name}: skipped (stream would exceed the reference's N*W buffer)")
            continue
        if prm["mode"] == "rel":
            stream = O.ref_compress_args(x, mode=1, rel=prm["rel"])
        else:
            stream = O.ref_compress(x, eb, bs)
        dec = O.ref_decompress(stream, x.size, x.dtype) if bs != 256 else np.zeros(0, x.dtype)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), x=x,
                            stream=np.frombuffer(stream, np.uint8), decoded=dec,
                            eb=np.float32(eb), bs=np.int32(bs),
                            rel=np.float32(prm.get("rel", 0.0)))
        manifest.append(f"{name}\tn={x.size}\t{x.dtype}\tbs={bs}\teb={eb!r}\tbytes={len(stream)}")
        print(manifest[-1])
    with open(os.path.join(HERE, "MANIFEST.txt"), "w") as f:
        f.write("# produced by make_golden.py from the compiled reference codec\n")
        f.write("\n".join(manifest) + "\n")


if __name__ == "__main__":
    main()
