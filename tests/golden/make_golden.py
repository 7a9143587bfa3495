"""Writes the golden vectors under tests/golden/ from the reference's own
protobuf-free code (huffman.cc, lloyds.cpp, quant.h), compiled in this
container into oracle/_ref/libref.so by `make -C oracle ref`.

The vectors are data only (inputs and the reference's outputs); they let the
CPU tests keep pinning the oracle where the reference tree is absent.
Run:  python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def main():
    r = O.ref()
    if r is None:
        raise SystemExit("build oracle/_ref/libref.so first (make -C oracle ref)")

    # Lloyd on a sorted Laplace tail (the shape computeLloydQuantization sees)
    rng = np.random.default_rng(20240601)
    x = np.sort(np.abs(rng.laplace(0, 30, 60000)).astype(np.float32))
    pts = x[-6000:-1].copy()
    cs = 63
    cb0 = np.empty(cs, np.float32)
    O.lib().or_initial_lloyd_codebook(_p(cb0), cs, C.c_float(x[-6001]),
                                      C.c_float(x[-2]))
    cb = cb0.copy()
    part = np.empty(cs - 1, np.float32)
    d, rel, it = C.c_float(), C.c_float(), C.c_int()
    r.ref_lloyd(_p(pts), len(pts), _p(cb), cs, _p(part), C.byref(d), C.byref(rel),
                C.c_float(1e-6), C.byref(it))
    np.savez_compressed(os.path.join(HERE, "lloyd_tail.npz"), points=pts,
                        codebook0=cb0, codebook=cb, partition=part,
                        iterations=np.int32(it.value))

    # Huffman with zero-bin runs
    bins, zb, n = 256, 127, 20000
    data = np.clip(rng.geometric(0.2, n) * rng.choice([-1, 1], n) + zb, 0, 255).astype(np.int32)
    pos = 0
    while pos < n:
        run = int(rng.integers(1, 400)) if rng.random() < 0.6 else 0
        data[pos:pos + run] = zb
        pos += run + int(rng.integers(1, 30))
    counts = np.bincount(data, minlength=bins).astype(np.int32)
    words = np.zeros(n * 2, np.uint32)
    dec = np.zeros(4 * bins + 8, np.int32)
    dlen, bits = C.c_int(), C.c_longlong()
    cl = np.zeros(bins + 1, np.int32)
    nw = r.ref_huffman(_p(counts), bins, zb, _p(data), n, _p(words),
                       C.c_longlong(len(words)), _p(dec), C.byref(dlen), len(dec),
                       _p(cl), C.byref(bits))
    np.savez_compressed(os.path.join(HERE, "huffman_bins.npz"), data=data,
                        bins=np.int32(bins), zero_bin=np.int32(zb),
                        words=words[:nw], decode=dec[:dlen.value],
                        bits=np.int64(bits.value))

    # log / uniform quantizer edges at the 2048^2 sample's parameters
    # (README.md:77: threshold 1.161753893, max 421519.75)
    q = np.concatenate([np.exp(rng.uniform(-2, 13, 4000)),
                        -np.exp(rng.uniform(-2, 13, 4000))]).astype(np.float32)
    outs = {}
    for alg, name in ((1, "uniform"), (2, "log")):
        o = np.empty(len(q), np.int32)
        r.ref_quantize(alg, 256, C.c_float(1.161753893), C.c_float(421519.75),
                       _p(q), _p(o), len(q))
        outs[name] = o
    np.savez_compressed(os.path.join(HERE, "quant_256.npz"), x=q,
                        threshold=np.float32(1.161753893),
                        max=np.float32(421519.75), **outs)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
