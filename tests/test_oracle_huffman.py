"""Pins the oracle's Huffman coder, duplicate-string encoding, bit stream and
decoder (CPU only).

  * bit-stream KATs from test_bit_stream.cc:93-127 ("0305070f",
    "030507ffffffff00") and the reference BitStreamWriter itself;
  * encoded words, decode table and code lengths bit-for-bit against the
    reference's huffman.cc compiled here (oracle/_ref/libref.so) and the
    committed golden vectors tests/golden/huffman_*.npz;
  * prefix-free codes, Kraft equality, optimal total length (an independent
    textbook Huffman cost), and decode(encode(x)) == x.
"""
import ctypes as C
import heapq
import os

import numpy as np
import pytest

import _oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _pack(values, nbits):
    v = np.array(values, np.uint32)
    n = np.array(nbits, np.int32)
    w = np.zeros(64, np.uint32)
    k = O.lib().or_bitstream_pack(_p(v), _p(n), len(v), _p(w), 64)
    return w[:k]


def test_bitstream_kats():
    assert _pack([3, 5, 7, 15], [8] * 4).tobytes().hex() == "0305070f"
    got = _pack([3, 5, 7, 0xffffffff], [8, 8, 8, 32]).tobytes().hex()
    assert got == "030507ffffffff00"


def test_bitstream_vs_reference_writer():
    r = O.ref()
    if r is None:
        pytest.skip("reference sources not compiled here")
    rng = np.random.default_rng(4)
    nb = rng.integers(0, 33, 2000).astype(np.int32)
    vals = rng.integers(0, 2 ** 32, 2000, dtype=np.uint64).astype(np.uint32)
    w = np.zeros(4000, np.uint32)
    k = r.ref_bitstream(_p(vals), _p(nb), len(vals), _p(w), 4000)
    assert np.array_equal(_pack_big(vals, nb), w[:k])


def _pack_big(vals, nb):
    w = np.zeros(len(vals) + 8, np.uint32)
    k = O.lib().or_bitstream_pack(_p(vals), _p(nb), len(vals), _p(w), len(w))
    return w[:k]


def _bins(n, bins, zero_frac, seed):
    rng = np.random.default_rng(seed)
    zb = (bins - 1) // 2
    x = rng.geometric(0.15, n) * rng.choice([-1, 1], n) + zb
    x = np.clip(x, 0, bins - 1).astype(np.int32)
    # long zero-bin runs of random length, like thresholded coefficients
    pos = 0
    while pos < n:
        run = int(rng.integers(1, 3000)) if rng.random() < zero_frac else 0
        x[pos:pos + run] = zb
        pos += run + int(rng.integers(1, 40))
    return x, zb


def _ref_huffman(data, bins, zb):
    r = O.ref()
    counts = np.bincount(data, minlength=bins).astype(np.int32)
    words = np.zeros(len(data) * 2 + 64, np.uint32)
    dec = np.zeros(4 * bins + 8, np.int32)
    dlen, bits = C.c_int(), C.c_longlong()
    cl = np.zeros(bins + 1, np.int32)
    n = r.ref_huffman(_p(counts), bins, zb, _p(data), len(data), _p(words),
                      C.c_longlong(len(words)), _p(dec), C.byref(dlen), len(dec),
                      _p(cl), C.byref(bits))
    return words[:n], dec[:dlen.value], cl, bits.value


@pytest.mark.parametrize("n,bins,zf,seed,noz", [(50_000, 256, 0.7, 1, False),
                                                (20_000, 1024, 0.3, 2, False),
                                                (10_000, 16, 0.9, 3, False),
                                                (30_000, 256, 0.7, 4, True)])
def test_huffman_vs_reference(n, bins, zf, seed, noz):
    if O.ref() is None:
        pytest.skip("reference sources not compiled here")
    data, zb = _bins(n, bins, zf, seed)
    if noz:
        zb = -1
    h = O.Huffman(data, bins, zb)
    words, bits = h.encode()
    rw, rdec, rcl, rbits = _ref_huffman(data, bins, zb)
    assert bits == rbits
    assert np.array_equal(words, rw)
    assert np.array_equal(h.decode_table(), rdec)
    nv = bins + (1 if zb >= 0 else 0)
    assert np.array_equal(h.code_lengths()[:nv], rcl[:nv] - np.r_[np.zeros(bins, np.int32), [3] * (nv - bins)])


def test_huffman_counts_path_equals_scan_path():
    data, zb = _bins(40_000, 256, 0.6, 7)
    counts = np.bincount(data, minlength=256).astype(np.int32)
    a = O.Huffman(data, 256, zb)
    b = O.Huffman(data, 256, zb, counts=counts)
    assert np.array_equal(a.counts(), b.counts())
    assert np.array_equal(a.encode()[0], b.encode()[0])


def _textbook_cost(counts):
    h = [int(c) for c in counts if c > 0]
    if len(h) == 1:
        return 0
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a, b = heapq.heappop(h), heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


def test_huffman_optimal_prefix_free_roundtrip():
    data, zb = _bins(60_000, 256, 0.5, 11)
    h = O.Huffman(data, 256, zb)
    counts = h.counts()
    lens = h.code_lengths()
    used = counts > 0
    # Kraft equality for a full binary tree
    assert abs(np.sum(2.0 ** -lens[used]) - 1.0) < 1e-12
    # optimal weighted length
    assert int(np.sum(lens[used] * counts[used])) == _textbook_cost(counts)
    words, bits = h.encode()
    back = O.huff_decode(h.decode_table(), 256, zb, words, len(data))
    assert np.array_equal(back, data)


def test_dup_run_statistics_closed_form():
    # a run of length L > 8 uses greedy powers 2^(e+3) <= 1024, rest singles
    zb = 5
    for L, (dups, singles) in {8: (0, 8), 9: (1, 1), 16: (1, 0), 17: (1, 1),
                               205: (3, 5), 1024: (1, 0), 1025: (1, 1),
                               3000: (7, 8)}.items():
        data = np.full(L, zb, np.int32)
        data = np.concatenate([[0], data, [0]]).astype(np.int32)
        c = O.Huffman(data, 16, zb).counts()
        assert c[16] == dups and c[zb] == singles, (L, c[16], c[zb])


def test_huffman_golden_vectors():
    path = os.path.join(GOLDEN, "huffman_bins.npz")
    if not os.path.exists(path):
        pytest.skip("golden vectors not generated")
    g = np.load(path)
    h = O.Huffman(g["data"], int(g["bins"]), int(g["zero_bin"]))
    words, bits = h.encode()
    assert np.array_equal(words, g["words"])
    assert np.array_equal(h.decode_table(), g["decode"])


def test_total_encoded_length_matches_written_payload():
    """totalEncodedLengthBytes (huffman.cc:241-252) against the bytes the
    coder actually writes: without dup strings every bit is a code bit, so
    the two agree exactly; with them the payload also holds 3 length bits
    per dup string, so it is at least as long."""
    rng = np.random.default_rng(4)
    vol = (rng.standard_normal((16, 24, 32)) * 30).astype(np.float32)
    for zeros in (False, True):
        res = O.compress(vol, 2, 2, 256, [3, 3, 3], 0.9, compress_zeros=zeros)
        t = res["res"].threshold
        size, _, _, _ = O.opt_test(vol, [3, 3, 3], 2, 100, zeros, t, 256, 2)
        _, cubes, _, _ = __import__("_cubeproto").parse_cube_file(res["file"])
        payload = len(cubes[0][2])
        if zeros:
            assert 0 < size <= payload
        else:
            assert size == payload
