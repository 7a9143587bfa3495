"""Pins the oracle's threshold selection, quantizers and Lloyd-Max training
(CPU only).

  * uniform / log / codebook quantizers and dequantizers: bit-for-bit against
    the reference's own quant.h compiled in this container
    (oracle/_ref/libref.so), on random values and every bin edge region;
  * Lloyd: against the reference's own lloyds.cpp (same library), and the
    committed golden vectors tests/golden/lloyd_*.npz that
    tests/golden/make_golden.py wrote from it;
  * threshold: brute force np.sort plus the float-index rule.
"""
import ctypes as C
import os

import numpy as np
import pytest

import _oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _ref_or_skip():
    r = O.ref()
    if r is None:
        pytest.skip("reference sources not compiled here")
    return r


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_threshold_float_index_rule():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100_003) * 10).astype(np.float32)
    for frac in [0.0, 0.5, 0.9, 0.95, 0.999]:
        t = O.thresh(x, frac)
        srt = np.sort(np.abs(x))
        idx = int(np.float32(frac) * np.float32(len(x)))
        assert t["idx"] == idx
        want = srt[idx] if srt[idx] != 0 else np.float32(1e-16)
        assert t["threshold"] == want
        assert t["nonzero"] == len(x) - idx - 1
        assert np.array_equal(t["sorted"], srt)
        assert t["max_abs"] == srt[-1]
        assert t["min"] == x.min() and t["max"] == x.max()


def test_threshold_float_index_differs_from_exact():
    # 512^3 with frac 0.95: float product 127506840 vs exact 127506841
    n = 512 ** 3
    assert int(np.float32(0.95) * np.float32(n)) == 127506840


def test_threshold_max_starts_at_flt_min():
    # Traverse starts max at FLT_MIN (thresh_cpu.cc:68)
    x = -np.abs(np.arange(1, 50, dtype=np.float32))
    t = O.thresh(x, 0.5)
    assert t["max"] == np.finfo(np.float32).tiny


def _edge_values(q_thresh, q_max, n, rng):
    x = np.concatenate([
        rng.uniform(-q_max, q_max, n),
        rng.uniform(-3 * q_thresh, 3 * q_thresh, n),
        np.exp(rng.uniform(np.log(q_thresh), np.log(q_max), n)),
        -np.exp(rng.uniform(np.log(q_thresh), np.log(q_max), n)),
        [0.0, -0.0, q_thresh, -q_thresh, q_max, -q_max, 2 * q_max, -2 * q_max],
    ]).astype(np.float32)
    # neighbours of the threshold and max in float steps
    t32, m32 = np.float32(q_thresh), np.float32(q_max)
    for v in (t32, m32):
        for k in range(-3, 4):
            y = np.float32(v)
            for _ in range(abs(k)):
                y = np.nextafter(y, np.float32(np.inf if k > 0 else 0))
            x = np.append(x, [y, -y]).astype(np.float32)
    return x


@pytest.mark.parametrize("alg", [O.UNIFORM, O.LOG])
@pytest.mark.parametrize("bins,thr,mx", [(256, 1.16175, 421519.75),
                                         (1024, 31.5, 40000.0),
                                         (255, 0.5, 10.0), (7, 1.0, 9.0),
                                         (2, 0.25, 4.0)])
def test_quantizer_vs_reference(alg, bins, thr, mx):
    r = _ref_or_skip()
    rng = np.random.default_rng(bins)
    x = _edge_values(thr, mx, 20000, rng)
    q = O.quantizer(alg, bins, thr, mx)
    got = O.quantize(q, x)
    want = np.empty(len(x), np.int32)
    r.ref_quantize(alg, bins, C.c_float(thr), C.c_float(mx), _p(x), _p(want), len(x))
    assert np.array_equal(got, want)
    b = np.arange(bins, dtype=np.int32)
    dq = O.dequantize(q, b)
    wdq = np.empty(bins, np.float32)
    r.ref_dequantize(alg, bins, C.c_float(thr), C.c_float(mx), _p(b), _p(wdq), bins)
    assert np.array_equal(dq.view(np.uint32), wdq.view(np.uint32))


def test_codebook_quantizer_vs_reference():
    r = _ref_or_skip()
    rng = np.random.default_rng(2)
    bounds = np.sort(rng.standard_normal(255).astype(np.float32) * 10)
    values = np.sort(rng.standard_normal(256).astype(np.float32) * 10)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 12,
                        bounds, np.nextafter(bounds, np.float32(-np.inf)),
                        np.nextafter(bounds, np.float32(np.inf))]).astype(np.float32)
    q = O.quantizer(O.LLOYD, 256, 0, 0, bounds, values)
    got = O.quantize(q, x)
    want = np.empty(len(x), np.int32)
    r.ref_quantize_codebook(_p(values), 256, _p(bounds), 255, _p(x), _p(want), len(x))
    assert np.array_equal(got, want)
    # lower_bound semantics with the lastBoundary rule (quant.h:515-526)
    np_want = np.searchsorted(bounds, x, side="left")
    np_want[x >= bounds[-1]] = 255
    assert np.array_equal(got, np_want)


def _lloyd_points(n, seed):
    rng = np.random.default_rng(seed)
    x = np.abs(rng.laplace(0, 30, n * 12)).astype(np.float32)
    x = np.sort(x)
    return x[-n:]   # a sorted tail like sorted[idx+1 ..]


@pytest.mark.parametrize("n,cs,seed", [(2000, 7, 1), (5000, 31, 2), (3000, 127, 3)])
def test_lloyd_vs_reference(n, cs, seed):
    r = _ref_or_skip()
    pts = _lloyd_points(n, seed)
    thr = pts[0]
    cb0 = np.empty(cs, np.float32)
    O.lib().or_initial_lloyd_codebook(_p(cb0), cs, C.c_float(thr), C.c_float(pts[-1]))
    cb, part, it, dist, rel = O.lloyd(pts, cb0)
    rcb = cb0.copy()
    rpart = np.empty(cs - 1, np.float32)
    rd, rr, ri = C.c_float(), C.c_float(), C.c_int()
    r.ref_lloyd(_p(pts), len(pts), _p(rcb), cs, _p(rpart), C.byref(rd), C.byref(rr),
                C.c_float(1e-6), C.byref(ri))
    assert it == ri.value
    assert np.array_equal(cb.view(np.uint32), rcb.view(np.uint32))
    assert np.array_equal(part.view(np.uint32), rpart.view(np.uint32))
    assert dist == np.float32(rd.value) and rel == np.float32(rr.value)


def test_lloyd_iteration_cap():
    pts = _lloyd_points(4000, 9)
    cb0 = np.geomspace(pts[0], pts[-2], 15).astype(np.float32)
    full = O.lloyd(pts, cb0)
    capped_big = O.lloyd(pts, cb0, max_iters=full[2] + 5)
    assert np.array_equal(full[0], capped_big[0]) and capped_big[2] == full[2]
    if full[2] >= 2:
        one = O.lloyd(pts, cb0, max_iters=1)
        assert one[2] == 1
    zero = O.lloyd(pts, cb0, max_iters=0)
    assert zero[2] == 0 and np.array_equal(zero[0], cb0)


def test_lloyd_golden_vectors():
    path = os.path.join(GOLDEN, "lloyd_tail.npz")
    if not os.path.exists(path):
        pytest.skip("golden vectors not generated")
    g = np.load(path)
    cb, part, it, dist, rel = O.lloyd(g["points"], g["codebook0"])
    assert it == int(g["iterations"])
    assert np.array_equal(cb.view(np.uint32), g["codebook"].view(np.uint32))


def test_set_bins_from_codebook_layout():
    cb = np.array([1.0, 2.0, 4.0], np.float32)
    b, v, _ = None, None, None
    vals = np.empty(8, np.float32)
    bnds = np.empty(8, np.float32)
    n = O.lib().or_set_bins_from_codebook(_p(vals), _p(bnds), 8, _p(cb), 3,
                                          C.c_float(0.5), C.c_float(-3.0), C.c_float(10.0))
    assert n == 8
    # even bin count, max >= |min| -> max gets the extra top bin
    assert list(vals) == [-4, -2, -1, 0, 1, 2, 4, 10]
    assert list(bnds[:7]) == [-3, -1.5, -0.5, 0.5, 1.5, 3, 7]


def test_quantizer_golden_vectors():
    path = os.path.join(GOLDEN, "quant_256.npz")
    if not os.path.exists(path):
        pytest.skip("golden vectors not generated")
    g = np.load(path)
    for alg, name in ((O.UNIFORM, "uniform"), (O.LOG, "log")):
        q = O.quantizer(alg, 256, g["threshold"], g["max"])
        assert np.array_equal(O.quantize(q, g["x"]), g[name])
