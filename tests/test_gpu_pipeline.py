"""GPU parity tests: libwv_hip.so on an MI355X against the CPU oracle.

The bar (integer / byte / index work bit-exact; the DWT is bit-exact too,
because the kernels keep the reference's operation order with contraction
off):
  * forward and inverse 3-D DWT coefficients: bit-identical to the oracle;
  * threshold, index, nonzero count, min / max / max |x|: identical;
  * bins: identical; histogram: identical to np.bincount of the bins;
  * Lloyd codebook (iteration-capped as the oracle): identical;
  * the .cube file of the one-call compress: byte-identical to the oracle's.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wv():
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    import wavelets_amd
    wavelets_amd.lib()
    return wavelets_amd


def _vol(shape_zyx, dtype, seed):
    rng = np.random.default_rng(seed)
    z, y, x = shape_zyx
    g = np.mgrid[0:z, 0:y, 0:x].astype(np.float32)
    base = (np.sin(g[2] * 0.21) * 40 + np.cos(g[1] * 0.13) * 30 + g[0] * 0.7)
    noise = rng.standard_normal(shape_zyx).astype(np.float32) * 12
    v = base + noise + 100
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype == np.int32:
        return np.clip(v * 10, 0, 4095).astype(np.int32)
    return v.astype(np.float32)


def _oracle_coef(vol, wavelet, steps):
    orig = [vol.shape[2], vol.shape[1], vol.shape[0]]
    st = O.set_steps(steps, orig)
    pad = O.pad_convert(vol, st)
    d, size = O.dwt3d(pad.ravel(), [pad.shape[2], pad.shape[1], pad.shape[0]], wavelet, st)
    return d, size, st


CASES = [
    ((16, 16, 16), np.float32, [3, 3, 3]),
    ((16, 64, 512), np.uint8, [3, 3, 3]),     # thin: merged rows in the z pass
    ((32, 32, 32), np.float32, [3, 3, 3]),    # short lines, 8 elements per thread
    ((64, 64, 64), np.float32, [3, 3, 3]),
    ((12, 48, 80), np.int32, [3, 3, 2]),      # padded, thin
    ((8, 32, 64), np.uint8, [-1, -1, -1]),
    ((5, 24, 40), np.int32, [3, 2, 1]),
    ((1, 96, 128), np.uint8, [-1, -1, 0]),
    ((1, 64, 2048), np.float32, [11, 6, 0]),
    ((33, 17, 70), np.float32, [2, 0, 3]),
    ((3, 7, 9), np.float32, [0, 2, 0]),
    ((20, 30, 40), np.float32, [0, 0, 0]),
    ((64, 64, 64), np.int32, [4, 4, 4]),
    ((16, 40, 48), np.float32, [1, 1, 1]),
]


@pytest.mark.parametrize("wavelet", [1, 2])
@pytest.mark.parametrize("shape,dtype,steps", CASES)
def test_dwt_forward_bit_exact(wv, wavelet, shape, dtype, steps):
    print("case", shape, dtype.__name__, steps, wavelet, flush=True)
    vol = _vol(shape, dtype, 1)
    want, size, st = _oracle_coef(vol, wavelet, steps)
    coef, g, stats = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), steps, wavelet)
    torch.cuda.synchronize()
    assert list(g.out_size) == size and list(g.steps) == st
    got = coef.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
        np.abs(got - want).max()


@pytest.mark.parametrize("wavelet", [1, 2])
@pytest.mark.parametrize("shape,dtype,steps", CASES[:11])
def test_dwt_inverse_bit_exact(wv, wavelet, shape, dtype, steps):
    vol = _vol(shape, dtype, 2)
    want_c, size, st = _oracle_coef(vol, wavelet, steps)
    g = wv.geom([shape[2], shape[1], shape[0]], steps, wavelet, wv.FLOAT32)
    back = wv.dwt3d_inverse(torch.from_numpy(want_c).cuda(), g)
    torch.cuda.synchronize()
    want, size2 = O.dwt3d(want_c, size, wavelet, st, inverse=True)
    got = back.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("frac", [0.0, 0.5, 0.9, 0.95, 0.999, 1.0])
def test_threshold_matches_oracle(wv, frac):
    vol = _vol((32, 48, 64), np.float32, 3)
    coef, g, stats = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [3, 3, 3], 2)
    res_d, r = wv.threshold(coef, frac, stats)
    c = coef.cpu().numpy()
    t = O.thresh(c, frac)
    assert r.max_abs == t["max_abs"]
    assert r.min_val == t["min"] and r.max_val == t["max"]
    if frac < 1.0:
        assert r.thresh_idx == t["idx"]
        assert r.threshold == t["threshold"]
        assert r.nonzero_count == t["nonzero"]
    else:
        assert r.threshold == t["max_abs"] and r.nonzero_count == 0


def test_threshold_with_standalone_stats_and_ties(wv):
    rng = np.random.default_rng(5)
    x = np.round(rng.standard_normal(300_001) * 4).astype(np.float32)   # many ties
    xd = torch.from_numpy(x).cuda()
    stats = wv.stats_compute(xd)
    for frac in (0.1, 0.5, 0.77):
        _, r = wv.threshold(xd, frac, stats)
        t = O.thresh(x, frac)
        assert r.threshold == t["threshold"] and r.thresh_idx == t["idx"]


@pytest.mark.parametrize("alg,bins", [(1, 1024), (2, 256), (1, 255), (2, 1023), (2, 7)])
def test_quantize_hist_matches_oracle(wv, alg, bins):
    vol = _vol((32, 32, 64), np.float32, 4)
    coef, g, stats = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [3, 3, 3], 2)
    _, r = wv.threshold(coef, 0.9, stats)
    q = wv.Quantizer(alg, bins, r.threshold, r.max_abs)
    b, counts = q.quantize(coef)
    torch.cuda.synchronize()
    oq = O.quantizer(alg, bins, r.threshold, r.max_abs)
    want = O.quantize(oq, coef.cpu().numpy())
    got = b.cpu().numpy()
    assert np.array_equal(got, want)
    assert np.array_equal(counts.cpu().numpy(), np.bincount(want, minlength=bins))
    dq = q.dequantize(b).cpu().numpy()
    assert np.array_equal(dq.view(np.uint32), O.dequantize(oq, want).view(np.uint32))


@pytest.mark.parametrize("bins", [7, 256, 1023, 4000])
def test_log_quantize_at_bin_edges(wv, bins):
    """Values on and one to three floats either side of every log bin edge
    (and of the threshold and max): the slot-table count (up to ~4000 bins)
    and the logarithm-guess count (4000 bins leave no LDS room for the slot
    tables) both give the oracle's QuantLog bins."""
    thr, mx = np.float32(0.37), np.float32(2345.6)
    oq = O.quantizer(2, bins, thr, mx)
    xs = []
    for scale, off in ((oq.pos_scale, oq.pos_offset), (oq.neg_scale, oq.neg_offset)):
        k = np.arange(-2, bins // 2 + 3, dtype=np.float64)
        e = np.exp((k - off) / scale).astype(np.float32)
        xs.append(e)
    e = np.concatenate(xs + [np.array([thr, mx, mx * 2, 1e-30, 3e38], np.float32)])
    e = e[np.isfinite(e)]
    near = [e]
    up, dn = e.copy(), e.copy()
    for _ in range(3):
        up = np.nextafter(up, np.float32(np.inf))
        dn = np.nextafter(dn, np.float32(0))
        near += [up, dn]
    pos = np.concatenate(near).astype(np.float32)
    rng = np.random.default_rng(bins)
    x = np.concatenate([pos, -pos, np.float32([0.0, -0.0]),
                        (rng.standard_normal(20000) * 300).astype(np.float32)])
    q = wv.Quantizer(2, bins, thr, mx)
    b, counts = q.quantize(torch.from_numpy(x).cuda())
    want = O.quantize(oq, x)
    assert np.array_equal(b.cpu().numpy(), want)
    assert np.array_equal(counts.cpu().numpy()[:bins], np.bincount(want, minlength=bins)[:bins])


def test_quantize_uniform_device_params(wv):
    vol = _vol((16, 32, 32), np.float32, 6)
    coef, g, stats = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [3, 3, 3], 1)
    res_d, r = wv.threshold(coef, 0.95, stats)
    b1, c1 = wv.quantize_uniform_dev(coef, 1024, res_d)
    q = wv.Quantizer(1, 1024, r.threshold, r.max_abs)
    b2, c2 = q.quantize(coef)
    assert torch.equal(b1, b2) and torch.equal(c1, c2)


def test_codebook_quantize(wv):
    rng = np.random.default_rng(8)
    x = (rng.standard_normal(100_003) * 50).astype(np.float32)
    bnd = np.sort(rng.standard_normal(255).astype(np.float32) * 40)
    val = np.sort(rng.standard_normal(256).astype(np.float32) * 40)
    x = np.concatenate([x, bnd]).astype(np.float32)
    q = wv.Quantizer(3, 256, 1.0, 0.0, bnd, val)
    b, counts = q.quantize(torch.from_numpy(x).cuda())
    oq = O.quantizer(3, 256, 1.0, 0.0, bnd, val)
    assert np.array_equal(b.cpu().numpy(), O.quantize(oq, x))


def _compress(wv, vol, wavelet, alg, bins, steps, frac, noz=False, iters=30):
    orig = (C.c_int * 3)(vol.shape[2], vol.shape[1], vol.shape[0])
    opt = wv.Options(wavelet, alg, bins, (C.c_int * 3)(*steps), C.c_float(frac),
                     0 if noz else 1, iters, 1)
    nbytes = wv.lib().wv_compress_work_bytes(orig, C.byref(opt))
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    vd = torch.from_numpy(vol).cuda()
    n = int(np.prod([O.padded_size(list(orig), O.set_steps(steps, list(orig)))]))
    cap = n * 8 + (1 << 20)
    fbuf = np.empty(cap, np.uint8)
    bins_h = np.empty(n, np.int32)
    info = wv.CompressInfo()
    wv.check(wv.lib().wv_compress(C.c_void_p(vd.data_ptr()), wv.dtype_id(vd), orig,
                                  C.byref(opt), C.c_void_p(work.data_ptr()), nbytes,
                                  bins_h.ctypes.data_as(C.c_void_p),
                                  fbuf.ctypes.data_as(C.c_void_p), cap, C.byref(info),
                                  wv._stream()))
    return fbuf[:info.file_bytes].tobytes(), bins_h, info


@pytest.mark.parametrize("shape,dtype,wavelet,alg,bins,steps,frac", [
    ((1, 128, 128), np.uint8, 2, 2, 256, [-1, -1, -1], 0.9),
    ((32, 32, 32), np.float32, 2, 1, 1024, [3, 3, 3], 0.95),
    ((16, 32, 48), np.int32, 1, 2, 256, [3, 3, 3], 0.9),
    ((24, 24, 24), np.float32, 1, 1, 255, [2, 3, 1], 0.5),
])
def test_compress_file_bytes_match_oracle(wv, shape, dtype, wavelet, alg, bins, steps, frac):
    vol = _vol(shape, dtype, 9)
    got, bins_h, info = _compress(wv, vol, wavelet, alg, bins, steps, frac)
    want = O.compress(vol, wavelet, alg, bins, steps, frac, want_bins=True)
    assert np.array_equal(bins_h, want["bins"])
    assert got == want["file"]


@pytest.mark.parametrize("shape,dtype,bins,frac", [((16, 32, 32), np.int32, 256, 0.9),
                                                   ((32, 32, 32), np.float32, 64, 0.8)])
def test_compress_lloyd_matches_oracle(wv, shape, dtype, bins, frac):
    vol = _vol(shape, dtype, 10)
    got, bins_h, info = _compress(wv, vol, 1, 3, bins, [4, 4, 4], frac, iters=30)
    want = O.compress(vol, 1, 3, bins, [4, 4, 4], frac, lloyd_max_iters=30, want_bins=True)
    assert info.lloyd_iters == want["res"].lloyd_iters
    assert np.array_equal(bins_h, want["bins"])
    assert got == want["file"]


@pytest.mark.parametrize("wavelet", [1, 2])
def test_int16_input_equals_widened_int32(wv, wavelet):
    vol = _vol((12, 20, 36), np.int32, 13).astype(np.int16)
    want, size, st = _oracle_coef(vol.astype(np.int32), wavelet, [3, 3, 2])
    coef, g, _ = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [3, 3, 2], wavelet,
                                  with_stats=False)
    got = coef.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    f16, b16, _ = _compress(wv, vol, wavelet, 2, 256, [3, 3, 2], 0.9)
    f32, b32, _ = _compress(wv, vol.astype(np.int32), wavelet, 2, 256, [3, 3, 2], 0.9)
    assert f16 == f32 and np.array_equal(b16, b32)


@pytest.mark.parametrize("zero_frac", [0.6, 0.95])
def test_threshold_zero_ties_need_no_fallback(wv, zero_frac):
    """Most keys exactly zero (and some -0.0): the zeros are counted, not
    listed, so a rank inside them or just above them needs no exact
    fallback, and the result is the oracle's."""
    n = 1_000_003
    rng = np.random.default_rng(21)
    x = (rng.standard_normal(n) * 30).astype(np.float32)
    z = rng.random(n) < zero_frac
    x[z] = 0.0
    x[z & (rng.random(n) < 0.3)] = -0.0
    xd = torch.from_numpy(x).cuda()
    for frac in (0.3, zero_frac - 0.01, zero_frac + 0.005, 0.99):
        _, r = wv.threshold(xd, frac)
        t = O.thresh(x, frac)
        assert r.threshold == t["threshold"] and r.thresh_idx == t["idx"], frac
        assert r.nonzero_count == t["nonzero"], frac
        assert r.n_candidates >= 0, frac      # the sampled window held the rank


def _sample_positions(n, m=16384):
    """The select's sample: one key per stratum [i n/m, (i+1) n/m), at the
    offset (h * width) >> 32 for the 32-bit mixer value h of i (lowbias32
    constants)."""
    M = np.uint64(0xffffffff)
    i = np.arange(m, dtype=np.uint64)
    h = i.copy()
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7feb352d)) & M
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846ca68b)) & M
    h ^= h >> np.uint64(16)
    lo = i * np.uint64(n) // np.uint64(m)
    hi = (i + np.uint64(1)) * np.uint64(n) // np.uint64(m)
    return (lo + ((h * (hi - lo)) >> np.uint64(32))).astype(np.int64)


def test_threshold_exact_fallback_when_sample_misleads(wv):
    """Values at the sampled positions lie far below the rest, so the
    sampled window misses the rank and the exact full-histogram path runs."""
    n = 1 << 20
    rng = np.random.default_rng(12)
    x = (rng.random(n).astype(np.float32) * 1000 + 100).astype(np.float32)
    x[_sample_positions(n)] = rng.random(16384).astype(np.float32) * 0.01
    xd = torch.from_numpy(x).cuda()
    for frac in (0.3, 0.9):
        _, r = wv.threshold(xd, frac)
        t = O.thresh(x, frac)
        assert r.n_candidates < 0            # the fallback ran
        assert r.threshold == t["threshold"] and r.thresh_idx == t["idx"]
        assert r.max_abs == t["max_abs"] and r.min_val == t["min"]


@pytest.mark.parametrize("n", [1, 7, 1000, 16384, 16385, 300_007])
def test_threshold_sizes(wv, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) * 10).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    for frac in (0.0, 0.5, 0.99):
        _, r = wv.threshold(xd, frac)
        t = O.thresh(x, frac)
        if t["idx"] < n:
            assert r.threshold == t["threshold"] and r.thresh_idx == t["idx"]
        assert r.max_abs == t["max_abs"]
