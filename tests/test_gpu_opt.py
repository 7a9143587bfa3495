"""GPU tests of the parameter optimisation (-opt): OptimizationData's
sorted |x| and findSorted, and testParameters' output size
(Huffman::totalEncodedLengthBytes) and error metrics, against the
oracle's restatement of one test point (optimize_gpu.cu:52-202 with the
CPU path's double error sums).  Sizes are exact; the error metrics are
float results of double sums taken in a different order, so they agree to
a relative 1e-6."""
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


def _vol(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    z, y, x = shape
    g = np.mgrid[0:z, 0:y, 0:x].astype(np.float32)
    v = (np.sin(g[2] * 0.21) * 70 + np.cos(g[1] * 0.13) * 40 + g[0] * 0.9 +
         rng.standard_normal(shape).astype(np.float32) * 12 + 120)
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype in (np.int32, np.int16):
        return np.clip(v * 11, 0, 4095).astype(dtype)
    return v.astype(np.float32)


def _close(a, b, tol=1e-6):
    return a == b or abs(a - b) <= tol * max(abs(a), abs(b))


CASES = [
    ((24, 40, 48), np.uint8, 2, [3, 3, 3], 0),
    ((20, 36, 28), np.float32, 1, [3, 2, 2], 4000),
    ((16, 32, 32), np.int16, 2, [2, 3, 3], 4095),
]


@pytest.mark.parametrize("shape,dtype,wave,steps,maxp", CASES)
def test_sorted_and_find_sorted(wv, shape, dtype, wave, steps, maxp):
    vol = _vol(shape, dtype, 2)
    o = wv.Opt(torch.from_numpy(vol).cuda(), steps, wave, maxp)
    ov = vol.astype(np.int32) if dtype == np.int16 else vol
    want = O.opt_sorted(ov, steps, wave)
    assert o.count() == want.size
    rng = np.random.default_rng(0)
    for i in list(rng.integers(0, want.size, 20)) + [0, want.size - 1]:
        assert o.sorted(i) == want[i]
    assert o.sorted(-1) == 0 and o.sorted(want.size) == 0
    for v in [want[5], want[want.size // 2], want[-1], want[-1] * 2, 0.0, -1.0]:
        assert o.find_sorted(v) == np.searchsorted(want, np.float32(v), side="left")
    mn, mx, mabs = o.stats()
    assert mabs == want[-1]
    o.close()


@pytest.mark.parametrize("shape,dtype,wave,steps,maxp", CASES)
@pytest.mark.parametrize("frac,bins,alg,zeros", [
    (0.9, 256, 2, True), (0.8, 1024, 1, True), (0.95, 100, 2, False), (0.7, 0, 2, True),
    (0.5, 37, 1, False), (0.99, 2000, 2, True)])
def test_point_matches_oracle(wv, shape, dtype, wave, steps, maxp, frac, bins, alg, zeros):
    vol = _vol(shape, dtype, 3)
    o = wv.Opt(torch.from_numpy(vol).cuda(), steps, wave, maxp, compress_zeros=zeros)
    t = o.sorted(int(np.float32(o.count()) * np.float32(frac)))
    r = o.test(t, bins, alg)
    ov = vol.astype(np.int32) if dtype == np.int16 else vol
    size, mse, psnr, acc = O.opt_test(ov, steps, wave, maxp, zeros, t, bins, alg)
    assert r.size_bytes == size
    assert (size == 0) == (bins <= 0)
    assert _close(r.mse, mse) and _close(r.psnr, psnr)
    assert _close(r.l1, np.float32(acc.sum_diff))
    assert _close(r.l2, np.float32(np.sqrt(acc.sum_diff_sq)))
    rel = np.float32(acc.sum_diff_sq / acc.sum_value_sq) if acc.sum_value_sq else 0
    assert _close(r.rel, rel)
    o.close()


def test_plot_all_rows_match_oracle(wv):
    """The reference's default sweep on a small volume: the first rows
    (fraction .5, bin counts 100 .. 2000) against the oracle's test points."""
    vol = _vol((16, 32, 32), np.uint8, 5)
    o = wv.Opt(torch.from_numpy(vol).cuda(), [3, 3, 3], 2)
    rows = wv.opt_plot_all(o)
    assert len(rows) == 50 * 32
    assert rows[0][:2] == (0.5, 100) and rows[1][1] == 110 and rows[31][1] == 1863
    for frac, b, size, psnr in rows[:32] + rows[-3:]:
        t = o.sorted(int(np.float32(o.count()) * np.float32(frac)))
        wsize, _, wpsnr, _ = O.opt_test(vol, [3, 3, 3], 2, 0, True, t, b, 2)
        assert size == wsize and _close(psnr, wpsnr), (frac, b)
    o.close()


@pytest.mark.parametrize("shape,dtype,wave,steps,maxp", CASES)
@pytest.mark.parametrize("frac,bins,zeros", [(0.9, 255, True), (0.6, 50, False),
                                             (0.95, 1001, True), (0.8, 64, True)])
def test_lloyd_point_matches_oracle(wv, shape, dtype, wave, steps, maxp, frac, bins, zeros):
    """A Lloyd test point (the fixedPSNR sweep's quantizer): the code book
    trained on the sorted |x| from findSorted(T), uncapped; iterations and
    size exact, errors to 1e-6."""
    vol = _vol(shape, dtype, 4)
    o = wv.Opt(torch.from_numpy(vol).cuda(), steps, wave, maxp, compress_zeros=zeros)
    t = o.sorted(int(np.float32(o.count()) * np.float32(frac)))
    r = o.test(t, bins, wv.LLOYD)
    ov = vol.astype(np.int32) if dtype == np.int16 else vol
    size, mse, psnr, acc, iters = O.opt_test(ov, steps, wave, maxp, zeros, t, bins, 3,
                                             want_iters=True)
    assert r.lloyd_iters == iters and iters > 0
    assert r.size_bytes == size
    assert _close(r.mse, mse) and _close(r.psnr, psnr)
    o.close()


def test_fixed_psnr_sweep_matches_oracle(wv):
    """fixedPSNR (optimize.cc:53-90) on a small image: every point it tests
    against the oracle's test point."""
    vol = _vol((1, 64, 96), np.uint8, 6)
    o = wv.Opt(torch.from_numpy(vol).cuda(), [4, 4, 0], 2)
    rows = wv.opt_fixed_psnr(o, 40.0)
    assert rows and rows[0][1] == 0
    for frac, b, size, psnr in rows[:6] + rows[-3:]:
        t = o.sorted(int(np.float32(o.count()) * np.float32(1 - np.float32(frac))))
        wsize, _, wpsnr, _ = O.opt_test(vol, [4, 4, 0], 2, 0, True, t, b, 3)
        assert size == wsize and _close(psnr, wpsnr), (frac, b)
    o.close()
