"""Pins the oracle's steps/padding, 1-D Haar and CDF 9/7 transforms, 3-D
rotations and 3-D drivers (CPU only).

What pins what:
  * dwt_cpu.cc:680 known-answer vector (6 significant digits, checked in the
    reference against Matlab fwt_1d);
  * test_haar_cpu.cc:135-169 KATs for countLeadingZeros / ceilLog2 /
    dwt_padded_length;
  * an independent FP64 *lifting* implementation of CDF 9/7 from the
    published lifting constants (Daubechies & Sweldens 1998), on a
    periodically-symmetric extension -- the reference's tap values are the
    same filter bank scaled by sqrt(2) (lowpass) and -1/sqrt(2) (highpass);
  * Haar in closed form; perfect reconstruction of fwd->inv;
  * the 3-D rotations against numpy transposes and the README layout
    (2048x2048x1 becomes 1x2048x2048, README.md:85).
"""
import numpy as np
import pytest

import _oracle as O

KAT_IN = [7, 2, 3, 4, 5, 3, 1, 5, 7, 1, 9, 2, 4, 8, 2, 6]
KAT_OUT = [7.00224, 3.52709, 6.82547, 2.87912, 7.34827, 7.39307, 6.14146,
           6.57857, 2.33178, -0.122068, -0.136087, -1.33848, 5.86312, 3.64358,
           -4.18374, -2.92383]


def test_kat_cdf97_one_level():
    out = O.row_fwd(O.CDF97, KAT_IN, 1)
    np.testing.assert_allclose(out, KAT_OUT, rtol=2e-5, atol=2e-5)


def test_kat_ceil_log2_padding():
    L = O.lib()
    for x, n in [(0, 32), (1, 31), (2, 30), (3, 30), (0x7fffffff, 1), (0xffffffff, 0)]:
        assert L.or_count_leading_zeros(x) == n
    for x, n in [(1, 0), (2, 1), (3, 2), (4, 2), (5, 3), (6, 3), (7, 3), (8, 3),
                 (9, 4), (511, 9), (512, 9), (513, 10)]:
        assert L.or_ceil_log2(x) == n
    for l, s, p in [(4, 1, 4), (5, 1, 6), (5, 2, 8), (6, 1, 6), (6, 2, 8)]:
        assert L.or_padded_length(l, s) == p


def test_set_steps_rules():
    # negative or too many -> ceilLog2 (test_compress_common.cc:362-374)
    assert O.set_steps([-1, -1, -1], [2048, 2048, 1]) == [11, 11, 0]
    assert O.set_steps([3, 3, 3], [512, 512, 512]) == [3, 3, 3]
    assert O.set_steps([20, 2, 5], [100, 7, 3]) == [7, 2, 2]


# ---- FP64 lifting oracle (independent of the reference taps) ----
ALPHA = -1.586134342059924
BETA = -0.052980118572961
GAMMA = 0.882911075530934
DELTA = 0.443506852043971
ZETA = 1.149604398860241


def _mirror_ext(x, k):
    n = len(x)
    idx = np.arange(-k, n + k)
    o = np.abs(idx)
    while (o >= n).any():
        o = np.where(o >= n, np.abs(2 * n - o - 2), o)
    return x[o]


def lifting_level64(x):
    """One CDF 9/7 analysis level in float64 by lifting on a symmetric
    extension.  With the published constants, the reference's filter bank is
    (zeta * s, -d / zeta): lowpass DC gain sqrt(2), highpass sign flipped."""
    n = len(x)
    K = 16
    e = _mirror_ext(np.asarray(x, np.float64), K)   # e[K+i] = x[i]
    s = e[0::2].copy()   # even positions (K even)
    d = e[1::2].copy()
    d[:-1] += ALPHA * (s[:-1] + s[1:])
    s[1:] += BETA * (d[:-1] + d[1:])
    d[:-1] += GAMMA * (s[:-1] + s[1:])
    s[1:] += DELTA * (d[:-1] + d[1:])
    h = K // 2
    return (s * ZETA)[h:h + n // 2], (-d / ZETA)[h:h + n // 2]


def _lifting_ref(x, steps):
    x = np.asarray(x, np.float64).copy()
    n = len(x)
    for _ in range(steps):
        lo, hi = lifting_level64(x[:n])
        x[:n // 2], x[n // 2:n] = lo, hi
        n //= 2
    return x


def test_lifting_matches_reference_taps_and_dc_gain():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(64)
    lo, hi = lifting_level64(x)
    out = O.row_fwd(O.CDF97, x.astype(np.float32), 1).astype(np.float64)
    scale = np.abs(out).max()
    assert np.abs(out[:32] - lo).max() < 1e-6 * scale
    assert np.abs(out[32:] - hi).max() < 1e-6 * scale
    dc = O.row_fwd(O.CDF97, np.ones(16, np.float32), 1)
    np.testing.assert_allclose(dc[:8], np.sqrt(2.0), rtol=1e-6)
    np.testing.assert_allclose(dc[8:], 0.0, atol=1e-6)


@pytest.mark.parametrize("n,steps", [(2, 1), (4, 2), (8, 3), (16, 4), (32, 5),
                                     (512, 3), (512, 9), (2048, 11), (96, 5)])
def test_cdf97_vs_fp64_lifting(n, steps):
    rng = np.random.default_rng(n * 31 + steps)
    x = (rng.standard_normal(n) * 100 + 50).astype(np.float32)
    got = O.row_fwd(O.CDF97, x, steps).astype(np.float64)
    want = _lifting_ref(x.astype(np.float64), steps)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-5, err


@pytest.mark.parametrize("n,steps", [(2, 1), (8, 3), (16, 2), (64, 6), (512, 3)])
def test_haar_closed_form(n, steps):
    rng = np.random.default_rng(n + steps)
    x = rng.integers(0, 256, n).astype(np.float32)
    got = O.row_fwd(O.HAAR, x, steps)
    r = np.float32(0.70710678118654752440)
    want = x.copy()
    m = n
    for _ in range(steps):
        a, b = want[0:m:2].copy(), want[1:m:2].copy()
        want[:m // 2] = (a + b) * r
        want[m // 2:m] = (a - b) * r
        m //= 2
    assert np.array_equal(got, want)


@pytest.mark.parametrize("wavelet", [O.HAAR, O.CDF97])
@pytest.mark.parametrize("n,steps", [(2, 1), (4, 2), (16, 4), (96, 5), (512, 9)])
def test_perfect_reconstruction(wavelet, n, steps):
    rng = np.random.default_rng(7 * n + steps)
    x = (rng.standard_normal(n) * 30).astype(np.float32)
    y = O.row_fwd(wavelet, x, steps)
    z = O.row_inv(wavelet, y, steps)
    assert np.abs(z - x).max() <= 2e-5 * max(1.0, np.abs(x).max())


def test_mirror_multiple_folds_match_periodic_symmetric():
    # n=2 needs several folds for taps at +-4: compare with a float32
    # evaluation over an explicitly mirrored extension
    x = np.array([3.0, -1.0], np.float32)
    got = O.row_fwd(O.CDF97, x, 1)
    e = _mirror_ext(x.astype(np.float32), 4)
    lp = np.array([.85269867900940, .377402855612650, -.110624404418420,
                   -.02384946501938, .037828455506995], np.float32)
    hp = np.array([-.788485616405660, .418092273222210, .040689417609558,
                   -.064538882628938], np.float32)
    c = 4
    lo = lp[0] * e[c]
    for k in range(1, 5):
        lo = np.float32(lo + lp[k] * (e[c + k] + e[c - k]))
    c = 5
    hi = hp[0] * e[c]
    for k in range(1, 4):
        hi = np.float32(hi + hp[k] * (e[c + k] + e[c - k]))
    assert got[0] == lo and got[1] == hi


def test_rotations_and_3d_layout():
    X, Y, Z = 8, 4, 2
    rng = np.random.default_rng(3)
    vol = rng.standard_normal((Z, Y, X)).astype(np.float32)
    # zero steps everywhere: a no-op without rotation
    d, s = O.dwt3d(vol.ravel(), [X, Y, Z], O.CDF97, [0, 0, 0])
    assert s == [X, Y, Z] and np.array_equal(d, vol.ravel())
    # steps only on x: the data ends in ZXY layout = numpy [y, x, z]
    d, s = O.dwt3d(vol.ravel(), [X, Y, Z], O.HAAR, [1, 0, 0])
    assert s == [Z, X, Y]
    rows = np.stack([O.row_fwd(O.HAAR, r, 1) for r in vol.reshape(-1, X)])
    want = rows.reshape(Z, Y, X).transpose(1, 2, 0)
    assert np.array_equal(d.reshape(Y, X, Z), want)


def test_3d_separable_equals_axis_passes():
    X, Y, Z = 16, 8, 4
    rng = np.random.default_rng(5)
    vol = (rng.standard_normal((Z, Y, X)) * 10).astype(np.float32)
    steps = [2, 3, 1]
    d, s = O.dwt3d(vol.ravel(), [X, Y, Z], O.CDF97, steps)
    a = np.apply_along_axis(lambda r: O.row_fwd(O.CDF97, r, steps[0]), 2, vol)
    a = np.apply_along_axis(lambda r: O.row_fwd(O.CDF97, r, steps[1]), 1, a)
    a = np.apply_along_axis(lambda r: O.row_fwd(O.CDF97, r, steps[2]), 0, a)
    assert s == [Z, X, Y]
    assert np.array_equal(d.reshape(Y, X, Z), a.transpose(1, 2, 0))
    back, s2 = O.dwt3d(d, s, O.CDF97, steps, inverse=True)
    assert s2 == [X, Y, Z]
    assert np.abs(back - vol.ravel()).max() < 1e-4


def test_readme_layout_2d_case():
    d, s = O.dwt3d(np.zeros(64 * 32, np.float32), [64, 32, 1], O.CDF97, [6, 5, 0])
    assert s == [1, 64, 32]


def test_repeat_padding():
    vol = np.arange(2 * 3 * 5, dtype=np.int32).reshape(2, 3, 5)
    out = O.pad_convert(vol, [3, 2, 1])     # pads to 8 x 4 x 2
    assert out.shape == (2, 4, 8)
    want = vol[np.minimum(np.arange(2), 1)][:, np.minimum(np.arange(4), 2)][:, :, np.minimum(np.arange(8), 4)]
    assert np.array_equal(out, want.astype(np.float32))


@pytest.mark.parametrize("wavelet", [O.HAAR, O.CDF97])
@pytest.mark.parametrize("shape,steps", [((8, 16, 32), [3, 2, 1]), ((4, 2, 64), [6, 1, 2]),
                                         ((16, 8, 8), [0, 3, 4]), ((2, 4, 16), [4, 2, 1])])
def test_fp64_3d_oracle_matches_numpy_lifting(wavelet, shape, steps):
    """or_dwt3d_f64 (the yardstick of the float transforms' error at the
    benchmark sizes) equals the numpy FP64 lifting above applied axis by
    axis, and the float oracle to float precision."""
    rng = np.random.default_rng(sum(shape) + steps[0])
    vol = (rng.standard_normal(shape) * 100 + 20).astype(np.float32)
    got = O.dwt3d_f64(vol, wavelet, steps)

    def line(r, s):
        if wavelet == O.CDF97:
            return _lifting_ref(r, s)
        x = np.asarray(r, np.float64).copy()
        m = len(x)
        for _ in range(s):
            a, b = x[0:m:2].copy(), x[1:m:2].copy()
            x[:m // 2], x[m // 2:m] = (a + b) / np.sqrt(2.0), (a - b) / np.sqrt(2.0)
            m //= 2
        return x

    a = vol.astype(np.float64)
    for axis, s in ((2, steps[0]), (1, steps[1]), (0, steps[2])):
        if s > 0:
            a = np.apply_along_axis(lambda r: line(r, s), axis, a)
    assert np.abs(got - a).max() <= 1e-12 * np.abs(a).max()
    d, _ = O.dwt3d(vol.ravel(), [shape[2], shape[1], shape[0]], wavelet, steps)
    assert O.rotated_rel_error(d, got) < 1e-6


def test_fp64_error_measure_catches_a_transposed_layout():
    """rotated_rel_error compares in the z,x,y layout: the float result
    read in the input's own layout (a transposed operand) fails it."""
    rng = np.random.default_rng(9)
    vol = (rng.standard_normal((8, 16, 32)) * 50).astype(np.float32)
    want = O.dwt3d_f64(vol, O.CDF97, [2, 2, 2])
    d, _ = O.dwt3d(vol.ravel(), [32, 16, 8], O.CDF97, [2, 2, 2])
    assert O.rotated_rel_error(d, want) < 1e-6
    assert O.rotated_rel_error(want.astype(np.float32).ravel(), want) > 1e-2
