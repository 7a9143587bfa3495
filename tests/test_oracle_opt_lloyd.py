"""The oracle's Lloyd test point of the parameter optimisation
(testParameters with QUANT_ALG_LLOYD, optimize_cpu.cc:84-96): its code book
is trained on the sorted |x| from findSorted(threshold) (std::lower_bound),
all but the largest, with no iteration cap.  Pinned against the
reference's own lloyd (lloyds.cpp, compiled into oracle/_ref) run on those
points: the iteration count the oracle reports must be the reference's."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O


def _vol(seed):
    rng = np.random.default_rng(seed)
    g = np.mgrid[0:12, 0:24, 0:20].astype(np.float32)
    v = np.sin(g[2] * 0.3) * 60 + np.cos(g[1] * 0.17) * 30 + g[0] + rng.standard_normal(
        g[0].shape).astype(np.float32) * 9 + 100
    return np.clip(v, 0, 255).astype(np.uint8)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("frac,bins", [(0.9, 255), (0.7, 51), (0.95, 1001)])
def test_lloyd_point_points_and_iterations(frac, bins):
    r = O.ref()
    if r is None:
        pytest.skip("reference sources not compiled here")
    vol = _vol(int(bins))
    steps, wave = [2, 2, 2], 2
    srt = O.opt_sorted(vol, steps, wave)
    t = srt[int(np.float32(srt.size) * np.float32(frac))]
    first = int(np.searchsorted(srt, t, side="left"))      # findSorted
    nonzero = np.ascontiguousarray(srt[first:])
    cs = (bins - 1) // 2
    cb = np.empty(cs, np.float32)
    O.lib().or_initial_lloyd_codebook(_p(cb), cs, C.c_float(t), C.c_float(nonzero[-2]))
    part = np.empty(cs - 1, np.float32)
    d, rel, it = C.c_float(), C.c_float(), C.c_int()
    r.ref_lloyd(_p(nonzero), nonzero.size - 1, _p(cb), cs, _p(part), C.byref(d), C.byref(rel),
                C.c_float(1e-6), C.byref(it))
    size, mse, psnr, acc, iters = O.opt_test(vol, steps, wave, 0, True, t, bins, 3,
                                             want_iters=True)
    assert iters == it.value and iters > 0
    assert size > 0 and psnr > 0
