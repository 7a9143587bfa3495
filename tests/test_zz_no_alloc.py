"""The allocation-free threshold entry point (wv_threshold_noalloc: the
select on the library's static per-GPU scratch) against the workspace entry
point wv_threshold on the same coefficients: identical result records on
calls in a row -- the second with a smaller input, over the state the first
left in the shared scratch -- including an exact-fallback case, and a clear
error beyond the scratch's size (raised on the host, before any launch)."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FIELDS = ("threshold", "min_val", "max_val", "max_abs", "thresh_idx", "nonzero_count",
          "threshold_key", "n_candidates")


@pytest.fixture(scope="module")
def wv():
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    import wavelets_amd
    wavelets_amd.lib()
    return wavelets_amd


def _same(a, b):
    return all(getattr(a, f) == getattr(b, f) for f in FIELDS)


def _sample_positions(n, m=16384):
    """The select's sampled positions (one per stratum, hashed offset)."""
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


def test_noalloc_threshold_equals_workspace_entry(wv):
    from wavelets_amd import synth
    assert wv.NOALLOC_MAX_N == 4194304
    img = synth.smooth_field_u8(2048, 2048, seed=1)           # 4,194,304 coefficients
    coef, _, _ = wv.dwt3d_forward(torch.from_numpy(img).cuda(), [11, 11, 0], wv.CDF97,
                                  with_stats=False)
    assert coef.numel() == wv.NOALLOC_MAX_N
    rng = np.random.default_rng(3)
    small = torch.from_numpy(np.round(rng.standard_normal(300_007) * 4).astype(np.float32)).cuda()
    n = 1 << 20
    fb = (rng.random(n).astype(np.float32) * 1000 + 100).astype(np.float32)
    fb[_sample_positions(n)] = rng.random(16384).astype(np.float32) * 0.01
    fb_d = torch.from_numpy(fb).cuda()
    calls = [(coef, 0.9), (small, 0.5), (fb_d, 0.3), (coef, 0.95), (small, 0.77)]
    for k, (x, frac) in enumerate(calls):
        print("call", k, x.numel(), frac, flush=True)
        _, got = wv.threshold_noalloc(x, frac)
        _, want = wv.threshold(x, frac)
        assert _same(got, want), (k, [(f, getattr(got, f), getattr(want, f)) for f in FIELDS])
        t = O.thresh(x.cpu().numpy(), frac)
        assert got.threshold == t["threshold"] and got.thresh_idx == t["idx"], k
        if x is fb_d:
            assert got.n_candidates < 0        # the exact fallback ran
        else:
            assert got.n_candidates >= 0


def test_noalloc_threshold_rejects_beyond_scratch(wv):
    x = torch.zeros(wv.NOALLOC_MAX_N + 1, dtype=torch.float32, device="cuda")
    with pytest.raises(wv.WvError, match="exceed"):
        wv.threshold_noalloc(x, 0.5)
    # and the entry point still serves a size that fits afterwards
    _, r = wv.threshold_noalloc(x[:1000], 0.5)
    assert r.threshold == np.float32(1e-16) and r.thresh_idx == 500
