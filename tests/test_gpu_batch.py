"""GPU parity tests of the batched path (wv_batch_*): every cubelet of a
mixed batch gets exactly what the single-cubelet path -- and the CPU
oracle -- give it alone: bit-identical coefficients, identical threshold
record, identical bins, histogram equal to np.bincount of the bins.
Graph replays are checked against an eager run, with other device work
and synchronisations between the replays."""
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
    base = np.sin(g[2] * 0.17 + seed) * 50 + np.cos(g[1] * 0.11) * 35 + g[0] * 0.9
    v = base + rng.standard_normal(shape_zyx).astype(np.float32) * 15 + 120
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype in (np.int32, np.int16):
        return np.clip(v * 10, 0, 4095).astype(dtype)
    return v.astype(np.float32)


# (shape z,y,x; dtype; wavelet; steps; quantizer; bins; thresh fraction)
MIX = [
    ((16, 16, 16), np.float32, 2, [3, 3, 3], 2, 256, 0.9),
    ((16, 16, 16), np.float32, 2, [3, 3, 3], 2, 256, 0.8),
    ((16, 16, 16), np.uint8, 1, [3, 3, 3], 1, 255, 0.9),
    ((32, 32, 32), np.int16, 2, [3, 3, 3], 2, 256, 0.9),
    ((32, 32, 32), np.int16, 2, [3, 3, 3], 2, 256, 0.95),
    ((8, 64, 64), np.uint8, 2, [3, 3, 3], 2, 256, 0.9),
    ((8, 64, 64), np.float32, 1, [3, 3, 3], 1, 1024, 0.95),
    ((5, 24, 40), np.int32, 2, [3, 2, 1], 2, 256, 0.9),     # padded shape
    ((5, 24, 40), np.int32, 2, [3, 2, 1], 1, 64, 0.5),
    ((1, 128, 96), np.uint8, 2, [-1, -1, -1], 2, 256, 0.9),  # depth 1
    ((12, 20, 36), np.float32, 1, [2, 2, 2], 1, 255, 0.0),
    ((64, 32, 16), np.float32, 2, [3, 3, 3], 2, 512, 0.99),
]


def _cubes(mix, seed0=100):
    cubes, vols = [], []
    for k, (shape, dt, wave, steps, alg, bins, frac) in enumerate(mix):
        v = _vol(shape, dt, seed0 + k)
        vols.append(v)
        cubes.append(dict(vol=torch.from_numpy(v).cuda(), wavelet=wave, steps=steps,
                          alg=alg, bins=bins, frac=frac))
    return cubes, vols


def _check_item(b, i, vol, c):
    ov = vol.astype(np.int32) if vol.dtype == np.int16 else vol
    want = O.compress(ov, c["wavelet"], c["alg"], c["bins"], c["steps"], c["frac"],
                      want_coef=True, want_bins=True, want_file=False)
    coef = b.coef[i].cpu().numpy()
    assert np.array_equal(coef.view(np.uint32), want["coef"].view(np.uint32)), i
    r, zb = b.result(i)
    w = want["res"]
    assert r.threshold == w.threshold and r.thresh_idx == w.thresh_idx, i
    assert r.max_abs == w.max_abs and r.min_val == w.min_val and r.max_val == w.max_val, i
    assert zb == w.zero_bin, i
    bins = b.bins[i].cpu().numpy()
    assert np.array_equal(bins, want["bins"]), i
    B = c["bins"]
    cnt = np.bincount(np.minimum(bins, B), minlength=B + 1)
    assert np.array_equal(b.counts[i].cpu().numpy().astype(np.int64), cnt), i


@pytest.mark.parametrize("flags", [0, 2])
def test_batch_items_match_oracle(wv, flags):
    cubes, vols = _cubes(MIX)
    b = wv.Batch(cubes, flags=flags)
    b.submit()
    b.wait()
    for i, c in enumerate(cubes):
        _check_item(b, i, vols[i], c)
    b.close()


def test_batch_item_equals_plain_path(wv):
    """The same cubelet alone through the plain single-cubelet calls."""
    cubes, vols = _cubes(MIX[:6], seed0=300)
    b = wv.Batch(cubes)
    b.submit()
    b.wait()
    for i, c in enumerate(cubes):
        coef, g, _ = wv.dwt3d_forward(c["vol"], c["steps"], c["wavelet"], with_stats=False)
        assert torch.equal(coef.view(torch.int32), b.coef[i].view(torch.int32))
        _, r = wv.threshold(coef, c["frac"])
        rb, _ = b.result(i)
        assert (r.threshold, r.thresh_idx, r.max_abs) == (rb.threshold, rb.thresh_idx, rb.max_abs)
    b.close()


def test_batch_graph_replays_match_eager(wv):
    cubes, vols = _cubes(MIX, seed0=500)
    eager = wv.Batch(cubes, flags=wv.BATCH_NO_GRAPH)
    eager.submit()
    eager.wait()
    ref = [(eager.coef[i].clone(), eager.bins[i].clone(), eager.counts[i].clone())
           for i in range(len(cubes))]
    eager.close()
    g = wv.Batch(cubes)
    scratch = torch.zeros(1 << 20, device="cuda")
    for rep in range(4):
        for i in range(len(cubes)):       # clobber the outputs between replays
            g.coef[i].fill_(float(rep + 1))
            g.bins[i].fill_(-7)
            g.counts[i].fill_(12345)
        scratch.add_(1.0)                 # unrelated eager work
        torch.cuda.synchronize()
        g.submit()
        g.wait()
        for i in range(len(cubes)):
            assert torch.equal(g.coef[i].view(torch.int32), ref[i][0].view(torch.int32)), (rep, i)
            assert torch.equal(g.bins[i], ref[i][1]), (rep, i)
            assert torch.equal(g.counts[i], ref[i][2]), (rep, i)
    g.close()


def test_batch_autotuned_results_unchanged(wv):
    cubes, vols = _cubes(MIX[:8], seed0=700)
    plain = wv.Batch(cubes)
    plain.submit()
    plain.wait()
    tuned = wv.Batch(cubes, flags=wv.BATCH_AUTOTUNE)
    tuned.submit()
    tuned.wait()
    assert len(tuned.groups()) == len(set((tuple(c["vol"].shape), str(c["vol"].dtype),
                                           c["wavelet"], tuple(c["steps"]))
                                          for c in cubes))
    for i in range(len(cubes)):
        assert torch.equal(plain.coef[i].view(torch.int32), tuned.coef[i].view(torch.int32))
        assert torch.equal(plain.bins[i], tuned.bins[i])
        assert torch.equal(plain.counts[i], tuned.counts[i])
    plain.close()
    tuned.close()


def test_batch_select_fallback_on_overflowing_window(wv):
    """Many equal |x| around the rank: the candidate window holds far more
    than 1/8 of the keys, the list overflows and the exact in-block select
    runs; the result still equals the oracle's."""
    shape = (64, 64, 64)
    v = np.full(shape, 1000.0, np.float32)
    v[::2] = 0.0
    v += (np.arange(v.size, dtype=np.float32).reshape(shape) % 3) * 1e-3
    cubes = [dict(vol=torch.from_numpy(v).cuda(), wavelet=1, steps=[0, 0, 0], alg=1,
                  bins=64, frac=0.7)]
    b = wv.Batch(cubes)
    b.submit()
    b.wait()
    _check_item(b, 0, v, cubes[0])
    r, _ = b.result(0)
    assert r.n_candidates < 0          # the exact path ran
    b.close()


def test_batch_select_zero_ties(wv):
    """Blocks of constant value: most Haar detail coefficients are exactly
    zero, the rank lies inside the zeros or just above them, and the zeros
    are counted rather than listed -- no fallback, the oracle's result."""
    shape = (32, 64, 64)
    rng = np.random.default_rng(3)
    tiles = rng.integers(0, 255, (4, 8, 8)).astype(np.uint8)
    v = np.kron(tiles, np.ones((8, 8, 8), np.uint8))
    v[::16, ::16, ::16] = rng.integers(0, 255, v[::16, ::16, ::16].shape).astype(np.uint8)
    # about 98.3% of the coefficients are exact zeros
    cubes = [dict(vol=torch.from_numpy(v).cuda(), wavelet=1, steps=[3, 3, 3], alg=a,
                  bins=256, frac=f)
             for a, f in ((2, 0.5), (2, 0.9), (1, 0.97), (2, 0.985), (1, 0.995))]
    b = wv.Batch(cubes)
    b.submit()
    b.wait()
    for i, c in enumerate(cubes):
        _check_item(b, i, v, c)
        r, _ = b.result(i)
        assert r.n_candidates >= 0, i
    b.close()


def test_batch_wait_on_null_stream_before_any_other_sync(wv):
    """Submitted on the null stream (handle 0), wait() must still wait:
    the threshold records read right after it, before any other
    synchronisation, are the final ones."""
    assert torch.cuda.current_stream().cuda_stream == 0
    cubes, vols = _cubes(MIX[:6], seed0=300)
    want = [O.compress(v.astype(np.int32) if v.dtype == np.int16 else v, c["wavelet"],
                       c["alg"], c["bins"], c["steps"], c["frac"], want_file=False)["res"]
            for v, c in zip(vols, cubes)]
    b = wv.Batch(cubes)
    for _ in range(2):
        b.submit()
        b.wait()
        got = [b.result(i)[0] for i in range(len(cubes))]
        for i, (r, w) in enumerate(zip(got, want)):
            assert r.threshold == w.threshold and r.max_abs == w.max_abs, i
    b.close()


def test_batch_write_stream_records_match_oracle(wv):
    """The whole batch coded on the device (grouped by bin count) into one
    .cube stream: every record equals the oracle's own file record for that
    cubelet (parameters message with the decode table, payload)."""
    import _cubeproto as P
    cubes, vols = _cubes(MIX, seed0=400)
    b = wv.Batch(cubes)
    b.submit()
    f = b.write_stream()
    _, recs, index, _ = P.parse_cube_file(f)
    assert len(recs) == len(cubes) == len(index.cubelets)
    for i, ((cb, raw, payload), v, c) in enumerate(zip(recs, vols, cubes)):
        ov = v.astype(np.int32) if v.dtype == np.int16 else v
        want = O.compress(ov, c["wavelet"], c["alg"], c["bins"], c["steps"], c["frac"])["file"]
        _, wrecs, _, _ = P.parse_cube_file(want)
        assert payload == wrecs[0][2], i
        assert raw == wrecs[0][1], i
    b.close()
