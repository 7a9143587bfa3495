"""GPU tests of the serving loop (wv_serve_*): every request's result --
bins, counts, threshold record and .cube bytes -- is the same alone and in
a mix of concurrent requests, equals the CPU oracle, and admission keeps
the device bytes of admitted requests within the budget."""
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
    v = (np.sin(g[2] * 0.19 + seed) * 45 + np.cos(g[1] * 0.07) * 30 + g[0] * 1.3 +
         rng.standard_normal(shape_zyx).astype(np.float32) * 14 + 110)
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype in (np.int16, np.int32):
        return np.clip(v * 9, 0, 4095).astype(dtype)
    return v.astype(np.float32)


# (shape; dtype; wavelet; steps; quantizer; bins; thresh)
REQS = [
    ((16, 32, 32), np.uint8, 1, [4, 4, 4], 3, 256, 0.9),
    ((24, 24, 40), np.int16, 1, [4, 4, 4], 3, 256, 0.85),
    ((32, 32, 32), np.float32, 2, [3, 3, 3], 1, 1024, 0.95),
    ((1, 96, 64), np.uint8, 2, [-1, -1, -1], 2, 256, 0.9),
    ((20, 48, 36), np.int32, 1, [4, 4, 4], 3, 64, 0.8),
    ((8, 64, 64), np.float32, 2, [3, 3, 3], 2, 256, 0.9),
    ((40, 16, 24), np.uint8, 1, [4, 4, 4], 3, 256, 0.98),
    ((16, 16, 16), np.int16, 2, [3, 3, 3], 1, 255, 0.5),
]


def _submit_all(srv, vols):
    ids = []
    for v, (shape, dt, wave, steps, alg, bins, frac) in zip(vols, REQS):
        ids.append(srv.submit(v, wave, steps, alg, bins, frac, want_bins=True))
    return ids


@pytest.mark.parametrize("lloyd_batch", [False, True])
def test_serve_results_same_alone_and_in_mix(wv, lloyd_batch):
    vols = [_vol(r[0], r[1], 40 + k) for k, r in enumerate(REQS)]
    alone = {}
    srv = wv.Server(slots=1)
    for k, v in enumerate(vols):
        shape, dt, wave, steps, alg, bins, frac = REQS[k]
        srv.submit(v, wave, steps, alg, bins, frac, want_bins=True)
        (r,) = srv.drain()
        assert r["status"] == 0
        alone[k] = r
    srv.close()
    mix = wv.Server(slots=6, lloyd_batch=lloyd_batch)
    ids = _submit_all(mix, vols)
    res = {r["id"]: r for r in mix.drain()}
    st = mix.stats()
    mix.close()
    assert st["peak_inflight"] > 1            # requests really overlapped
    for k, i in enumerate(ids):
        a, m = alone[k], res[i]
        assert m["status"] == 0
        assert np.array_equal(a["bins"], m["bins"]), k
        assert np.array_equal(a["counts"], m["counts"]), k
        assert a["file"] == m["file"], k
        assert a["thresh"].threshold == m["thresh"].threshold, k
        assert a["lloyd_iters"] == m["lloyd_iters"], k


def test_serve_matches_oracle_files(wv):
    vols = [_vol(r[0], r[1], 70 + k) for k, r in enumerate(REQS)]
    srv = wv.Server(slots=4)
    ids = _submit_all(srv, vols)
    res = {r["id"]: r for r in srv.drain()}
    srv.close()
    for k, i in enumerate(ids):
        shape, dt, wave, steps, alg, bins, frac = REQS[k]
        v = vols[k].astype(np.int32) if vols[k].dtype == np.int16 else vols[k]
        want = O.compress(v, wave, alg, bins, steps, frac, lloyd_max_iters=30, want_bins=True)
        assert np.array_equal(res[i]["bins"], want["bins"]), k
        assert res[i]["file"] == want["file"], k


def test_serve_admission_respects_budget(wv):
    vols = [_vol((32, 32, 32), np.float32, 90 + k) for k in range(10)]
    probe = wv.Server(slots=1)
    probe.submit(vols[0], 2, [3, 3, 3], 1, 256, 0.9)
    (r,) = probe.drain()
    probe.close()
    fp = r["device_bytes"]
    budget = 2 * fp + fp // 2                  # room for two requests at a time
    srv = wv.Server(slots=6, hbm_budget=budget)
    for v in vols:
        srv.submit(v, 2, [3, 3, 3], 1, 256, 0.9)
    out = srv.drain()
    st = srv.stats()
    srv.close()
    assert len(out) == 10 and all(o["status"] == 0 for o in out)
    assert st["peak_bytes"] <= budget
    assert st["peak_inflight"] <= 2


def test_serve_rejects_request_over_budget(wv):
    srv = wv.Server(slots=2, hbm_budget=1 << 20)
    with pytest.raises(wv.WvError):
        srv.submit(_vol((32, 32, 32), np.float32, 5), 2, [3, 3, 3], 1, 256, 0.9)
    srv.close()


# Lloyd requests of different sizes, codebook sizes and iteration caps: with
# lloyd_batch their windows share launches (runs join and leave at window
# boundaries, a capped run's last window is shorter than the others')
LLOYD_REQS = [
    ((16, 32, 32), np.uint8, 256, 30),
    ((24, 40, 40), np.int16, 64, 3),
    ((32, 32, 32), np.float32, 1024, 30),
    ((8, 64, 48), np.int16, 256, 9),
    ((40, 24, 24), np.float32, 256, -1),
    ((16, 16, 16), np.uint8, 128, 17),
    ((32, 48, 32), np.int32, 512, 30),
    ((12, 36, 52), np.float32, 256, 1),
]


def test_serve_lloyd_batch_same_as_alone(wv):
    vols = [_vol(r[0], r[1], 120 + k) for k, r in enumerate(LLOYD_REQS)]
    alone = {}
    srv = wv.Server(slots=1, lloyd_batch=False)
    for k, v in enumerate(vols):
        shape, dt, bins, iters = LLOYD_REQS[k]
        srv.submit(v, 1, [4, 4, 4], 3, bins, 0.85, lloyd_iters=iters, want_bins=True)
        (r,) = srv.drain()
        assert r["status"] == 0
        alone[k] = r
    srv.close()
    mix = wv.Server(slots=8, lloyd_batch=True)
    ids = [mix.submit(v, 1, [4, 4, 4], 3, r[2], 0.85, lloyd_iters=r[3], want_bins=True)
           for v, r in zip(vols, LLOYD_REQS)]
    res = {r["id"]: r for r in mix.drain()}
    st = mix.stats()
    mix.close()
    for k, i in enumerate(ids):
        a, m = alone[k], res[i]
        assert m["status"] == 0, k
        assert a["lloyd_iters"] == m["lloyd_iters"], k
        assert np.array_equal(a["bins"], m["bins"]), k
        assert a["file"] == m["file"], k
    cap = LLOYD_REQS[1][3]
    assert res[ids[1]]["lloyd_iters"] <= cap
    print("lloyd windows %d serving %d runs" % (st["lloyd_windows"], st["lloyd_run_windows"]))
    assert st["lloyd_windows"] > 0
    assert st["lloyd_run_windows"] > st["lloyd_windows"]     # windows were shared
