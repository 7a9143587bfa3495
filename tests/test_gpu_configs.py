"""GPU parity at the benchmark workloads' real sizes, against the CPU oracle
on the same inputs (test_compress_cpu.cc:68-184 restated in oracle/).

  1. 2048 x 2048 x 1 uint8 image, CDF 9/7 11,11,0, keep 10%, log 256 bins;
  2. 512^3 float32 volume, CDF 9/7 3,3,3, keep 5%, uniform 1024 bins;
  3. 256^3 int16 volume, Haar 4,4,4, keep 10%, Lloyd 256 bins (30 iterations);
  4. one batch holding every size class of the mixed stream (32^3, 64^3,
     128^3, 256x256x64, 512x512x16) x uint8 / int16 / float32 x both
     wavelets, log 256 bins, keep 10%;
  5. twenty serving requests from the serving workload's generator (edges
     lognormal in [32, 512], depth in [1, 512], keep U[0.02, 0.2], Haar
     4,4,4, Lloyd 256), each the same alone, in the mix and in the oracle.

The bar: coefficients bit-identical to the oracle (same operation order, no
contraction); threshold record, bins, counts, Lloyd code book and .cube
bytes identical.  The float coefficients' error against an FP64 lifting
filter bank (oracle/or_dwt64.c) is printed and must stay below 1e-5 of the
largest coefficient, the tolerance the float transform is held to.  Round
trips: the GPU decompression of the GPU's file equals the oracle's
reconstruction of it, so the PSNR is the same (checked to 0.01 dB).

Oracle work runs on host threads (the C calls release the GIL) beside the
GPU work.
"""
import concurrent.futures as cf
import ctypes as C

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FP64_REL_TOL = 1e-5
POOL = cf.ThreadPoolExecutor(max_workers=12)


@pytest.fixture(scope="module")
def wv():
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    import wavelets_amd
    wavelets_amd.lib()
    return wavelets_amd


def _oracle_input(vol):
    """int16 volumes are INT32 cubelets on the reference side."""
    return vol.astype(np.int32) if vol.dtype == np.int16 else vol


def _compress_gpu(wv, vol, wavelet, alg, bins, steps, frac, iters=30):
    """The one-call compress (device stages, device coding, container):
    (.cube bytes, host bins, CompressInfo)."""
    orig = (C.c_int * 3)(vol.shape[2], vol.shape[1], vol.shape[0])
    opt = wv.Options(wavelet, alg, bins, (C.c_int * 3)(*steps), C.c_float(frac), 1, iters, 1)
    L = wv.lib()
    nbytes = L.wv_compress_work_bytes(orig, C.byref(opt))
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    vd = torch.from_numpy(np.ascontiguousarray(vol)).cuda()
    st = O.set_steps(steps, list(orig))
    n = int(np.prod(O.padded_size(list(orig), st)))
    cap = n * 5 + (1 << 20)
    fbuf = np.empty(cap, np.uint8)
    bins_h = np.empty(n, np.int32)
    info = wv.CompressInfo()
    wv.check(L.wv_compress(C.c_void_p(vd.data_ptr()), wv.dtype_id(vd), orig, C.byref(opt),
                           C.c_void_p(work.data_ptr()), nbytes,
                           bins_h.ctypes.data_as(C.c_void_p),
                           fbuf.ctypes.data_as(C.c_void_p), cap, C.byref(info), wv._stream()))
    return fbuf[:info.file_bytes].tobytes(), bins_h, info


def _fp64_error(vol, wavelet, steps, coef_gpu):
    st = O.set_steps(steps, [vol.shape[2], vol.shape[1], vol.shape[0]])
    want = O.dwt3d_f64(O.pad_convert(_oracle_input(vol), st), wavelet, st)
    return O.rotated_rel_error(coef_gpu, want)


def _psnr(orig, restored, max_value):
    d = orig.astype(np.float64) - restored.astype(np.float64)
    mse = float(np.mean(d * d))
    return 10.0 * np.log10(max_value * max_value / mse)


def _check_thresh(r, w):
    assert r.threshold == w.threshold and r.thresh_idx == w.thresh_idx
    assert r.nonzero_count == w.nonzero_count
    assert r.min_val == w.min_val and r.max_val == w.max_val and r.max_abs == w.max_abs


def _compressed_records(f):
    import _cubeproto as P
    _, recs, _, _ = P.parse_cube_file(f)
    return recs


# ------------------------------------------------------------------ cfg 1

def test_cfg1_image_2048_u8_cdf97_log256(wv):
    from wavelets_amd import synth
    from test_gpu_decompress import oracle_decompress
    img = synth.smooth_field_u8(2048, 2048, seed=1)        # [1, 2048, 2048]
    steps = [11, 11, 0]
    want_f = POOL.submit(O.compress, img, O.CDF97, O.LOG, 256, steps, 0.9,
                         want_coef=True, want_bins=True)
    coef, g, _ = wv.dwt3d_forward(torch.from_numpy(img).cuda(), steps, wv.CDF97,
                                  with_stats=False)
    coef_h = coef.cpu().numpy()
    err_f = POOL.submit(_fp64_error, img, O.CDF97, steps, coef_h)
    f, bins_h, info = _compress_gpu(wv, img, wv.CDF97, wv.LOG, 256, steps, 0.9)
    want = want_f.result()
    assert list(g.out_size) == [1, 2048, 2048]
    assert np.array_equal(coef_h.view(np.uint32), want["coef"].view(np.uint32))
    _check_thresh(info.thresh, want["res"])
    assert info.zero_bin == want["res"].zero_bin
    assert np.array_equal(bins_h, want["bins"])
    assert f == want["file"]
    err = err_f.result()
    print("cfg1: %d-byte .cube identical; max |coef - fp64| / max |fp64| = %.3g"
          % (len(f), err), flush=True)
    assert err < FP64_REL_TOL
    got = wv.decompress(f)
    rec = oracle_decompress(f)
    assert np.array_equal(got, rec)
    p_gpu, p_or = _psnr(img, got, 255.0), _psnr(img, rec, 255.0)
    print("cfg1: round-trip PSNR %.4f dB (oracle %.4f dB)" % (p_gpu, p_or), flush=True)
    assert abs(p_gpu - p_or) <= 0.01


# ------------------------------------------------------------------ cfg 2

@pytest.fixture(scope="module")
def cfg2(wv):
    from wavelets_amd import synth
    vol = synth.ellipsoids_f32(512, seed=2)
    want = O.compress(vol, O.CDF97, O.UNIFORM, 1024, [3, 3, 3], 0.95, want_coef=True,
                      want_bins=True)
    return vol, want


def test_cfg2_volume_512_coefficients_and_fp64_error(wv, cfg2):
    vol, want = cfg2
    coef, g, stats = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [3, 3, 3], wv.CDF97)
    coef_h = coef.cpu().numpy()
    err_f = POOL.submit(_fp64_error, vol, O.CDF97, [3, 3, 3], coef_h)
    assert list(g.out_size) == [512, 512, 512]
    assert np.array_equal(coef_h.view(np.uint32), want["coef"].view(np.uint32))
    err = err_f.result()
    print("cfg2: 134217728 coefficients bit-identical; max |coef - fp64| / max |fp64| = %.3g"
          % err, flush=True)
    assert err < FP64_REL_TOL


def test_cfg2_volume_512_threshold_bins_counts(wv, cfg2):
    vol, want = cfg2
    coef, g, stats = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [3, 3, 3], wv.CDF97)
    res_d, r = wv.threshold(coef, 0.95, stats)
    _check_thresh(r, want["res"])
    assert r.n_candidates >= 0                   # no exact fallback was needed
    bins, counts = wv.quantize_uniform_dev(coef, 1024, res_d)
    b = bins.cpu().numpy()
    assert np.array_equal(b, want["bins"])
    assert np.array_equal(counts.cpu().numpy().astype(np.int64),
                          np.bincount(want["bins"], minlength=1024))


def test_cfg2_volume_512_file_and_round_trip(wv, cfg2):
    from test_gpu_decompress import oracle_decompress
    vol, want = cfg2
    f, bins_h, info = _compress_gpu(wv, vol, wv.CDF97, wv.UNIFORM, 1024, [3, 3, 3], 0.95)
    rec_f = POOL.submit(oracle_decompress, f)
    assert np.array_equal(bins_h, want["bins"])
    assert f == want["file"]
    got = wv.decompress(f)
    rec = rec_f.result()
    assert np.array_equal(got.view(np.uint32), rec.view(np.uint32))
    p_gpu, p_or = _psnr(vol, got, 4096.0), _psnr(vol, rec, 4096.0)
    print("cfg2: %d-byte .cube identical; round-trip PSNR %.4f dB (oracle %.4f dB)"
          % (len(f), p_gpu, p_or), flush=True)
    assert abs(p_gpu - p_or) <= 0.01


# ------------------------------------------------------------------ cfg 3

def test_cfg3_volume_256_int16_haar_lloyd256(wv):
    from wavelets_amd import synth
    vol = synth.lattice_i16(256, seed=3)
    want_f = POOL.submit(O.compress, vol.astype(np.int32), O.HAAR, O.LLOYD, 256, [4, 4, 4],
                         0.9, lloyd_max_iters=30, want_bins=True)
    coef, g, _ = wv.dwt3d_forward(torch.from_numpy(vol).cuda(), [4, 4, 4], wv.HAAR,
                                  with_stats=False)
    err_f = POOL.submit(_fp64_error, vol, O.HAAR, [4, 4, 4], coef.cpu().numpy())
    f, bins_h, info = _compress_gpu(wv, vol, wv.HAAR, wv.LLOYD, 256, [4, 4, 4], 0.9, iters=30)
    want = want_f.result()
    _check_thresh(info.thresh, want["res"])
    assert info.lloyd_iters == want["res"].lloyd_iters
    ginfo, _, gb, gv = wv.cube_parse(f)
    winfo, _, wb, wvals = wv.cube_parse(want["file"])
    assert np.array_equal(gb.view(np.uint32), wb.view(np.uint32))      # code book
    assert np.array_equal(gv.view(np.uint32), wvals.view(np.uint32))
    assert np.array_equal(bins_h, want["bins"])
    assert f == want["file"]
    err = err_f.result()
    print("cfg3: Lloyd %d iterations, code book and %d-byte .cube identical; "
          "max |coef - fp64| / max |fp64| = %.3g" % (info.lloyd_iters, len(f), err), flush=True)
    assert err < FP64_REL_TOL


# ------------------------------------------------------------------ cfg 4

def test_cfg4_batch_every_size_class_dtype_wavelet(wv):
    """One batch over 5 size classes x 3 dtypes x 2 wavelets: every item's
    coefficients, threshold record, bins and counts equal the oracle's, and
    the batch's .cube stream holds, record for record, the oracle's file
    record of each cubelet.  Its records average well over 128 KB, so the
    stream writer copies every payload from the device straight into the
    file image (the batch tests of small cubelets cover the staged copy)."""
    from wavelets_amd import synth
    specs, vols = [], []
    k = 0
    for size in synth.MIXED_SIZES:
        for dt in ("uint8", "int16", "float32"):
            for wave in (wv.HAAR, wv.CDF97):
                vols.append(synth.cubelet(size, dt, 9000 + k))
                specs.append(dict(wavelet=wave, steps=synth.stream_steps(size), alg=wv.LOG,
                                  bins=256, frac=0.9))
                k += 1
    futs = [POOL.submit(O.compress, _oracle_input(v), s["wavelet"], s["alg"], s["bins"],
                        s["steps"], s["frac"], want_coef=True, want_bins=True)
            for v, s in zip(vols, specs)]
    cubes = [dict(vol=torch.from_numpy(v).cuda(), **s) for v, s in zip(vols, specs)]
    b = wv.Batch(cubes)
    b.submit()
    b.wait()
    f = b.write_stream()
    recs = _compressed_records(f)
    assert len(recs) == len(cubes)
    sizes = []
    for i, (fut, s) in enumerate(zip(futs, specs)):
        want = fut.result()
        assert np.array_equal(b.coef[i].cpu().numpy().view(np.uint32),
                              want["coef"].view(np.uint32)), i
        r, zb = b.result(i)
        _check_thresh(r, want["res"])
        assert zb == want["res"].zero_bin, i
        bins = b.bins[i].cpu().numpy()
        assert np.array_equal(bins, want["bins"]), i
        cnt = np.bincount(np.minimum(bins, 256), minlength=257)
        assert np.array_equal(b.counts[i].cpu().numpy().astype(np.int64), cnt), i
        wrec = _compressed_records(want["file"])[0]
        assert recs[i][1] == wrec[1], i          # parameters message
        assert recs[i][2] == wrec[2], i          # payload
        sizes.append(len(wrec[2]))
    # the writer's choice is made from the average payload: the direct copy
    assert sum(sizes) >= len(sizes) * (128 << 10)
    b.close()


# ------------------------------------------------------------------ cfg 5

def test_cfg5_serving_requests_alone_mixed_and_oracle(wv):
    from wavelets_amd import synth
    reqs = synth.serving_requests(20, seed=5)
    vols = [synth.cubelet(size, dt, 5000 + i) for i, (size, dt, keep) in enumerate(reqs)]
    thr = [float(np.float32(1.0 - keep)) for (_, _, keep) in reqs]
    futs = [POOL.submit(O.compress, _oracle_input(v), O.HAAR, O.LLOYD, 256, [4, 4, 4], t,
                        lloyd_max_iters=30, want_bins=True)
            for v, t in zip(vols, thr)]
    alone = []
    srv = wv.Server(slots=1)
    for v, t in zip(vols, thr):
        srv.submit(v, wv.HAAR, [4, 4, 4], wv.LLOYD, 256, t, lloyd_iters=30, want_bins=True)
        (r,) = srv.drain()
        assert r["status"] == 0
        alone.append(r)
    srv.close()
    mix = wv.Server(slots=8)
    ids = [mix.submit(v, wv.HAAR, [4, 4, 4], wv.LLOYD, 256, t, lloyd_iters=30, want_bins=True)
           for v, t in zip(vols, thr)]
    res = {r["id"]: r for r in mix.drain()}
    st = mix.stats()
    mix.close()
    assert st["peak_inflight"] > 1
    for k, i in enumerate(ids):
        a, m = alone[k], res[i]
        want = futs[k].result()
        assert m["status"] == 0
        assert a["file"] == m["file"], k
        assert np.array_equal(a["bins"], m["bins"]), k
        assert a["lloyd_iters"] == m["lloyd_iters"] == want["res"].lloyd_iters, k
        assert np.array_equal(m["bins"], want["bins"]), k
        assert m["file"] == want["file"], k
    print("cfg5: %d requests (%s), identical alone, in the mix (peak %d in flight) and to "
          "the oracle" % (len(reqs), ", ".join("%dx%dx%d" % s for s, _, _ in reqs),
                          st["peak_inflight"]), flush=True)
