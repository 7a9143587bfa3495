"""GPU tests of decompression (decompressFile, test_compress_cpu.cc:187-236):
a .cube written by the oracle's compress is read, Huffman-decoded on the device,
dequantized and inverted on the GPU and restored to the original type
(translateCubeDataToOriginal, test_compress_common.cc:699-747).  The result
equals the oracle's own reconstruction of the same file bit for bit: oracle
decode, dequantize and inverse transform, then the reference's restore rules
(u8: <= 0 -> 0, >= 255 -> 255, else truncation; int32: (int)(f + 0.5f))."""
import numpy as np
import pytest

import _cubeproto as P
import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import wavelets_amd
    return wavelets_amd


def _vol(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    z, y, x = shape
    g = np.mgrid[0:z, 0:y, 0:x].astype(np.float32)
    v = (np.sin(g[2] * 0.21) * 70 + np.cos(g[1] * 0.13) * 40 + g[0] * 0.9 +
         rng.standard_normal(shape).astype(np.float32) * 12 + 120)
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype == np.int32:
        return np.clip(v * 11, 0, 4095).astype(np.int32)
    return v.astype(np.float32)


def oracle_decompress(f):
    """The oracle's reconstruction of a .cube image (the first compressed
    cubelet), in its original type, indexed [z, y, x]."""
    _, cubes, _, _ = P.parse_cube_file(f)
    cb, _raw, payload = cubes[0]
    p = cb.compression_parameters
    n = cb.width * cb.height * cb.depth
    dup_key = p.huffman_dup_key if p.HasField("huffman_dup_key") else -1
    dup_val = p.huffman_dup_value if p.HasField("huffman_dup_value") else -1
    bins = O.huff_decode(np.array(p.huffman_decode_table, np.int32), dup_key, dup_val,
                         np.frombuffer(payload, np.uint32), n)
    assert bins.size == n
    q_id = {1: O.LOG, 2: O.UNIFORM, 3: O.LLOYD}[p.quantization_algorithm]
    if q_id == O.LLOYD:
        q = O.quantizer(q_id, p.quant_bin_count, p.threshold_value, 0.0,
                        np.array(p.quant_bin_boundaries, np.float32),
                        np.array(p.quant_bin_values, np.float32))
    else:
        q = O.quantizer(q_id, p.quant_bin_count, p.threshold_value, p.quant_max_value)
    coef = O.dequantize(q, bins)
    steps = [p.wavelet_transform_steps_x, p.wavelet_transform_steps_y,
             p.wavelet_transform_steps_z]
    back, size = O.dwt3d(coef, [cb.width, cb.height, cb.depth], p.wavelet_algorithm, steps,
                         inverse=True)
    ox, oy, oz = p.original_width, p.original_height, p.original_depth
    f32 = back.reshape(size[2], size[1], size[0])[:oz, :oy, :ox]
    dt = p.original_data_type
    if dt == 1:
        return np.where(f32 <= 0, 0, np.where(f32 >= 255, 255, np.clip(f32, 0, 255))).astype(np.uint8)
    if dt == 2:
        return (f32 + np.float32(0.5)).astype(np.int32)
    return f32.astype(np.float32)


CASES = [
    ((16, 32, 48), np.uint8, O.CDF97, O.LOG, 256, [3, 3, 3], 0.9),
    ((24, 24, 24), np.float32, O.HAAR, O.UNIFORM, 1024, [3, 3, 3], 0.95),
    ((20, 36, 28), np.int32, O.CDF97, O.LOG, 256, [3, 2, 0], 0.8),
    ((16, 16, 32), np.int32, O.HAAR, O.LLOYD, 64, [4, 4, 4], 0.9),
    ((1, 40, 72), np.uint8, O.CDF97, O.UNIFORM, 256, [3, 3, 0], 0.7),
    ((33, 17, 9), np.float32, O.CDF97, O.LOG, 512, [2, 2, 2], 0.5),
    ((64, 64, 64), np.float32, O.CDF97, O.UNIFORM, 1024, [3, 3, 3], 0.95),
]


@pytest.mark.parametrize("shape,dtype,wave,qalg,bins,steps,frac", CASES)
def test_decompress_matches_oracle(wv, shape, dtype, wave, qalg, bins, steps, frac):
    vol = _vol(shape, dtype, 5)
    res = O.compress(vol, wave, qalg, bins, steps, frac, lloyd_max_iters=20)
    f = res["file"]
    want = oracle_decompress(f)
    got, info, _ = wv.decompress(f, with_info=True)
    assert got.dtype == want.dtype and got.shape == want.shape == vol.shape
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_gpu_round_trip(wv):
    """GPU compress (one call) then GPU decompress of its file: the oracle's
    reconstruction of that file."""
    import ctypes as C
    vol = _vol((48, 40, 56), np.uint8, 8)
    vol_d = torch.from_numpy(vol).cuda()
    opt = wv.Options(wv.CDF97, wv.LOG, 256, (C.c_int * 3)(3, 3, 3), C.c_float(0.9), 1, 30, 1)
    orig = (C.c_int * 3)(56, 40, 48)
    L = wv.lib()
    wb = L.wv_compress_work_bytes(orig, C.byref(opt))
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    cap = vol.size * 8 + (1 << 20)
    fbuf = np.empty(cap, np.uint8)
    info = wv.CompressInfo()
    wv.check(L.wv_compress(C.c_void_p(vol_d.data_ptr()), wv.UINT8, orig, C.byref(opt),
                           C.c_void_p(work.data_ptr()), wb, None,
                           fbuf.ctypes.data_as(C.c_void_p), cap, C.byref(info), wv._stream()))
    f = fbuf[:info.file_bytes].tobytes()
    got = wv.decompress(f)
    want = oracle_decompress(f)
    assert np.array_equal(got, want)
    # and it is close to the input (lossy): mean abs error well under the range
    assert np.abs(got.astype(np.float32) - vol).mean() < 10


def test_decompress_rejects_non_cube(wv):
    with pytest.raises(wv.WvError):
        wv.decompress(b"not a cube file" + b"\0" * 80)


def _stream_bins(n, nb, seed, p_zero=0.85, max_run=3000, wide=False):
    rng = np.random.default_rng(seed)
    zb = (nb - 1) // 2
    x = np.clip(rng.geometric(0.2, n) * rng.choice([-1, 1], n) + zb, 0, nb - 1).astype(np.int32)
    pos = 0
    while pos < n:
        run = int(rng.integers(1, max_run)) if rng.random() < p_zero else 0
        x[pos:pos + run] = zb
        pos += run + int(rng.integers(1, 8))
    if wide:
        sel = rng.random(n) < 0.1
        x[sel] = rng.integers(0, nb, int(sel.sum()))
    return x, zb


@pytest.mark.parametrize("n,nb,dup,wide,seed", [
    (50_000, 256, True, False, 0), (50_000, 256, False, False, 1),
    (1_234_567, 1024, True, False, 2), (300_000, 65_536, True, True, 4),
    (9000, 2, True, False, 5), (37, 16, False, True, 6), (2_000_000, 1024, True, True, 7)])
def test_device_decoder_matches_oracle(wv, n, nb, dup, wide, seed):
    """Parallel decoding with self-synchronising subsequences gives the
    oracle decoder's values (runs of the dup value over piece boundaries,
    codes longer than the 12-bit lookup, no dup key)."""
    x, zb = _stream_bins(n, nb, seed, wide=wide)
    d = zb if dup else -1
    h = O.Huffman(x, nb, d)
    words, _ = h.encode()
    table = h.decode_table()
    key = nb if dup else -1
    want = O.huff_decode(table, key, d, words, n)
    assert np.array_equal(want, x)
    got, rounds = wv.huff_decode_device(table, key, d, words, n)
    assert np.array_equal(got, want)
    assert rounds <= 16, rounds   # synchronisation takes a few rounds


def test_device_decoder_one_symbol(wv):
    """One symbol only: its code has no bits, the stream is empty, and like
    the reference decoder (and the oracle) no value can be read."""
    x = np.full(3, 7, np.int32)
    h = O.Huffman(x, 16, -1)
    words, _ = h.encode()
    assert O.huff_decode(h.decode_table(), -1, -1, words, 3).size == 0
    with pytest.raises(wv.WvError):
        wv.huff_decode_device(h.decode_table(), -1, -1, words, 3)


def test_device_decoder_short_stream_is_an_error(wv):
    x, zb = _stream_bins(20_000, 256, 9)
    h = O.Huffman(x, 256, zb)
    words, _ = h.encode()
    with pytest.raises(wv.WvError):
        wv.huff_decode_device(h.decode_table(), 256, zb, words[: len(words) // 2], x.size)


def test_device_decoder_rejects_leaf_outside_range(wv):
    """The device decoder's lookup entries keep 27 bits of a leaf value:
    a table whose leaf exceeds the largest bin count is refused on the host
    before any launch."""
    x = np.array([0, 1, 1, 2, 2, 2, 3, 3] * 64, np.int32)
    h = O.Huffman(x, 4, -1)
    words, _ = h.encode()
    table = h.decode_table().copy()
    got, _ = wv.huff_decode_device(table, -1, -1, words, x.size)
    assert np.array_equal(got, x)
    leaf = next(k for k in range(0, table.size, 2) if table[k] == 0)
    table[leaf + 1] = 1 << 28
    with pytest.raises(wv.WvError, match="malformed"):
        wv.huff_decode_device(table, -1, -1, words, x.size)
