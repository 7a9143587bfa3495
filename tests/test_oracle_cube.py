"""Pins the oracle's .cube writer and the whole CPU pipeline (CPU only).

  * the metadata bytes re-parse with the Python protobuf runtime against the
    schema of wavelet_compress.proto and re-serialise to the same bytes
    (canonical field order), with the field values the reference sets
    (wavelet.cc:283-343, cubelet_file.cc:18-84,259-327);
  * header text and layout (cubelet_file.cc:19-51), the footer
    (0xFFFFFFFF, index, length, "cube");
  * decode(file) reproduces the quantized bins, and the round trip gives the
    PSNR the error metric defines.
"""
import numpy as np

import _cubeproto
import _oracle as O


def _smooth_u8(nx, ny, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    img = np.zeros((ny, nx))
    for _ in range(8):
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
        s = rng.uniform(4, 30)
        img += rng.uniform(-1, 1) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    img += 0.05 * rng.standard_normal(img.shape)
    img = (img - img.min()) / (img.max() - img.min()) * 255
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


def _check_file(out, vol, wavelet, alg, bins, frac, footer):
    res = out["res"]
    header, cubes, index, raw_index = _cubeproto.parse_cube_file(out["file"])
    assert header[:24] == b"SCU cubelets 1.0\n" + b"\0" * 7
    assert header[24:] == b"\0" * 40
    assert len(cubes) == 1
    cb, raw, payload = cubes[0]
    assert cb.SerializeToString() == raw          # canonical encoding
    assert [cb.width, cb.height, cb.depth] == list(res.size)
    assert cb.data_type == 2 and cb.compression_algorithm == 2
    assert cb.byte_count == res.compressed_bytes == len(payload)
    assert cb.x_offset == 0 and cb.HasField("x_offset")
    assert not cb.HasField("maximum_value")
    p = cb.compression_parameters
    assert [p.original_width, p.original_height, p.original_depth] == \
        [vol.shape[2], vol.shape[1], vol.shape[0]]
    assert [p.wavelet_transform_steps_x, p.wavelet_transform_steps_y,
            p.wavelet_transform_steps_z] == list(res.steps)
    assert p.wavelet_algorithm == wavelet
    assert p.threshold_fraction == np.float32(frac)
    assert p.threshold_value == res.threshold
    assert p.quant_bin_count == bins
    assert p.quantization_algorithm == {O.UNIFORM: 2, O.LOG: 1, O.LLOYD: 3}[alg]
    assert p.HasField("transform_2d") and not p.transform_2d
    if alg in (O.UNIFORM, O.LOG):
        assert p.quant_max_value == res.max_abs
    else:
        assert len(p.quant_bin_boundaries) == bins - 1
        assert len(p.quant_bin_values) == bins
    if footer:
        assert index is not None and len(index.cubelets) == 1
        c2 = index.cubelets[0]
        assert c2.data_file_offset == 64 + 4 + len(raw)
        assert index.SerializeToString() == raw_index
    else:
        assert index is None
    # decode the payload with the table from the file
    words = np.frombuffer(payload, np.uint32)
    n = int(np.prod(res.size))
    key = p.huffman_dup_key if p.HasField("huffman_dup_key") else -1
    val = p.huffman_dup_value if p.HasField("huffman_dup_value") else -1
    dec = O.huff_decode(np.array(p.huffman_decode_table, np.int32), key, val, words, n)
    return p, dec


def test_pipeline_2d_log_cdf97():
    img = _smooth_u8(128, 96, 1)[None]          # [z=1, y, x]
    out = O.compress(img, O.CDF97, O.LOG, 256, [-1, -1, -1], 0.9, want_bins=True,
                     want_coef=True)
    p, dec = _check_file(out, img, O.CDF97, O.LOG, 256, 0.9, True)
    assert list(out["res"].size) == [1, 128, 128]   # 96 rows pad to 2^7
    assert np.array_equal(dec, out["bins"])
    assert p.huffman_dup_key == 256 and p.huffman_dup_value == 127


def test_pipeline_3d_uniform_haar_nofooter_noz():
    rng = np.random.default_rng(3)
    vol = (rng.standard_normal((20, 24, 40)) * 100).astype(np.float32)
    out = O.compress(vol, O.HAAR, O.UNIFORM, 1024, [3, 3, 3], 0.95,
                     compress_zeros=False, with_footer=False, want_bins=True)
    p, dec = _check_file(out, vol, O.HAAR, O.UNIFORM, 1024, 0.95, False)
    assert not p.HasField("huffman_dup_key")
    assert np.array_equal(dec, out["bins"])
    assert list(out["res"].size) == [24, 40, 24]   # padded 40x24x24, rotated


def test_pipeline_lloyd_int32():
    rng = np.random.default_rng(4)
    vol = rng.integers(0, 4096, (16, 32, 32)).astype(np.int32)
    out = O.compress(vol, O.HAAR, O.LLOYD, 256, [4, 4, 4], 0.9, lloyd_max_iters=30,
                     want_bins=True)
    p, dec = _check_file(out, vol, O.HAAR, O.LLOYD, 256, 0.9, True)
    assert np.array_equal(dec, out["bins"])
    assert 0 < out["res"].lloyd_iters <= 30
    # codebook layout: values ascending, zero bin at the threshold boundaries
    v = np.array(p.quant_bin_values, np.float32)
    assert (np.diff(v) >= 0).all()


def test_roundtrip_psnr_u8():
    img = _smooth_u8(64, 64, 7)[None]
    out = O.compress(img, O.CDF97, O.UNIFORM, 256, [-1, -1, -1], 0.5, want_bins=True)
    res = out["res"]
    q = O.quantizer(O.UNIFORM, 256, res.threshold, res.max_abs)
    coef = O.dequantize(q, out["bins"])
    back, size = O.dwt3d(coef, list(res.size), O.CDF97, list(res.steps), inverse=True)
    psnr, mse = O.error_psnr(img, back, size, 0)
    # the formula of test_compress_common.h:148-153 on the same data
    b = np.clip(back.reshape(img.shape), 0, 255)
    px = np.where(back.reshape(img.shape) <= 0, 0,
                  np.where(back.reshape(img.shape) >= 255, 255, np.trunc(b))).astype(np.float32)
    d = np.abs(img.astype(np.float32) - px)
    mse_np = np.float32((d.astype(np.float32) ** 2).astype(np.float64).sum() / d.size)
    assert abs(mse - mse_np) <= 1e-4 * max(1.0, mse_np)
    assert 20 < psnr < 60
