"""CPU tests of the decompression host side: the host Huffman decoder against
the oracle's decoder (HuffmanDecoder::decodeFromStream, huffman.h:96-131) on
the oracle coder's streams, the .cube reader against files the oracle writes
(readQuantData's rule, test_compress_common.cc:850-894), and the uncompressed
cubelet writer against messages built by the protobuf runtime from the
.proto descriptor (CubeletStreamWriter::addCubelet / close,
cubelet_file.cc:259-327).  No GPU calls."""
import numpy as np
import pytest

import _cubeproto as P
import _oracle as O


@pytest.fixture(scope="module")
def wv():
    import wavelets_amd
    if not __import__("os").path.exists(wavelets_amd.LIB_PATH):
        pytest.skip("library not built")
    return wavelets_amd


def _bins(n, nb, seed, p_zero=0.85, max_run=3000):
    rng = np.random.default_rng(seed)
    zb = (nb - 1) // 2
    x = np.clip(rng.geometric(0.2, n) * rng.choice([-1, 1], n) + zb, 0, nb - 1).astype(np.int32)
    pos = 0
    while pos < n:
        run = int(rng.integers(1, max_run)) if rng.random() < p_zero else 0
        x[pos:pos + run] = zb
        pos += run + int(rng.integers(1, 8))
    return x, zb


@pytest.mark.parametrize("n,nb,dup,seed", [(50_000, 256, True, 0), (50_000, 256, False, 1),
                                            (123_457, 1024, True, 2), (7, 16, True, 3),
                                            (200_000, 65_536, True, 4), (9000, 2, True, 5)])
def test_host_decoder_matches_oracle(wv, n, nb, dup, seed):
    x, zb = _bins(n, nb, seed)
    if nb > 4096:   # wide code book: long codes exercise the bit-by-bit tail
        rng = np.random.default_rng(seed)
        sel = rng.random(n) < 0.1
        x[sel] = rng.integers(0, nb, int(sel.sum()))
    d = zb if dup else -1
    h = O.Huffman(x, nb, d)
    words, nbits = h.encode()
    table = h.decode_table()
    key = nb if dup else -1
    want = O.huff_decode(table, key, d, words, n)
    got = wv.huff_decode_host(table, key, d, words, n)
    assert np.array_equal(got, want)
    if n >= 100:
        assert np.array_equal(want, x)   # the round trip itself
    else:
        # one symbol only: its code has no bits, the stream is empty and the
        # reference decoder returns no values (decodeFromStream's isEmpty)
        assert want.size == 0


def test_host_decoder_stops_at_stream_end(wv):
    x, zb = _bins(20_000, 256, 9)
    h = O.Huffman(x, 256, zb)
    words, _ = h.encode()
    table = h.decode_table()
    cut = words[: len(words) // 2]
    want = O.huff_decode(table, 256, zb, cut, x.size)
    got = wv.huff_decode_host(table, 256, zb, cut, x.size)
    assert got.size == want.size < x.size
    assert np.array_equal(got, want)


def _vol(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    z, y, x = shape
    g = np.mgrid[0:z, 0:y, 0:x].astype(np.float32)
    v = np.sin(g[2] * 0.2) * 60 + g[1] * 0.7 + rng.standard_normal(shape).astype(np.float32) * 9 + 100
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype == np.int32:
        return np.clip(v * 10, 0, 4095).astype(np.int32)
    return v.astype(np.float32)


@pytest.mark.parametrize("q", [O.UNIFORM, O.LOG, O.LLOYD])
def test_cube_parse_matches_oracle_file(wv, q):
    vol = _vol((12, 20, 28), np.int32, 1)
    res = O.compress(vol, O.CDF97, q, 64, [2, 2, 1], 0.85, lloyd_max_iters=10, want_bins=True)
    f = res["file"]
    info, table, bnd, val = wv.cube_parse(f)
    _, cubes, _, _ = P.parse_cube_file(f)
    cb, _raw, payload = cubes[0]
    p = cb.compression_parameters
    assert [info.size[i] for i in range(3)] == [cb.width, cb.height, cb.depth]
    assert [info.orig[i] for i in range(3)] == [28, 20, 12]
    assert [info.steps[i] for i in range(3)] == [2, 2, 1]
    assert info.wavelet == O.CDF97 and info.quant_alg == q and info.bin_count == 64
    assert np.float32(info.threshold) == np.float32(p.threshold_value)
    assert info.orig_dtype == 2
    assert list(table) == list(p.huffman_decode_table)
    assert info.dup_key == p.huffman_dup_key and info.dup_value == p.huffman_dup_value
    assert f[info.payload_offset:info.payload_offset + info.payload_bytes] == payload
    if q == O.LLOYD:
        assert np.array_equal(bnd, np.array(p.quant_bin_boundaries, np.float32))
        assert np.array_equal(val, np.array(p.quant_bin_values, np.float32))
    else:
        assert np.float32(info.max_val) == np.float32(p.quant_max_value)
    # the payload decodes to the compressed bins
    got = wv.huff_decode_host(table, info.dup_key, info.dup_value,
                              np.frombuffer(payload, np.uint32), res["bins"].size)
    assert np.array_equal(got, res["bins"])


@pytest.mark.parametrize("dtype,shape", [(np.uint8, (3, 5, 7)), (np.int32, (4, 4, 9)),
                                         (np.float32, (2, 6, 3))])
def test_write_cube_raw_matches_protobuf(wv, dtype, shape):
    vol = _vol(shape, dtype, 2)
    got = wv.write_cube_raw(vol)
    m = P.messages()
    cb = m["CubeletBuffer"]()
    cb.width, cb.height, cb.depth = shape[2], shape[1], shape[0]
    cb.x_offset = cb.y_offset = cb.z_offset = 0
    cb.data_type = {np.uint8: 1, np.int32: 2, np.float32: 3}[dtype]
    cb.byte_count = vol.nbytes
    cb.compression_algorithm = 0
    meta = cb.SerializeToString()
    import struct
    head = b"SCU cubelets 1.0\n".ljust(64, b"\0") + struct.pack("<I", len(meta)) + meta
    cb.data_file_offset = len(head)
    index = m["CubeletIndexBuffer"]()
    index.cubelets.add().CopyFrom(cb)
    ib = index.SerializeToString()
    want = head + vol.tobytes() + struct.pack("<i", -1) + ib + struct.pack("<i", len(ib)) + b"cube"
    assert got == want


@pytest.mark.parametrize("bad", [-1, 65, 1 << 27, 1 << 30])
def test_cube_parse_rejects_decode_leaf_outside_bins(wv, bad):
    """A decode-table leaf holds a bin (0..bin_count-1) or the dup key
    (bin_count); the reader rejects any other value, so the host and the
    device decoders never see a table they would read differently."""
    import struct
    vol = _vol((8, 16, 16), np.int32, 3)
    f = O.compress(vol, O.HAAR, O.LOG, 64, [2, 2, 2], 0.8)["file"]
    wv.cube_parse(f)                                  # the oracle's file is fine
    mlen = struct.unpack_from("<I", f, 64)[0]
    cb = P.messages()["CubeletBuffer"].FromString(f[68:68 + mlen])
    t = cb.compression_parameters.huffman_decode_table
    leaf = next(h for h in range(0, len(t), 2) if t[h] == 0)
    t[leaf + 1] = bad
    meta = cb.SerializeToString()
    g = f[:64] + struct.pack("<I", len(meta)) + meta + f[68 + mlen:]
    with pytest.raises(wv.WvError, match="leaf"):
        wv.cube_parse(g)


def test_parse_and_host_decode_survive_corrupted_files(wv):
    """Mutated and truncated .cube images: the reader and the host decoder
    either reject them (WvError) or return a result -- never read or write
    out of bounds (the host-sanitized build runs this too)."""
    rng = np.random.default_rng(11)
    vol = _vol((8, 16, 24), np.int32, 4)
    base = O.compress(vol, O.CDF97, O.LOG, 64, [2, 2, 2], 0.7)["file"]
    outcomes = {"rejected": 0, "parsed": 0}
    for k in range(300):
        f = bytearray(base)
        if k % 3 == 0:
            f = f[:int(rng.integers(70, len(f)))]
        for _ in range(int(rng.integers(1, 6))):
            pos = int(rng.integers(64, min(len(f), 64 + 400)))
            f[pos] = int(rng.integers(0, 256))
        try:
            info, table, _, _ = wv.cube_parse(bytes(f))
        except wv.WvError:
            outcomes["rejected"] += 1
            continue
        outcomes["parsed"] += 1
        end = min(len(f), info.payload_offset + info.payload_bytes)
        words = np.frombuffer(bytes(f[info.payload_offset:end - (end - info.payload_offset) % 4]),
                              np.uint32)
        count = int(min(1 << 20, max(1, info.size[0]) * max(1, info.size[1]) *
                        max(1, info.size[2])))
        try:
            wv.huff_decode_host(table, info.dup_key, info.dup_value, words, count)
        except wv.WvError:
            pass
    assert outcomes["rejected"] > 0 and outcomes["parsed"] > 0
