"""GPU tests of the device Huffman coder (wv_encode_device): its payload words
equal the oracle coder's (huffScanDupsOnly + computeHuffmanCoding +
encodeToStream, test_compress_common.cc:927-1022, huffman.cc, huffman.h), and
the .cube around them equals the host writer's bytes.  The cases cover zero
runs at every split rule of dupEncodeString (lengths 8, 9, 16, 24, 1023,
1024, 1032, runs over chunk boundaries and runs of 10^5+ bins), the end of
the array, unaligned bins, no duplicate code (-noz), code books too large
for LDS, one-symbol trees and the bins of a real compress."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import wavelets_amd
    return wavelets_amd


def _meta(bins, zb, noz=False):
    return dict(orig=[64, 64, 64], steps=[3, 3, 3], out_size=[64, 64, 64], wavelet=2,
                quant_alg=1, bin_count=bins, orig_dtype=3, thresh_frac=0.9,
                threshold=1.5, max_val=1000.0, zero_bin=-1 if noz else zb, with_footer=1)


def _device_encode(wv, bins_np, nb, dup, offset=0):
    """Encodes bins_np on the device (optionally starting `offset` elements into
    the buffer, so the bins are not 16-byte aligned).  Returns (words, book)."""
    buf = torch.zeros(bins_np.size + offset, dtype=torch.int32, device="cuda")
    buf[offset:] = torch.from_numpy(bins_np).cuda()
    b = buf[offset:]
    counts = np.bincount(np.clip(bins_np, 0, nb), minlength=nb + 1).astype(np.uint32)
    counts_d = torch.from_numpy(counts.view(np.int32)).cuda()
    words, book = wv.encode_device(b, counts_d, nb, dup)
    torch.cuda.synchronize()
    return words.cpu().numpy().view(np.uint32), book, counts


def _check(wv, bins_np, nb, dup, offset=0, container=True):
    words, book, counts = _device_encode(wv, bins_np, nb, dup, offset)
    want, nbits = O.Huffman(bins_np, nb, dup).encode()
    want = np.asarray(want, np.uint32)
    assert book.payload_bits == nbits
    assert words.size == want.size
    bad = np.nonzero(words != want)[0]
    assert bad.size == 0, (bad.size, bad[:5], words[bad[:3]], want[bad[:3]])
    if container:
        meta = _meta(nb, dup, noz=dup < 0)
        got = wv.write_cube_payload(dict(meta), book, words)
        assert got == wv.write_cube(dict(meta), bins_np, counts[:nb])
        # the container around the payload still on the device
        keep = []
        d = wv._cube_desc(dict(meta), keep)
        wd = torch.from_numpy(words.view(np.int32).copy()).cuda()
        buf = np.empty(len(got) + 64, np.uint8)
        nbytes = wv.check(wv.lib().wv_write_cube_device(
            C.byref(d), book.handle, C.c_void_p(wd.data_ptr()) if words.size else None,
            words.size, buf.ctypes.data_as(C.c_void_p), buf.size, wv._stream()))
        assert buf[:nbytes].tobytes() == got
    book.close()


def _runs(n, nb, seed, run_lengths=(), p_run=0.5, max_run=3000):
    """Bins around the zero bin with zero runs of chosen and random lengths."""
    rng = np.random.default_rng(seed)
    zb = (nb - 1) // 2
    x = np.clip(rng.geometric(0.15, n) * rng.choice([-1, 1], n) + zb, 0, nb - 1).astype(np.int32)
    x[x == zb] = zb + 1
    pos = int(rng.integers(0, 50))
    lens = list(run_lengths)
    while pos < n:
        if lens:
            run = lens.pop(0)
        else:
            run = int(rng.integers(1, max_run)) if rng.random() < p_run else 0
        x[pos:pos + run] = zb
        pos += run + int(rng.integers(1, 6))
    return x, zb


@pytest.mark.parametrize("seed", [0, 1])
def test_split_rules(wv, seed):
    lens = [1, 7, 8, 9, 15, 16, 17, 23, 24, 25, 31, 32, 1016, 1023, 1024, 1025, 1031, 1032,
            1033, 2047, 2048, 2056, 4095, 4096, 4097, 5119, 5120, 5121, 9000]
    x, zb = _runs(200_000, 256, seed, run_lengths=lens * 2)
    _check(wv, x, 256, zb)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 4095, 4096, 4097, 8191, 12_289])
def test_sizes(wv, n):
    x, zb = _runs(n, 64, n, p_run=0.6, max_run=700)
    _check(wv, x, 64, zb)


def test_sizes_all_zero_and_tail_runs(wv):
    zb = 127
    for n, tail in [(10, 10), (4096, 4096), (50_000, 50_000), (70_001, 33_333), (9, 9)]:
        x = np.full(n, zb + 3, np.int32)
        x[n - tail:] = zb
        _check(wv, x, 256, zb)


def test_long_runs_over_chunks(wv):
    # runs of 10^5 .. 10^6 bins span many 4096-bin chunks and leave a
    # periodic string of 1024-bin dup codes for the block to write
    rng = np.random.default_rng(5)
    n = 3_000_000
    zb = 511
    x = np.clip(rng.integers(0, 1024, n), 0, 1023).astype(np.int32)
    x[x == zb] = 0
    for start, ln in [(1000, 100_000), (200_000, 1_000_000 + 7), (1_300_003, 8 * 1024),
                      (1_400_000, 777_777), (n - 300_000, 300_000)]:
        x[start:start + ln] = zb
    _check(wv, x, 1024, zb)


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_unaligned_bins(wv, offset):
    x, zb = _runs(50_001, 256, 10 + offset)
    _check(wv, x, 256, zb, offset=offset)


def test_no_dup_code(wv):
    x, zb = _runs(100_000, 256, 3)
    _check(wv, x, 256, -1)


@pytest.mark.parametrize("nb", [4096, 4097, 20_000, 65_536])
def test_large_code_books(wv, nb):
    # code books of more than 4097 entries stay in global memory
    rng = np.random.default_rng(nb)
    n = 300_000
    zb = (nb - 1) // 2
    x = np.clip(np.round(rng.laplace(zb, nb / 40, n)), 0, nb - 1).astype(np.int32)
    x[rng.random(n) < 0.5] = zb
    _check(wv, x, nb, zb)


def test_long_run_codes(wv):
    """Rare zero bin and rare dup key: their codes are long, the codes of a
    run of 1023 pass 120 bits, and the writer codes runs code by code."""
    rng = np.random.default_rng(21)
    n = 400_000
    zb = 511
    x = rng.integers(0, 1024, n).astype(np.int32)
    x[x == zb] = 3
    pos = 100
    for ln in [1023, 9, 1015, 17, 1031, 2047, 8, 7, 4096 + 1023] * 3:
        x[pos:pos + ln] = zb
        pos += ln + 5000
    _check(wv, x, 1024, zb)


def test_single_symbol_trees(wv):
    # one leaf: its code has no bits
    _check(wv, np.full(77_777, 5, np.int32), 16, 7)
    _check(wv, np.full(77_777, 5, np.int32), 16, -1)
    # only zero runs of 8k: only dup codes (3 bits each, a 0-bit key code)
    x = np.full(8 * 1024 * 13, 7, np.int32)
    _check(wv, x, 16, 7)


def test_words_cap_too_small(wv):
    x, zb = _runs(100_000, 256, 8)
    b = torch.from_numpy(x).cuda()
    counts = np.bincount(x, minlength=257).astype(np.uint32)
    cd = torch.from_numpy(counts.view(np.int32)).cuda()
    L = wv.lib()
    ws = torch.empty(L.wv_encode_ws_bytes(x.size, 256), dtype=torch.uint8, device="cuda")
    words = torch.empty(16, dtype=torch.int32, device="cuda")
    nw = C.c_int64()
    rc = L.wv_encode_device(C.c_void_p(b.data_ptr()), x.size, C.c_void_p(cd.data_ptr()), 256,
                            zb, C.c_void_p(ws.data_ptr()), C.c_void_p(words.data_ptr()), 16,
                            C.byref(nw), None, wv._stream())
    assert rc == -4   # WV_ERR_NOMEM
    _, nbits = O.Huffman(x, 256, zb).encode()
    assert nw.value == (nbits + 31) // 32


def test_pipeline_bins(wv):
    """Bins of the quantize + histogram pass on a real volume: the device
    payload is the oracle's, with the counts the pass produced."""
    rng = np.random.default_rng(11)
    g = np.mgrid[0:96, 0:80, 0:64].astype(np.float32)
    vol = (np.sin(g[2] * 0.11) * 400 + g[1] * 3 + rng.standard_normal(g[0].shape) * 30)
    vol = torch.from_numpy(vol.astype(np.float32)).cuda()
    coef, gm, _ = wv.dwt3d_forward(vol, [3, 3, 3], wv.CDF97, with_stats=False)
    _, res = wv.threshold(coef, 0.9)
    q = wv.Quantizer(wv.LOG, 256, res.threshold, res.max_abs)
    bins, counts = q.quantize(coef)
    torch.cuda.synchronize()
    counts_full = torch.zeros(257, dtype=torch.int32, device="cuda")
    counts_full[:256] = counts[:256]
    words, book = wv.encode_device(bins, counts_full, 256, q.zero_bin)
    torch.cuda.synchronize()
    bins_np = bins.cpu().numpy()
    want, nbits = O.Huffman(bins_np, 256, q.zero_bin).encode()
    assert book.payload_bits == nbits
    assert np.array_equal(words.cpu().numpy().view(np.uint32), np.asarray(want, np.uint32))


def test_full_size_payload_matches_host_coder(wv):
    """512^3 bins (134 M, 32768 chunks): the device payload equals the host
    coder's (the library's C++ encoder, which the oracle pins)."""
    n = 512 ** 3
    gen = torch.Generator(device="cuda").manual_seed(3)
    zb = 511
    u = torch.rand(n, device="cuda", generator=gen)
    bins = torch.where(u < 0.93, torch.full_like(u, zb), (u * 1e6) % 1024).to(torch.int32)
    # a few long runs
    bins[1_000_000:3_500_000] = zb
    bins[n - 777_777:] = zb
    counts = torch.bincount(bins, minlength=1025).to(torch.int32)
    words, book = wv.encode_device(bins, counts, 1024, zb)
    torch.cuda.synchronize()
    got = words.cpu().numpy().view(np.uint32)
    bins_np = bins.cpu().numpy()
    meta = _meta(1024, zb)
    host = wv.write_cube(dict(meta), bins_np, counts.cpu().numpy()[:1024])
    dev = wv.write_cube_payload(dict(meta), book, got)
    assert len(dev) == len(host)
    assert dev == host
