"""CPU tests of libwv_hip.so: it loads, exports every function the C ABI
header declares, and its host-side logic (geometry, quantizer parameters and
log bin-edge tables, Huffman + .cube writer) matches the oracle bit for bit.
No GPU is touched here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def wv():
    import wavelets_amd
    if not os.path.exists(wavelets_amd.LIB_PATH):
        pytest.fail("libwv_hip.so not built (python -m wavelets_amd.build)")
    wavelets_amd.lib()
    return wavelets_amd


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "wv_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wv_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(wv):
    names = _header_functions()
    assert len(names) > 15
    missing = [n for n in names if not hasattr(wv.lib(), n)]
    assert not missing, missing


def test_geometry_matches_oracle(wv):
    for orig, steps in [([2048, 2048, 1], [-1, -1, -1]), ([512, 512, 512], [3, 3, 3]),
                        ([100, 37, 5], [9, 2, -1]), ([33, 64, 17], [3, 3, 3]),
                        ([7, 1, 1], [0, 0, 0])]:
        g = wv.geom(orig, steps, wv.CDF97, wv.FLOAT32)
        st = O.set_steps(steps, orig)
        assert list(g.steps) == st
        assert list(g.padded) == O.padded_size(orig, st)
        d, s = O.dwt3d(np.zeros(g.count, np.float32), list(g.padded), O.CDF97, st)
        assert list(g.out_size) == s


def test_forward_pass_plan(wv):
    # planes of at most 32 x 32 with x and y levels: one plane pass (x and y
    # in LDS) plus the z pass; otherwise one pass per axis, a depth-1 plane
    # needs no z pass, and no levels at all is one conversion pass
    for orig, steps, n in [([32, 32, 32], [3, 3, 3], 2), ([33, 20, 17], [3, 3, 3], 2),
                           ([32, 32, 1], [3, 3, 0], 1), ([32, 32, 32], [0, 3, 3], 2),
                           ([64, 64, 64], [3, 3, 3], 3), ([512, 512, 16], [3, 3, 3], 3),
                           ([2048, 2048, 1], [-1, -1, -1], 2), ([16, 16, 16], [0, 0, 0], 1)]:
        g = wv.geom(orig, steps, wv.CDF97, wv.FLOAT32)
        assert wv.lib().wv_dwt3d_num_passes(C.byref(g)) == n, (orig, steps)


@pytest.mark.parametrize("alg", [1, 2])
@pytest.mark.parametrize("bins,thr,mx", [(256, 1.161753893, 421519.75), (1024, 31.5, 40000.0),
                                         (255, 0.5, 10.0), (7, 1.0, 9.0), (2, 0.25, 4.0)])
def test_quant_params_and_host_quant_match_oracle(wv, alg, bins, thr, mx):
    q = wv.Quantizer.__new__(wv.Quantizer)
    p = wv.QuantParams()
    wv.check(wv.lib().wv_quant_params_init(C.byref(p), alg, bins, C.c_float(thr),
                                           C.c_float(mx), None, None))
    oq = O.quantizer(alg, bins, thr, mx)
    for f in ("neg_scale", "neg_inv_scale", "neg_offset", "pos_scale", "pos_inv_scale",
              "pos_offset"):
        assert np.float32(getattr(p, f)) == np.float32(getattr(oq, f)), f
    rng = np.random.default_rng(bins)
    x = np.concatenate([rng.uniform(-mx, mx, 3000),
                        np.exp(rng.uniform(np.log(thr), np.log(mx), 3000)) *
                        rng.choice([-1, 1], 3000)]).astype(np.float32)
    want = O.quantize(oq, x)
    got = np.array([wv.lib().wv_quant_host(C.byref(p), C.c_float(v)) for v in x])
    assert np.array_equal(got, want)
    assert p.zero_bin == O.lib().or_quant(C.byref(oq), C.c_float(0.0))
    wv.lib().wv_quant_params_free(C.byref(p))


def _edge_quant(pe, ne, zb, thr, x, clamp, bins):
    """The device's comparison-only log quantizer, in numpy."""
    a = np.abs(x)
    pos = zb + np.searchsorted(pe, x, side="right")
    pos = np.where(x >= clamp, bins - 1, pos)
    neg = zb - np.searchsorted(ne, a, side="right")
    out = np.where(x < 0, neg, pos)
    return np.where(a <= thr, zb, out).astype(np.int32)


@pytest.mark.parametrize("bins,thr,mx", [(256, 1.161753893, 421519.75),
                                         (1024, 31.5, 40000.0), (256, 0.0625, 7.0e6),
                                         (5, 2.0, 3.0)])
def test_log_edges_reproduce_host_logf_exactly(wv, bins, thr, mx):
    """Every float within 2^12 ulps of every bin edge (where a comparison
    quantizer could differ from the logf one), plus a sweep of the whole
    range: the edge tables give the host quantizer's bin."""
    p = wv.QuantParams()
    wv.check(wv.lib().wv_quant_params_init(C.byref(p), 2, bins, C.c_float(thr),
                                           C.c_float(mx), None, None))
    pe = np.ctypeslib.as_array(p.pos_edges, (p.n_pos_edges,)).copy()
    ne = np.ctypeslib.as_array(p.neg_edges, (p.n_neg_edges,)).copy()
    assert (np.diff(pe) >= 0).all() and (np.diff(ne) >= 0).all()
    oq = O.quantizer(2, bins, thr, mx)
    e = np.concatenate([pe, ne])
    e = e[np.isfinite(e)].view(np.uint32).astype(np.int64)
    off = np.arange(-4096, 4097, dtype=np.int64)
    u = (e[:, None] + off[None, :]).ravel()
    u = u[(u > 0) & (u < 0x7f800000)].astype(np.uint32)
    x = u.view(np.float32)
    sweep = np.linspace(np.float32(thr).view(np.uint32).astype(np.int64),
                        np.float32(mx * 1.5).view(np.uint32).astype(np.int64),
                        400_000).astype(np.uint32).view(np.float32)
    x = np.concatenate([x, sweep])
    x = np.concatenate([x, -x])
    want = O.quantize(oq, x)
    got = _edge_quant(pe, ne, p.zero_bin, np.float32(thr), x, np.float32(p.pos_clamp), bins)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (bad.size, x[bad[:5]], got[bad[:5]], want[bad[:5]])
    wv.lib().wv_quant_params_free(C.byref(p))


def _bins_like(n, bins, seed):
    rng = np.random.default_rng(seed)
    zb = (bins - 1) // 2
    x = np.clip(rng.geometric(0.2, n) * rng.choice([-1, 1], n) + zb, 0, bins - 1)
    x = x.astype(np.int32)
    pos = 0
    while pos < n:
        run = int(rng.integers(1, 2000)) if rng.random() < 0.6 else 0
        x[pos:pos + run] = zb
        pos += run + int(rng.integers(1, 30))
    return x, zb


@pytest.mark.parametrize("alg,footer,noz", [(1, True, False), (2, False, False),
                                            (2, True, True), (3, True, False)])
def test_write_cube_matches_oracle_bytes(wv, alg, footer, noz):
    bins, zb = _bins_like(60_000, 256, alg)
    counts = np.bincount(bins, minlength=256).astype(np.uint32)
    bnd = np.linspace(-100, 100, 255).astype(np.float32)
    val = np.linspace(-101, 101, 256).astype(np.float32)
    meta = dict(orig=[50, 40, 30], steps=[3, 3, 2], out_size=[32, 56, 40], wavelet=2,
                quant_alg=alg, bin_count=256, orig_dtype=3, thresh_frac=0.9,
                threshold=1.25, max_val=4000.5, zero_bin=-1 if noz else zb,
                with_footer=int(footer))
    if alg == 3:
        meta.update(boundaries=bnd, values=val)
    got = wv.write_cube(dict(meta), bins, counts)
    # the oracle's encoder + writer on the same bins
    h = O.Huffman(bins, 256, -1 if noz else zb)
    words, _ = h.encode()
    dec = h.decode_table()
    m = O.lib()
    import _cubeproto
    hdr, cubes, index, _raw = _cubeproto.parse_cube_file(got)
    cb, raw, payload = cubes[0]
    assert payload == words.tobytes()
    p = cb.compression_parameters
    assert list(p.huffman_decode_table) == list(dec)
    assert [cb.width, cb.height, cb.depth] == [32, 56, 40]
    assert cb.SerializeToString() == raw
    assert (index is not None) == footer
    if alg == 3:
        assert np.array_equal(np.array(p.quant_bin_boundaries, np.float32), bnd)
    assert p.HasField("huffman_dup_key") == (not noz)
