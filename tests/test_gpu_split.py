"""GPU tests of split + batched compression (wv_split_compress): a
device-resident volume cut into overlapping boxes (cube split,
cubelet_convert.cc:424-553), every box compressed by the batched path
reading the volume in place, coded on the device and written as one .cube
stream.  Each record equals, byte for byte, the record the oracle writes
when it compresses that box's samples alone (the box offsets aside), and
batch items that are views into a larger volume give exactly what dense
copies of them give."""
import numpy as np
import pytest

import _cubeproto as P
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


def _vol(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    z, y, x = shape
    g = np.mgrid[0:z, 0:y, 0:x].astype(np.float32)
    v = (np.sin(g[2] * 0.19) * 60 + np.cos(g[1] * 0.23) * 40 + g[0] * 1.3 +
         rng.standard_normal(shape).astype(np.float32) * 10 + 120)
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype == np.int16:
        return np.clip(v * 10, 0, 4095).astype(np.int16)
    return v.astype(np.float32)


def _fields(msg):
    """The message as text, one field per line."""
    from google.protobuf import text_format
    return text_format.MessageToString(msg).splitlines()


@pytest.mark.parametrize("shape,dtype,cub,overlap,wave,alg,bins,frac,zeros", [
    ((28, 36, 40), np.uint8, [16, 16, 16], 4, 2, 2, 256, 0.9, True),
    ((20, 24, 33), np.float32, [16, 12, 8], 3, 1, 1, 1024, 0.95, True),
    ((9, 30, 30), np.int16, [16, 16, 6], 0, 2, 2, 256, 0.8, True),
    ((24, 40, 40), np.float32, [32, 32, 24], 6, 2, 1, 256, 0.9, False),   # no dup strings
    ((16, 48, 64), np.float32, [64, 48, 16], 5, 2, 1, 8192, 0.5, True),   # code book past LDS
])
def test_split_compress_records_match_oracle(wv, shape, dtype, cub, overlap, wave, alg, bins,
                                             frac, zeros):
    vol = _vol(shape, dtype, 4)
    print("case", shape, dtype.__name__, cub, overlap, bins, zeros, flush=True)
    f, info = wv.split_compress(torch.from_numpy(vol).cuda(), cub, overlap, wave, [3, 3, 3],
                                alg, bins, frac, compress_zeros=zeros)
    ov = vol.astype(np.int32) if dtype == np.int16 else vol
    boxes, datas = O.split([shape[2], shape[1], shape[0]], cub, overlap, ov)
    _, cubes, index, _ = P.parse_cube_file(f)
    assert info.cubelets == len(boxes) == len(cubes) == len(index.cubelets)
    for k, ((cb, raw, payload), (off, s), d) in enumerate(zip(cubes, boxes, datas)):
        assert (cb.x_offset, cb.y_offset, cb.z_offset) == off, k
        want = O.compress(d, wave, alg, bins, [3, 3, 3], frac, compress_zeros=zeros)["file"]
        _, wcubes, _, _ = P.parse_cube_file(want)
        wcb, wraw, wpayload = wcubes[0]
        assert payload == wpayload, k
        cb.x_offset = cb.y_offset = cb.z_offset = 0
        assert _fields(cb) == _fields(wcb), k
        assert cb.SerializeToString() == wraw, k
        ix = index.cubelets[k]
        assert f[ix.data_file_offset:ix.data_file_offset + ix.byte_count] == payload


def test_batch_views_equal_dense_copies(wv):
    """Items that are views of boxes inside one volume (read through its
    row and slice strides) against contiguous copies of the same boxes."""
    vol = torch.from_numpy(_vol((40, 48, 64), np.float32, 7)).cuda()
    boxes = [((0, 0, 0), (32, 32, 16)), ((24, 8, 4), (32, 32, 16)), ((40, 16, 24), (24, 32, 16)),
             ((3, 5, 7), (20, 17, 9))]
    views, dense = [], []
    for (x, y, z), (w, h, d) in boxes:
        v = vol[z:z + d, y:y + h, x:x + w]
        for lst, t in ((views, v), (dense, v.contiguous())):
            lst.append(dict(vol=t, wavelet=2, steps=[3, 3, 3], alg=2, bins=256, frac=0.9))
    a = wv.Batch(views)
    b = wv.Batch(dense)
    for bb in (a, b):
        bb.submit()
        bb.wait()
    for i in range(len(boxes)):
        assert torch.equal(a.coef[i], b.coef[i]), i
        assert torch.equal(a.bins[i], b.bins[i]), i
        assert torch.equal(a.counts[i], b.counts[i]), i
        ra, _ = a.result(i)
        rb, _ = b.result(i)
        assert ra.threshold == rb.threshold and ra.max_abs == rb.max_abs, i
    a.close()
    b.close()
