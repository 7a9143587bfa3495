"""CPU tests of cube split with overlap (cube split, cube_split_typed
cubelet_convert.cc:368-553): the library's box plan against the oracle's
restatement of the reference's frame loop, the oracle itself against what
the split must satisfy (every sample covered, neighbouring boxes sharing
exactly `overlap` samples, box counts from the closed form, no-overlap
splits tiling the volume exactly, a worked example), and the uncompressed
cubelet stream against messages parsed by the protobuf runtime.  No GPU
calls."""
import itertools

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


def test_oracle_worked_example():
    """10 x 10 x 11 into 4^3 with overlap 1: x / y boxes at 0, 3, 6 (the
    last one ends exactly at 10); z chunks at 0, 3, 6 of 4 frames, then the
    two frames 9..10 left at the end."""
    b = O.split([10, 10, 11], [4, 4, 4], 1)
    assert len(b) == 36
    xy = [(0, 0), (3, 0), (6, 0), (0, 3), (3, 3), (6, 3), (0, 6), (3, 6), (6, 6)]
    want = [((x, y, z), (4, 4, d)) for z, d in ((0, 4), (3, 4), (6, 4), (9, 2))
            for x, y in xy]
    assert b == want


def _coverage(size, boxes):
    cov = np.zeros(size[::-1], np.int32)
    for off, s in boxes:
        cov[off[2]:off[2] + s[2], off[1]:off[1] + s[1], off[0]:off[0] + s[0]] += 1
    return cov


@pytest.mark.parametrize("size,cub,overlap", [
    ([40, 36, 28], [16, 16, 16], 0), ([40, 36, 28], [16, 16, 16], 4),
    ([33, 17, 9], [8, 5, 3], 2), ([7, 7, 7], [8, 8, 8], 3), ([64, 64, 1], [16, 16, 16], 0),
    ([20, 20, 20], [6, 7, 5], 4), ([9, 9, 30], [9, 9, 10], 9 - 1)])
def test_oracle_invariants(size, cub, overlap):
    boxes = O.split(size, cub, overlap)
    # box counts: (size - overlap - 1) // (cub - overlap) + 1 per x / y axis
    nx = (size[0] - overlap - 1) // (cub[0] - overlap) + 1 if size[0] > overlap else 1
    ny = (size[1] - overlap - 1) // (cub[1] - overlap) + 1 if size[1] > overlap else 1
    assert len(boxes) % (nx * ny) == 0
    for off, s in boxes:
        assert all(0 <= off[a] and 0 < s[a] and off[a] + s[a] <= size[a] for a in range(3))
        assert all(s[a] <= cub[a] for a in range(3))
    cov = _coverage(size, boxes)
    assert cov.min() >= 1                      # every sample is in some box
    if overlap == 0:
        assert cov.max() == 1                  # no overlap: an exact tiling
    # neighbours along x share exactly `overlap` samples
    xs = sorted({off[0] for off, _ in boxes})
    for a, b in zip(xs, xs[1:]):
        w = [s[0] for off, s in boxes if off[0] == a][0]
        assert a + w - b == overlap
    zs = sorted({(off[2], s[2]) for off, s in boxes})
    for (a, da), (b, _) in zip(zs, zs[1:]):
        assert a + da - b == overlap


def test_plan_matches_oracle(wv):
    rng = np.random.default_rng(3)
    cases = 0
    for _ in range(300):
        size = [int(v) for v in rng.integers(1, 70, 3)]
        cub = [int(v) for v in rng.integers(1, 40, 3)]
        overlap = int(rng.integers(0, min(cub)))
        assert wv.split_plan(size, cub, overlap) == O.split(size, cub, overlap), \
            (size, cub, overlap)
        cases += 1
    assert cases == 300


@pytest.mark.parametrize("cub,overlap", [([4, 4, 4], 4), ([4, 4, 4], 5), ([4, 0, 4], 0),
                                         ([4, 4, 4], -1)])
def test_plan_rejects_bad_splits(wv, cub, overlap):
    with pytest.raises(wv.WvError):
        wv.split_plan([16, 16, 16], cub, overlap)


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float32])
def test_split_raw_stream(wv, dtype):
    """The uncompressed stream: one record per box in write order, the
    boxes' offsets and sizes, their samples, and an index footer whose
    data_file_offset points at each record's data."""
    rng = np.random.default_rng(1)
    vol = (rng.random((11, 13, 17)) * 200).astype(dtype)
    f = wv.split_raw(vol, [6, 5, 4], 2)
    boxes, datas = O.split([17, 13, 11], [6, 5, 4], 2, vol)
    _, cubes, index, _ = P.parse_cube_file(f)
    assert len(cubes) == len(boxes) == len(index.cubelets)
    dt = {np.uint8: 1, np.int32: 2, np.float32: 3}[dtype]
    for (cb, _raw, payload), (off, s), d, ix in zip(cubes, boxes, datas, index.cubelets):
        assert (cb.width, cb.height, cb.depth) == s
        assert (cb.x_offset, cb.y_offset, cb.z_offset) == off
        assert cb.data_type == dt and cb.byte_count == d.nbytes
        assert payload == d.tobytes()
        assert f[ix.data_file_offset:ix.data_file_offset + ix.byte_count] == d.tobytes()
        assert (ix.x_offset, ix.y_offset, ix.z_offset) == off
