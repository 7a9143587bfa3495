"""GPU tests of the measurement helpers: the streaming copy that gives the
HBM ceiling copies exactly (whole chunks and a tail), and the batch's
serialised stage-timed run leaves the same outputs as a submit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def wv():
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    import wavelets_amd
    wavelets_amd.lib()
    return wavelets_amd


@pytest.mark.parametrize("nfloat", [4, 16384, 16384 * 3 + 36, (1 << 22) + 4])
@pytest.mark.parametrize("nt,bpc,vpt", [(0, 2, 4), (1, 8, 16), (1, 1, 8)])
def test_copy_stream_copies_exactly(wv, nfloat, nt, bpc, vpt):
    src = torch.arange(nfloat, dtype=torch.float32, device="cuda")
    dst = torch.full_like(src, -1.0)
    wv.copy_stream(src, dst, nt, bpc, vpt)
    torch.cuda.synchronize()
    assert torch.equal(src, dst)


def test_copy_stream_rejects_unaligned(wv):
    src = torch.zeros(64, dtype=torch.float32, device="cuda")
    with pytest.raises(wv.WvError):
        wv.copy_stream(src[1:5], torch.zeros(4, dtype=torch.float32, device="cuda"))


def test_batch_stage_times_outputs_equal_submit(wv):
    from wavelets_amd import synth
    cubes = []
    for k, size in enumerate([(32, 32, 32), (64, 64, 64), (64, 32, 16)]):
        for dt in ("uint8", "float32"):
            v = synth.cubelet(size, dt, 300 + k)
            cubes.append(dict(vol=torch.from_numpy(v).cuda(), wavelet=1 + (k & 1),
                              steps=synth.stream_steps(size), alg=wv.LOG, bins=256, frac=0.9))
    b = wv.Batch(cubes)
    b.submit()
    b.wait()
    ref = [(b.coef[i].clone(), b.bins[i].clone(), b.counts[i].clone(), b.result(i)[0].threshold)
           for i in range(len(cubes))]
    for i in range(len(cubes)):
        b.bins[i].fill_(-3)
        b.counts[i].fill_(7)
    st = b.stage_times()
    assert st["transform_ms"] > 0 and st["select_ms"] > 0 and st["quantize_ms"] > 0
    # one launch per pass of every group (a group: one shape and input type);
    # planes of at most 128 x 128 take two passes, the others three
    npass = {}
    for c in cubes:
        v = c["vol"]
        g = wv.geom([v.shape[2], v.shape[1], v.shape[0]], c["steps"], c["wavelet"], wv.dtype_id(v))
        npass[(tuple(v.shape), str(v.dtype))] = wv.lib().wv_dwt3d_num_passes(C.byref(g))
    assert len(npass) == len(b.groups())
    assert st["transform_launches"] == sum(npass.values())
    for i in range(len(cubes)):
        assert torch.equal(b.coef[i].view(torch.int32), ref[i][0].view(torch.int32)), i
        assert torch.equal(b.bins[i], ref[i][1]), i
        assert torch.equal(b.counts[i], ref[i][2]), i
        assert b.result(i)[0].threshold == ref[i][3], i
    b.close()
