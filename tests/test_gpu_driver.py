"""GPU tests of the drop-in driver wv_compress_gpu (test_compress_gpu's
compress path and flags): the .cube it writes for an input .cube equals the
oracle's file byte for byte, and its -err round trip (dequantize + inverse
transform on the GPU) gives the oracle's PSNR within 0.01 dB."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import _cubeproto as P
import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "wavelets_amd", "wv_compress_gpu")
PROTO_DTYPE = {np.uint8: 1, np.int32: 2, np.float32: 3}


def _write_input_cube(path, vol, max_value=0):
    """An uncompressed cubelet file: the 64-byte header, then one record
    (cubelet_file.cc:259-300)."""
    m = P.messages()
    cb = m["CubeletBuffer"]()
    cb.width, cb.height, cb.depth = vol.shape[2], vol.shape[1], vol.shape[0]
    cb.data_type = PROTO_DTYPE[vol.dtype.type]
    cb.byte_count = vol.nbytes
    if max_value:
        cb.maximum_value = max_value
    meta = cb.SerializeToString()
    header = b"SCU cubelets 1.0\n".ljust(24, b"\0") + b"\0" * 40
    with open(path, "wb") as f:
        f.write(header + struct.pack("<I", len(meta)) + meta + vol.tobytes())


def _vol(shape_zyx, dtype, seed):
    rng = np.random.default_rng(seed)
    z, y, x = shape_zyx
    g = np.mgrid[0:z, 0:y, 0:x].astype(np.float32)
    v = (np.sin(g[2] * 0.23) * 50 + np.cos(g[1] * 0.09) * 40 + g[0] * 0.8 +
         rng.standard_normal(shape_zyx).astype(np.float32) * 10 + 120)
    if dtype == np.uint8:
        return np.clip(v, 0, 255).astype(np.uint8)
    if dtype == np.int32:
        return np.clip(v * 12, 0, 4095).astype(np.int32)
    return v.astype(np.float32)


CASES = [
    ((16, 32, 48), np.uint8, "cdf97", "log", 256, "-1", 0.9, 0),
    ((24, 24, 24), np.float32, "haar", "uniform", 1024, "3", 0.95, 4096),
    ((20, 36, 28), np.int32, "cdf97", "log", 256, "3,2", 0.8, 4095),
    ((16, 16, 32), np.int32, "haar", "lloyd", 64, "4,4,4", 0.9, 4095),
]


@pytest.mark.parametrize("shape,dtype,wave,qalg,bins,steps,frac,maxv", CASES)
def test_driver_file_matches_oracle(tmp_path, shape, dtype, wave, qalg, bins, steps, frac, maxv):
    assert os.path.exists(DRIVER), "driver not built"
    vol = _vol(shape, dtype, 3)
    src, dst = str(tmp_path / "in.cube"), str(tmp_path / "out.cube")
    _write_input_cube(src, vol, maxv)
    args = [DRIVER, "-steps", steps, "-thresh", str(frac), "-qcount", str(bins),
            "-wave", wave, "-qalg", qalg, "-err"]
    if qalg == "lloyd":
        args += ["-lloyd-iters", "30"]
    r = subprocess.run(args + [src, dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = open(dst, "rb").read()
    st = [int(s) for s in steps.split(",")]
    st = st * 3 if len(st) == 1 else (st + [0] if len(st) == 2 else st)
    wv_id = {"haar": O.HAAR, "cdf97": O.CDF97}[wave]
    q_id = {"uniform": O.UNIFORM, "log": O.LOG, "lloyd": O.LLOYD}[qalg]
    want = O.compress(vol, wv_id, q_id, bins, st, frac, lloyd_max_iters=30,
                      max_possible=maxv, want_bins=True)
    assert got == want["file"]
    # -err: the oracle's round trip and PSNR formula on the same bins
    res = want["res"]
    if q_id == O.LLOYD:
        cbuf = P.parse_cube_file(want["file"])[1][0][0].compression_parameters
        q = O.quantizer(q_id, bins, res.threshold, 0.0,
                        np.array(cbuf.quant_bin_boundaries, np.float32),
                        np.array(cbuf.quant_bin_values, np.float32))
    else:
        q = O.quantizer(q_id, bins, res.threshold, res.max_abs)
    coef = O.dequantize(q, want["bins"])
    back, size = O.dwt3d(coef, list(res.size), wv_id, list(res.steps), inverse=True)
    psnr, mse = O.error_psnr(vol, back, size, maxv)
    m = re.search(r"PSNR ([0-9.]+) dB", r.stdout)
    assert m, r.stdout
    assert abs(float(m.group(1)) - psnr) <= 0.01, (m.group(1), psnr)


@pytest.mark.parametrize("shape,dtype,wave,qalg,bins,steps,frac,maxv", CASES[:3])
def test_driver_decompress_matches_oracle(tmp_path, shape, dtype, wave, qalg, bins, steps, frac,
                                          maxv):
    """-d on the oracle's compressed file: an uncompressed cubelet holding the
    oracle's reconstruction in the original type (decompressFile,
    test_compress_cpu.cc:187-236)."""
    import wavelets_amd as wv
    from test_gpu_decompress import oracle_decompress
    vol = _vol(shape, dtype, 7)
    st = [int(s) for s in steps.split(",")]
    st = st * 3 if len(st) == 1 else (st + [0] if len(st) == 2 else st)
    wv_id = {"haar": O.HAAR, "cdf97": O.CDF97}[wave]
    q_id = {"uniform": O.UNIFORM, "log": O.LOG, "lloyd": O.LLOYD}[qalg]
    comp = O.compress(vol, wv_id, q_id, bins, st, frac, lloyd_max_iters=30)["file"]
    src, dst = str(tmp_path / "c.cube"), str(tmp_path / "d.cube")
    open(src, "wb").write(comp)
    r = subprocess.run([DRIVER, "-d", src, dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = wv.write_cube_raw(oracle_decompress(comp))
    assert open(dst, "rb").read() == want


def test_driver_opt_table_matches_oracle(tmp_path):
    """-opt prints plotAll's table (optimize.cc:99-122): threshold fraction,
    bin count, size and pSNR of the log quantizer; no output file.  Rows
    spot-checked against the oracle's test points."""
    vol = _vol((8, 32, 48), np.uint8, 11)
    src, dst = str(tmp_path / "in.cube"), str(tmp_path / "out.cube")
    _write_input_cube(src, vol)
    r = subprocess.run([DRIVER, "-opt", "-steps", "3", src, dst], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert not os.path.exists(dst)
    lines = r.stdout.strip().splitlines()
    head = lines.index("Threshold\tBins\tsize\tpSNR")
    rows = [l.split("\t") for l in lines[head + 1:]]
    assert len(rows) == 50 * 32
    srt = O.opt_sorted(vol, [3, 3, 3], O.CDF97)
    n = np.float32(srt.size)
    frac = np.float32(0.5)
    fracs = []
    while frac < np.float32(0.999):
        fracs.append(frac)
        frac = np.float32(frac + np.float32(0.01))
    for k in (0, 5, 31, 32 * 17 + 3, len(rows) - 1):
        f = fracs[k // 32]
        assert rows[k][0] == "%.3f" % f
        b = int(rows[k][1])
        t = srt[int(n * f)]
        size, _, psnr, _ = O.opt_test(vol, [3, 3, 3], O.CDF97, 0, True, t, b, O.LOG)
        assert int(rows[k][2]) == size, k
        assert abs(float(rows[k][3]) - psnr) <= 1e-4 * abs(psnr), k


def _codes_from_decode_table(table):
    """Every leaf's code, root first (0 = left child), from a decode table
    (pairs; a leaf is (0, value); internal pairs hold child ids)."""
    codes, stack = {}, [(0, "")]
    while stack:
        node, path = stack.pop()
        a, b = int(table[2 * node]), int(table[2 * node + 1])
        if a == 0:
            codes[b] = path
        else:
            stack.append((a, path + "0"))
            stack.append((b, path + "1"))
    return codes


@pytest.mark.parametrize("noz", [False, True])
def test_driver_enc_prints_the_encoding(tmp_path, noz):
    """-enc prints Huffman::printEncoding's table (huffman.cc:100-107): the
    values that occur, by decreasing count (ties by value), as
    "value (count): length bits"; checked against the oracle's counts and
    the codes read off the decode table of the file written."""
    vol = _vol((16, 24, 32), np.uint8, 5)
    src, dst = str(tmp_path / "in.cube"), str(tmp_path / "out.cube")
    _write_input_cube(src, vol)
    args = [DRIVER, "-q", "-enc", "-steps", "3", "-thresh", "0.8", "-qcount", "64"]
    if noz:
        args.append("-noz")
    r = subprocess.run(args + [src, dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = O.compress(vol, O.CDF97, O.LOG, 64, [3, 3, 3], 0.8, compress_zeros=not noz,
                      want_bins=True)
    h = O.Huffman(want["bins"], 64, -1 if noz else want["res"].zero_bin)
    counts = h.counts()
    codes = _codes_from_decode_table(h.decode_table())
    leaves = [(v, int(counts[v])) for v in range(65 if not noz else 64) if counts[v] > 0]
    leaves.sort(key=lambda t: (-t[1], t[0]))
    lines = ["%3d (%8d): %d %s" % (v, c, len(codes[v]), codes[v]) for v, c in leaves]
    got = [l for l in r.stdout.splitlines() if re.match(r"^\s*\d+ \(\s*\d+\): \d+ [01]*$", l)]
    assert got == lines
    assert open(dst, "rb").read() == want["file"]


@pytest.mark.parametrize("flags", [["-bq", "before.dat", "-aq", "after.dat", "-pb", "-v"],
                                   ["-experiment"]])
def test_driver_accepts_reference_flags(tmp_path, flags):
    """Options the reference GPU driver parses without effect are accepted
    and leave the file unchanged; -experiment implies -q."""
    vol = _vol((8, 16, 16), np.uint8, 6)
    src, dst = str(tmp_path / "in.cube"), str(tmp_path / "out.cube")
    _write_input_cube(src, vol)
    r = subprocess.run([DRIVER] + flags + ["-thresh", "0.9", src, dst], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert open(dst, "rb").read() == O.compress(vol, O.CDF97, O.LOG, 256, [-1, -1, -1],
                                                0.9)["file"]
    assert not os.path.exists(tmp_path / "before.dat")
    if flags == ["-experiment"]:
        assert r.stdout.strip() == ""


def test_driver_rejects_2d_like_the_reference(tmp_path):
    vol = _vol((8, 16, 16), np.uint8, 7)
    src, dst = str(tmp_path / "in.cube"), str(tmp_path / "out.cube")
    _write_input_cube(src, vol)
    r = subprocess.run([DRIVER, "-2d", src, dst], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "-2d option not implemented for test_compress_gpu yet." in r.stderr
    assert not os.path.exists(dst)
