"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was
built, to oracle/_ref/libref.so (the reference's own protobuf-free sources).

Test infrastructure only: the oracle is the checker, never the thing measured
or shipped.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

UNIFORM, LOG, LLOYD = 1, 2, 3
HAAR, CDF97 = 1, 2
U8, I32, F32 = 1, 2, 3

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run make -C oracle")
        _lib = C.CDLL(ORACLE_SO)
        _declare(_lib)
    return _lib


def ref():
    """The reference's own code, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = C.CDLL(REF_SO)
    return _ref


P = C.c_void_p
I = C.c_int
F = C.c_float


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


class Quantizer(C.Structure):
    _fields_ = [("alg", I), ("bin_count", I), ("threshold", F), ("max_val", F),
                ("neg_scale", F), ("neg_inv_scale", F), ("neg_offset", F),
                ("pos_scale", F), ("pos_inv_scale", F), ("pos_offset", F),
                ("n_boundaries", I), ("boundaries", P), ("values", P)]


class Options(C.Structure):
    _fields_ = [("wavelet", I), ("quant_alg", I), ("bin_count", I),
                ("steps", I * 3), ("thresh_frac", F), ("compress_zeros", I),
                ("lloyd_max_iters", I), ("with_footer", I)]


class Result(C.Structure):
    _fields_ = [("size", I * 3), ("steps", I * 3), ("threshold", F),
                ("min_val", F), ("max_val", F), ("max_abs", F),
                ("thresh_idx", I), ("nonzero_count", I), ("zero_bin", I),
                ("lloyd_iters", I), ("compressed_bytes", C.c_int64),
                ("file_bytes", C.c_int64), ("t_dwt", C.c_double),
                ("t_thresh", C.c_double), ("t_quant", C.c_double),
                ("t_huff", C.c_double), ("t_total", C.c_double)]


class ErrorAcc(C.Structure):
    _fields_ = [("sum_value", C.c_double), ("sum_value_sq", C.c_double),
                ("sum_diff", C.c_double), ("sum_diff_sq", C.c_double),
                ("max_value", F), ("count", I)]


def _declare(L):
    L.or_ceil_log2.restype = C.c_uint
    L.or_ceil_log2.argtypes = [C.c_uint]
    L.or_count_leading_zeros.restype = C.c_uint
    L.or_count_leading_zeros.argtypes = [C.c_uint]
    L.or_padded_length.argtypes = [I, I]
    L.or_set_steps.argtypes = [P, P]
    L.or_split.argtypes = [P, P, I, P, I, P, I, P]
    L.or_opt_test.restype = C.c_int64
    L.or_opt_test.argtypes = [P, I, P, P, I, I, I, F, I, I, C.POINTER(ErrorAcc),
                              C.POINTER(I)]
    L.or_opt_sorted.argtypes = [P, I, P, P, I, P, C.c_int64]
    L.or_pad_convert.argtypes = [P, I, P, P, P]
    for f in ("or_cdf97_fwd", "or_cdf97_inv", "or_haar_fwd", "or_haar_inv"):
        getattr(L, f).argtypes = [I, P, I, P]
    L.or_dwt3d.argtypes = [P, P, P, I, P, I]
    L.or_dwt3d_f64.argtypes = [P, P, I, P, P]
    L.or_thresh.restype = F
    L.or_thresh.argtypes = [P, I, F, P, P, P, P, P, P]
    L.or_quant_init.argtypes = [C.POINTER(Quantizer), I, I, F, F, P, P]
    L.or_quant.argtypes = [C.POINTER(Quantizer), F]
    L.or_quantize.argtypes = [C.POINTER(Quantizer), P, P, I]
    L.or_dequantize.argtypes = [C.POINTER(Quantizer), P, P, I]
    L.or_lloyd.argtypes = [P, C.c_uint, P, C.c_uint, P, P, P, P, F, P, I]
    L.or_compute_lloyd_quantization.argtypes = [P, I, I, F, F, F, P, P, I, P]
    L.or_huff_create_scan.restype = P
    L.or_huff_create_scan.argtypes = [P, I, I, I]
    L.or_huff_create_counts.restype = P
    L.or_huff_create_counts.argtypes = [P, P, I, I, I]
    L.or_huff_free.argtypes = [P]
    L.or_huff_get_counts.argtypes = [P, P]
    L.or_huff_code_length.argtypes = [P, I]
    L.or_huff_longest.argtypes = [P]
    L.or_huff_decode_table.argtypes = [P, P, I]
    L.or_huff_encode.restype = C.c_int64
    L.or_huff_encode.argtypes = [P, P, I, P, C.c_int64, P]
    L.or_huff_decode.argtypes = [P, I, I, I, P, C.c_int64, P, I]
    L.or_bitstream_pack.argtypes = [P, P, I, P, I]
    L.or_error_compare.argtypes = [C.POINTER(ErrorAcc), P, I, P, P, P, I]
    L.or_error_mse.restype = F
    L.or_error_mse.argtypes = [C.POINTER(ErrorAcc)]
    L.or_error_psnr.restype = F
    L.or_error_psnr.argtypes = [C.POINTER(ErrorAcc)]
    L.or_compress.argtypes = [P, I, P, I, C.POINTER(Options),
                              C.POINTER(Result), P, P, P, C.c_int64]
    L.or_now.restype = C.c_double


# ---------------------------------------------------------------- helpers

def i3(v):
    return np.ascontiguousarray(np.array(v, dtype=np.int32))


def dtype_id(a):
    return {np.uint8: U8, np.int32: I32, np.float32: F32}[a.dtype.type]


def set_steps(steps, size):
    s = i3(steps)
    lib().or_set_steps(_p(s), _p(i3(size)))
    return [int(v) for v in s]


def padded_size(size, steps):
    return [lib().or_padded_length(int(n), int(s)) for n, s in zip(size, steps)]


def pad_convert(vol, steps):
    """vol is indexed [z, y, x]; returns padded float32 [Z, Y, X]."""
    orig = [vol.shape[2], vol.shape[1], vol.shape[0]]
    pad = padded_size(orig, steps)
    out = np.empty((pad[2], pad[1], pad[0]), np.float32)
    v = np.ascontiguousarray(vol)
    lib().or_pad_convert(_p(v), dtype_id(v), _p(i3(orig)), _p(out), _p(i3(pad)))
    return out


def row_fwd(wavelet, row, steps):
    d = np.array(row, dtype=np.float32)
    tmp = np.empty_like(d)
    f = lib().or_cdf97_fwd if wavelet == CDF97 else lib().or_haar_fwd
    f(len(d), _p(d), steps, _p(tmp))
    return d


def row_inv(wavelet, row, steps):
    d = np.array(row, dtype=np.float32)
    tmp = np.empty_like(d)
    f = lib().or_cdf97_inv if wavelet == CDF97 else lib().or_haar_inv
    f(len(d), _p(d), steps, _p(tmp))
    return d


def dwt3d(data, size, wavelet, steps, inverse=False):
    """data: flat float32 in the current layout; size [x,y,z] of that
    layout.  Returns (flat result, rotated size)."""
    d = np.array(data, dtype=np.float32).ravel()
    tmp = np.empty_like(d)
    s = i3(size)
    lib().or_dwt3d(_p(d), _p(tmp), _p(s), wavelet, _p(i3(steps)), int(inverse))
    return d, [int(v) for v in s]


def dwt3d_f64(padded_vol, wavelet, steps):
    """FP64 lifting transform of a padded float32 volume [Z, Y, X]; the
    result is float64 [Z, Y, X] in the input's layout (no rotation)."""
    v = np.ascontiguousarray(padded_vol, np.float32)
    out = np.empty(v.shape, np.float64)
    size = i3([v.shape[2], v.shape[1], v.shape[0]])
    lib().or_dwt3d_f64(_p(v), _p(size), wavelet, _p(i3(steps)), _p(out))
    return out


def rotated_rel_error(coef_flat, f64_zyx):
    """max |coef - f64| / max |f64| for float coefficients in the rotated
    z,x,y layout (flat, as the transform writes them) against the FP64
    result in [z, y, x]: the layout maps [z, y, x] -> [y, x, z]."""
    Z, Y, X = f64_zyx.shape
    got = np.asarray(coef_flat, np.float32).reshape(Y, X, Z)
    want = f64_zyx.transpose(1, 2, 0)
    err = 0.0
    for y in range(0, Y, 64):          # bounded temporaries
        d = np.abs(got[y:y + 64].astype(np.float64) - want[y:y + 64])
        err = max(err, float(d.max()))
    return err / float(np.abs(f64_zyx).max())


def thresh(coef, frac):
    c = np.ascontiguousarray(coef, dtype=np.float32).ravel()
    nz, mx, mn, mabs, idx = (C.c_int(), C.c_float(), C.c_float(), C.c_float(),
                             C.c_int())
    srt = np.empty_like(c)
    t = lib().or_thresh(_p(c), len(c), F(frac), C.byref(nz), C.byref(mx),
                        C.byref(mn), C.byref(mabs), C.byref(idx), _p(srt))
    return dict(threshold=np.float32(t), nonzero=nz.value, max=np.float32(mx.value),
                min=np.float32(mn.value), max_abs=np.float32(mabs.value),
                idx=idx.value, sorted=srt)


def quantizer(alg, bins, threshold, max_val, boundaries=None, values=None):
    q = Quantizer()
    b = None if boundaries is None else np.ascontiguousarray(boundaries, np.float32)
    v = None if values is None else np.ascontiguousarray(values, np.float32)
    lib().or_quant_init(C.byref(q), alg, bins, F(threshold), F(max_val), _p(b), _p(v))
    q._keep = (b, v)
    return q


def quantize(q, x):
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    out = np.empty(len(x), np.int32)
    lib().or_quantize(C.byref(q), _p(x), _p(out), len(x))
    return out


def dequantize(q, bins):
    b = np.ascontiguousarray(bins, dtype=np.int32).ravel()
    out = np.empty(len(b), np.float32)
    lib().or_dequantize(C.byref(q), _p(b), _p(out), len(b))
    return out


def lloyd(points, codebook, stop=np.float32(1e-6), max_iters=-1):
    pts = np.ascontiguousarray(points, np.float32)
    cb = np.array(codebook, np.float32)
    part = np.empty(max(len(cb) - 1, 1), np.float32)
    dist, rel, it = C.c_float(), C.c_float(), C.c_int()
    lib().or_lloyd(_p(pts), len(pts), _p(cb), len(cb), _p(part), C.byref(dist),
                   C.byref(rel), None, F(stop), C.byref(it), max_iters)
    return cb, part, it.value, np.float32(dist.value), np.float32(rel.value)


def lloyd_quantization(nonzero, bins, min_val, max_val, threshold, max_iters=-1):
    nz = np.ascontiguousarray(nonzero, np.float32)
    b = np.empty(bins, np.float32)
    v = np.empty(bins, np.float32)
    it = C.c_int()
    n = lib().or_compute_lloyd_quantization(_p(nz), len(nz), bins, F(min_val),
                                            F(max_val), F(threshold), _p(b), _p(v),
                                            max_iters, C.byref(it))
    assert n == bins
    return b[:bins - 1], v, it.value


class Huffman:
    def __init__(self, data, bin_count, zero_bin, counts=None):
        self.data = np.ascontiguousarray(data, np.int32)
        self.bin_count = bin_count
        L = lib()
        if counts is None:
            self.h = L.or_huff_create_scan(_p(self.data), len(self.data), bin_count, zero_bin)
        else:
            c = np.ascontiguousarray(counts, np.int32)
            self.h = L.or_huff_create_counts(_p(c), _p(self.data), len(self.data),
                                             bin_count, zero_bin)

    def __del__(self):
        try:
            lib().or_huff_free(self.h)
        except Exception:
            pass

    def counts(self):
        out = np.empty(self.bin_count + 1, np.int64)
        lib().or_huff_get_counts(self.h, _p(out))
        return out

    def code_lengths(self):
        return np.array([lib().or_huff_code_length(self.h, v)
                         for v in range(self.bin_count + 1)], np.int32)

    def decode_table(self):
        n = lib().or_huff_decode_table(self.h, None, 0)
        t = np.empty(n, np.int32)
        lib().or_huff_decode_table(self.h, _p(t), n)
        return t

    def encode(self):
        cap = len(self.data) * 2 + 1024
        w = np.empty(cap, np.uint32)
        bits = C.c_int64()
        n = lib().or_huff_encode(self.h, _p(self.data), len(self.data), _p(w), cap,
                                 C.byref(bits))
        return w[:n].copy(), bits.value


def huff_decode(table, dup_key, dup_value, words, count):
    t = np.ascontiguousarray(table, np.int32)
    w = np.ascontiguousarray(words, np.uint32)
    out = np.zeros(count, np.int32)
    n = lib().or_huff_decode(_p(t), len(t), dup_key, dup_value, _p(w), len(w),
                             _p(out), count)
    return out[:n]


def opt_test(vol, steps, wavelet, max_possible, compress_zeros, threshold, bins, alg,
             want_iters=False):
    """One parameter-optimisation test point: (size bytes, mse, psnr, acc)
    (and the Lloyd iterations with want_iters)."""
    v = np.ascontiguousarray(vol)
    orig = [v.shape[2], v.shape[1], v.shape[0]]
    acc = ErrorAcc()
    it = I(0)
    nb = lib().or_opt_test(_p(v), dtype_id(v), _p(i3(orig)), _p(i3(steps)), wavelet,
                           max_possible, int(compress_zeros), F(threshold), bins, alg,
                           C.byref(acc), C.byref(it))
    out = (nb, lib().or_error_mse(C.byref(acc)), lib().or_error_psnr(C.byref(acc)), acc)
    return out + (it.value,) if want_iters else out


def opt_sorted(vol, steps, wavelet):
    """The ascending |x| of the transformed volume."""
    v = np.ascontiguousarray(vol)
    orig = [v.shape[2], v.shape[1], v.shape[0]]
    cap = int(np.prod(padded_size(orig, set_steps(steps, orig))))
    out = np.empty(cap, np.float32)
    n = lib().or_opt_sorted(_p(v), dtype_id(v), _p(i3(orig)), _p(i3(steps)), wavelet,
                            _p(out), cap)
    return out[:n]


def split(size, cub, overlap, vol=None):
    """cube split of one large cubelet: list of (offset xyz, size xyz) in
    write order and, with vol ([z, y, x] numpy), the boxes' samples."""
    sz = (I * 3)(*size)
    cb = (I * 3)(*cub)
    n = lib().or_split(sz, cb, overlap, None, 0, None, 0, None)
    boxes = np.zeros(6 * max(n, 1), np.int32)
    lib().or_split(sz, cb, overlap, _p(boxes), n, None, 0, None)
    if vol is not None:
        v = np.ascontiguousarray(vol)
        total = int(sum(int(np.prod(boxes[6 * k + 3:6 * k + 6])) for k in range(n)))
        packed = np.empty(max(total, 1), v.dtype)
        lib().or_split(sz, cb, overlap, _p(boxes), n, _p(v), v.itemsize, _p(packed))
    b = [(tuple(int(x) for x in boxes[6 * k:6 * k + 3]), tuple(int(x) for x in boxes[6 * k + 3:6 * k + 6]))
         for k in range(n)]
    if vol is None:
        return b
    out, pos = [], 0
    for off, s3 in b:
        m = s3[0] * s3[1] * s3[2]
        out.append(packed[pos:pos + m].reshape(s3[2], s3[1], s3[0]))
        pos += m
    return b, out


def compress(vol, wavelet, quant_alg, bins, steps, frac, compress_zeros=True,
             lloyd_max_iters=-1, with_footer=True, max_possible=0,
             want_coef=False, want_bins=False, want_file=True):
    """vol indexed [z, y, x] (uint8 / int32 / float32)."""
    v = np.ascontiguousarray(vol)
    orig = [v.shape[2], v.shape[1], v.shape[0]]
    st = set_steps(steps, orig)
    pad = padded_size(orig, st)
    n = pad[0] * pad[1] * pad[2]
    opt = Options(wavelet, quant_alg, bins, (I * 3)(*steps), F(frac),
                  int(compress_zeros), lloyd_max_iters, int(with_footer))
    res = Result()
    coef = np.empty(n, np.float32) if want_coef else None
    b = np.empty(n, np.int32) if want_bins else None
    cap = n * 8 + (1 << 20)
    f = np.empty(cap, np.uint8) if want_file else None
    lib().or_compress(_p(v), dtype_id(v), _p(i3(orig)), max_possible,
                      C.byref(opt), C.byref(res), _p(coef), _p(b), _p(f),
                      cap if want_file else 0)
    out = dict(res=res, coef=coef, bins=b)
    if want_file:
        out["file"] = f[:res.file_bytes].tobytes()
    return out


def error_psnr(orig_vol, restored_padded, padded, max_possible):
    v = np.ascontiguousarray(orig_vol)
    orig = [v.shape[2], v.shape[1], v.shape[0]]
    r = np.ascontiguousarray(restored_padded, np.float32)
    acc = ErrorAcc()
    lib().or_error_compare(C.byref(acc), _p(v), dtype_id(v), _p(i3(orig)), _p(r),
                           _p(i3(padded)), max_possible)
    return float(lib().or_error_psnr(C.byref(acc))), float(lib().or_error_mse(C.byref(acc)))
