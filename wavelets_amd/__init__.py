"""wavelets_amd -- MI355X (gfx950) implementation of the Wavelets compression
hot path (oshkosher/Wavelets, test_compress_gpu): 3-D CDF 9/7 / Haar DWT,
threshold selection, uniform / log / Lloyd quantization with a fused bin
histogram, and the Huffman + .cube writer, behind the C ABI of
libwv_hip.so (include/wv_hip.h).

This module is the Python mirror of that ABI (ctypes).  torch is used only
for device memory and streams.  There is no CPU fallback: every call runs
the HIP library and fails loudly if it is not built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WV_LIB names another build of the same library (the host-sanitized one)
LIB_PATH = os.environ.get("WV_LIB") or os.path.join(HERE, "libwv_hip.so")

HAAR, CDF97 = 1, 2
UNIFORM, LOG, LLOYD = 1, 2, 3
UINT8, INT32, FLOAT32, INT16 = 1, 2, 3, 4
STATS_BYTES = 8256

_lib = None


class Geom(C.Structure):
    _fields_ = [("orig", C.c_int * 3), ("steps", C.c_int * 3),
                ("padded", C.c_int * 3), ("out_size", C.c_int * 3),
                ("wavelet", C.c_int), ("dtype", C.c_int),
                ("count", C.c_int64), ("ws_floats", C.c_int64)]


class ThreshResult(C.Structure):
    _fields_ = [("threshold", C.c_float), ("min_val", C.c_float),
                ("max_val", C.c_float), ("max_abs", C.c_float),
                ("thresh_idx", C.c_int32), ("nonzero_count", C.c_int32),
                ("threshold_key", C.c_uint32), ("n_candidates", C.c_int32)]


class QuantParams(C.Structure):
    _fields_ = [("alg", C.c_int), ("bin_count", C.c_int), ("zero_bin", C.c_int),
                ("threshold", C.c_float), ("max_val", C.c_float),
                ("neg_scale", C.c_float), ("neg_inv_scale", C.c_float),
                ("neg_offset", C.c_float), ("pos_scale", C.c_float),
                ("pos_inv_scale", C.c_float), ("pos_offset", C.c_float),
                ("n_pos_edges", C.c_int), ("n_neg_edges", C.c_int),
                ("pos_clamp", C.c_float), ("n_boundaries", C.c_int),
                ("pos_edges", C.POINTER(C.c_float)),
                ("neg_edges", C.POINTER(C.c_float)),
                ("boundaries", C.POINTER(C.c_float)),
                ("values", C.POINTER(C.c_float)),
                ("table_d", C.c_void_p), ("priv", C.c_void_p)]


class CubeDesc(C.Structure):
    _fields_ = [("orig", C.c_int * 3), ("steps", C.c_int * 3),
                ("out_size", C.c_int * 3), ("wavelet", C.c_int),
                ("quant_alg", C.c_int), ("bin_count", C.c_int),
                ("orig_dtype", C.c_int), ("thresh_frac", C.c_float),
                ("threshold", C.c_float), ("max_val", C.c_float),
                ("boundaries", C.c_void_p), ("values", C.c_void_p),
                ("zero_bin", C.c_int), ("with_footer", C.c_int)]


class Options(C.Structure):
    _fields_ = [("wavelet", C.c_int), ("quant_alg", C.c_int),
                ("bin_count", C.c_int), ("steps", C.c_int * 3),
                ("thresh_frac", C.c_float), ("compress_zeros", C.c_int),
                ("lloyd_max_iters", C.c_int), ("with_footer", C.c_int)]


class CompressInfo(C.Structure):
    _fields_ = [("thresh", ThreshResult), ("out_size", C.c_int * 3),
                ("steps", C.c_int * 3), ("zero_bin", C.c_int),
                ("lloyd_iters", C.c_int), ("payload_bytes", C.c_int64),
                ("file_bytes", C.c_int64), ("ms_device", C.c_double),
                ("ms_total", C.c_double)]


class BatchItem(C.Structure):
    _fields_ = [("in_d", C.c_void_p), ("dtype", C.c_int), ("orig", C.c_int * 3),
                ("wavelet", C.c_int), ("steps", C.c_int * 3), ("quant_alg", C.c_int),
                ("bin_count", C.c_int), ("thresh_frac", C.c_float),
                ("coef_d", C.c_void_p), ("bins_d", C.c_void_p), ("counts_d", C.c_void_p),
                ("in_stride", C.c_int64 * 2)]


BATCH_AUTOTUNE, BATCH_NO_GRAPH = 1, 2


class SplitBox(C.Structure):
    _fields_ = [("offset", C.c_int * 3), ("size", C.c_int * 3)]


class OptResult(C.Structure):
    _fields_ = [("size_bytes", C.c_int64), ("l1", C.c_float), ("l2", C.c_float),
                ("mse", C.c_float), ("psnr", C.c_float), ("rel", C.c_float),
                ("zero_bin", C.c_int), ("lloyd_iters", C.c_int)]


class SplitInfo(C.Structure):
    _fields_ = [("cubelets", C.c_int), ("need", C.c_int64), ("ms_plan", C.c_double),
                ("ms_batch", C.c_double), ("ms_encode", C.c_double), ("ms_total", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int), ("hbm_budget_bytes", C.c_int64), ("encode", C.c_int),
                ("host_threads", C.c_int), ("lloyd_batch", C.c_int)]

P = C.c_void_p
I64 = C.c_int64


class ServeRequest(C.Structure):
    _fields_ = [("in_host", C.c_void_p), ("dtype", C.c_int), ("orig", C.c_int * 3),
                ("opt", Options), ("bins_out", C.c_void_p), ("counts_out", C.c_void_p),
                ("file_out", C.c_void_p), ("file_cap", C.c_int64)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("in_bytes", C.c_int64),
                ("device_bytes", C.c_int64), ("file_bytes", C.c_int64),
                ("thresh", ThreshResult), ("zero_bin", C.c_int), ("lloyd_iters", C.c_int)]


class CubeInfo(C.Structure):
    _fields_ = [("size", C.c_int * 3), ("orig", C.c_int * 3), ("orig_dtype", C.c_int),
                ("steps", C.c_int * 3), ("wavelet", C.c_int), ("quant_alg", C.c_int),
                ("bin_count", C.c_int), ("thresh_frac", C.c_float), ("threshold", C.c_float),
                ("max_val", C.c_float), ("dup_key", C.c_int), ("dup_value", C.c_int),
                ("maximum_value", C.c_int), ("n_decode", C.c_int), ("n_boundaries", C.c_int),
                ("n_values", C.c_int), ("payload_offset", C.c_int64),
                ("payload_bytes", C.c_int64)]


def _declare(L):
    L.wv_last_error.restype = C.c_char_p
    L.wv_version.restype = C.c_char_p
    L.wv_geom_init.argtypes = [C.POINTER(Geom), P, P, C.c_int, C.c_int]
    L.wv_stats_clear.argtypes = [P, P]
    L.wv_stats_compute.argtypes = [P, I64, P, P]
    L.wv_dwt3d_forward.argtypes = [C.POINTER(Geom), P, P, P, P, P]
    L.wv_dwt3d_inverse.argtypes = [C.POINTER(Geom), P, P, P, P]
    L.wv_dwt3d_forward_ev.argtypes = [C.POINTER(Geom), P, P, P, P, P, P]
    L.wv_dwt3d_num_passes.argtypes = [C.POINTER(Geom)]
    L.wv_event_create.argtypes = [C.POINTER(C.c_void_p)]
    L.wv_event_destroy.argtypes = [P]
    L.wv_event_record.argtypes = [P, P]
    L.wv_event_elapsed_ms.argtypes = [P, P, C.POINTER(C.c_float)]
    L.wv_threshold_ws_bytes.restype = I64
    L.wv_threshold_ws_bytes.argtypes = [I64]
    L.wv_threshold.argtypes = [P, I64, C.c_float, P, P, P, P]
    if hasattr(L, "wv_threshold_noalloc"):
        L.wv_threshold_noalloc.argtypes = [P, I64, C.c_float, P, P]
    L.wv_quant_params_init.argtypes = [C.POINTER(QuantParams), C.c_int, C.c_int,
                                       C.c_float, C.c_float, P, P]
    L.wv_quant_params_free.argtypes = [C.POINTER(QuantParams)]
    L.wv_quant_table_bytes.restype = I64
    L.wv_quant_table_bytes.argtypes = [C.POINTER(QuantParams)]
    L.wv_quant_params_upload.argtypes = [C.POINTER(QuantParams), P, P]
    L.wv_quant_host.argtypes = [C.POINTER(QuantParams), C.c_float]
    L.wv_dequant_host.restype = C.c_float
    L.wv_dequant_host.argtypes = [C.POINTER(QuantParams), C.c_int]
    L.wv_quantize_hist.argtypes = [P, I64, C.POINTER(QuantParams), P, P, P]
    L.wv_quantize_hist_uniform_dev.argtypes = [P, I64, C.c_int, P, P, P, P]
    L.wv_dequantize.argtypes = [P, I64, C.POINTER(QuantParams), P, P]
    L.wv_write_cube.restype = I64
    L.wv_write_cube.argtypes = [C.POINTER(CubeDesc), P, I64, P, P, I64, P]
    if hasattr(L, "wv_cube_parse"):
        L.wv_cube_parse.argtypes = [P, I64, C.POINTER(CubeInfo), P, C.c_int, P, P, C.c_int]
        L.wv_huff_decode_host.restype = I64
        L.wv_huff_decode_host.argtypes = [P, C.c_int, C.c_int, C.c_int, P, I64, P, I64]
        L.wv_write_cube_raw.restype = I64
        L.wv_write_cube_raw.argtypes = [P, C.c_int, P, P, I64]
    if hasattr(L, "wv_split_plan"):
        L.wv_split_plan.argtypes = [P, P, C.c_int, C.POINTER(SplitBox), C.c_int]
        L.wv_cube_stream_open.restype = P
        L.wv_cube_stream_add_raw.argtypes = [P, P, C.c_int, P, P]
        L.wv_cube_stream_bytes.restype = I64
        L.wv_cube_stream_bytes.argtypes = [P, C.c_int]
        L.wv_cube_stream_finish.restype = I64
        L.wv_cube_stream_finish.argtypes = [P, C.c_int, P, I64]
        L.wv_cube_stream_free.argtypes = [P]
        L.wv_batch_write_stream.restype = I64
        L.wv_batch_write_stream.argtypes = [P, P, C.c_int, C.c_int, P, I64,
                                            C.POINTER(C.c_int64), P]
        L.wv_split_compress.restype = I64
        L.wv_split_compress.argtypes = [P, C.c_int, P, P, C.c_int, C.POINTER(Options), C.c_int,
                                        P, I64, C.POINTER(SplitInfo), P]
    if hasattr(L, "wv_opt_create"):
        L.wv_opt_create.argtypes = [P, C.c_int, P, P, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_void_p), P]
        L.wv_opt_count.restype = I64
        L.wv_opt_count.argtypes = [P]
        L.wv_opt_stats.argtypes = [P, P, P, P]
        L.wv_opt_sorted.restype = C.c_float
        L.wv_opt_sorted.argtypes = [P, I64]
        L.wv_opt_find_sorted.restype = I64
        L.wv_opt_find_sorted.argtypes = [P, C.c_float]
        L.wv_opt_test.argtypes = [P, C.c_float, C.c_int, C.c_int, C.POINTER(OptResult), P]
        L.wv_opt_destroy.argtypes = [P]
    if hasattr(L, "wv_huff_decode_device"):
        L.wv_huff_decode_ws_bytes.restype = I64
        L.wv_huff_decode_ws_bytes.argtypes = [I64, C.c_int]
        L.wv_huff_decode_device.argtypes = [P, C.c_int, C.c_int, C.c_int, P, I64, P, I64, P,
                                            C.POINTER(C.c_int), P]
    if hasattr(L, "wv_decompress"):
        L.wv_decompress_work_bytes.restype = I64
        L.wv_decompress_work_bytes.argtypes = [C.POINTER(CubeInfo)]
        L.wv_decompress_device.argtypes = [C.POINTER(CubeInfo), P, P, P, P, I64, P, P]
        L.wv_decompress.restype = I64
        L.wv_decompress.argtypes = [P, I64, P, I64, P, I64, C.POINTER(CubeInfo),
                                    C.POINTER(C.c_double), P]
    if hasattr(L, "wv_encode_device"):
        L.wv_encode_ws_bytes.restype = I64
        L.wv_encode_ws_bytes.argtypes = [I64, C.c_int]
        L.wv_encode_words_bound.restype = I64
        L.wv_encode_words_bound.argtypes = [I64, C.c_int]
        L.wv_encode_device.argtypes = [P, I64, P, C.c_int, C.c_int, P, P, I64,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_void_p), P]
        L.wv_huff_payload_bits.restype = I64
        L.wv_huff_payload_bits.argtypes = [P]
        L.wv_write_cube_payload.restype = I64
        L.wv_write_cube_payload.argtypes = [C.POINTER(CubeDesc), P, P, I64, P, I64]
        L.wv_write_cube_device.restype = I64
        L.wv_write_cube_device.argtypes = [C.POINTER(CubeDesc), P, P, I64, P, I64, P]
        L.wv_huff_free.restype = None
        L.wv_huff_free.argtypes = [P]
    if hasattr(L, "wv_compress"):
        L.wv_compress.argtypes = [P, C.c_int, P, C.POINTER(Options), P, I64, P,
                                  P, I64, C.POINTER(CompressInfo), P]
        L.wv_compress_work_bytes.restype = I64
        L.wv_compress_work_bytes.argtypes = [P, C.POINTER(Options)]
        L.wv_compress_device.argtypes = [P, C.c_int, P, C.POINTER(Options), P, I64, P,
                                         C.POINTER(CompressInfo), P]
        L.wv_compress_counts_offset.restype = I64
        L.wv_compress_counts_offset.argtypes = [P, C.POINTER(Options)]
    if hasattr(L, "wv_batch_plan"):
        L.wv_batch_plan.argtypes = [C.POINTER(BatchItem), C.c_int, C.c_int,
                                    C.POINTER(C.c_void_p)]
        L.wv_batch_submit.argtypes = [P, P]
        L.wv_batch_poll.argtypes = [P]
        L.wv_batch_wait.argtypes = [P]
        L.wv_batch_result.argtypes = [P, C.c_int, C.POINTER(ThreshResult),
                                      C.POINTER(C.c_int)]
        L.wv_batch_size.argtypes = [P]
        L.wv_batch_groups.argtypes = [P]
        L.wv_batch_group_tuning.argtypes = [P, C.c_int, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int)]
        L.wv_batch_destroy.argtypes = [P]
        if hasattr(L, "wv_batch_stage_times"):
            L.wv_batch_stage_times.argtypes = [P, P, C.POINTER(C.c_double)]
    if hasattr(L, "wv_copy_stream"):
        L.wv_copy_stream.argtypes = [P, P, I64, C.c_int, C.c_int, C.c_int, P]
    if hasattr(L, "wv_serve_create"):
        L.wv_serve_create.argtypes = [C.POINTER(ServeConfig), C.POINTER(C.c_void_p)]
        L.wv_serve_submit.argtypes = [P, C.POINTER(ServeRequest), C.POINTER(C.c_int64)]
        L.wv_serve_poll.argtypes = [P, C.POINTER(ServeResult), C.c_int]
        L.wv_serve_drain.argtypes = [P]
        L.wv_serve_stats.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
        L.wv_serve_destroy.argtypes = [P]
        L.wv_serve_lloyd_stats.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    if hasattr(L, "wv_lloyd"):
        L.wv_lloyd_ws_bytes.restype = I64
        L.wv_lloyd_ws_bytes.argtypes = [I64, C.c_int]
        L.wv_lloyd.argtypes = [P, I64, C.POINTER(ThreshResult), C.c_int, C.c_int,
                               P, P, P, P, P]


def lib():
    """The loaded libwv_hip.so.  torch is imported first so the HIP runtime
    it carries is the one the library binds to."""
    global _lib
    if _lib is None:
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libwv_hip.so is not built: run python -m wavelets_amd.build")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class WvError(RuntimeError):
    pass


def check(rc):
    if rc < 0:
        raise WvError("wv_hip error %d: %s" % (rc, lib().wv_last_error().decode()))
    return rc


class Event:
    """A hipEvent_t owned by the library (timing on any stream)."""

    def __init__(self):
        self.h = C.c_void_p()
        check(lib().wv_event_create(C.byref(self.h)))

    def record(self, stream=None):
        check(lib().wv_event_record(self.h, _stream(stream)))

    def elapsed_ms(self, other):
        ms = C.c_float()
        check(lib().wv_event_elapsed_ms(self.h, other.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().wv_event_destroy(self.h)
        except Exception:
            pass


def _i3(v):
    return (C.c_int * 3)(*[int(a) for a in v])


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


DTYPE_ID = {"uint8": UINT8, "int32": INT32, "float32": FLOAT32, "int16": INT16}


def dtype_id(t):
    name = str(t.dtype).replace("torch.", "")
    if name not in DTYPE_ID:
        raise ValueError("unsupported input dtype %s (uint8, int16, int32, float32)" % name)
    return DTYPE_ID[name]


def geom(orig, steps, wavelet, dtype):
    g = Geom()
    check(lib().wv_geom_init(C.byref(g), _i3(orig), _i3(steps), wavelet, dtype))
    return g


def dwt3d_forward(vol, steps, wavelet, with_stats=True, stream=None):
    """vol: cuda tensor [z, y, x] of uint8 / int32 / float32.
    Returns (coef, geom, stats) with coef flat float32 in the rotated
    layout of geom.out_size."""
    import torch
    orig = [vol.shape[2], vol.shape[1], vol.shape[0]]
    g = geom(orig, steps, wavelet, dtype_id(vol))
    dev = vol.device
    coef = torch.empty(g.count, dtype=torch.float32, device=dev)
    ws = torch.empty(max(g.ws_floats, 1), dtype=torch.float32, device=dev)
    stats = None
    s = _stream(stream)
    if with_stats:
        stats = torch.empty(STATS_BYTES, dtype=torch.uint8, device=dev)
        check(lib().wv_stats_clear(_ptr(stats), s))
    check(lib().wv_dwt3d_forward(C.byref(g), _ptr(vol.contiguous()), _ptr(coef), _ptr(ws),
                                 _ptr(stats), s))
    return coef, g, stats


def dwt3d_inverse(coef, g, stream=None):
    import torch
    out = torch.empty(g.count, dtype=torch.float32, device=coef.device)
    ws = torch.empty(max(g.ws_floats, 1), dtype=torch.float32, device=coef.device)
    check(lib().wv_dwt3d_inverse(C.byref(g), _ptr(coef), _ptr(out), _ptr(ws), _stream(stream)))
    return out


def stats_compute(coef, stream=None):
    import torch
    stats = torch.empty(STATS_BYTES, dtype=torch.uint8, device=coef.device)
    check(lib().wv_stats_compute(_ptr(coef), coef.numel(), _ptr(stats), _stream(stream)))
    return stats


def threshold(coef, frac, stats=None, stream=None, sync=True):
    """Returns a device tensor holding a wv_thresh_result and, with sync,
    the host ThreshResult."""
    import torch
    n = coef.numel()
    ws = torch.empty(lib().wv_threshold_ws_bytes(n), dtype=torch.uint8, device=coef.device)
    res = torch.zeros(C.sizeof(ThreshResult), dtype=torch.uint8, device=coef.device)
    check(lib().wv_threshold(_ptr(coef), n, C.c_float(frac), _ptr(stats) if stats is not None
                             else None, _ptr(ws), _ptr(res),
                             _stream(stream)))
    if not sync:
        return res, None
    host = ThreshResult.from_buffer_copy(res.cpu().numpy().tobytes())
    return res, host


NOALLOC_MAX_N = 4194304     # WV_NOALLOC_MAX_N


def threshold_noalloc(coef, frac, stream=None, sync=True):
    """wv_threshold_noalloc: the select on the library's static scratch (no
    workspace, no allocation inside; at most NOALLOC_MAX_N coefficients).
    Returns the device result record and, with sync, the host ThreshResult."""
    import torch
    res = torch.zeros(C.sizeof(ThreshResult), dtype=torch.uint8, device=coef.device)
    check(lib().wv_threshold_noalloc(_ptr(coef), coef.numel(), C.c_float(frac), _ptr(res),
                                     _stream(stream)))
    if not sync:
        return res, None
    return res, ThreshResult.from_buffer_copy(res.cpu().numpy().tobytes())


def copy_stream(src, dst, nontemporal=True, blocks_per_cu=8, vec_per_thread=16,
                stream=None):
    """wv_copy_stream: device-to-device copy of src into dst (same byte
    size, 16-byte multiples) with 128-bit accesses (measurement helper)."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("copy_stream: sizes differ")
    check(lib().wv_copy_stream(_ptr(src), _ptr(dst), nbytes, int(nontemporal),
                               int(blocks_per_cu), int(vec_per_thread), _stream(stream)))


class Quantizer:
    """Host parameters + device tables of one quantizer."""

    def __init__(self, alg, bin_count, threshold, max_val, boundaries=None, values=None,
                 device="cuda", stream=None):
        import torch
        self.p = QuantParams()
        b = None if boundaries is None else np.ascontiguousarray(boundaries, np.float32)
        v = None if values is None else np.ascontiguousarray(values, np.float32)
        self._keep = (b, v)
        check(lib().wv_quant_params_init(
            C.byref(self.p), alg, bin_count, C.c_float(threshold), C.c_float(max_val),
            b.ctypes.data_as(P) if b is not None else None,
            v.ctypes.data_as(P) if v is not None else None))
        nbytes = lib().wv_quant_table_bytes(C.byref(self.p))
        self.table = torch.empty(max(nbytes, 4), dtype=torch.uint8, device=device)
        check(lib().wv_quant_params_upload(C.byref(self.p), _ptr(self.table), _stream(stream)))

    def __del__(self):
        try:
            lib().wv_quant_params_free(C.byref(self.p))
        except Exception:
            pass

    @property
    def zero_bin(self):
        return self.p.zero_bin

    def host_quant(self, x):
        return lib().wv_quant_host(C.byref(self.p), C.c_float(x))

    def edges(self):
        pe = np.ctypeslib.as_array(self.p.pos_edges, (self.p.n_pos_edges,)).copy() \
            if self.p.n_pos_edges else np.zeros(0, np.float32)
        ne = np.ctypeslib.as_array(self.p.neg_edges, (self.p.n_neg_edges,)).copy() \
            if self.p.n_neg_edges else np.zeros(0, np.float32)
        return pe, ne

    def quantize(self, coef, out=None, stream=None):
        import torch
        n = coef.numel()
        bins = out if out is not None else torch.empty(n, dtype=torch.int32, device=coef.device)
        counts = torch.empty(self.p.bin_count + 64, dtype=torch.int32, device=coef.device)
        check(lib().wv_quantize_hist(_ptr(coef), n, C.byref(self.p), _ptr(bins), _ptr(counts),
                                     _stream(stream)))
        return bins, counts[:self.p.bin_count]

    def dequantize(self, bins, stream=None):
        import torch
        out = torch.empty(bins.numel(), dtype=torch.float32, device=bins.device)
        check(lib().wv_dequantize(_ptr(bins), bins.numel(), C.byref(self.p), _ptr(out),
                                  _stream(stream)))
        return out


def quantize_uniform_dev(coef, bin_count, res_dev, stream=None):
    import torch
    n = coef.numel()
    bins = torch.empty(n, dtype=torch.int32, device=coef.device)
    counts = torch.empty(((bin_count + 63) // 64) * 64 + 64, dtype=torch.int32,
                         device=coef.device)
    check(lib().wv_quantize_hist_uniform_dev(_ptr(coef), n, bin_count, _ptr(res_dev),
                                             _ptr(bins), _ptr(counts), _stream(stream)))
    return bins, counts[:bin_count]


def _cube_desc(desc_kwargs, keep):
    d = CubeDesc()
    for k, v in desc_kwargs.items():
        if k in ("orig", "steps", "out_size"):
            setattr(d, k, _i3(v))
        elif k in ("boundaries", "values"):
            arr = np.ascontiguousarray(v, np.float32)
            keep.append(arr)
            setattr(d, k, arr.ctypes.data_as(P))
        else:
            setattr(d, k, v)
    return d


def write_cube(desc_kwargs, bins_host, counts_host):
    """bins_host / counts_host: numpy int32 arrays.  Returns the .cube bytes."""
    keep = []
    d = _cube_desc(desc_kwargs, keep)
    b = np.ascontiguousarray(bins_host, np.int32)
    c = np.ascontiguousarray(counts_host, np.uint32)
    cap = b.size * 8 + (1 << 20)
    buf = np.empty(cap, np.uint8)
    pay = C.c_int64()
    n = check(lib().wv_write_cube(C.byref(d), b.ctypes.data_as(P), b.size, c.ctypes.data_as(P),
                                  buf.ctypes.data_as(P), cap, C.byref(pay)))
    return buf[:n].tobytes()


class HuffBook:
    """Code book of one device-coded payload (wv_huff)."""

    def __init__(self, handle):
        self.handle = handle

    @property
    def payload_bits(self):
        return lib().wv_huff_payload_bits(self.handle)

    def close(self):
        if self.handle:
            lib().wv_huff_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode_device(bins, counts, bin_count, dup_value, stream=None):
    """Device Huffman coding of bins (int32 device tensor) with its histogram
    counts (bin_count + 1 entries, device).  Returns (words, book): words is
    an int32 device tensor of the LSB-first payload words."""
    import torch
    n = bins.numel()
    dev = bins.device
    ws = torch.empty(lib().wv_encode_ws_bytes(n, bin_count), dtype=torch.uint8, device=dev)
    cap = lib().wv_encode_words_bound(n, bin_count)
    words = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
    nw = C.c_int64()
    h = C.c_void_p()
    check(lib().wv_encode_device(_ptr(bins), n, _ptr(counts), bin_count, dup_value, _ptr(ws),
                                 _ptr(words), cap, C.byref(nw), C.byref(h), _stream(stream)))
    book = HuffBook(h)
    book._ws = ws     # the payload kernels may still run on the stream
    return words[:nw.value], book


def write_cube_payload(desc_kwargs, book, words_host):
    """The .cube bytes around a device-coded payload (numpy words)."""
    keep = []
    d = _cube_desc(desc_kwargs, keep)
    w = np.ascontiguousarray(words_host).view(np.uint32)
    cap = w.size * 4 + (4 << 20)
    buf = np.empty(cap, np.uint8)
    n = check(lib().wv_write_cube_payload(C.byref(d), book.handle, w.ctypes.data_as(P), w.size,
                                          buf.ctypes.data_as(P), cap))
    return buf[:n].tobytes()


def cube_parse(file_bytes):
    """The first compressed cubelet of a .cube image: (CubeInfo, decode table,
    boundaries, values) -- the arrays as numpy."""
    buf = np.frombuffer(file_bytes, np.uint8)
    info = CubeInfo()
    check(lib().wv_cube_parse(buf.ctypes.data_as(P), buf.size, C.byref(info), None, 0, None,
                              None, 0))
    table = np.empty(max(info.n_decode, 1), np.int32)
    cap = max(info.n_boundaries, info.n_values, 1)
    bnd = np.empty(cap, np.float32)
    val = np.empty(cap, np.float32)
    check(lib().wv_cube_parse(buf.ctypes.data_as(P), buf.size, C.byref(info),
                              table.ctypes.data_as(P), table.size, bnd.ctypes.data_as(P),
                              val.ctypes.data_as(P), cap))
    return info, table[:info.n_decode], bnd[:info.n_boundaries], val[:info.n_values]


def huff_decode_host(table, dup_key, dup_value, words, count):
    t = np.ascontiguousarray(table, np.int32)
    w = np.ascontiguousarray(words).view(np.uint32)
    out = np.zeros(count, np.int32)
    n = lib().wv_huff_decode_host(t.ctypes.data_as(P), t.size, dup_key, dup_value,
                                  w.ctypes.data_as(P), w.size, out.ctypes.data_as(P), count)
    check(n if n < 0 else 0)
    return out[:n]


def huff_decode_device(table, dup_key, dup_value, words, count, stream=None):
    """Device decoding of a payload (numpy words); returns (values, rounds)."""
    import torch
    t = np.ascontiguousarray(table, np.int32)
    w = np.ascontiguousarray(words).view(np.uint32)
    wd = torch.from_numpy(w.view(np.int32).copy()).cuda() if w.size else None
    out = torch.empty(max(count, 1), dtype=torch.int32, device="cuda")
    ws = torch.empty(check_size(lib().wv_huff_decode_ws_bytes(w.size, t.size)), dtype=torch.uint8,
                     device="cuda")
    rounds = C.c_int()
    check(lib().wv_huff_decode_device(t.ctypes.data_as(P), t.size, dup_key, dup_value,
                                      _ptr(wd), w.size, _ptr(out), count, _ptr(ws),
                                      C.byref(rounds), _stream(stream)))
    return out[:count].cpu().numpy(), rounds.value


OUT_NP = {UINT8: np.uint8, INT32: np.int32, FLOAT32: np.float32}


def decompress(file_bytes, stream=None, with_info=False):
    """GPU decompression of a .cube image: the reconstructed volume, indexed
    [z, y, x], in the original type."""
    import torch
    buf = np.frombuffer(file_bytes, np.uint8)
    info, _, _, _ = cube_parse(file_bytes)
    wb = check_size(lib().wv_decompress_work_bytes(C.byref(info)))
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    dt = info.orig_dtype if info.orig_dtype in (UINT8, INT32) else FLOAT32
    out = np.empty((info.orig[2], info.orig[1], info.orig[0]), OUT_NP[dt])
    ms = C.c_double()
    n = lib().wv_decompress(buf.ctypes.data_as(P), buf.size, _ptr(work), wb,
                            out.ctypes.data_as(P), out.nbytes, C.byref(info), C.byref(ms),
                            _stream(stream))
    check(n if n < 0 else 0)
    return (out, info, ms.value) if with_info else out


def check_size(n):
    if n < 0:
        check(int(n))
    return n


def write_cube_raw(vol):
    """An uncompressed cubelet file holding vol ([z, y, x]; uint8 / int32 / float32)."""
    v = np.ascontiguousarray(vol)
    dt = {np.dtype(np.uint8): UINT8, np.dtype(np.int32): INT32,
          np.dtype(np.float32): FLOAT32}[v.dtype]
    size = (C.c_int * 3)(v.shape[2], v.shape[1], v.shape[0])
    cap = v.nbytes + (1 << 16)
    buf = np.empty(cap, np.uint8)
    n = lib().wv_write_cube_raw(size, dt, v.ctypes.data_as(P), buf.ctypes.data_as(P), cap)
    check(n if n < 0 else 0)
    return buf[:n].tobytes()


class Opt:
    """Parameter optimisation data (wv_opt_*): OptimizationData and
    testParameters of the reference's -opt (optimize.h, optimize_gpu.cu).
    vol: cuda tensor [z, y, x], kept alive by this object."""

    def __init__(self, vol, steps, wavelet, max_possible=0, compress_zeros=True, stream=None):
        if vol.dim() != 3 or not vol.is_contiguous():
            raise ValueError("Opt: needs a contiguous [z, y, x] volume")
        self.vol = vol
        self.h = C.c_void_p()
        orig = [vol.shape[2], vol.shape[1], vol.shape[0]]
        check(lib().wv_opt_create(_ptr(vol), dtype_id(vol), _i3(orig), _i3(steps), wavelet,
                                  max_possible, int(compress_zeros), C.byref(self.h),
                                  _stream(stream)))

    def count(self):
        return check_size(lib().wv_opt_count(self.h))

    def stats(self):
        v = [C.c_float() for _ in range(3)]
        check(lib().wv_opt_stats(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def sorted(self, index):
        """getSorted: element index of the ascending |x| (0 out of range)."""
        return lib().wv_opt_sorted(self.h, int(index))

    def find_sorted(self, value):
        """findSorted: first index whose |x| >= value."""
        return check_size(lib().wv_opt_find_sorted(self.h, C.c_float(value)))

    def test(self, threshold, bins, alg=LOG, stream=None):
        """testParameters: OptResult (size_bytes, l1, l2, mse, psnr, rel,
        zero_bin, lloyd_iters); alg LLOYD trains the code book on the sorted
        |x| above the threshold."""
        r = OptResult()
        check(lib().wv_opt_test(self.h, C.c_float(threshold), int(bins), int(alg), C.byref(r),
                                _stream(stream)))
        return r

    def close(self):
        if self.h:
            lib().wv_opt_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def opt_plot_all(o, alg=LOG, stream=None):
    """plotAll (optimize.cc:99-122), the reference's default optimisation:
    for threshold fractions .5, .51, ... (float steps, below .999) and bin
    counts 100, 110, 121, ... (x1.1 truncated, up to 2000) the size and PSNR
    of each point; rows (frac, bins, size, psnr)."""
    rows = []
    n = o.count()
    frac = np.float32(0.5)
    while frac < np.float32(0.999):
        t = o.sorted(int(np.float32(n) * frac))
        b = 100
        while b <= 2000:
            r = o.test(t, b, alg, stream)
            rows.append((float(frac), b, int(r.size_bytes), float(r.psnr)))
            b = int(b * 1.1)
        frac = np.float32(frac + np.float32(0.01))
    return rows


def opt_fixed_psnr(o, target_psnr, stream=None):
    """fixedPSNR (optimize.cc:53-90), the reference's Lloyd sweep: for kept
    fractions .40, .35, ... (float steps, above .04) the threshold-only
    PSNR at threshold getSorted(count * (1 - frac)); where that reaches the
    target, Lloyd bin counts 50, 55, 60, ... (x1.1 truncated, up to 30000)
    until one exceeds it.  Rows (kept fraction, bins (0: threshold only),
    size, psnr) in the order tested."""
    rows = []
    n = o.count()
    frac = np.float32(0.40)
    while frac > np.float32(0.04):
        t = o.sorted(int(np.float32(n) * np.float32(np.float32(1) - frac)))
        r = o.test(t, 0, LLOYD, stream)
        rows.append((float(frac), 0, int(r.size_bytes), float(r.psnr)))
        if r.psnr >= target_psnr:
            b = 50
            while b <= 30000:
                r = o.test(t, b, LLOYD, stream)
                rows.append((float(frac), b, int(r.size_bytes), float(r.psnr)))
                if r.psnr > target_psnr:
                    break
                b = int(b * 1.1)
        frac = np.float32(np.float64(frac) - 0.05)   # float -= double literal
    return rows


def split_plan(size, cub, overlap):
    """Boxes of `cube split` (wv_split_plan): list of (offset xyz, size xyz)."""
    n = lib().wv_split_plan(_i3(size), _i3(cub), overlap, None, 0)
    check(n if n < 0 else 0)
    boxes = (SplitBox * max(1, n))()
    check(min(0, lib().wv_split_plan(_i3(size), _i3(cub), overlap, boxes, n)))
    return [(tuple(b.offset), tuple(b.size)) for b in boxes[:n]]


def split_raw(vol, cub, overlap, with_footer=True):
    """`cube split` of one large volume (numpy [z, y, x]; uint8 / int32 /
    float32) into an uncompressed cubelet stream with the boxes' offsets."""
    v = np.ascontiguousarray(vol)
    dt = {np.dtype(np.uint8): UINT8, np.dtype(np.int32): INT32,
          np.dtype(np.float32): FLOAT32}[v.dtype]
    L = lib()
    w = L.wv_cube_stream_open()
    try:
        for off, sz in split_plan([v.shape[2], v.shape[1], v.shape[0]], cub, overlap):
            box = np.ascontiguousarray(v[off[2]:off[2] + sz[2], off[1]:off[1] + sz[1],
                                         off[0]:off[0] + sz[0]])
            check(L.wv_cube_stream_add_raw(w, _i3(sz), dt, box.ctypes.data_as(P), _i3(off)))
        n = check_size(L.wv_cube_stream_bytes(w, int(with_footer)))
        buf = np.empty(n, np.uint8)
        check_size(L.wv_cube_stream_finish(w, int(with_footer), buf.ctypes.data_as(P), n))
        return buf.tobytes()
    finally:
        L.wv_cube_stream_free(w)


def split_compress(vol, cub, overlap, wavelet, steps, alg, bins, frac, flags=0,
                   compress_zeros=True, with_footer=True, stream=None):
    """Split a device volume (cuda tensor [z, y, x]) into boxes and compress
    them all through the batched path into one .cube stream
    (wv_split_compress).  Returns (file bytes, SplitInfo)."""
    if vol.dim() != 3 or not vol.is_contiguous():
        raise ValueError("split_compress: needs a contiguous [z, y, x] volume")
    opt = Options(wavelet, alg, bins, _i3(steps), C.c_float(frac), int(compress_zeros), 0,
                  int(with_footer))
    size = _i3([vol.shape[2], vol.shape[1], vol.shape[0]])
    info = SplitInfo()
    L = lib()
    cap = vol.numel() * vol.element_size() * 2 + (1 << 20)
    for _ in range(2):   # a second run only when the first buffer was short
        buf = np.empty(cap, np.uint8)
        n = L.wv_split_compress(_ptr(vol), dtype_id(vol), size, _i3(cub), overlap,
                                C.byref(opt), flags, buf.ctypes.data_as(P), cap,
                                C.byref(info), _stream(stream))
        if n != -4 or info.need <= cap:   # WV_ERR_NOMEM with a known size
            break
        cap = info.need
    check_size(n)
    return buf[:n].tobytes(), info


class Batch:
    """Many cubelets compressed as one batch (wv_batch_*).

    cubes: list of dicts with keys vol (cuda tensor [z, y, x] of uint8 /
    int16 / int32 / float32), wavelet, steps, alg (UNIFORM or LOG), bins,
    frac.  Outputs per item after wait(): coef[i] (flat float32, z,x,y
    layout), bins[i] (int32), counts[i] (bin_count + 1), result(i)."""

    def __init__(self, cubes, flags=0):
        import torch
        self.cubes = cubes
        self.geoms, self.coef, self.bins, self.counts = [], [], [], []
        items = (BatchItem * len(cubes))()
        for i, c in enumerate(cubes):
            vol = c["vol"]
            # a view of a box inside a larger volume is read in place
            if vol.dim() != 3 or vol.stride(2) != 1:
                raise ValueError("batch input %d: needs [z, y, x] with x contiguous" % i)
            orig = [vol.shape[2], vol.shape[1], vol.shape[0]]
            g = geom(orig, c["steps"], c["wavelet"], dtype_id(vol))
            self.geoms.append(g)
            dev = vol.device
            self.coef.append(torch.empty(g.count, dtype=torch.float32, device=dev))
            self.bins.append(torch.empty(g.count, dtype=torch.int32, device=dev))
            self.counts.append(torch.empty(c["bins"] + 1, dtype=torch.int32, device=dev))
            it = items[i]
            it.in_d = vol.data_ptr()
            it.dtype = dtype_id(vol)
            it.orig = _i3(orig)
            it.wavelet = c["wavelet"]
            it.steps = _i3(c["steps"])
            it.quant_alg = c["alg"]
            it.bin_count = c["bins"]
            it.thresh_frac = c["frac"]
            it.coef_d = self.coef[-1].data_ptr()
            it.bins_d = self.bins[-1].data_ptr()
            it.counts_d = self.counts[-1].data_ptr()
            if not vol.is_contiguous():
                it.in_stride[0] = vol.stride(1)
                it.in_stride[1] = vol.stride(0)
        self.h = C.c_void_p()
        check(lib().wv_batch_plan(items, len(cubes), flags, C.byref(self.h)))

    def submit(self, stream=None):
        check(lib().wv_batch_submit(self.h, _stream(stream)))

    def wait(self):
        check(lib().wv_batch_wait(self.h))

    def poll(self):
        return check(lib().wv_batch_poll(self.h)) == 1

    def result(self, i):
        r = ThreshResult()
        zb = C.c_int()
        check(lib().wv_batch_result(self.h, i, C.byref(r), C.byref(zb)))
        return r, zb.value

    def write_stream(self, compress_zeros=True, with_footer=True, stream=None):
        """Every item coded on the device into one .cube stream (bytes)."""
        need = C.c_int64()
        cap = 2 * sum(int(g.count) for g in self.geoms) + (1 << 20)
        for _ in range(2):
            buf = np.empty(cap, np.uint8)
            n = lib().wv_batch_write_stream(self.h, None, int(compress_zeros), int(with_footer),
                                            buf.ctypes.data_as(P), cap, C.byref(need),
                                            _stream(stream))
            if n != -4 or need.value <= cap:
                break
            cap = need.value
        check_size(n)
        return buf[:n].tobytes()

    def stage_times(self, stream=None):
        """One eager run serialised on one stream with events between the
        stages (wv_batch_stage_times): dict of transform / select / quantize
        milliseconds and the number of transform pass launches."""
        ms = (C.c_double * 4)()
        check(lib().wv_batch_stage_times(self.h, _stream(stream), ms))
        return dict(transform_ms=ms[0], select_ms=ms[1], quantize_ms=ms[2],
                    transform_launches=int(ms[3]))

    def groups(self):
        out = []
        for g in range(lib().wv_batch_groups(self.h)):
            v = [C.c_int() for _ in range(4)]
            check(lib().wv_batch_group_tuning(self.h, g, *[C.byref(x) for x in v]))
            out.append(dict(tile_cap=v[0].value, ept0=v[1].value, bpc=v[2].value,
                            ncub=v[3].value))
        return out

    def close(self):
        if self.h:
            lib().wv_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NP_DTYPE_ID = {"uint8": UINT8, "int16": INT16, "int32": INT32, "float32": FLOAT32}


class Server:
    """The serving loop (wv_serve_*): requests are host arrays [z, y, x];
    results come back with the bins, counts and .cube image of each."""

    def __init__(self, slots=8, hbm_budget=200 << 30, encode=True, host_threads=0,
                 lloyd_batch=False):
        cfg = ServeConfig(slots, hbm_budget, 1 if encode else 0, host_threads,
                          1 if lloyd_batch else 0)
        self.h = C.c_void_p()
        check(lib().wv_serve_create(C.byref(cfg), C.byref(self.h)))
        self.live = {}
        self.encode = encode

    def submit(self, vol, wavelet, steps, alg, bins, frac, compress_zeros=True,
               lloyd_iters=30, footer=True, want_bins=False):
        vol = np.ascontiguousarray(vol)
        orig = [vol.shape[2], vol.shape[1], vol.shape[0]]
        g = Geom()
        check(lib().wv_geom_init(C.byref(g), _i3(orig), _i3(steps), wavelet,
                                 NP_DTYPE_ID[str(vol.dtype)]))
        out_bins = np.empty(g.count, np.int32) if want_bins else None
        counts = np.empty(bins + 1, np.uint32)
        cap = g.count * 5 + (1 << 20)
        fbuf = np.empty(cap, np.uint8) if self.encode else None
        r = ServeRequest()
        r.in_host = vol.ctypes.data
        r.dtype = NP_DTYPE_ID[str(vol.dtype)]
        r.orig = _i3(orig)
        r.opt = Options(wavelet, alg, bins, _i3(steps), C.c_float(frac),
                        1 if compress_zeros else 0, lloyd_iters, 1 if footer else 0)
        r.bins_out = out_bins.ctypes.data if out_bins is not None else None
        r.counts_out = counts.ctypes.data
        r.file_out = fbuf.ctypes.data if fbuf is not None else None
        r.file_cap = cap if fbuf is not None else 0
        rid = C.c_int64()
        check(lib().wv_serve_submit(self.h, C.byref(r), C.byref(rid)))
        self.live[rid.value] = (vol, out_bins, counts, fbuf)
        return rid.value

    def poll(self, max_results=1024):
        arr = (ServeResult * max_results)()
        n = check(lib().wv_serve_poll(self.h, arr, max_results))
        out = []
        for i in range(n):
            r = arr[i]
            vol, b, counts, fbuf = self.live.pop(r.id)
            out.append(dict(id=r.id, status=r.status, t_submit=r.t_submit, t_admit=r.t_admit,
                            t_done=r.t_done, in_bytes=r.in_bytes, device_bytes=r.device_bytes,
                            thresh=r.thresh, zero_bin=r.zero_bin, lloyd_iters=r.lloyd_iters,
                            bins=b, counts=counts,
                            file=None if fbuf is None else fbuf[:r.file_bytes].tobytes()))
        return out

    def drain(self):
        check(lib().wv_serve_drain(self.h))
        return self.poll(1 << 16)

    def stats(self):
        pb, pi, now = C.c_int64(), C.c_int(), C.c_double()
        check(lib().wv_serve_stats(self.h, C.byref(pb), C.byref(pi), C.byref(now)))
        w, rw = C.c_int64(), C.c_int64()
        check(lib().wv_serve_lloyd_stats(self.h, C.byref(w), C.byref(rw)))
        return dict(peak_bytes=pb.value, peak_inflight=pi.value, now=now.value,
                    lloyd_windows=w.value, lloyd_run_windows=rw.value)

    def close(self):
        if self.h:
            lib().wv_serve_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

