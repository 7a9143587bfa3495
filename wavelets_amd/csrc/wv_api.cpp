// wv_api.cpp -- the extern "C" entry points of libwv_hip.so: geometry,
// transform pass planning, threshold, quantization, and the one-call
// compress that strings them together the way compressFile does
// (test_compress_gpu.cu:150-436).
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

static_assert(sizeof(WvStats) == WV_STATS_BYTES, "stats block size");

namespace wv {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

static int hip_fail(hipError_t e, const char *what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return WV_ERR_HIP;
}

#define HIPCHK(expr, what)                              \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hip_fail(_e, what);    \
  } while (0)

static unsigned ceil_log2(unsigned x) {   // dwt_cpu.cc:75-78
  if (x == 0) return 0;
  unsigned r = 0;
  while ((1u << r) < x) r++;
  return r;
}


// host parameter block -> the device form the kernels read
QuantDev quant_dev_of(const wv_quant_params *p) {
  QuantDev q{};
  q.alg = p->alg;
  q.bin_count = p->bin_count;
  q.zero_bin = p->zero_bin;
  q.threshold = p->threshold;
  q.neg_scale = p->neg_scale;
  q.neg_offset = p->neg_offset;
  q.pos_scale = p->pos_scale;
  q.pos_offset = p->pos_offset;
  if (p->alg == WV_QUANT_LOG) {
    q.n_pos = p->n_pos_edges;
    q.n_neg = p->n_neg_edges;
    q.pos_clamp = p->pos_clamp;
  }
  else if (p->alg == WV_QUANT_LLOYD) { q.n_pos = p->n_boundaries; q.n_neg = 0; }
  q.table = p->table_d;
  return q;
}

}  // namespace wv

using namespace wv;

extern "C" {

const char *wv_last_error(void) { return g_err.c_str(); }
const char *wv_version(void) { return "wv_hip 1.0 gfx950"; }

int wv_geom_init(wv_geom *g, const int orig[3], const int steps[3],
                 int wavelet, int dtype) {
  if (!g || !orig || !steps) { set_error("null argument"); return WV_ERR_ARG; }
  std::memset(g, 0, sizeof *g);
  if (wavelet != WV_WAVELET_HAAR && wavelet != WV_WAVELET_CDF97) {
    set_error("unknown wavelet %d", wavelet);
    return WV_ERR_ARG;
  }
  if (dtype < WV_DTYPE_UINT8 || dtype > WV_DTYPE_INT16) {
    set_error("unknown dtype %d", dtype);
    return WV_ERR_ARG;
  }
  int64_t count = 1;
  for (int a = 0; a < 3; a++) {
    if (orig[a] < 1) { set_error("size must be positive"); return WV_ERR_ARG; }
    g->orig[a] = orig[a];
    int m = (int)ceil_log2((unsigned)orig[a]);          // setWaveletSteps
    int s = steps[a];
    if (s < 0 || s > m) s = m;
    g->steps[a] = s;
    const int mod = 1 << s;                              // dwt_padded_length
    g->padded[a] = mod * ((orig[a] - 1) / mod + 1);
    count *= g->padded[a];
    if (g->padded[a] > WV_MAX_LINE) {
      set_error("axis %d length %d exceeds %d", a, g->padded[a], WV_MAX_LINE);
      return WV_ERR_UNSUPPORTED;
    }
  }
  if (count > 0x7fffffffLL) {   // Cube index math is int (wavelet.h:386-391)
    set_error("cubelet has more than 2^31-1 elements");
    return WV_ERR_UNSUPPORTED;
  }
  g->wavelet = wavelet;
  g->dtype = dtype;
  g->count = count;
  if (g->steps[0] || g->steps[1] || g->steps[2]) {
    g->out_size[0] = g->padded[2];   // rotated twice: (z, x, y)
    g->out_size[1] = g->padded[0];
    g->out_size[2] = g->padded[1];
  } else {
    std::memcpy(g->out_size, g->padded, sizeof g->padded);
  }
  // intermediates need one workspace of the padded size (forward: the
  // passes before the last layout change; inverse: none, but callers size
  // the workspace the same way)
  g->ws_floats = count;
  return WV_OK;
}

int wv_stats_clear(void *stats_d, void *stream) {
  HIPCHK(stats_clear(stats_d, (hipStream_t)stream), "stats clear");
  return WV_OK;
}

int wv_stats_compute(const float *coef_d, int64_t n, void *stats_d,
                     void *stream) {
  if (!coef_d || !stats_d || n <= 0) { set_error("bad arguments"); return WV_ERR_ARG; }
  HIPCHK(stats_compute(coef_d, n, stats_d, (hipStream_t)stream), "stats");
  return WV_OK;
}

int wv_dwt3d_forward(const wv_geom *g, const void *in_d, float *out_d,
                     float *ws_d, void *stats_d, void *stream) {
  return wv_dwt3d_forward_ev(g, in_d, out_d, ws_d, stats_d, nullptr, stream);
}

int wv_dwt3d_forward_ev(const wv_geom *g, const void *in_d, float *out_d,
                        float *ws_d, void *stats_d, void **events,
                        void *stream) {
  if (!g || !in_d || !out_d) { set_error("null argument"); return WV_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  PassSpec spec[3];
  const int n = forward_specs(g->padded, g->steps, spec);
  int dst_out[3];
  pass_buffers(spec, n, dst_out);
  for (int i = 0; i < n; i++)
    if (!dst_out[i] && !ws_d) { set_error("workspace required"); return WV_ERR_NOMEM; }
  const void *src = in_d;
  for (int i = 0; i < n; i++) {
    DwtPass P{};
    P.X = g->padded[0]; P.Y = g->padded[1]; P.Z = g->padded[2];
    P.ox = g->orig[0]; P.oy = g->orig[1]; P.oz = g->orig[2];
    P.src = src;
    P.dst = dst_out[i] ? out_d : ws_d;
    const bool clamp = i == 0;
    const int ntiles = plan_pass(P, spec[i], clamp);
    if (ntiles < 0) { set_error("pass %d: unsupported geometry", i); return WV_ERR_UNSUPPORTED; }
    if (events && i == 0) HIPCHK(hipEventRecord((hipEvent_t)events[0], s), "event");
    HIPCHK(launch_pass(P, ntiles, g->wavelet, false, clamp ? g->dtype : WV_DTYPE_FLOAT32,
                       clamp, false, s), "dwt pass");
    if (events) HIPCHK(hipEventRecord((hipEvent_t)events[i + 1], s), "event");
    src = P.dst;
  }
  if (stats_d) HIPCHK(stats_compute(out_d, g->count, stats_d, s), "stats");
  return WV_OK;
}

int wv_dwt3d_inverse(const wv_geom *g, const float *coef_d, float *out_d,
                     float *ws_d, void *stream) {
  if (!g || !coef_d || !out_d) { set_error("null argument"); return WV_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  PassSpec spec[3];
  const int n = inverse_specs(g->padded, g->steps, spec);
  if (n == 0) {
    HIPCHK(hipMemcpyAsync(out_d, coef_d, g->count * 4, hipMemcpyDeviceToDevice, s),
           "copy");
    return WV_OK;
  }
  int dst_out[3];
  pass_buffers(spec, n, dst_out);
  const float *src = coef_d;
  for (int i = 0; i < n; i++) {
    if (!dst_out[i] && !ws_d) { set_error("workspace required"); return WV_ERR_NOMEM; }
    DwtPass P{};
    P.X = g->padded[0]; P.Y = g->padded[1]; P.Z = g->padded[2];
    P.ox = P.X; P.oy = P.Y; P.oz = P.Z;
    P.src = src;
    P.dst = dst_out[i] ? out_d : ws_d;
    const int ntiles = plan_pass(P, spec[i], false);
    if (ntiles < 0) { set_error("inverse pass %d: unsupported geometry", i); return WV_ERR_UNSUPPORTED; }
    HIPCHK(launch_pass(P, ntiles, g->wavelet, true, WV_DTYPE_FLOAT32, false, false, s),
           "inverse pass");
    src = P.dst;
  }
  return WV_OK;
}

int wv_dwt3d_num_passes(const wv_geom *g) {
  PassSpec spec[3];
  return forward_specs(g->padded, g->steps, spec);
}

int wv_event_create(void **ev) {
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e), "event create");
  *ev = (void *)e;
  return WV_OK;
}

int wv_event_destroy(void *ev) {
  HIPCHK(hipEventDestroy((hipEvent_t)ev), "event destroy");
  return WV_OK;
}

int wv_event_record(void *ev, void *stream) {
  HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "event record");
  return WV_OK;
}

int wv_event_elapsed_ms(void *a, void *b, float *ms) {
  HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "event elapsed");
  return WV_OK;
}

int64_t wv_threshold_ws_bytes(int64_t n) { return threshold_ws_bytes(n); }

int wv_threshold(const float *coef_d, int64_t n, float frac,
                 const void *stats_d, void *ws_d, wv_thresh_result *result_d,
                 void *stream) {
  if (!coef_d || !ws_d || !result_d || n <= 0 || n > 0x7fffffffLL) {
    set_error("bad arguments");
    return WV_ERR_ARG;
  }
  HIPCHK(threshold_select(coef_d, n, frac, stats_d, ws_d, result_d,
                          (hipStream_t)stream), "threshold");
  return WV_OK;
}

const float *wv_quant_params_dequant_table(const wv_quant_params *p);

int64_t wv_quant_table_bytes(const wv_quant_params *p) {
  int64_t t = 0;
  if (p->alg == WV_QUANT_LOG) t = p->n_pos_edges + p->n_neg_edges;
  else if (p->alg == WV_QUANT_LLOYD) t = p->n_boundaries;
  return (t + p->bin_count) * 4;
}

int wv_quant_params_upload(wv_quant_params *p, void *table_d, void *stream) {
  if (!p || !table_d) { set_error("bad arguments"); return WV_ERR_ARG; }
  std::vector<float> t;
  if (p->alg == WV_QUANT_LOG) {
    t.insert(t.end(), p->pos_edges, p->pos_edges + p->n_pos_edges);
    t.insert(t.end(), p->neg_edges, p->neg_edges + p->n_neg_edges);
  } else if (p->alg == WV_QUANT_LLOYD) {
    t.insert(t.end(), p->boundaries, p->boundaries + p->n_boundaries);
  }
  const float *dq = wv_quant_params_dequant_table(p);
  t.insert(t.end(), dq, dq + p->bin_count);
  HIPCHK(hipMemcpyAsync(table_d, t.data(), t.size() * 4, hipMemcpyHostToDevice,
                        (hipStream_t)stream), "table upload");
  // the staging vector dies here: make sure the copy has consumed it
  HIPCHK(hipStreamSynchronize((hipStream_t)stream), "table upload sync");
  p->table_d = (const float *)table_d;
  return WV_OK;
}


int wv_quantize_hist(const float *coef_d, int64_t n, const wv_quant_params *p,
                     int32_t *bins_d, uint32_t *counts_d, void *stream) {
  if (!coef_d || !p || !bins_d || !counts_d || n <= 0) { set_error("bad arguments"); return WV_ERR_ARG; }
  if (p->alg != WV_QUANT_UNIFORM && !p->table_d) {
    set_error("parameter tables not uploaded");
    return WV_ERR_STATE;
  }
  HIPCHK(quantize_hist(coef_d, n, quant_dev_of(p), bins_d, counts_d, (hipStream_t)stream),
         "quantize");
  return WV_OK;
}

int wv_quantize_hist_uniform_dev(const float *coef_d, int64_t n, int bin_count,
                                 const wv_thresh_result *result_d,
                                 int32_t *bins_d, uint32_t *counts_d,
                                 void *stream) {
  if (!coef_d || !result_d || !bins_d || !counts_d || n <= 0 || bin_count < 2) {
    set_error("bad arguments");
    return WV_ERR_ARG;
  }
  HIPCHK(quantize_hist_uniform_dev(coef_d, n, bin_count, result_d, bins_d, counts_d,
                                   (hipStream_t)stream), "quantize");
  return WV_OK;
}

int wv_dequantize(const int32_t *bins_d, int64_t n, const wv_quant_params *p,
                  float *out_d, void *stream) {
  if (!bins_d || !p || !p->table_d || !out_d) { set_error("bad arguments"); return WV_ERR_ARG; }
  const float *dq = p->table_d + (wv_quant_table_bytes(p) / 4 - p->bin_count);
  HIPCHK(dequantize_table(bins_d, n, dq, p->bin_count, out_d, (hipStream_t)stream),
         "dequantize");
  return WV_OK;
}

}  // extern "C"
