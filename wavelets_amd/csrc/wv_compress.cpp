// wv_compress.cpp -- the one-call drop-in for compressFile
// (test_compress_gpu.cu:150-436) on a device-resident cubelet:
//   forward DWT (statistics fused) -> radix-select threshold -> [Lloyd] ->
//   quantize + histogram -> bins and counts to the host -> Huffman + .cube.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstring>
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

using namespace wv;

namespace {

inline int64_t align256(int64_t v) { return (v + 255) & ~(int64_t)255; }

struct Layout {
  int64_t coef, ws, stats, result, counts, table, lloyd, total;
};

Layout layout_for(const wv_geom &g, const wv_options *opt) {
  Layout L{};
  int64_t off = 0;
  L.coef = off; off += align256(g.count * 4);
  int64_t ws = g.ws_floats * 4;
  int64_t tws = wv_threshold_ws_bytes(g.count);
  if (tws > ws) ws = tws;
  L.ws = off; off += align256(ws);
  L.stats = off; off += align256(WV_STATS_BYTES);
  L.result = off; off += align256(sizeof(wv_thresh_result));
  L.counts = off; off += align256(((int64_t)opt->bin_count + 128) * 4);
  L.table = off; off += align256(((int64_t)opt->bin_count * 3 + 16) * 4);
  L.lloyd = off;
  if (opt->quant_alg == WV_QUANT_LLOYD) off += align256(wv_lloyd_ws_bytes(g.count, opt->bin_count));
  L.total = off;
  return L;
}

}  // namespace

extern "C" int64_t wv_compress_work_bytes(const int orig[3], const wv_options *opt) {
  wv_geom g;
  if (wv_geom_init(&g, orig, opt->steps, opt->wavelet, WV_DTYPE_FLOAT32) != WV_OK) return -1;
  return layout_for(g, opt).total;
}

#define CHK(call)                         \
  do {                                    \
    int _rc = (call);                     \
    if (_rc != WV_OK) return _rc;         \
  } while (0)
#define HCHK(expr, what)                                          \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) {                                       \
      set_error("%s: %s", what, hipGetErrorString(_e));           \
      return WV_ERR_HIP;                                          \
    }                                                             \
  } while (0)

extern "C" int wv_compress(const void *in_d, int dtype, const int orig[3],
                           const wv_options *opt, void *work_d,
                           int64_t work_bytes, int32_t *bins_host,
                           uint8_t *file_buf, int64_t file_cap,
                           wv_compress_info *info, void *stream) {
  if (!in_d || !orig || !opt || !work_d || !info) { set_error("null argument"); return WV_ERR_ARG; }
  if (opt->bin_count < 2 || opt->bin_count > WV_MAX_BINS) { set_error("bin count"); return WV_ERR_ARG; }
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = (hipStream_t)stream;
  std::memset(info, 0, sizeof *info);
  wv_geom g;
  CHK(wv_geom_init(&g, orig, opt->steps, opt->wavelet, dtype));
  const Layout L = layout_for(g, opt);
  if (work_bytes < L.total) { set_error("work buffer %lld < %lld", (long long)work_bytes, (long long)L.total); return WV_ERR_NOMEM; }
  char *w = (char *)work_d;
  float *coef = (float *)(w + L.coef);
  float *ws = (float *)(w + L.ws);
  void *stats = nullptr;   // the threshold select needs no precomputed statistics
  wv_thresh_result *res_d = (wv_thresh_result *)(w + L.result);
  uint32_t *counts_d = (uint32_t *)(w + L.counts);
  void *table_d = w + L.table;

  hipEvent_t e0, e1;
  HCHK(hipEventCreate(&e0), "event");
  HCHK(hipEventCreate(&e1), "event");
  HCHK(hipEventRecord(e0, s), "event record");

  CHK(wv_dwt3d_forward(&g, in_d, coef, ws, nullptr, s));
  CHK(wv_threshold(coef, g.count, opt->thresh_frac, stats, ws, res_d, s));
  wv_thresh_result res;
  HCHK(hipMemcpyAsync(&res, res_d, sizeof res, hipMemcpyDeviceToHost, s), "D2H threshold");
  HCHK(hipStreamSynchronize(s), "sync");

  std::vector<float> bounds, values;
  int lloyd_iters = 0;
  wv_quant_params qp;
  if (opt->quant_alg == WV_QUANT_LLOYD) {
    bounds.resize(opt->bin_count - 1);
    values.resize(opt->bin_count);
    CHK(wv_lloyd(coef, g.count, &res, opt->bin_count, opt->lloyd_max_iters, w + L.lloyd,
                 bounds.data(), values.data(), &lloyd_iters, s));
    CHK(wv_quant_params_init(&qp, WV_QUANT_LLOYD, opt->bin_count, res.threshold, 0,
                             bounds.data(), values.data()));
  } else {
    // QuantUniform / QuantLog with maxValue = max |x| (test_compress_cpu.cc:262-265)
    CHK(wv_quant_params_init(&qp, opt->quant_alg, opt->bin_count, res.threshold,
                             res.max_abs, nullptr, nullptr));
  }
  int rc = wv_quant_params_upload(&qp, table_d, s);
  if (rc == WV_OK) rc = wv_quantize_hist(coef, g.count, &qp, (int32_t *)coef, counts_d, s);
  if (rc != WV_OK) { wv_quant_params_free(&qp); return rc; }
  const int zero_bin = qp.zero_bin;
  HCHK(hipEventRecord(e1, s), "event record");

  std::vector<uint32_t> counts(opt->bin_count);
  std::vector<int32_t> own_bins;
  int32_t *bins = bins_host;
  if (!bins) { own_bins.resize(g.count); bins = own_bins.data(); }
  HCHK(hipMemcpyAsync(counts.data(), counts_d, counts.size() * 4, hipMemcpyDeviceToHost, s), "D2H counts");
  HCHK(hipMemcpyAsync(bins, coef, g.count * 4, hipMemcpyDeviceToHost, s), "D2H bins");
  HCHK(hipStreamSynchronize(s), "sync");
  float ms = 0;
  HCHK(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);

  info->thresh = res;
  std::memcpy(info->out_size, g.out_size, sizeof g.out_size);
  std::memcpy(info->steps, g.steps, sizeof g.steps);
  info->zero_bin = zero_bin;
  info->lloyd_iters = lloyd_iters;
  info->ms_device = ms;

  if (file_buf) {
    wv_cube_desc d{};
    std::memcpy(d.orig, g.orig, sizeof d.orig);
    std::memcpy(d.steps, g.steps, sizeof d.steps);
    std::memcpy(d.out_size, g.out_size, sizeof d.out_size);
    d.wavelet = opt->wavelet;
    d.quant_alg = opt->quant_alg;
    d.bin_count = opt->bin_count;
    // 16-bit data is stored as an INT32 cubelet (shortImagesToCubelets.cc:9-27)
    d.orig_dtype = dtype == WV_DTYPE_INT16 ? WV_DTYPE_INT32 : dtype;
    d.thresh_frac = opt->thresh_frac;
    d.threshold = res.threshold;
    d.max_val = res.max_abs;
    d.boundaries = bounds.empty() ? nullptr : bounds.data();
    d.values = values.empty() ? nullptr : values.data();
    d.zero_bin = opt->compress_zeros ? zero_bin : -1;
    d.with_footer = opt->with_footer;
    int64_t pay = 0;
    const int64_t nb = wv_write_cube(&d, bins, g.count, counts.data(), file_buf, file_cap, &pay);
    if (nb < 0) { wv_quant_params_free(&qp); set_error("file buffer too small"); return (int)nb; }
    info->file_bytes = nb;
    info->payload_bytes = pay;
  }
  wv_quant_params_free(&qp);
  info->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return WV_OK;
}
