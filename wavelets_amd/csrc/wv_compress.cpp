// wv_compress.cpp -- the one-call drop-in for compressFile
// (test_compress_gpu.cu:150-436) on a device-resident cubelet:
//   forward DWT -> radix-select threshold -> [Lloyd] -> quantize +
//   histogram -> Huffman coding on the device -> payload to the host ->
//   .cube container.
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
  int64_t coef, ws, stats, result, counts, table, enc, lloyd, total;
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
  L.enc = off; off += align256(wv_encode_ws_bytes(g.count, opt->bin_count));
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

namespace {

// The device-resident stages of one cubelet: transform, threshold select
// (one synchronisation for its result), [Lloyd], quantizer tables, fused
// quantize + histogram into the work buffer (bins over the coefficients,
// counts at L.counts).  The quantize is left running on the stream.
struct Stages {
  wv_geom g;
  Layout L;
  wv_thresh_result res;
  wv_quant_params qp;
  bool have_qp = false;
  std::vector<float> bounds, values;
  int lloyd_iters = 0;
  ~Stages() {
    if (have_qp) wv_quant_params_free(&qp);
  }
};

int device_stages(const void *in_d, int dtype, const int orig[3], const wv_options *opt,
                  void *work_d, int64_t work_bytes, hipStream_t s, Stages &st,
                  LloydBatcher *lb = nullptr) {
  if (!in_d || !orig || !opt || !work_d) { set_error("null argument"); return WV_ERR_ARG; }
  if (opt->bin_count < 2 || opt->bin_count > WV_MAX_BINS) { set_error("bin count"); return WV_ERR_ARG; }
  CHK(wv_geom_init(&st.g, orig, opt->steps, opt->wavelet, dtype));
  const wv_geom &g = st.g;
  st.L = layout_for(g, opt);
  const Layout &L = st.L;
  if (work_bytes < L.total) {
    set_error("work buffer %lld < %lld", (long long)work_bytes, (long long)L.total);
    return WV_ERR_NOMEM;
  }
  char *w = (char *)work_d;
  float *coef = (float *)(w + L.coef);
  float *ws = (float *)(w + L.ws);
  wv_thresh_result *res_d = (wv_thresh_result *)(w + L.result);
  uint32_t *counts_d = (uint32_t *)(w + L.counts);
  void *table_d = w + L.table;

  CHK(wv_dwt3d_forward(&g, in_d, coef, ws, nullptr, s));
  // the threshold select needs no precomputed statistics
  CHK(wv_threshold(coef, g.count, opt->thresh_frac, nullptr, ws, res_d, s));
  const wv_thresh_result &res = st.res;
  if (opt->quant_alg == WV_QUANT_UNIFORM) {
    // parameters from the threshold record on the device (the float
    // expressions of QuantUniform::init): the quantize follows the select
    // with no host round trip; the record comes back after it
    CHK(wv_quantize_hist_uniform_dev(coef, g.count, opt->bin_count, res_d, (int32_t *)coef,
                                     counts_d, s));
    HCHK(hipMemcpyAsync(&st.res, res_d, sizeof st.res, hipMemcpyDeviceToHost, s), "D2H threshold");
    HCHK(hipStreamSynchronize(s), "sync");
    CHK(wv_quant_params_init(&st.qp, WV_QUANT_UNIFORM, opt->bin_count, res.threshold,
                             res.max_abs, nullptr, nullptr));
    st.have_qp = true;
    return WV_OK;
  }
  HCHK(hipMemcpyAsync(&st.res, res_d, sizeof st.res, hipMemcpyDeviceToHost, s), "D2H threshold");
  HCHK(hipStreamSynchronize(s), "sync");
  if (opt->quant_alg == WV_QUANT_LLOYD) {
    st.bounds.resize(opt->bin_count - 1);
    st.values.resize(opt->bin_count);
    CHK(lloyd_run(coef, g.count, &res, opt->bin_count, opt->lloyd_max_iters, w + L.lloyd,
                  st.bounds.data(), st.values.data(), &st.lloyd_iters, s, lb));
    CHK(wv_quant_params_init(&st.qp, WV_QUANT_LLOYD, opt->bin_count, res.threshold, 0,
                             st.bounds.data(), st.values.data()));
  } else {
    // QuantUniform / QuantLog with maxValue = max |x| (test_compress_cpu.cc:262-265)
    CHK(wv_quant_params_init(&st.qp, opt->quant_alg, opt->bin_count, res.threshold,
                             res.max_abs, nullptr, nullptr));
  }
  st.have_qp = true;
  CHK(wv_quant_params_upload(&st.qp, table_d, s));
  CHK(wv_quantize_hist(coef, g.count, &st.qp, (int32_t *)coef, counts_d, s));
  return WV_OK;
}

}  // namespace

namespace wv {

// Device Huffman coding of the bins device_stages left in the work buffer
// (over the coefficients); the payload words land in the DWT workspace,
// which holds at least 4 bytes per bin (*words_d, *nwords).
int encode_work_device(const wv_geom &g, const wv_options *opt, int zero_bin, void *work_d,
                       const uint32_t **words_d, int64_t *nwords, wv_huff **book,
                       hipStream_t s) {
  const Layout L = layout_for(g, opt);
  char *w = (char *)work_d;
  const int64_t cap = (L.stats - L.ws) / 4;
  CHK(wv_encode_device((const int32_t *)(w + L.coef), g.count, (const uint32_t *)(w + L.counts),
                       opt->bin_count, opt->compress_zeros ? zero_bin : -1, w + L.enc,
                       (uint32_t *)(w + L.ws), cap, nwords, book, s));
  *words_d = (const uint32_t *)(w + L.ws);
  return WV_OK;
}

// ... and the payload copied to words_host
int encode_work(const wv_geom &g, const wv_options *opt, int zero_bin, void *work_d,
                std::vector<uint32_t> &words_host, wv_huff **book, hipStream_t s) {
  const uint32_t *wd = nullptr;
  int64_t nw = 0;
  CHK(encode_work_device(g, opt, zero_bin, work_d, &wd, &nw, book, s));
  words_host.resize((size_t)nw);
  if (nw > 0)
    HCHK(hipMemcpyAsync(words_host.data(), wd, nw * 4, hipMemcpyDeviceToHost, s), "D2H payload");
  HCHK(hipStreamSynchronize(s), "sync");
  return WV_OK;
}

void cube_desc_of(const wv_geom &g, const wv_options *opt, int dtype, const wv_thresh_result &res,
                  const std::vector<float> &bounds, const std::vector<float> &values,
                  int zero_bin, wv_cube_desc &d) {
  d = wv_cube_desc{};
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
}

}  // namespace wv

extern "C" int64_t wv_compress_counts_offset(const int orig[3], const wv_options *opt) {
  wv_geom g;
  if (wv_geom_init(&g, orig, opt->steps, opt->wavelet, WV_DTYPE_FLOAT32) != WV_OK) return -1;
  return layout_for(g, opt).counts;
}

namespace wv {

// wv_compress_device with an optional Lloyd batcher (the serving loop's)
int compress_device(const void *in_d, int dtype, const int orig[3], const wv_options *opt,
                    void *work_d, int64_t work_bytes, float *codebook_out,
                    wv_compress_info *info, hipStream_t stream, LloydBatcher *lb) {
  if (!info) { set_error("null argument"); return WV_ERR_ARG; }
  std::memset(info, 0, sizeof *info);
  Stages st;
  CHK(device_stages(in_d, dtype, orig, opt, work_d, work_bytes, stream, st, lb));
  if (codebook_out && opt->quant_alg == WV_QUANT_LLOYD) {
    std::memcpy(codebook_out, st.bounds.data(), st.bounds.size() * sizeof(float));
    std::memcpy(codebook_out + st.bounds.size(), st.values.data(),
                st.values.size() * sizeof(float));
  }
  info->thresh = st.res;
  std::memcpy(info->out_size, st.g.out_size, sizeof st.g.out_size);
  std::memcpy(info->steps, st.g.steps, sizeof st.g.steps);
  info->zero_bin = st.qp.zero_bin;
  info->lloyd_iters = st.lloyd_iters;
  return WV_OK;
}

}  // namespace wv

extern "C" int wv_compress_device(const void *in_d, int dtype, const int orig[3],
                                  const wv_options *opt, void *work_d, int64_t work_bytes,
                                  float *codebook_out, wv_compress_info *info,
                                  void *stream) {
  return compress_device(in_d, dtype, orig, opt, work_d, work_bytes, codebook_out, info,
                         (hipStream_t)stream, nullptr);
}

extern "C" int wv_compress(const void *in_d, int dtype, const int orig[3],
                           const wv_options *opt, void *work_d,
                           int64_t work_bytes, int32_t *bins_host,
                           uint8_t *file_buf, int64_t file_cap,
                           wv_compress_info *info, void *stream) {
  if (!info) { set_error("null argument"); return WV_ERR_ARG; }
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = (hipStream_t)stream;
  std::memset(info, 0, sizeof *info);
  Stages st;
  CHK(device_stages(in_d, dtype, orig, opt, work_d, work_bytes, s, st));
  const wv_geom &g = st.g;
  const Layout &L = st.L;
  char *w = (char *)work_d;
  float *coef = (float *)(w + L.coef);
  uint32_t *counts_d = (uint32_t *)(w + L.counts);
  const wv_thresh_result &res = st.res;
  const std::vector<float> &bounds = st.bounds, &values = st.values;
  const int lloyd_iters = st.lloyd_iters;
  const int zero_bin = st.qp.zero_bin;
  HCHK(hipStreamSynchronize(s), "sync");
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

  if (bins_host)
    HCHK(hipMemcpyAsync(bins_host, coef, g.count * 4, hipMemcpyDeviceToHost, s), "D2H bins");

  info->thresh = res;
  std::memcpy(info->out_size, g.out_size, sizeof g.out_size);
  std::memcpy(info->steps, g.steps, sizeof g.steps);
  info->zero_bin = zero_bin;
  info->lloyd_iters = lloyd_iters;
  info->ms_device = ms;

  if (file_buf) {
    // the payload goes from the device straight into the file image
    const uint32_t *words_d = nullptr;
    int64_t nw = 0;
    wv_huff *book = nullptr;
    CHK(encode_work_device(g, opt, zero_bin, work_d, &words_d, &nw, &book, s));
    wv_cube_desc d;
    cube_desc_of(g, opt, dtype, res, bounds, values, zero_bin, d);
    const int64_t nb = wv_write_cube_device(&d, book, words_d, nw, file_buf, file_cap, s);
    wv_huff_free(book);
    if (nb < 0) {
      if (nb == WV_ERR_NOMEM) set_error("file buffer too small");
      return (int)nb;
    }
    info->file_bytes = nb;
    info->payload_bytes = nw * 4;
  } else {
    HCHK(hipStreamSynchronize(s), "sync");
  }
  info->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return WV_OK;
}
