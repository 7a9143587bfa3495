// wv_opt.hip -- parameter optimisation (-opt): OptimizationData and
// testParameters (optimize.h:6-145, optimize_gpu.cu:15-202), the
// interface the reference's error-control sweeps (optimize.cc:4-133) are
// written against.
//
// The forward transform runs once; the sorted |x| the sweeps index
// (getSorted / findSorted) is one device radix sort of the 31-bit keys.
// A test point then stays on the device: quantize + histogram with the
// fused pass (or the plain threshold when no bins are asked for), the
// zero-run statistics of the coder, dequantize, inverse transform, and the
// error sums against the original over its own extent with the per-type
// restore rules of computeErrorRatesAfterDequant
// (test_compress_common.cc:616-694).  Only the histogram, two run counts
// and five sums cross PCIe; the Huffman code lengths are built on the host
// as for a compression.  The sums are deterministic: fixed per-block
// ranges reduced in a fixed order, in double like ErrorAccumulator
// (test_compress_common.h:90-160).
#include <cmath>
#include <cstring>
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define OPT_ERR_BLOCKS 2048

__global__ void abs_keys_kernel(const float *__restrict__ x, int64_t n, uint32_t *__restrict__ k) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    k[i] = __builtin_bit_cast(uint32_t, x[i]) & 0x7fffffffu;
}

// applyThresholdKernel (optimize_gpu.cu:38-47): |x| <= T -> 0
__global__ void apply_threshold_kernel(const float *__restrict__ x, int64_t n, float t,
                                       float *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float f = x[i];
    if (fabsf(f) <= t) f = 0;
    out[i] = f;
  }
}

// ErrorAccumulator::add terms (float products, double sums) of one
// original sample a against its restored value b
struct ErrSums {
  double v, vv, d, dd;
};

template <typename T>
__device__ inline float restored(float f);
template <> __device__ inline float restored<uint8_t>(float f) {
  // ByteInputData::floatToByte: <= 0 -> 0, >= 255 -> 255, else truncation
  return f <= 0.f ? 0.f : (f >= 255.f ? 255.f : (float)(uint8_t)f);
}
template <> __device__ inline float restored<int32_t>(float f) { return (float)(int)(f + 0.5f); }
template <> __device__ inline float restored<int16_t>(float f) { return (float)(int)(f + 0.5f); }
// float input is compared against (int)restored (test_compress_common.cc:679-686)
template <> __device__ inline float restored<float>(float f) { return (float)(int)f; }

// block b sums rows b, b + G, ... (row = (z, y) of the original extent);
// partial[b] holds its four sums
template <typename T>
__global__ void __launch_bounds__(256)
error_partials_kernel(const T *__restrict__ orig, const float *__restrict__ rec, int ox, int oy,
                      int oz, int px, int py, double *__restrict__ partial) {
  ErrSums s{0, 0, 0, 0};
  const int64_t rows = (int64_t)oy * oz;
  for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    const int64_t z = r / oy, y = r - z * oy;
    const T *o = orig + r * ox;
    const float *q = rec + (z * py + y) * (int64_t)px;
    for (int x = threadIdx.x; x < ox; x += blockDim.x) {
      const float a = (float)o[x];
      const float b = restored<T>(q[x]);
      s.v += a;
      s.vv += a * a;
      const float diff = fabsf(a - b);
      s.d += diff;
      s.dd += diff * diff;
    }
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = s.v;
  red[1][threadIdx.x] = s.vv;
  red[2][threadIdx.x] = s.d;
  red[3][threadIdx.x] = s.dd;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// one block: the partials in a fixed order
__global__ void __launch_bounds__(256) error_final_kernel(double *partial, int nblk) {
  __shared__ double red[4][256];
  double s[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int k = 0; k < 4; k++) s[k] += partial[b * 4 + k];
  for (int k = 0; k < 4; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[(int64_t)nblk * 4 + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace wv

using namespace wv;

struct wv_opt {
  wv_geom g;
  int dtype, max_possible, compress_zeros;
  const void *in_d;
  char *arena = nullptr;
  float *coef, *work, *rec, *ws;
  uint32_t *sorted;
  int32_t *bins;
  uint32_t *counts;
  char *enc_ws = nullptr, *sort_tmp, *table;
  int64_t enc_cap = 0;   // enc_ws: its own buffer, grown to what a test needs
  double *partial;
  double *sums_h = nullptr;   // pinned: the four error sums
  char *lloyd_ws = nullptr;   // Lloyd test points: wv_lloyd_ws_bytes(n, bins), grown
  int64_t lloyd_cap = 0;
  size_t sort_bytes;
  int64_t table_cap;
  float min_val, max_val, max_abs;
};

namespace {

inline int64_t al256(int64_t v) { return (v + 255) & ~(int64_t)255; }

#define OCHK(call)                      \
  do {                                  \
    int _rc = (call);                   \
    if (_rc != WV_OK) return _rc;       \
  } while (0)
#define OHIP(expr, what)                                          \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) {                                       \
      set_error("%s: %s", what, hipGetErrorString(_e));           \
      return WV_ERR_HIP;                                          \
    }                                                             \
  } while (0)

unsigned grid_of(int64_t n) {
  int64_t b = (n + 255) / 256;
  const int64_t cap = 8 * (int64_t)cu_count() * 8;
  return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

// the error sums of o->rec against the original -> o->sums_h (pinned),
// enqueued only: the caller synchronises
int error_sums(wv_opt *o, hipStream_t s) {
  const wv_geom &g = o->g;
  const int ox = g.orig[0], oy = g.orig[1], oz = g.orig[2];
  const int64_t rows = (int64_t)oy * oz;
  const int nblk = (int)(rows < OPT_ERR_BLOCKS ? rows : OPT_ERR_BLOCKS);
  const int px = g.padded[0], py = g.padded[1];
  switch (o->dtype) {
    case WV_DTYPE_UINT8:
      hipLaunchKernelGGL(error_partials_kernel<uint8_t>, dim3(nblk), dim3(256), 0, s,
                         (const uint8_t *)o->in_d, o->rec, ox, oy, oz, px, py, o->partial);
      break;
    case WV_DTYPE_INT16:
      hipLaunchKernelGGL(error_partials_kernel<int16_t>, dim3(nblk), dim3(256), 0, s,
                         (const int16_t *)o->in_d, o->rec, ox, oy, oz, px, py, o->partial);
      break;
    case WV_DTYPE_INT32:
      hipLaunchKernelGGL(error_partials_kernel<int32_t>, dim3(nblk), dim3(256), 0, s,
                         (const int32_t *)o->in_d, o->rec, ox, oy, oz, px, py, o->partial);
      break;
    default:
      hipLaunchKernelGGL(error_partials_kernel<float>, dim3(nblk), dim3(256), 0, s,
                         (const float *)o->in_d, o->rec, ox, oy, oz, px, py, o->partial);
  }
  hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(256), 0, s, o->partial, nblk);
  OHIP(hipGetLastError(), "wv_opt: error sums");
  OHIP(hipMemcpyAsync(o->sums_h, o->partial + (int64_t)nblk * 4, 4 * sizeof(double),
                      hipMemcpyDeviceToHost, s),
       "wv_opt: error sums D2H");
  return WV_OK;
}

}  // namespace

extern "C" int wv_opt_create(const void *in_d, int dtype, const int orig[3], const int steps[3],
                             int wavelet, int max_possible, int compress_zeros, wv_opt **out,
                             void *stream) {
  if (!in_d || !orig || !steps || !out) {
    set_error("wv_opt_create: null argument");
    return WV_ERR_ARG;
  }
  *out = nullptr;
  wv_opt *o = new wv_opt();
  int rc = wv_geom_init(&o->g, orig, steps, wavelet, dtype);
  if (rc != WV_OK) { delete o; return rc; }
  o->dtype = dtype;
  o->max_possible = max_possible;
  o->compress_zeros = compress_zeros;
  o->in_d = in_d;
  const int64_t n = o->g.count;
  hipStream_t s = (hipStream_t)stream;
  if (n > 0x7fffffffll) {
    delete o;
    set_error("wv_opt_create: sort of %lld keys", (long long)n);
    return WV_ERR_ARG;
  }
  // the sort's key count (device) and its per-block digit histograms
  const size_t sort_bytes = 256 + (size_t)radix_sort_hist_bytes(n);
  hipError_t e = hipSuccess;
  o->sort_bytes = sort_bytes;
  o->table_cap = al256(((int64_t)WV_MAX_BINS * 3 + 16) * 4);
  const int64_t enc = wv_encode_ws_bytes(n, 256);
  const int64_t thr = wv_threshold_ws_bytes(n);
  o->enc_cap = al256(enc > thr ? enc : thr);
  int64_t off = 0;
  const int64_t o_coef = off; off += al256(n * 4);
  const int64_t o_work = off; off += al256(n * 4);
  const int64_t o_rec = off; off += al256(n * 4);
  const int64_t o_ws = off; off += al256((o->g.ws_floats > 0 ? o->g.ws_floats : 1) * 4);
  const int64_t o_sorted = off; off += al256(n * 4);
  const int64_t o_bins = off; off += al256(n * 4);
  const int64_t o_counts = off; off += al256(((int64_t)WV_MAX_BINS + 64) * 4);
  const int64_t o_sort = off; off += al256((int64_t)sort_bytes);
  const int64_t o_table = off; off += o->table_cap;
  const int64_t o_part = off; off += al256((OPT_ERR_BLOCKS + 1) * 4 * 8);
  if (hipMalloc((void **)&o->arena, (size_t)off) != hipSuccess ||
      hipMalloc((void **)&o->enc_ws, (size_t)o->enc_cap) != hipSuccess ||
      hipHostMalloc((void **)&o->sums_h, 4 * sizeof(double), 0) != hipSuccess) {
    if (o->arena) (void)hipFree(o->arena);
    if (o->enc_ws) (void)hipFree(o->enc_ws);
  if (o->lloyd_ws) (void)hipFree(o->lloyd_ws);
    if (o->sums_h) (void)hipHostFree(o->sums_h);
    delete o;
    set_error("wv_opt_create: %lld device bytes", (long long)off);
    return WV_ERR_NOMEM;
  }
  char *a = o->arena;
  o->coef = (float *)(a + o_coef);
  o->work = (float *)(a + o_work);
  o->rec = (float *)(a + o_rec);
  o->ws = (float *)(a + o_ws);
  o->sorted = (uint32_t *)(a + o_sorted);
  o->bins = (int32_t *)(a + o_bins);
  o->counts = (uint32_t *)(a + o_counts);
  o->sort_tmp = a + o_sort;
  o->table = a + o_table;
  o->partial = (double *)(a + o_part);
  // forward transform; min / max / max |x| from a select; the sorted keys
  rc = wv_dwt3d_forward(&o->g, in_d, o->coef, o->ws, nullptr, stream);
  wv_thresh_result *res_d = (wv_thresh_result *)(a + o_part);
  if (rc == WV_OK) rc = wv_threshold(o->coef, n, 0.5f, nullptr, o->enc_ws, res_d, stream);
  wv_thresh_result res;
  if (rc == WV_OK &&
      (hipMemcpyAsync(&res, res_d, sizeof res, hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)) {
    set_error("wv_opt_create: statistics");
    rc = WV_ERR_HIP;
  }
  if (rc == WV_OK) {
    hipLaunchKernelGGL(abs_keys_kernel, dim3(grid_of(n)), dim3(256), 0, s, o->coef, n,
                       o->sorted);
    const uint32_t nkeys = (uint32_t)n;
    uint32_t *count_d = (uint32_t *)o->sort_tmp;
    e = hipMemcpyAsync(count_d, &nkeys, 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = radix_sort_u32(o->sorted, (uint32_t *)o->work, (uint32_t *)(o->sort_tmp + 256), count_d,
                         n, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("wv_opt_create: sort: %s", hipGetErrorString(e));
      rc = WV_ERR_HIP;
    }
  }
  if (rc != WV_OK) {
    (void)hipFree(o->arena);
    (void)hipFree(o->enc_ws);
    (void)hipHostFree(o->sums_h);
    delete o;
    return rc;
  }
  o->min_val = res.min_val;
  o->max_val = res.max_val;
  o->max_abs = res.max_abs;
  *out = o;
  return WV_OK;
}

extern "C" int64_t wv_opt_count(const wv_opt *o) { return o ? o->g.count : WV_ERR_ARG; }

extern "C" int wv_opt_stats(const wv_opt *o, float *min_val, float *max_val, float *max_abs) {
  if (!o) return WV_ERR_ARG;
  if (min_val) *min_val = o->min_val;
  if (max_val) *max_val = o->max_val;
  if (max_abs) *max_abs = o->max_abs;
  return WV_OK;
}

// OptimizationData::getSorted (optimize_gpu.cu:15-27): one element of the
// sorted |x| (out of range: 0)
extern "C" float wv_opt_sorted(const wv_opt *o, int64_t index) {
  if (!o || index < 0 || index >= o->g.count) return 0.f;
  uint32_t k = 0;
  if (hipMemcpy(&k, o->sorted + index, 4, hipMemcpyDeviceToHost) != hipSuccess) return 0.f;
  return __builtin_bit_cast(float, k);
}

// OptimizationData::findSorted (optimize_gpu.cu:31-35): first index whose
// value is >= value (count when none is) -- a binary search with one
// element read per step
extern "C" int64_t wv_opt_find_sorted(const wv_opt *o, float value) {
  if (!o) return WV_ERR_ARG;
  int64_t lo = 0, len = o->g.count;
  while (len > 0) {
    const int64_t half = len >> 1;
    if (wv_opt_sorted(o, lo + half) < value) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

// testParameters (optimize_gpu.cu:52-202) for the uniform and log
// quantizers (max = max |x|, as compress does), the Lloyd code book trained
// on the sorted |x| above the threshold, or no quantization when
// bin_count <= 0
extern "C" int wv_opt_test(wv_opt *o, float threshold, int bin_count, int quant_alg,
                           wv_opt_result *r, void *stream) {
  if (!o || !r) {
    set_error("wv_opt_test: null argument");
    return WV_ERR_ARG;
  }
  std::memset(r, 0, sizeof *r);
  r->zero_bin = -1;
  const int64_t n = o->g.count;
  hipStream_t s = (hipStream_t)stream;
  const float *inv_in = o->work;
  if (bin_count <= 0) {
    hipLaunchKernelGGL(apply_threshold_kernel, dim3(grid_of(n)), dim3(256), 0, s, o->coef, n,
                       threshold, o->work);
    OHIP(hipGetLastError(), "wv_opt_test: threshold");
    OCHK(wv_dwt3d_inverse(&o->g, inv_in, o->rec, o->ws, stream));
    OCHK(error_sums(o, s));
    OHIP(hipStreamSynchronize(s), "wv_opt_test: sync");
  } else {
    if (quant_alg != WV_QUANT_UNIFORM && quant_alg != WV_QUANT_LOG &&
        quant_alg != WV_QUANT_LLOYD) {
      set_error("wv_opt_test: quantizer %d", quant_alg);
      return WV_ERR_UNSUPPORTED;
    }
    if (bin_count > WV_MAX_BINS) {
      set_error("wv_opt_test: %d bins", bin_count);
      return WV_ERR_ARG;
    }
    wv_quant_params p;
    if (quant_alg == WV_QUANT_LLOYD) {
      // computeLloydQuantization over the sorted |x| from findSorted(T)
      // (optimize_cpu.cc:90-96): all but the largest, no iteration cap
      const int64_t first = wv_opt_find_sorted(o, threshold);
      const int64_t need = wv_lloyd_ws_bytes(n, bin_count);
      if (need > o->lloyd_cap) {
        if (hipStreamSynchronize(s) != hipSuccess) {
          set_error("wv_opt_test: sync");
          return WV_ERR_HIP;
        }
        (void)hipFree(o->lloyd_ws);
        o->lloyd_ws = nullptr;
        o->lloyd_cap = 0;
        if (hipMalloc((void **)&o->lloyd_ws, (size_t)need) != hipSuccess) {
          set_error("wv_opt_test: Lloyd workspace");
          return WV_ERR_NOMEM;
        }
        o->lloyd_cap = need;
      }
      std::vector<float> bounds((size_t)bin_count - 1), values((size_t)bin_count);
      OCHK(lloyd_run_sorted((const float *)(o->sorted + first), n - first, threshold, bin_count,
                            -1, o->lloyd_ws, n, o->min_val, o->max_val, bounds.data(),
                            values.data(), &r->lloyd_iters, s));
      OCHK(wv_quant_params_init(&p, quant_alg, bin_count, threshold, 0.f, bounds.data(),
                                values.data()));
    } else {
      OCHK(wv_quant_params_init(&p, quant_alg, bin_count, threshold, o->max_abs, nullptr,
                                nullptr));
    }
    int rc = wv_quant_table_bytes(&p) <= o->table_cap ? WV_OK : WV_ERR_ARG;
    if (rc == WV_OK) rc = wv_quant_params_upload(&p, o->table, stream);
    if (rc == WV_OK) rc = wv_quantize_hist(o->coef, n, &p, o->bins, o->counts, stream);
    const int zb = wv_quant_host(&p, 0.f);
    r->zero_bin = zb;
    // the code book's frequencies: histogram, and with zero strings the
    // singles / dup strings of the zero bin (huffScanDupsOnly)
    const int64_t need = wv_encode_ws_bytes(n, bin_count);
    if (rc == WV_OK && need > o->enc_cap) {
      // the run statistics' workspace for this book size (grown once)
      if (hipStreamSynchronize(s) != hipSuccess) rc = WV_ERR_HIP;
      (void)hipFree(o->enc_ws);
      o->enc_ws = nullptr;
      o->enc_cap = 0;
      if (rc == WV_OK && hipMalloc((void **)&o->enc_ws, (size_t)need) != hipSuccess)
        rc = WV_ERR_NOMEM;
      else if (rc == WV_OK)
        o->enc_cap = need;
    }
    std::vector<uint32_t> counts((size_t)bin_count + 1);
    unsigned long long sd[2] = {0, 0};
    const int dup = o->compress_zeros ? zb : -1;
    if (rc == WV_OK &&
        hipMemcpyAsync(counts.data(), o->counts, counts.size() * 4, hipMemcpyDeviceToHost, s) !=
            hipSuccess)
      rc = WV_ERR_HIP;
    if (rc == WV_OK &&
        enc_count_runs(o->bins, n, dup, bin_count, o->enc_ws, sd, s) != hipSuccess)
      rc = WV_ERR_HIP;
    // the device goes on (dequantize, inverse, error sums) while the host
    // builds the code book for the size
    if (rc == WV_OK) rc = wv_dequantize(o->bins, n, &p, o->work, stream);
    if (rc == WV_OK) rc = wv_dwt3d_inverse(&o->g, inv_in, o->rec, o->ws, stream);
    if (rc == WV_OK) rc = error_sums(o, s);
    if (rc == WV_OK)
      r->size_bytes = huff_total_encoded_bytes(counts.data(), bin_count, dup, (int64_t)sd[0],
                                               (int64_t)sd[1]);
    if (hipStreamSynchronize(s) != hipSuccess && rc == WV_OK) rc = WV_ERR_HIP;
    wv_quant_params_free(&p);   // its tables were copied by the stream
    if (rc != WV_OK) {
      if (rc == WV_ERR_HIP) set_error("wv_opt_test: quantize / dequantize / inverse / error");
      return rc;
    }
  }
  const double *sums = o->sums_h;
  // ErrorAccumulator getters (test_compress_common.h:135-159)
  const int64_t cnt = (int64_t)o->g.orig[0] * o->g.orig[1] * o->g.orig[2];
  float maxv = 0.f;
  if (o->dtype == WV_DTYPE_UINT8) maxv = o->max_possible > 0 ? (float)o->max_possible : 255.f;
  else if (o->max_possible > 0) maxv = (float)o->max_possible;
  r->l1 = (float)sums[2];
  r->l2 = (float)std::sqrt(sums[3]);
  r->mse = cnt ? (float)(sums[3] / (double)cnt) : 0.f;
  r->psnr = (maxv == 0.f || r->mse == 0.f) ? 0.f : (float)(10 * log10(maxv * maxv / r->mse));
  r->rel = sums[1] == 0 ? 0.f : (float)(sums[3] / sums[1]);
  return WV_OK;
}

extern "C" int wv_opt_destroy(wv_opt *o) {
  if (!o) return WV_OK;
  if (o->arena) (void)hipFree(o->arena);
  if (o->enc_ws) (void)hipFree(o->enc_ws);
  if (o->lloyd_ws) (void)hipFree(o->lloyd_ws);
  if (o->sums_h) (void)hipHostFree(o->sums_h);
  delete o;
  return WV_OK;
}
