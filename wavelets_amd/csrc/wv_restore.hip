// wv_restore.hip -- the device half of decompression (decompressFile,
// test_compress_cpu.cc:187-236; the reference GPU decompress is a stub,
// test_compress_gpu.cu:438-501): dequantize (dequantize,
// test_compress_common.cc:567-584), inverse transform (waveletTransform
// inverse, test_compress_common.cc:492-564), then the original data type
// over the original size (translateCubeDataToOriginal,
// test_compress_common.cc:699-747, with ByteInputDataUnchanged::floatToByte
// and IntInputDataUnchanged::floatToInt, wavelet.h:149-218).
#include <chrono>
#include <cstring>
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

// padded x,y,z floats -> original x,y,z in the original type; block
// (bx, y, z) covers x in [bx * WV_BLOCK, ...) of row (y, z)
template <int DT>
__global__ void __launch_bounds__(WV_BLOCK)
restore_kernel(const float *__restrict__ src, int px, int py, int ox, int oy, int oz,
               void *__restrict__ dst) {
  const int x = blockIdx.x * WV_BLOCK + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x < ox) {
    const int64_t i = ((int64_t)z * oy + y) * ox + x;
    const float f = src[((int64_t)z * py + y) * px + x];
    if (DT == WV_DTYPE_UINT8) {
      // floatToByte: <= 0 -> 0, >= 255 -> 255, else truncation
      ((uint8_t *)dst)[i] = f <= 0.f ? (uint8_t)0 : (f >= 255.f ? (uint8_t)255 : (uint8_t)f);
    } else if (DT == WV_DTYPE_INT32) {
      ((int32_t *)dst)[i] = (int32_t)(f + 0.5f);   // floatToInt: (int)(f + 0.5f)
    } else {
      ((float *)dst)[i] = f;
    }
  }
}

}  // namespace wv

using namespace wv;

namespace {

inline int64_t al256(int64_t v) { return (v + 255) & ~(int64_t)255; }

struct DLayout {
  int64_t table, coef, rec, ws, bins, out, words, dws, total;
};

int out_dtype_of(const wv_cube_info *I) {
  return I->orig_dtype == WV_DTYPE_UINT8 || I->orig_dtype == WV_DTYPE_INT32 ? I->orig_dtype
                                                                            : WV_DTYPE_FLOAT32;
}

int geom_of(const wv_cube_info *I, wv_geom &g) {
  int rc = wv_geom_init(&g, I->orig, I->steps, I->wavelet, WV_DTYPE_FLOAT32);
  if (rc != WV_OK) return rc;
  if (g.out_size[0] != I->size[0] || g.out_size[1] != I->size[1] || g.out_size[2] != I->size[2]) {
    set_error("stored size %dx%dx%d is not the transform of %dx%dx%d with steps %d,%d,%d",
              I->size[0], I->size[1], I->size[2], I->orig[0], I->orig[1], I->orig[2],
              I->steps[0], I->steps[1], I->steps[2]);
    return WV_ERR_ARG;
  }
  return WV_OK;
}

DLayout dlayout(const wv_geom &g, const wv_cube_info *I) {
  DLayout L;
  int64_t o = 0;
  L.table = o; o += al256(((int64_t)I->bin_count * 3 + 16) * 4);
  L.coef = o; o += al256(g.count * 4);
  L.rec = o; o += al256(g.count * 4);
  L.ws = o; o += al256(g.ws_floats * 4);
  L.bins = o; o += al256(g.count * 4);
  const int64_t on = (int64_t)I->orig[0] * I->orig[1] * I->orig[2];
  L.out = o; o += al256(on * (out_dtype_of(I) == WV_DTYPE_UINT8 ? 1 : 4));
  L.words = o; o += al256(I->payload_bytes);
  L.dws = o; o += al256(wv_huff_decode_ws_bytes(I->payload_bytes / 4, I->n_decode));
  L.total = o;
  return L;
}

#define DCHK(call)                      \
  do {                                  \
    int _rc = (call);                   \
    if (_rc != WV_OK) return _rc;       \
  } while (0)
#define DHIP(expr, what)                                          \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) {                                       \
      set_error("%s: %s", what, hipGetErrorString(_e));           \
      return WV_ERR_HIP;                                          \
    }                                                             \
  } while (0)

}  // namespace

extern "C" int64_t wv_decompress_work_bytes(const wv_cube_info *info) {
  if (!info) return WV_ERR_ARG;
  wv_geom g;
  if (geom_of(info, g) != WV_OK) return WV_ERR_ARG;
  return dlayout(g, info).total;
}

extern "C" int wv_decompress_device(const wv_cube_info *info, const int32_t *bins_d,
                                    const float *boundaries, const float *values, void *work_d,
                                    int64_t work_bytes, void *out_d, void *stream) {
  if (!info || !bins_d || !work_d || !out_d) {
    set_error("wv_decompress_device: null argument");
    return WV_ERR_ARG;
  }
  if (info->quant_alg < WV_QUANT_UNIFORM || info->quant_alg > WV_QUANT_LLOYD ||
      info->bin_count < 2 || info->bin_count > WV_MAX_BINS) {
    set_error("wv_decompress_device: quantizer %d with %d bins", info->quant_alg,
              info->bin_count);
    return WV_ERR_ARG;
  }
  if (info->quant_alg == WV_QUANT_LLOYD &&
      (!boundaries || !values || info->n_boundaries != info->bin_count - 1 ||
       info->n_values != info->bin_count)) {
    set_error("wv_decompress_device: Lloyd codebook of %d / %d entries for %d bins",
              info->n_boundaries, info->n_values, info->bin_count);
    return WV_ERR_ARG;
  }
  wv_geom g;
  DCHK(geom_of(info, g));
  const DLayout L = dlayout(g, info);
  if (work_bytes < L.total) {
    set_error("wv_decompress_device: work buffer %lld < %lld", (long long)work_bytes,
              (long long)L.total);
    return WV_ERR_NOMEM;
  }
  hipStream_t s = (hipStream_t)stream;
  char *w = (char *)work_d;
  // Quantizer::dequant with the stored parameters (createQuantizer, quant.cc)
  wv_quant_params p;
  DCHK(wv_quant_params_init(&p, info->quant_alg, info->bin_count, info->threshold,
                            info->max_val, boundaries, values));
  int rc = wv_quant_params_upload(&p, w + L.table, s);
  if (rc == WV_OK) rc = wv_dequantize(bins_d, g.count, &p, (float *)(w + L.coef), s);
  // the host tables may go once the upload has run
  if (rc == WV_OK && hipStreamSynchronize(s) != hipSuccess) rc = WV_ERR_HIP;
  wv_quant_params_free(&p);
  DCHK(rc);
  const int dt = out_dtype_of(info);
  // float32 with no padding: the inverse transform writes the output itself
  const bool direct = dt == WV_DTYPE_FLOAT32 && g.padded[0] == info->orig[0] &&
                      g.padded[1] == info->orig[1] && g.padded[2] == info->orig[2];
  DCHK(wv_dwt3d_inverse(&g, (const float *)(w + L.coef), direct ? (float *)out_d : (float *)(w + L.rec),
                        (float *)(w + L.ws), s));
  if (direct) return WV_OK;
  const dim3 blocks((unsigned)((info->orig[0] + WV_BLOCK - 1) / WV_BLOCK), (unsigned)info->orig[1],
                    (unsigned)info->orig[2]);
  const float *rec = (const float *)(w + L.rec);
  if (dt == WV_DTYPE_UINT8)
    hipLaunchKernelGGL(restore_kernel<WV_DTYPE_UINT8>, blocks, dim3(WV_BLOCK), 0,
                       s, rec, g.padded[0], g.padded[1], info->orig[0], info->orig[1],
                       info->orig[2], out_d);
  else if (dt == WV_DTYPE_INT32)
    hipLaunchKernelGGL(restore_kernel<WV_DTYPE_INT32>, blocks, dim3(WV_BLOCK), 0,
                       s, rec, g.padded[0], g.padded[1], info->orig[0], info->orig[1],
                       info->orig[2], out_d);
  else
    hipLaunchKernelGGL(restore_kernel<WV_DTYPE_FLOAT32>, blocks, dim3(WV_BLOCK),
                       0, s, rec, g.padded[0], g.padded[1], info->orig[0], info->orig[1],
                       info->orig[2], out_d);
  DHIP(hipGetLastError(), "restore");
  return WV_OK;
}

extern "C" int64_t wv_decompress(const uint8_t *file, int64_t len, void *work_d,
                                 int64_t work_bytes, uint8_t *out_host, int64_t out_cap,
                                 wv_cube_info *info_out, double *ms_decode, void *stream) {
  wv_cube_info I;
  DCHK(wv_cube_parse(file, len, &I, nullptr, 0, nullptr, nullptr, 0));
  std::vector<int32_t> table((size_t)I.n_decode);
  std::vector<float> bounds((size_t)I.n_boundaries), values((size_t)I.n_values);
  const int cap = I.n_boundaries > I.n_values ? I.n_boundaries : I.n_values;
  DCHK(wv_cube_parse(file, len, &I, table.data(), I.n_decode, bounds.data(), values.data(), cap));
  if (info_out) *info_out = I;
  wv_geom g;
  DCHK(geom_of(&I, g));
  const DLayout L = dlayout(g, &I);
  if (work_bytes < L.total) {
    set_error("wv_decompress: work buffer %lld < %lld", (long long)work_bytes, (long long)L.total);
    return WV_ERR_NOMEM;
  }
  const int dt = out_dtype_of(&I);
  const int64_t out_bytes = (int64_t)I.orig[0] * I.orig[1] * I.orig[2] * (dt == WV_DTYPE_UINT8 ? 1 : 4);
  if (!out_host || out_cap < out_bytes) {
    set_error("wv_decompress: output buffer %lld < %lld", (long long)out_cap, (long long)out_bytes);
    return WV_ERR_NOMEM;
  }
  // readQuantData: the payload words, decoded on the device
  hipStream_t s = (hipStream_t)stream;
  char *w = (char *)work_d;
  const int64_t nwords = I.payload_bytes / 4;
  if (nwords)
    DHIP(hipMemcpyAsync(w + L.words, file + I.payload_offset, nwords * 4, hipMemcpyHostToDevice, s),
         "H2D payload");
  const auto t0 = std::chrono::steady_clock::now();
  DCHK(wv_huff_decode_device(table.data(), I.n_decode, I.dup_key, I.dup_value,
                             (const uint32_t *)(w + L.words), nwords, (int32_t *)(w + L.bins),
                             g.count, w + L.dws, nullptr, stream));
  if (ms_decode)
    *ms_decode = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  DCHK(wv_decompress_device(&I, (const int32_t *)(w + L.bins),
                            bounds.empty() ? nullptr : bounds.data(),
                            values.empty() ? nullptr : values.data(), work_d, work_bytes,
                            w + L.out, stream));
  DHIP(hipMemcpyAsync(out_host, w + L.out, out_bytes, hipMemcpyDeviceToHost, s), "D2H output");
  DHIP(hipStreamSynchronize(s), "sync");
  return out_bytes;
}
