/*
 * or_opt.c -- ORACLE ONLY (test infrastructure).  One test point of the
 * reference's parameter optimisation (testParameters,
 * optimize_gpu.cu:52-202) restated on the CPU from the oracle's pieces:
 * transform, threshold or quantize / histogram / Huffman size / dequantize,
 * inverse transform, and the ErrorAccumulator comparison over the
 * original extent (computeErrorRatesAfterDequant,
 * test_compress_common.cc:616-694).  The reference GPU path sums the error
 * terms in float with atomics (computeErrorKernel); the sums here are the
 * CPU path's doubles.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "wv_oracle.h"

static float *forward(const void *in, int dtype, const int orig[3], const int steps_in[3],
                      int wavelet, int steps[3], int size[3], size_t *n_out) {
  steps[0] = steps_in[0]; steps[1] = steps_in[1]; steps[2] = steps_in[2];
  or_set_steps(steps, orig);
  for (int a = 0; a < 3; a++) size[a] = or_padded_length(orig[a], steps[a]);
  size_t n = (size_t)size[0] * size[1] * size[2];
  float *data = (float *)malloc(sizeof(float) * n);
  float *tmp = (float *)malloc(sizeof(float) * n);
  or_pad_convert(in, dtype, orig, data, size);
  or_dwt3d(data, tmp, size, wavelet, steps, 0);
  free(tmp);
  *n_out = n;
  return data;
}

static int cmp_float(const void *a, const void *b) {
  float x = *(const float *)a, y = *(const float *)b;
  return x < y ? -1 : x > y ? 1 : 0;
}

int or_opt_sorted(const void *in, int dtype, const int orig[3], const int steps_in[3],
                  int wavelet, float *sorted_out, int64_t cap) {
  int steps[3], size[3];
  size_t n;
  float *data = forward(in, dtype, orig, steps_in, wavelet, steps, size, &n);
  if ((int64_t)n > cap) { free(data); return -1; }
  for (size_t i = 0; i < n; i++) sorted_out[i] = fabsf(data[i]);
  qsort(sorted_out, n, sizeof(float), cmp_float);
  free(data);
  return (int)n;
}

int64_t or_opt_test(const void *in, int dtype, const int orig[3], const int steps_in[3],
                    int wavelet, int max_possible, int compress_zeros, float threshold,
                    int bin_count, int quant_alg, or_error_acc *acc, int *lloyd_iters) {
  int steps[3], size[3];
  size_t n;
  float *data = forward(in, dtype, orig, steps_in, wavelet, steps, size, &n);
  int64_t bytes = 0;
  if (bin_count <= 0) {
    /* applyThresholdKernel, optimize_gpu.cu:38-47 */
    for (size_t i = 0; i < n; i++)
      if (fabsf(data[i]) <= threshold) data[i] = 0;
  } else {
    float mabs = 0, mn = 3.40282347e+38f, mx = 1.17549435e-38f;
    for (size_t i = 0; i < n; i++) {
      if (fabsf(data[i]) > mabs) mabs = fabsf(data[i]);
      if (data[i] < mn) mn = data[i];
      if (data[i] > mx) mx = data[i];
    }
    or_quantizer q;
    float *bounds = NULL, *values = NULL;
    if (quant_alg == OR_QUANT_LLOYD) {
      /* optimize_cpu.cc:90-96: the sorted |x| from findSorted(threshold)
         (std::lower_bound, optimize_cpu.cc:16-18) to computeLloydQuantization,
         with the transformed data's min and max (thresh_cpu.cc:67-68 starts
         them at FLT_MAX / FLT_MIN) */
      float *sorted = (float *)malloc(sizeof(float) * n);
      for (size_t i = 0; i < n; i++) sorted[i] = fabsf(data[i]);
      qsort(sorted, n, sizeof(float), cmp_float);
      size_t lo = 0, len = n;
      while (len > 0) {
        size_t half = len >> 1;
        if (sorted[lo + half] < threshold) { lo += half + 1; len -= half + 1; }
        else len = half;
      }
      bounds = (float *)malloc(sizeof(float) * bin_count);
      values = (float *)malloc(sizeof(float) * bin_count);
      int iters = 0;
      or_compute_lloyd_quantization(sorted + lo, (int)(n - lo), bin_count, mn, mx, threshold,
                                    bounds, values, -1, &iters);
      if (lloyd_iters) *lloyd_iters = iters;
      free(sorted);
      or_quant_init(&q, quant_alg, bin_count, threshold, 0, bounds, values);
    } else {
      or_quant_init(&q, quant_alg, bin_count, threshold, mabs, NULL, NULL);
    }
    int zero_bin = or_quant(&q, 0);
    int32_t *bins = (int32_t *)malloc(sizeof(int32_t) * n);
    or_quantize(&q, data, bins, (int)n);
    or_huffman *h = or_huff_create_scan(bins, (int)n, bin_count, compress_zeros ? zero_bin : -1);
    bytes = or_huff_total_encoded_bytes(h);
    or_huff_free(h);
    or_dequantize(&q, bins, data, (int)n);
    free(bins);
    free(bounds);
    free(values);
  }
  float *tmp = (float *)malloc(sizeof(float) * n);
  or_dwt3d(data, tmp, size, wavelet, steps, 1);
  free(tmp);
  or_error_compare(acc, in, dtype, orig, data, size, max_possible);
  free(data);
  return bytes;
}
