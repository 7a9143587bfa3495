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
                    int bin_count, int quant_alg, or_error_acc *acc) {
  int steps[3], size[3];
  size_t n;
  float *data = forward(in, dtype, orig, steps_in, wavelet, steps, size, &n);
  int64_t bytes = 0;
  if (bin_count <= 0) {
    /* applyThresholdKernel, optimize_gpu.cu:38-47 */
    for (size_t i = 0; i < n; i++)
      if (fabsf(data[i]) <= threshold) data[i] = 0;
  } else {
    float mabs = 0;
    for (size_t i = 0; i < n; i++) if (fabsf(data[i]) > mabs) mabs = fabsf(data[i]);
    or_quantizer q;
    or_quant_init(&q, quant_alg, bin_count, threshold, mabs, NULL, NULL);
    int zero_bin = or_quant(&q, 0);
    int32_t *bins = (int32_t *)malloc(sizeof(int32_t) * n);
    or_quantize(&q, data, bins, (int)n);
    or_huffman *h = or_huff_create_scan(bins, (int)n, bin_count, compress_zeros ? zero_bin : -1);
    bytes = or_huff_total_encoded_bytes(h);
    or_huff_free(h);
    or_dequantize(&q, bins, data, (int)n);
    free(bins);
  }
  float *tmp = (float *)malloc(sizeof(float) * n);
  or_dwt3d(data, tmp, size, wavelet, steps, 1);
  free(tmp);
  or_error_compare(acc, in, dtype, orig, data, size, max_possible);
  free(data);
  return bytes;
}
