/*
 * or_pipeline.c -- ORACLE ONLY (test infrastructure and the reported CPU
 * baseline).  The whole single-threaded CPU compress path of the reference
 * (test_compress_cpu.cc:68-184) built from the restated pieces, plus the
 * error accumulator used for round-trip PSNR.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "wv_oracle.h"

double or_now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ErrorAccumulator::add, test_compress_common.h:123-130 */
static void acc_add(or_error_acc *a, float x, float y) {
  a->sum_value += x;
  a->sum_value_sq += x * x;
  float diff = fabsf(x - y);
  a->sum_diff += diff;
  a->sum_diff_sq += diff * diff;
  a->count++;
}

/* computeErrorRatesAfterDequant, test_compress_common.cc:616-694, after
   the inverse transform: compare over the original extent with the
   per-type restore rules (wavelet.h:155-163,210-212; FLOAT32 compares
   against (int)restored, :679-686). */
void or_error_compare(or_error_acc *acc, const void *orig, int dtype,
                      const int osz[3], const float *restored,
                      const int padded[3], int max_possible) {
  memset(acc, 0, sizeof *acc);
  if (dtype == OR_DTYPE_UINT8) acc->max_value = max_possible > 0 ? (float)max_possible : 255.f;
  else if (max_possible > 0) acc->max_value = (float)max_possible;
  else return;
  for (int z = 0; z < osz[2]; z++)
    for (int y = 0; y < osz[1]; y++) {
      const float *r = restored + ((size_t)z * padded[1] + y) * padded[0];
      size_t o = ((size_t)z * osz[1] + y) * osz[0];
      for (int x = 0; x < osz[0]; x++) {
        if (dtype == OR_DTYPE_UINT8) {
          float f = r[x];
          unsigned char px = f <= 0 ? 0 : f >= 255 ? 255 : (unsigned char)f;
          acc_add(acc, ((const uint8_t *)orig)[o + x], px);
        } else if (dtype == OR_DTYPE_INT32) {
          acc_add(acc, (float)((const int32_t *)orig)[o + x], (float)(int)(r[x] + 0.5f));
        } else {
          acc_add(acc, ((const float *)orig)[o + x], (float)(int)r[x]);
        }
      }
    }
}

/* getMeanSquaredError / getPeakSignalToNoiseRatio,
   test_compress_common.h:143-153 */
float or_error_mse(const or_error_acc *a) {
  if (a->count == 0) return 0;
  return (float)(a->sum_diff_sq / a->count);
}
float or_error_psnr(const or_error_acc *a) {
  if (a->max_value == 0) return 0;
  float mse = or_error_mse(a);
  if (mse == 0) return 0;
  return (float)(10 * log10(a->max_value * a->max_value / mse));
}

int or_compress(const void *in, int dtype, const int orig[3], int max_possible,
                const or_options *opt, or_result *res, float *coef_out,
                int32_t *bins_out, uint8_t *file, int64_t file_cap) {
  memset(res, 0, sizeof *res);
  double t0 = or_now();
  int steps[3] = { opt->steps[0], opt->steps[1], opt->steps[2] };
  or_set_steps(steps, orig);
  int size[3];
  for (int a = 0; a < 3; a++) size[a] = or_padded_length(orig[a], steps[a]);
  const size_t n = (size_t)size[0] * size[1] * size[2];
  float *data = (float *)malloc(sizeof(float) * n);
  float *tmp = (float *)malloc(sizeof(float) * n);
  or_pad_convert(in, dtype, orig, data, size);

  double t1 = or_now();
  or_dwt3d(data, tmp, size, opt->wavelet, steps, 0);
  double t2 = or_now();
  if (coef_out) memcpy(coef_out, data, sizeof(float) * n);

  float *sorted = tmp;   /* reuse */
  float tval, mn, mx, mabs;
  int nz, idx;
  tval = or_thresh(data, (int)n, opt->thresh_frac, &nz, &mx, &mn, &mabs, &idx,
                   sorted);
  double t3 = or_now();

  or_quantizer q;
  float *bounds = NULL, *values = NULL;
  int lloyd_iters = 0;
  if (opt->quant_alg == OR_QUANT_LLOYD) {
    bounds = (float *)malloc(sizeof(float) * (size_t)opt->bin_count);
    values = (float *)malloc(sizeof(float) * (size_t)opt->bin_count);
    or_compute_lloyd_quantization(sorted + idx + 1, nz, opt->bin_count, mn, mx,
                                  tval, bounds, values, opt->lloyd_max_iters,
                                  &lloyd_iters);
    or_quant_init(&q, OR_QUANT_LLOYD, opt->bin_count, tval, 0, bounds, values);
  } else {
    or_quant_init(&q, opt->quant_alg, opt->bin_count, tval, mabs, NULL, NULL);
  }
  int zero_bin = or_quant(&q, 0);
  int32_t *bins = bins_out ? bins_out : (int32_t *)malloc(sizeof(int32_t) * n);
  or_quantize(&q, data, bins, (int)n);
  double t4 = or_now();

  int huff_zero = opt->compress_zeros ? zero_bin : -1;
  or_huffman *h = or_huff_create_scan(bins, (int)n, opt->bin_count, huff_zero);
  int64_t cap_words = (int64_t)n * 2 + 1024;
  uint32_t *words = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cap_words);
  int64_t bits = 0;
  int64_t nw = or_huff_encode(h, bins, (int)n, words, cap_words, &bits);
  double t5 = or_now();

  int ndec = or_huff_decode_table(h, NULL, 0);
  int32_t *dec = (int32_t *)malloc(sizeof(int32_t) * (size_t)ndec);
  or_huff_decode_table(h, dec, ndec);
  or_cube_meta m;
  memset(&m, 0, sizeof m);
  m.width = size[0]; m.height = size[1]; m.depth = size[2];
  m.orig_width = orig[0]; m.orig_height = orig[1]; m.orig_depth = orig[2];
  m.orig_dtype = dtype;
  memcpy(m.steps, steps, sizeof steps);
  m.wavelet = opt->wavelet;
  m.threshold_fraction = opt->thresh_frac;
  m.threshold_value = tval;
  m.bin_count = opt->bin_count;
  m.quant_alg = opt->quant_alg;
  m.max_value = mabs;
  if (opt->quant_alg == OR_QUANT_LLOYD) {
    m.n_boundaries = opt->bin_count - 1; m.boundaries = bounds;
    m.n_values = opt->bin_count; m.values = values;
  }
  m.n_decode = ndec; m.decode_table = dec;
  m.dup_key = huff_zero >= 0 ? opt->bin_count : -1;
  m.dup_value = huff_zero >= 0 ? huff_zero : -1;
  m.byte_count = nw * 4;
  m.max_possible = 0;
  (void)max_possible;
  int64_t fbytes = 0;
  if (file) fbytes = or_cube_write(&m, words, opt->with_footer, file, file_cap);
  double t6 = or_now();

  memcpy(res->size, size, sizeof size);
  memcpy(res->steps, steps, sizeof steps);
  res->threshold = tval; res->min_val = mn; res->max_val = mx; res->max_abs = mabs;
  res->thresh_idx = idx; res->nonzero_count = nz; res->zero_bin = zero_bin;
  res->lloyd_iters = lloyd_iters;
  res->compressed_bytes = nw * 4;
  res->file_bytes = fbytes;
  res->t_dwt = t2 - t1; res->t_thresh = t3 - t2; res->t_quant = t4 - t3;
  res->t_huff = t5 - t4; res->t_total = t6 - t0;

  or_huff_free(h);
  free(words); free(dec); free(bounds); free(values);
  if (!bins_out) free(bins);
  free(data); free(tmp);
  return 0;
}
