/*
 * or_quant.c -- ORACLE ONLY (test infrastructure).  CPU restatement of the
 * reference threshold selection, the uniform / log / codebook quantizers and
 * the CPU Lloyd-Max training.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "wv_oracle.h"

/* LSD radix sort of the bit patterns of |x|.  For non-negative floats the
   order of the uint32 patterns is the order of the values, so this yields
   the same (unique) ascending array as std::sort at thresh_cpu.cc:94. */
void or_sort_abs(const float *data, int count, float *sorted_abs) {
  uint32_t *a = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)count);
  uint32_t *b = (uint32_t *)sorted_abs;
  for (int i = 0; i < count; i++) {
    uint32_t u;
    memcpy(&u, &data[i], 4);
    a[i] = u & 0x7fffffffu;
  }
  size_t cnt[2048];
  for (int pass = 0; pass < 3; pass++) {
    int shift = pass * 11;
    memset(cnt, 0, sizeof cnt);
    for (int i = 0; i < count; i++) cnt[(a[i] >> shift) & 2047]++;
    size_t s = 0;
    for (int d = 0; d < 2048; d++) { size_t c = cnt[d]; cnt[d] = s; s += c; }
    for (int i = 0; i < count; i++) b[cnt[(a[i] >> shift) & 2047]++] = a[i];
    uint32_t *t = a; a = b; b = t;
  }
  /* after 3 passes the result is in 'a' */
  if ((void *)a != (void *)sorted_abs) {
    memcpy(sorted_abs, a, sizeof(uint32_t) * (size_t)count);
    free(a);
  } else {
    free(b);
  }
}

/* thresh_cpu(CubeFloat*) thresh_cpu.cc:80-109 with Traverse :60-77.
   min starts at FLT_MAX and max at FLT_MIN (sic, :67-68).  The index is
   computed in float: (int)(compRatio * count).  A zero threshold becomes
   1e-16 (:98).  compRatio >= 1 reads past the array in the reference; here
   it follows the GPU driver (test_compress_gpu.cu:299-301): T = max |x|. */
float or_thresh(const float *data, int count, float frac, int *nonzero_count,
                float *max_val, float *min_val, float *max_abs,
                int *thresh_idx, float *sorted_abs) {
  float mn = FLT_MAX, mx = FLT_MIN;
  for (int i = 0; i < count; i++) {
    float x = data[i];
    if (x < mn) mn = x;
    if (x > mx) mx = x;
  }
  *min_val = mn;
  *max_val = mx;
  float *s = sorted_abs ? sorted_abs
                        : (float *)malloc(sizeof(float) * (size_t)count);
  or_sort_abs(data, count, s);
  *max_abs = s[count - 1];
  int idx = (int)(frac * count);
  float t;
  if (idx >= count) {
    idx = count;
    t = s[count - 1];
    *nonzero_count = 0;
  } else {
    t = s[idx];
    if (t == 0) t = 1e-16f;
    *nonzero_count = count - idx - 1;
  }
  *thresh_idx = idx;
  if (!sorted_abs) free(s);
  return t;
}

/* QuantUniform::init quant.h:196-242, QuantLog::init quant.h:337-372,
   QuantCodebook::init quant.h:487-492.  Integer bin counts convert to float
   in the divisions exactly as in the reference expressions. */
void or_quant_init(or_quantizer *q, int alg, int bin_count, float threshold,
                   float max_val, const float *boundaries,
                   const float *values) {
  memset(q, 0, sizeof *q);
  q->alg = alg;
  q->bin_count = bin_count;
  q->threshold = threshold;
  q->max_val = max_val;
  if (alg == OR_QUANT_UNIFORM) {
    int neg_bins = (bin_count - 1) / 2;
    q->neg_scale = (neg_bins - 1) / (max_val - threshold);
    int pos_bins = bin_count / 2;
    q->pos_scale = (pos_bins - 1) / (max_val - threshold);
    q->neg_inv_scale = q->neg_scale == 0 ? 0 : 1 / q->neg_scale;
    q->pos_inv_scale = q->pos_scale == 0 ? 0 : 1 / q->pos_scale;
    q->neg_offset = 1 - threshold * q->neg_scale;
    q->pos_offset = 1 - threshold * q->pos_scale;
  } else if (alg == OR_QUANT_LOG) {
    int neg_bins = (bin_count - 1) / 2;
    q->neg_scale = (neg_bins - 1) / logf(max_val / threshold);
    q->neg_inv_scale = 1 / q->neg_scale;
    q->neg_offset = 1 - q->neg_scale * logf(threshold);
    int pos_bins = bin_count / 2;
    q->pos_scale = (pos_bins - 1) / logf(max_val / threshold);
    q->pos_inv_scale = 1 / q->pos_scale;
    q->pos_offset = 1 - q->pos_scale * logf(threshold);
  } else {
    q->n_boundaries = bin_count - 1;
    q->boundaries = boundaries;
    q->values = values;
  }
}

/* QuantUniform::quant quant.h:244-276, QuantLog::quant quant.h:374-406,
   QuantCodebook::quant quant.h:515-526 (host variant, with the
   lastBoundary test, then std::lower_bound). */
int or_quant(const or_quantizer *q, float x) {
  const int zero_bin = (q->bin_count - 1) >> 1;
  if (q->alg == OR_QUANT_UNIFORM) {
    if (x < 0) {
      if (x >= -q->threshold) return zero_bin;
      x = x * q->neg_scale - q->neg_offset;
      int bin = (int)x + zero_bin;
      return bin < 0 ? 0 : bin;
    } else {
      if (x <= q->threshold) return zero_bin;
      x = x * q->pos_scale + q->pos_offset + zero_bin;
      if (x > q->bin_count - 1) return q->bin_count - 1;
      return (int)x;
    }
  } else if (q->alg == OR_QUANT_LOG) {
    if (x < 0) {
      x = -x;
      if (x <= q->threshold) return zero_bin;
      x = zero_bin - q->neg_scale * logf(x) - q->neg_offset;
      if (x < 0) return 0;
      return (int)(x + 0.5f);
    } else {
      if (x <= q->threshold) return zero_bin;
      x = zero_bin + q->pos_scale * logf(x) + q->pos_offset;
      if (x >= q->bin_count) return q->bin_count - 1;
      return (int)(x + 0.5f);
    }
  } else {
    const int n = q->n_boundaries;
    if (x >= q->boundaries[n - 1]) return n;
    int lo = 0, len = n;   /* lower_bound: first boundary >= x */
    while (len > 0) {
      int half = len / 2;
      if (q->boundaries[lo + half] < x) { lo += half + 1; len -= half + 1; }
      else len = half;
    }
    return lo;
  }
}

/* dequant: quant.h:278-291, 409-437, 552-556 */
float or_dequant(const or_quantizer *q, int x) {
  const int zero_bin = (q->bin_count - 1) >> 1;
  if (q->alg == OR_QUANT_UNIFORM) {
    if (x < zero_bin)
      return (x - 0.5f + q->neg_offset - zero_bin) * q->neg_inv_scale;
    else if (x > zero_bin)
      return (x + 0.5f - q->pos_offset - zero_bin) * q->pos_inv_scale;
    return 0;
  } else if (q->alg == OR_QUANT_LOG) {
    float sign, inv_scale, offset;
    if (x < zero_bin) { sign = -1; offset = q->neg_offset; inv_scale = q->neg_inv_scale; }
    else { sign = 1; offset = q->pos_offset; inv_scale = q->pos_inv_scale; }
    if (x == zero_bin) return 0;
    return sign * expf((sign * (x - zero_bin) - offset) * inv_scale);
  } else {
    return q->values[x];
  }
}

void or_quantize(const or_quantizer *q, const float *in, int32_t *out,
                 int count) {
  for (int i = 0; i < count; i++) out[i] = or_quant(q, in[i]);
}

void or_dequantize(const or_quantizer *q, const int32_t *in, float *out,
                   int count) {
  for (int i = 0; i < count; i++) out[i] = or_dequant(q, in[i]);
}

/* lloyd(), lloyds.cpp:30-141, statement for statement: sequential float
   sums, fabsf nearest-codebook with strict '<' (first minimum wins),
   signed distortion, mean-of-interval partition update, empty-cell
   fallbacks.  max_iters >= 0 adds an iteration cap that the reference
   lacks (the loop also stops once *iterations reaches it). */
void or_lloyd(const float *points, unsigned psize, float *codebook,
              unsigned csize, float *partition, float *dist_out,
              float *reldist_out, unsigned *groups, float stop_criteria,
              int *iterations, int max_iters) {
  float *sum = (float *)calloc(csize, sizeof(float));
  unsigned *count = (unsigned *)calloc(csize, sizeof(unsigned));
  for (unsigned i = 0; i < csize - 1; i++)
    partition[i] = (codebook[i] + codebook[i + 1]) / 2;
  float pmax = points[0], pmin = points[0];
  float dist = 0, reldist;
  int iters = 0;
  for (unsigned i = 0; i < psize; i++) {
    float p = points[i];
    pmax = (p > pmax) ? p : pmax;
    pmin = (p < pmin) ? p : pmin;
    float d = fabsf(p - codebook[0]);
    float mn = d;
    unsigned idx = 0;
    for (unsigned j = 1; j < csize; j++) {
      d = fabsf(p - codebook[j]);
      if (d < mn) { idx = j; mn = d; }
    }
    if (groups) groups[i] = idx;
    sum[idx] += p;
    count[idx]++;
    dist += codebook[idx] - p;
  }
  dist /= psize;
  reldist = fabsf(dist);
  while (reldist > stop_criteria && (max_iters < 0 || iters < max_iters)) {
    ++iters;
    for (unsigned i = 0; i < csize; i++) {
      if (count[i] != 0) {
        codebook[i] = sum[i] / count[i];
        count[i] = 0;
      } else if (i == 0) {
        codebook[i] = (partition[0] + pmin) / 2;
      } else if (i == csize - 1) {
        codebook[i] = (partition[i - 1] + pmax) / 2;
      } else {
        codebook[i] = (partition[i - 1] + partition[i]) / 2;
      }
      sum[i] = 0;
    }
    for (unsigned i = 0; i < psize; i++) {
      for (unsigned j = 0; j < csize - 1; j++) {
        float p = points[i];
        if ((p >= codebook[j]) && (p < codebook[j + 1])) {
          count[j]++;
          sum[j] += p;
          break;
        }
      }
    }
    for (unsigned i = 0; i < csize - 1; i++) {
      if (count[i] != 0) {
        partition[i] = sum[i] / count[i];
        count[i] = 0;
      } else {
        partition[i] = (codebook[i] + codebook[i + 1]) / 2;
      }
      sum[i] = 0;
    }
    reldist = dist;
    dist = 0;
    for (unsigned i = 0; i < psize; i++) {
      float p = points[i];
      float d = fabsf(p - codebook[0]);
      float mn = d;
      unsigned idx = 0;
      for (unsigned j = 1; j < csize; j++) {
        d = fabsf(p - codebook[j]);
        if (d < mn) { idx = j; mn = d; }
      }
      if (groups) groups[i] = idx;
      sum[idx] += p;
      count[idx]++;
      dist += codebook[idx] - p;
    }
    dist /= psize;
    reldist = fabsf(reldist - dist);
  }
  *dist_out = dist;
  *reldist_out = reldist;
  if (iterations) *iterations = iters;
  free(sum);
  free(count);
}

/* initialLloydCodebook, test_compress_common.cc:749-779 */
void or_initial_lloyd_codebook(float *codebook, int cs, float min_abs,
                               float max_abs) {
  float log_scale = logf(max_abs / min_abs) / cs;
  for (int i = 0; i < cs; i++) codebook[i] = min_abs * expf(i * log_scale);
}

/* setBinsFromCodebook, test_compress_common.cc:784-846.  Returns the number
   of bin values written (== bin_count when consistent). */
int or_set_bins_from_codebook(float *v, float *b, int bin_count,
                              const float *c, int cs, float threshold,
                              float min_val, float max_val) {
  int nv = 0, nb = 0;
  if ((bin_count & 1) == 0 && fabsf(min_val) > max_val) {
    v[nv++] = min_val;
    b[nb++] = (min_val + -c[cs - 1]) / 2;
  }
  v[nv++] = -c[cs - 1];
  for (int i = cs - 2; i >= 0; i--) {
    b[nb++] = (-c[i] + -c[i + 1]) / 2;
    v[nv++] = -c[i];
  }
  b[nb++] = -threshold;
  v[nv++] = 0;
  b[nb++] = threshold;
  for (int i = 0; i < cs - 1; i++) {
    v[nv++] = c[i];
    b[nb++] = (c[i] + c[i + 1]) / 2;
  }
  v[nv++] = c[cs - 1];
  if ((bin_count & 1) == 0 && max_val >= fabsf(min_val)) {
    v[nv++] = max_val;
    b[nb++] = (max_val + c[cs - 1]) / 2;
  }
  return (nb + 1 == nv) ? nv : -1;
}

/* computeLloydQuantization, test_compress_cpu.cc:404-459: points are the
   sorted tail without its largest entry, the initial codebook is
   log-spaced between the threshold and the second-largest |x|, stop
   criterion DEFAULT_LLOYD_STOP_CRITERIA = 10e-7 (lloyds.h:1) as a float. */
int or_compute_lloyd_quantization(const float *nonzero, int count,
                                  int bin_count, float min_val, float max_val,
                                  float threshold, float *bin_boundaries,
                                  float *bin_values, int max_iters,
                                  int *iters_out) {
  int cs = (bin_count - 1) / 2;
  float max_abs = nonzero[count - 2];
  float *codebook = (float *)malloc(sizeof(float) * cs);
  float *part = (float *)malloc(sizeof(float) * (cs > 1 ? cs - 1 : 1));
  or_initial_lloyd_codebook(codebook, cs, threshold, max_abs);
  float dist, reldist;
  int iters = 0;
  or_lloyd(nonzero, (unsigned)(count - 1), codebook, (unsigned)cs, part,
           &dist, &reldist, NULL, (float)10e-7, &iters, max_iters);
  if (iters_out) *iters_out = iters;
  int nv = or_set_bins_from_codebook(bin_values, bin_boundaries, bin_count,
                                     codebook, cs, threshold, min_val,
                                     max_val);
  free(codebook);
  free(part);
  return nv;
}
