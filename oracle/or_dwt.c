/*
 * or_dwt.c -- ORACLE ONLY (test infrastructure, never shipped in the product
 * path).  CPU restatement of the reference's steps/padding, 1-D Haar and
 * CDF 9/7 (convolution form) transforms, 3-D rotations and the 3-D drivers.
 */
#include <string.h>
#include <stdlib.h>
#include "wv_oracle.h"

/* Taps exactly as the float initialisers of dwt_cpu.cc:15-43. */
static const float LP[5] = {
  .85269867900940f, .377402855612650f, -.110624404418420f,
  -.02384946501938f, .037828455506995f };
static const float HP[4] = {
  -.788485616405660f, .418092273222210f, .040689417609558f,
  -.064538882628938f };
static const float SLP[4] = {
  .788485616405660f, .377402855612650f, -.040689417609558f,
  -.02384946501938f };
static const float SHP[5] = {
  -.85269867900940f, .418092273222210f, .110624404418420f,
  -.064538882628938f, -.037828455506995f };
/* dwt_cpu.cc:11 */
#define OR_INV_SQRT2 0.70710678118654752440f

/* dwt_cpu.cc:61-70 */
unsigned or_count_leading_zeros(unsigned x) {
  unsigned y, n = 32;
  y = x >> 16; if (y != 0) { n -= 16; x = y; }
  y = x >> 8;  if (y != 0) { n -= 8;  x = y; }
  y = x >> 4;  if (y != 0) { n -= 4;  x = y; }
  y = x >> 2;  if (y != 0) { n -= 2;  x = y; }
  y = x >> 1;  if (y != 0) return n - 2;
  return n - x;
}

/* dwt_cpu.cc:75-78 */
unsigned or_ceil_log2(unsigned x) {
  if (x == 0) return 0;
  return 32 - or_count_leading_zeros(x - 1);
}

/* dwt_cpu.cc:212-219 (powerOfTwo == false, the only form on the path) */
int or_padded_length(int length, int steps) {
  int mod = 1 << steps;
  return mod * ((length - 1) / mod + 1);
}

/* test_compress_common.cc:362-374 */
void or_set_steps(int steps[3], const int size[3]) {
  for (int a = 0; a < 3; a++) {
    int m = (int)or_ceil_log2((unsigned)size[a]);
    if (steps[a] < 0 || steps[a] > m) steps[a] = m;
  }
}

/* Conversion (wavelet.h:149-164,204-218: the "Unchanged" policies, so
   u8 -> (float)b and i32 -> (float)b) followed by REPEAT padding
   (test_compress_common.cc:386-423): rows, then rows of the last row,
   then copies of the last layer. */
void or_pad_convert(const void *in, int dtype, const int orig[3],
                    float *out, const int padded[3]) {
  const int X = padded[0], Y = padded[1], Z = padded[2];
  const int ox = orig[0], oy = orig[1], oz = orig[2];
  for (int z = 0; z < oz; z++)
    for (int y = 0; y < oy; y++) {
      float *dst = out + ((size_t)z * Y + y) * X;
      size_t src_off = ((size_t)z * oy + y) * ox;
      for (int x = 0; x < ox; x++) {
        float v;
        if (dtype == OR_DTYPE_UINT8)
          v = (float)((const uint8_t *)in)[src_off + x];
        else if (dtype == OR_DTYPE_INT32)
          v = (float)((const int32_t *)in)[src_off + x];
        else
          v = ((const float *)in)[src_off + x];
        dst[x] = v;
      }
    }
  if (X == ox && Y == oy && Z == oz) return;   /* :388 */
  for (int z = 0; z < oz; z++) {
    for (int y = 0; y < oy; y++) {
      float *row = out + ((size_t)z * Y + y) * X;
      float value = row[ox - 1];
      for (int x = ox; x < X; x++) row[x] = value;
    }
    const float *last = out + ((size_t)z * Y + (oy - 1)) * X;
    for (int y = oy; y < Y; y++)
      memcpy(out + ((size_t)z * Y + y) * X, last, sizeof(float) * X);
  }
  const float *layer = out + (size_t)(oz - 1) * Y * X;
  for (int z = oz; z < Z; z++)
    memcpy(out + (size_t)z * Y * X, layer, sizeof(float) * X * Y);
}

/* MirroredArray::getOffset, dwt_cpu.h:162-176 */
static inline int mirror(int o, int n) {
  if (o < 0) o = -o;
  while (o >= n) {
    o = n * 2 - o - 2;
    if (o < 0) o = -o;
  }
  return o;
}

/* cdf97_tmpl, dwt_cpu.cc:683-739: symmetric 9/7 convolution, lowpass at
   even positions into temp[0..n/2), highpass at odd positions into
   temp[n/2..n), then copy back.  The accumulation order (centre tap first,
   then k=1..4 each as F[k]*(a[i+k]+a[i-k])) is the reference's. */
void or_cdf97_fwd(int length, float *data, int steps, float *tmp) {
  int n = length;
  for (int s = 0; s < steps; s++, n /= 2) {
    int lo = 0, hi = n / 2, i = 0;
    while (i < n) {
      float sum = LP[0] * data[mirror(i, n)];
      for (int k = 1; k <= 4; k++)
        sum += LP[k] * (data[mirror(i + k, n)] + data[mirror(i - k, n)]);
      tmp[lo++] = sum;
      i++;
      sum = HP[0] * data[mirror(i, n)];
      for (int k = 1; k <= 3; k++)
        sum += HP[k] * (data[mirror(i + k, n)] + data[mirror(i - k, n)]);
      tmp[hi++] = sum;
      i++;
    }
    memcpy(data, tmp, sizeof(float) * n);
  }
}

/* cdf97_inverse_tmpl, dwt_cpu.cc:757-799: interleave 01234567->04152636
   into temp, then evens with synthesis taps S0..S3 and odds with H0..H4,
   mirroring on the interleaved array, accumulating into data[i]. */
void or_cdf97_inv(int length, float *data, int steps, float *tmp) {
  int n = length >> (steps - 1);
  for (int s = 0; s < steps; s++, n *= 2) {
    int half = n >> 1;
    for (int j = 0; j < half; j++) {
      tmp[j * 2] = data[j];
      tmp[j * 2 + 1] = data[half + j];
    }
    int i = 0;
    while (i < n) {
      data[i] = tmp[mirror(i, n)] * SLP[0];
      for (int j = 1; j <= 3; j++)
        data[i] += (tmp[mirror(i + j, n)] + tmp[mirror(i - j, n)]) * SLP[j];
      i++;
      data[i] = tmp[mirror(i, n)] * SHP[0];
      for (int j = 1; j <= 4; j++)
        data[i] += (tmp[mirror(i + j, n)] + tmp[mirror(i - j, n)]) * SHP[j];
      i++;
    }
  }
}

/* haar_internal, dwt_cpu.cc:139-170 (orthonormal: sums and differences
   scaled by 1/sqrt(2)). */
void or_haar_fwd(int length, float *data, int steps, float *tmp) {
  unsigned n = (unsigned)length;
  while (n > 1 && steps > 0) {
    int half = (int)(n >> 1);
    steps--;
    float *s = tmp, *d = tmp + half;
    for (int i = 0; i < half; i++) {
      d[i] = (data[2 * i] - data[2 * i + 1]) * OR_INV_SQRT2;
      s[i] = (data[2 * i] + data[2 * i + 1]) * OR_INV_SQRT2;
    }
    memcpy(data, tmp, sizeof(float) * n);
    n = (unsigned)half;
  }
}

/* haar_inv_internal, dwt_cpu.cc:174-203 */
void or_haar_inv(int length, float *data, int steps, float *tmp) {
  int n = length >> (steps - 1);
  while (n <= length) {
    int half = n >> 1;
    const float *s = data, *d = data + half;
    for (int i = 0; i < half; i++) {
      tmp[2 * i] = OR_INV_SQRT2 * (s[i] + d[i]);
      tmp[2 * i + 1] = OR_INV_SQRT2 * (s[i] - d[i]);
    }
    memcpy(data, tmp, sizeof(float) * n);
    n <<= 1;
  }
}

/* transpose3dFwd, wavelet.h:864-877: (x,y,z) -> (y,z,x); the data move
   is a 2-D transpose of an X by (Y*Z) matrix. */
void or_transpose3d_fwd(const float *src, float *dst, const int size[3]) {
  const size_t w = (size_t)size[0], h = (size_t)size[1] * size[2];
  for (size_t r = 0; r < h; r++)
    for (size_t c = 0; c < w; c++)
      dst[c * h + r] = src[r * w + c];
}

/* transpose3dBack, wavelet.h:899-912: (x,y,z) -> (z,x,y); a 2-D transpose
   of an (X*Y) by Z matrix. */
void or_transpose3d_back(const float *src, float *dst, const int size[3]) {
  const size_t w = (size_t)size[0] * size[1], h = (size_t)size[2];
  for (size_t r = 0; r < h; r++)
    for (size_t c = 0; c < w; c++)
      dst[c * h + r] = src[r * w + c];
}

/* *_3d_one_axis: every row along x (CubeNum::visitRows, wavelet.h:604-611),
   only when steps > 0 (dwt_cpu.cc:614-620,902-907). */
static void one_axis(float *data, const int size[3], int wavelet, int steps,
                     int inverse, float *rowtmp) {
  if (steps <= 0) return;
  const size_t rows = (size_t)size[1] * size[2];
  const int len = size[0];
  for (size_t r = 0; r < rows; r++) {
    float *row = data + r * len;
    if (wavelet == OR_WAVELET_CDF97) {
      if (inverse) or_cdf97_inv(len, row, steps, rowtmp);
      else or_cdf97_fwd(len, row, steps, rowtmp);
    } else {
      if (inverse) or_haar_inv(len, row, steps, rowtmp);
      else or_haar_fwd(len, row, steps, rowtmp);
    }
  }
}

static void rotate_fwd(int s[3]) { int t = s[0]; s[0] = s[1]; s[1] = s[2]; s[2] = t; }
static void rotate_back(int s[3]) { int t = s[2]; s[2] = s[1]; s[1] = s[0]; s[0] = t; }

/* cdf97_3d (dwt_cpu.cc:943-985) and haar_3d (dwt_cpu.cc:624-648), guarded
   by the all-zero-steps no-op of waveletTransform
   (test_compress_common.cc:496). */
void or_dwt3d(float *data, float *tmp, int size[3], int wavelet,
              const int steps[3], int inverse) {
  if (steps[0] == 0 && steps[1] == 0 && steps[2] == 0) return;
  const size_t count = (size_t)size[0] * size[1] * size[2];
  int maxlen = size[0];
  if (size[1] > maxlen) maxlen = size[1];
  if (size[2] > maxlen) maxlen = size[2];
  float *rowtmp = (float *)malloc(sizeof(float) * maxlen);
  if (!inverse) {
    one_axis(data, size, wavelet, steps[0], 0, rowtmp);
    or_transpose3d_fwd(data, tmp, size); rotate_fwd(size);
    memcpy(data, tmp, sizeof(float) * count);
    one_axis(data, size, wavelet, steps[1], 0, rowtmp);
    or_transpose3d_fwd(data, tmp, size); rotate_fwd(size);
    memcpy(data, tmp, sizeof(float) * count);
    one_axis(data, size, wavelet, steps[2], 0, rowtmp);
  } else {
    one_axis(data, size, wavelet, steps[2], 1, rowtmp);
    or_transpose3d_back(data, tmp, size); rotate_back(size);
    memcpy(data, tmp, sizeof(float) * count);
    one_axis(data, size, wavelet, steps[1], 1, rowtmp);
    or_transpose3d_back(data, tmp, size); rotate_back(size);
    memcpy(data, tmp, sizeof(float) * count);
    one_axis(data, size, wavelet, steps[0], 1, rowtmp);
  }
  free(rowtmp);
}
