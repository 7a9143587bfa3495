/*
 * or_dwt64.c -- ORACLE ONLY (test infrastructure, never shipped in the
 * product path).  An FP64 *lifting* CDF 9/7 (and FP64 Haar) over a 3-D
 * volume: the yardstick the float32 transforms' error is stated against.
 *
 * It does not follow the reference's convolution (dwt_cpu.cc:683-739); it
 * is the same filter bank computed another way, in double, from the
 * published lifting constants (Daubechies & Sweldens 1998):
 *   d += a (s_i + s_{i+1});  s += b (d_{i-1} + d_i);
 *   d += g (s_i + s_{i+1});  s += e (d_{i-1} + d_i);
 *   lo = zeta s,  hi = -d / zeta,
 * over the whole-sample symmetric extension the reference's MirroredArray
 * gives (dwt_cpu.h:162-176: o < 0 -> -o; o >= n -> 2n - o - 2, repeated).
 * With these scalings the bank equals the reference's taps (lowpass DC
 * gain sqrt(2), highpass sign flipped).  Haar is (a + b) / sqrt(2),
 * (a - b) / sqrt(2) in double (dwt_cpu.cc:139-170).
 *
 * Pinned in tests/test_oracle_dwt.py against the numpy FP64 lifting there
 * (itself pinned to the reference's taps, the dwt_cpu.cc:680 known-answer
 * vector and the DC gain) and against the float oracle at 1e-6.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "wv_oracle.h"

#define K 16   /* extension on each side; lifting reaches 4-5 samples */

static const double ALPHA = -1.586134342059924;
static const double BETA = -0.052980118572961;
static const double GAMMA = 0.882911075530934;
static const double DELTA = 0.443506852043971;
static const double ZETA = 1.149604398860241;

static int mirror(int o, int n) {
  if (o < 0) o = -o;
  while (o >= n) {
    o = 2 * n - o - 2;
    if (o < 0) o = -o;
  }
  return o;
}

/* one analysis level of x[0..n) (n even, n >= 2) in place: lo then hi */
static void cdf97_level64(double *x, int n, double *ext, double *s, double *d) {
  const int m = n + 2 * K;           /* extended length, K even */
  for (int i = 0; i < m; i++) ext[i] = x[mirror(i - K, n)];
  const int h = m / 2;
  for (int i = 0; i < h; i++) { s[i] = ext[2 * i]; d[i] = ext[2 * i + 1]; }
  for (int i = 0; i + 1 < h; i++) d[i] += ALPHA * (s[i] + s[i + 1]);
  for (int i = 1; i < h; i++) s[i] += BETA * (d[i - 1] + d[i]);
  for (int i = 0; i + 1 < h; i++) d[i] += GAMMA * (s[i] + s[i + 1]);
  for (int i = 1; i < h; i++) s[i] += DELTA * (d[i - 1] + d[i]);
  const int o = K / 2;
  for (int i = 0; i < n / 2; i++) {
    x[i] = s[o + i] * ZETA;
    x[n / 2 + i] = -d[o + i] / ZETA;
  }
}

static void haar_level64(double *x, int n, double *t) {
  const double r = 1.0 / sqrt(2.0);
  for (int i = 0; i < n / 2; i++) {
    t[i] = (x[2 * i] + x[2 * i + 1]) * r;
    t[n / 2 + i] = (x[2 * i] - x[2 * i + 1]) * r;
  }
  memcpy(x, t, sizeof(double) * (size_t)n);
}

/* levels of one line of length n (a power-of-two multiple of 2^steps) */
static void line64(double *x, int n, int steps, int wavelet, double *w) {
  double *ext = w, *s = w + n + 2 * K, *d = s + (n + 2 * K) / 2 + 1;
  for (int l = 0, len = n; l < steps && len >= 2; l++, len /= 2) {
    if (wavelet == OR_WAVELET_HAAR) haar_level64(x, len, ext);
    else cdf97_level64(x, len, ext, s, d);
  }
}

/* Forward transform in double of a padded float volume in[z][y][x]
   (size[] = X, Y, Z): steps[0] levels along x, steps[1] along y,
   steps[2] along z.  out[z][y][x] holds the result, in the input's own
   layout (no rotation).  Lines are gathered LB at a time, neighbours along
   the fastest other axis, so the strided axes read whole cache lines. */
#define LB 64
void or_dwt3d_f64(const float *in, const int size[3], int wavelet, const int steps[3],
                  double *out) {
  const int X = size[0], Y = size[1], Z = size[2];
  const size_t n = (size_t)X * Y * Z;
  for (size_t i = 0; i < n; i++) out[i] = in[i];
  int maxlen = X > Y ? X : Y;
  if (Z > maxlen) maxlen = Z;
  double *lines = (double *)malloc(sizeof(double) * (size_t)maxlen * LB);
  double *w = (double *)malloc(sizeof(double) * (size_t)(3 * (maxlen + 2 * K) + 8));
  const int len[3] = { X, Y, Z };
  const size_t stride[3] = { 1, (size_t)X, (size_t)X * Y };
  for (int a = 0; a < 3; a++) {
    if (steps[a] <= 0) continue;
    /* the two other axes enumerate the lines; b is the faster of them */
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const int L = len[a];
    for (int j = 0; j < len[c]; j++)
      for (int i0 = 0; i0 < len[b]; i0 += LB) {
        const int nb = len[b] - i0 < LB ? len[b] - i0 : LB;
        double *base = out + (size_t)i0 * stride[b] + (size_t)j * stride[c];
        for (int p = 0; p < L; p++)
          for (int i = 0; i < nb; i++)
            lines[(size_t)i * L + p] = base[i * stride[b] + p * stride[a]];
        for (int i = 0; i < nb; i++) line64(lines + (size_t)i * L, L, steps[a], wavelet, w);
        for (int p = 0; p < L; p++)
          for (int i = 0; i < nb; i++)
            base[i * stride[b] + p * stride[a]] = lines[(size_t)i * L + p];
      }
  }
  free(lines);
  free(w);
}
