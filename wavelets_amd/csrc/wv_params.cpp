// wv_params.cpp -- host side of the quantizers: parameters exactly as
// QuantUniform::init / QuantLog::init (quant.h:196-242,337-372) with the
// host libm, the host quantizer (Quantizer::quant), the bin-edge tables that
// let the device reproduce the log quantizer by comparison, and the
// dequantization table.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../../include/wv_hip.h"

namespace {

struct ParamStore {
  std::vector<float> pos_edges, neg_edges, boundaries, values, dequant;
};

inline float bits2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f2bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// QuantLog::quant for x > threshold on each side (quant.h:374-406), with
// the reference's evaluation order: zeroBin + posScale*logf(x) + posOffset
// and zeroBin - negScale*logf(x) - negOffset.
float log_pos_v(const wv_quant_params *p, float x) {
  return p->zero_bin + p->pos_scale * logf(x) + p->pos_offset;
}
int log_pos(const wv_quant_params *p, float x) {
  float v = log_pos_v(p, x);
  if (v >= p->bin_count) return p->bin_count - 1;
  return (int)(v + 0.5f);
}
int log_neg(const wv_quant_params *p, float a) {   // a = |x| > threshold
  float v = p->zero_bin - p->neg_scale * logf(a) - p->neg_offset;
  if (v < 0) return 0;
  return (int)(v + 0.5f);
}

// Smallest float bit pattern u in (lo, hi] with pred(u) true, where pred is
// monotone (false ... false true ... true); returns hi+1 when none.  Starts
// from a guess and gallops outward so a good guess costs a few evaluations.
template <class Pred>
uint32_t first_true(uint32_t lo, uint32_t hi, uint32_t guess, Pred pred) {
  if (guess <= lo) guess = lo + 1;
  if (guess > hi) guess = hi;
  uint32_t a, b;   // invariant: pred(a) false (or a == lo), pred(b) true
  if (pred(guess)) {
    b = guess;
    uint32_t step = 1;
    for (;;) {
      uint32_t c = (b - lo > step) ? b - step : lo;
      if (c == lo) { a = lo; break; }
      if (!pred(c)) { a = c; break; }
      b = c;
      step <<= 1;
    }
  } else {
    a = guess;
    uint32_t step = 1;
    for (;;) {
      if (a == hi) return hi + 1;
      uint32_t c = (hi - a > step) ? a + step : hi;
      if (pred(c)) { b = c; break; }
      a = c;
      step <<= 1;
    }
  }
  while (b - a > 1) {
    uint32_t m = a + (b - a) / 2;
    if (pred(m)) b = m; else a = m;
  }
  return b;
}

void build_log_edges(wv_quant_params *p, ParamStore *s) {
  const uint32_t lo = f2bits(p->threshold);       // edges are > threshold
  const uint32_t hi = 0x7f7fffffu;                 // FLT_MAX
  const int zb = p->zero_bin;
  // Positive side: QuantLog::quant is B-1 once v >= B, else (int)(v+0.5f).
  // Just below that clamp, v+0.5f can round to B: the reference then returns
  // bin B (only for |x| above the max it was built with).  Both conditions
  // are monotone in x, so the table holds the first x of every
  // unclamped bin zb+1 .. B plus the first x of the clamp.
  const int npos = p->bin_count - zb;
  s->pos_edges.resize(npos);
  for (int k = 1; k <= npos; k++) {
    const int target = zb + k;
    // analytic guess: v = target - 0.5 => x = exp((target-0.5-zb-pO)/pS)
    double g = std::exp(((double)target - 0.5 - zb - p->pos_offset) / p->pos_scale);
    uint32_t gb = (std::isfinite(g) && g < 3.4e38) ? f2bits((float)g) : hi;
    uint32_t e = first_true(lo, hi, gb, [&](uint32_t u) {
      return log_pos_v(p, bits2f(u)) + 0.5f >= (float)target; });
    s->pos_edges[k - 1] = e > hi ? INFINITY : bits2f(e);
  }
  {
    double g = std::exp(((double)p->bin_count - zb - p->pos_offset) / p->pos_scale);
    uint32_t gb = (std::isfinite(g) && g < 3.4e38) ? f2bits((float)g) : hi;
    uint32_t e = first_true(lo, hi, gb, [&](uint32_t u) {
      return log_pos_v(p, bits2f(u)) >= (float)p->bin_count; });
    p->pos_clamp = e > hi ? INFINITY : bits2f(e);
  }
  const int nneg = zb;
  s->neg_edges.resize(nneg);
  for (int k = 1; k <= nneg; k++) {
    const int target = zb - k;
    double g = std::exp(((double)zb - target - 0.5 - p->neg_offset) / p->neg_scale);
    uint32_t gb = (std::isfinite(g) && g < 3.4e38) ? f2bits((float)g) : hi;
    uint32_t e = first_true(lo, hi, gb, [&](uint32_t u) {
      return log_neg(p, bits2f(u)) <= target; });
    s->neg_edges[k - 1] = e > hi ? INFINITY : bits2f(e);
  }
}

}  // namespace

extern "C" {

int wv_quant_host(const wv_quant_params *p, float x) {
  const int zb = p->zero_bin;
  if (p->alg == WV_QUANT_UNIFORM) {
    // QuantUniform::quant quant.h:244-276
    if (x < 0) {
      if (x >= -p->threshold) return zb;
      x = x * p->neg_scale - p->neg_offset;
      int bin = (int)x + zb;
      return bin < 0 ? 0 : bin;
    }
    if (x <= p->threshold) return zb;
    x = x * p->pos_scale + p->pos_offset + zb;
    if (x > p->bin_count - 1) return p->bin_count - 1;
    return (int)x;
  } else if (p->alg == WV_QUANT_LOG) {
    if (x < 0) {
      x = -x;
      if (x <= p->threshold) return zb;
      return log_neg(p, x);
    }
    if (x <= p->threshold) return zb;
    return log_pos(p, x);
  } else {
    // QuantCodebook::quant quant.h:515-526
    const int m = p->n_boundaries;
    const float *b = p->boundaries;
    if (x >= b[m - 1]) return m;
    int lo = 0, len = m;
    while (len > 0) {
      int half = len / 2;
      if (b[lo + half] < x) { lo += half + 1; len -= half + 1; }
      else len = half;
    }
    return lo;
  }
}

float wv_dequant_host(const wv_quant_params *p, int x) {
  const int zb = p->zero_bin;
  if (p->alg == WV_QUANT_UNIFORM) {
    // quant.h:278-291
    if (x < zb) return (x - 0.5f + p->neg_offset - zb) * p->neg_inv_scale;
    if (x > zb) return (x + 0.5f - p->pos_offset - zb) * p->pos_inv_scale;
    return 0;
  } else if (p->alg == WV_QUANT_LOG) {
    // quant.h:409-437
    float sign, inv, off;
    if (x < zb) { sign = -1; off = p->neg_offset; inv = p->neg_inv_scale; }
    else { sign = 1; off = p->pos_offset; inv = p->pos_inv_scale; }
    if (x == zb) return 0;
    return sign * expf((sign * (x - zb) - off) * inv);
  }
  return p->values[x];
}

int wv_quant_params_init(wv_quant_params *p, int alg, int bin_count,
                         float threshold, float max_val,
                         const float *boundaries, const float *values) {
  std::memset(p, 0, sizeof *p);
  if (bin_count < 2 || bin_count > WV_MAX_BINS) return WV_ERR_ARG;
  ParamStore *s = new ParamStore();
  p->priv = s;
  p->alg = alg;
  p->bin_count = bin_count;
  p->zero_bin = (bin_count - 1) >> 1;
  p->threshold = threshold;
  p->max_val = max_val;
  if (alg == WV_QUANT_UNIFORM) {
    if (!(max_val > threshold) || threshold < 0) return WV_ERR_ARG;
    int negBinCount = (bin_count - 1) / 2;
    p->neg_scale = (negBinCount - 1) / (max_val - threshold);
    int posBinCount = bin_count / 2;
    p->pos_scale = (posBinCount - 1) / (max_val - threshold);
    p->neg_inv_scale = p->neg_scale == 0 ? 0 : 1 / p->neg_scale;
    p->pos_inv_scale = p->pos_scale == 0 ? 0 : 1 / p->pos_scale;
    p->neg_offset = 1 - threshold * p->neg_scale;
    p->pos_offset = 1 - threshold * p->pos_scale;
  } else if (alg == WV_QUANT_LOG) {
    if (!(max_val > threshold) || !(threshold > 0)) return WV_ERR_ARG;
    int negBinCount = (bin_count - 1) / 2;
    p->neg_scale = (negBinCount - 1) / logf(max_val / threshold);
    p->neg_inv_scale = 1 / p->neg_scale;
    p->neg_offset = 1 - p->neg_scale * logf(threshold);
    int posBinCount = bin_count / 2;
    p->pos_scale = (posBinCount - 1) / logf(max_val / threshold);
    p->pos_inv_scale = 1 / p->pos_scale;
    p->pos_offset = 1 - p->pos_scale * logf(threshold);
    build_log_edges(p, s);
    p->n_pos_edges = (int)s->pos_edges.size();
    p->n_neg_edges = (int)s->neg_edges.size();
    p->pos_edges = s->pos_edges.data();
    p->neg_edges = s->neg_edges.data();
  } else if (alg == WV_QUANT_LLOYD) {
    if (!boundaries || !values) return WV_ERR_ARG;
    s->boundaries.assign(boundaries, boundaries + bin_count - 1);
    s->values.assign(values, values + bin_count);
    p->n_boundaries = bin_count - 1;
    p->boundaries = s->boundaries.data();
    p->values = s->values.data();
    p->zero_bin = wv_quant_host(p, 0.0f);
  } else {
    return WV_ERR_ARG;
  }
  // zeroBin = quant(0) (test_compress_cpu.cc:280)
  p->zero_bin = (alg == WV_QUANT_LLOYD) ? wv_quant_host(p, 0.0f) : p->zero_bin;
  s->dequant.resize(bin_count);
  for (int b = 0; b < bin_count; b++) s->dequant[b] = wv_dequant_host(p, b);
  return WV_OK;
}

void wv_quant_params_free(wv_quant_params *p) {
  delete (ParamStore *)p->priv;
  std::memset(p, 0, sizeof *p);
}

// dequantization table of a parameter set (bin_count floats)
const float *wv_quant_params_dequant_table(const wv_quant_params *p) {
  return ((const ParamStore *)p->priv)->dequant.data();
}

}  // extern "C"
