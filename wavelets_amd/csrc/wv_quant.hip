// wv_quant.hip -- quantization fused with the bin histogram, and
// dequantization, for gfx950.
//
// One pass reads the coefficients (16 B vector loads), writes int32 bins
// (16 B vector stores) and accumulates the bin histogram in LDS (the zero
// bin, which holds most values after thresholding, is counted in a register
// per thread), so the stage moves 8 B per coefficient.
//
//  uniform : the float expressions of QuantUniform::quant (quant.h:244-276),
//            contraction off, so bins equal the host's;
//  log     : the bin is the number of host-tabulated bin edges (the
//            smallest float of every bin under the host logf) at or below
//            |x|; this reproduces QuantLog::quant (quant.h:374-406) exactly,
//            whatever the device logarithm's rounding.  The count starts
//            from a per-block LDS table indexed by the top bits of |x|'s
//            float pattern (the edge count at the start of each slot of
//            keys, up to 1024 slots per sign between the threshold and the
//            last edge) and steps up over the edges inside the slot -- no
//            logarithm, usually one or two LDS reads; when the edges leave
//            no room for the slot table, from a device-logarithm guess;
//  codebook: lower_bound over the boundaries plus the lastBoundary rule of
//            the host QuantCodebook::quant (quant.h:515-526).
#include <type_traits>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define WV_LDS_TABLE 4096    // floats of edge / boundary table kept in LDS
#define WV_LDS_HIST 4096     // bins counted in LDS (more: global atomics)

// number of entries of e[0..m) that are <= x (e ascending), from a
// starting guess g: steps down while e[g-1] > x and up
// while e[g] <= x.  The result depends only on the table (ascending), not on
// the guess; a guess within a bin or two costs two or three LDS reads.
__device__ inline int count_le_from(const float *e, int m, float x, int g) {
  g = g < 0 ? 0 : (g > m ? m : g);
  while (g > 0 && e[g - 1] > x) g--;
  while (g < m && e[g] <= x) g++;
  return g;
}

// first index with e[i] >= x (std::lower_bound)
__device__ inline int lower_bound_f(const float *e, int m, float x) {
  int lo = 0, len = m;
  while (len > 0) {
    const int half = len >> 1;
    if (e[lo + half] < x) { lo += half + 1; len -= half + 1; }
    else len = half;
  }
  return lo;
}

// The slot table of one sign of the log quantizer: slot s covers the float
// keys [(base + s) << sh, (base + s + 1) << sh) and holds the number of
// edges <= the slot's first key; keys >= k1 (the last edge's) count all m.
// The fields are block-uniform (scalar registers).
struct LogSide {
  uint32_t k1, base, sh;
};
#define LOG_LUT_MAX 1024

__device__ inline int log_count(const float *e, int m, const int *lut, const LogSide &L,
                                float a) {
  const uint32_t key = __builtin_bit_cast(uint32_t, a);
  // NaN counts 0 edges, like the logarithm-guess path
  if (key >= L.k1) return key > 0x7f800000u ? 0 : m;
  int c = lut[(key >> L.sh) - L.base];
  while (c < m && e[c] <= a) c++;
  return c;
}

// log quantizer through the slot tables lut (positive side) and lut + per:
// one count for either sign (the sign selects the tables), so a wave with
// both signs runs the count once, not once per side
__device__ inline int quant_log_lut(const QuantDev &q, const float *tab, const int *lut,
                                   int per, const LogSide &P, const LogSide &N, float x) {
  const int zb = q.zero_bin;
  const bool neg = x < 0;
  const float a = fabsf(x);
  if (a <= q.threshold) return zb;
  if (!neg && x >= q.pos_clamp) return q.bin_count - 1;
  LogSide L;
  L.k1 = neg ? N.k1 : P.k1;
  L.base = neg ? N.base : P.base;
  L.sh = neg ? N.sh : P.sh;
  const int c = log_count(neg ? tab + q.n_pos : tab, neg ? q.n_neg : q.n_pos,
                          neg ? lut + per : lut, L, a);
  return neg ? zb - c : zb + c;
}

// Builds the slot table of one side (edges e[0..m), per slots at most) into
// l; every thread of the block takes part and gets the same parameters.
__device__ inline LogSide build_log_side(float threshold, const float *e, int m, int *l,
                                         int per) {
  const uint32_t k0 = __builtin_bit_cast(uint32_t, threshold) & 0x7fffffffu;
  const uint32_t k1 = (uint32_t)__builtin_amdgcn_readfirstlane(
      m ? (int)__builtin_bit_cast(uint32_t, e[m - 1]) : 0);
  uint32_t sh = 0, nslot = 0;
  if (k1 > k0) {
    while (((k1 >> sh) - (k0 >> sh) + 1u) > (uint32_t)per) sh++;
    nslot = (k1 >> sh) - (k0 >> sh) + 1u;
  }
  const uint32_t base = k0 >> sh;
  for (uint32_t sl = threadIdx.x; sl < nslot; sl += WV_BLOCK) {
    const float sf = __builtin_bit_cast(float, (base + sl) << sh);
    int lo = 0, len = m;   // edges <= sf (upper bound)
    while (len > 0) {
      const int half = len >> 1;
      if (e[lo + half] <= sf) { lo += half + 1; len -= half + 1; }
      else len = half;
    }
    l[sl] = lo;
  }
  LogSide L;
  L.k1 = k1;
  L.base = base;
  L.sh = sh;
  return L;
}

template <int ALG>
__device__ inline int quant_one(const QuantDev &q, const float *tab, float x) {
  const int zb = q.zero_bin;
  if (ALG == WV_QUANT_UNIFORM) {
    if (x < 0) {
      if (x >= -q.threshold) return zb;
      x = x * q.neg_scale - q.neg_offset;
      const int bin = (int)x + zb;
      return bin < 0 ? 0 : bin;
    } else {
      if (x <= q.threshold) return zb;
      x = x * q.pos_scale + q.pos_offset + (float)zb;
      if (x > (float)(q.bin_count - 1)) return q.bin_count - 1;
      return (int)x;
    }
  } else if (ALG == WV_QUANT_LOG) {
    // The bin is the number of host-tabulated edges <= |x| (each edge is the
    // smallest float of its bin under the host logf); the device logarithm
    // only supplies the starting guess, scale*ln|x| + offset, of that count.
    // One count for either sign (the sign selects table and guess), so a
    // wave holding both signs runs the search once.
    constexpr float LN2 = 0.69314718055994531f;
    const bool neg = x < 0;
    const float a = fabsf(x);
    if (a <= q.threshold) return zb;
    if (!neg && x >= q.pos_clamp) return q.bin_count - 1;
    const float sc = neg ? q.neg_scale : q.pos_scale;
    const float off = neg ? q.neg_offset : q.pos_offset + 0.5f;
    const int g = (int)(sc * (__log2f(a) * LN2) + off);
    const int c = count_le_from(neg ? tab + q.n_pos : tab, neg ? q.n_neg : q.n_pos, a, g);
    return neg ? zb - c : zb + c;
  } else {
    const int m = q.n_pos;   // boundaries
    if (x >= tab[m - 1]) return m;
    return lower_bound_f(tab, m, x);
  }
}

// One block's share (block bidx of nblk) of the fused quantize + histogram
// pass over coef[0, n).
// tab / h are the caller's LDS arrays of tab_cap / hist_cap entries
// (declared once per kernel, so instantiations for several quantizers in
// one kernel share them).
template <int ALG>
__device__ __forceinline__ void quant_hist_body(const float *coef, int64_t n, const QuantDev &q,
                                int32_t *bins, uint32_t *__restrict__ counts,
                                int bidx, int nblk, float *tab, uint32_t *h,
                                int tab_cap = WV_LDS_TABLE, int hist_cap = WV_LDS_HIST) {
  const int ntab = q.n_pos + (ALG == WV_QUANT_LOG ? q.n_neg : 0);
  const bool tab_in_lds = ntab <= tab_cap;
  if (ALG != WV_QUANT_UNIFORM && tab_in_lds)
    for (int i = threadIdx.x; i < ntab; i += WV_BLOCK) tab[i] = gptr(q.table)[i];
  const bool hist_in_lds = q.bin_count <= hist_cap;
  if (hist_in_lds)
    for (int i = threadIdx.x; i < q.bin_count; i += WV_BLOCK) h[i] = 0;
  __syncthreads();
  const float *t = tab_in_lds ? tab : q.table;
  // log: slot tables after the edges when they leave room
  LogSide LP{}, LN{};
  const int per = tab_in_lds ? min((tab_cap - ntab) / 2, LOG_LUT_MAX) : 0;
  const bool use_lut = ALG == WV_QUANT_LOG && per >= 64;
  int *lut = (int *)(tab + ntab);
  if (use_lut) {
    LP = build_log_side(q.threshold, tab, q.n_pos, lut, per);
    LN = build_log_side(q.threshold, tab + q.n_pos, q.n_neg, lut + per, per);
    __syncthreads();
  }
  unsigned zcount = 0;
  const int zb = q.zero_bin;

  // counts[bin_count] collects out-of-range bins (the log quantizer's bin
  // B above the max it was built with, quant.h:398-403)
  auto count = [&](int b) {
    if (b == zb) zcount++;
    else if (b >= q.bin_count) atomicAdd(&counts[q.bin_count], 1u);
    else if (hist_in_lds) atomicAdd(&h[b], 1u);
    else atomicAdd(&counts[b], 1u);
  };

  // the pass, with the quantizer qf (one loop body per quantizer form).
  // ROLL keeps four copies of qf in the code, not sixteen (the loop over
  // the four loaded vectors is rolled, registers shifted through, no
  // indexed array): the log quantizers have loops of their own.
  auto pass = [&](auto qf, auto roll) {
    constexpr bool ROLL = decltype(roll)::value;
    const bool aligned = (((uintptr_t)coef | (uintptr_t)bins) & 15) == 0;
    const int64_t nvec = aligned ? (n >> 2) : 0;
    // global views: the batch passes these pointers through its job table
    const WV_GLOBAL wv_f4 *c4 = gptr((const wv_f4 *)coef);
    WV_GLOBAL wv_i4 *b4 = gptr((wv_i4 *)bins);
    const WV_GLOBAL float *cg = gptr(coef);
    WV_GLOBAL int32_t *bg = gptr(bins);
    auto one4 = [&](wv_f4 v, int64_t i) {
      wv_i4 r;
      r.x = qf(v.x);
      r.y = qf(v.y);
      r.z = qf(v.z);
      r.w = qf(v.w);
      count(r.x); count(r.y); count(r.z); count(r.w);
      b4[i] = r;
    };
    // four independent 16-byte loads per thread in flight per iteration
    const int64_t stride = (int64_t)nblk * WV_BLOCK;
    int64_t i0 = (int64_t)bidx * WV_BLOCK + threadIdx.x;
    for (; i0 + 3 * stride < nvec; i0 += 4 * stride) {
      wv_f4 v0 = c4[i0], v1 = c4[i0 + stride], v2 = c4[i0 + 2 * stride],
            v3 = c4[i0 + 3 * stride];
      if constexpr (ROLL) {
#pragma unroll 1
        for (int u = 0; u < 4; u++) {
          one4(v0, i0 + u * stride);
          v0 = v1; v1 = v2; v2 = v3;
        }
      } else {
        one4(v0, i0);
        one4(v1, i0 + stride);
        one4(v2, i0 + 2 * stride);
        one4(v3, i0 + 3 * stride);
      }
    }
    for (int64_t i = i0; i < nvec; i += stride) one4(c4[i], i);
    // tail
    for (int64_t i = nvec * 4 + (int64_t)bidx * WV_BLOCK + threadIdx.x; i < n;
         i += (int64_t)nblk * WV_BLOCK) {
      const int r = qf(cg[i]);
      bg[i] = r;
      count(r);
    }
  };
  using Unrolled = std::integral_constant<bool, false>;
  using Rolled = std::integral_constant<bool, true>;
  if constexpr (ALG == WV_QUANT_LOG) {
    if (use_lut)
      pass([&](float x) { return quant_log_lut(q, tab, lut, per, LP, LN, x); }, Rolled());
    else
      pass([&](float x) { return quant_one<ALG>(q, t, x); }, Rolled());
  } else {
    pass([&](float x) { return quant_one<ALG>(q, t, x); }, Unrolled());
  }
  for (int off = 32; off > 0; off >>= 1)
    zcount += (unsigned)__shfl_xor((int)zcount, off);
  if ((threadIdx.x & 63) == 0 && zcount) {
    if (hist_in_lds) atomicAdd(&h[zb], zcount);
    else atomicAdd(&counts[zb], zcount);
  }
  __syncthreads();
  if (hist_in_lds)
    for (int i = threadIdx.x; i < q.bin_count; i += WV_BLOCK)
      if (h[i]) atomicAdd(&counts[i], h[i]);
}

template <int ALG>
__global__ void __launch_bounds__(WV_BLOCK)
quant_hist_kernel(const float *coef, int64_t n, QuantDev q,
                  int32_t *bins, uint32_t *__restrict__ counts) {
  __shared__ float tab[WV_LDS_TABLE];
  __shared__ uint32_t h[WV_LDS_HIST];
  quant_hist_body<ALG>(coef, n, q, bins, counts, blockIdx.x, gridDim.x, tab, h);
}

__global__ void zero_u32(uint32_t *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0;
}

static unsigned grid_for(int64_t n) {
  int64_t b = (n / 4 + WV_BLOCK * 16 - 1) / (WV_BLOCK * 16);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

hipError_t quantize_hist(const float *coef, int64_t n, const QuantDev &q,
                         int32_t *bins, uint32_t *counts, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32, dim3((q.bin_count + 256) / 256), dim3(256), 0, s,
                     counts, q.bin_count + 1);
  const unsigned g = grid_for(n);
  if (q.alg == WV_QUANT_UNIFORM)
    hipLaunchKernelGGL(quant_hist_kernel<WV_QUANT_UNIFORM>, dim3(g), dim3(WV_BLOCK), 0, s,
                       coef, n, q, bins, counts);
  else if (q.alg == WV_QUANT_LOG)
    hipLaunchKernelGGL(quant_hist_kernel<WV_QUANT_LOG>, dim3(g), dim3(WV_BLOCK), 0, s,
                       coef, n, q, bins, counts);
  else
    hipLaunchKernelGGL(quant_hist_kernel<WV_QUANT_LLOYD>, dim3(g), dim3(WV_BLOCK), 0, s,
                       coef, n, q, bins, counts);
  return hipGetLastError();
}

// Uniform parameters from the device-resident threshold result: the same
// float expressions as QuantUniform::init (quant.h:227-239) evaluated by one
// thread, then the fused pass.
__global__ void uniform_params_kernel(const wv_thresh_result *res, int bins,
                                      QuantDev *q) {
  const float threshold = res->threshold, maxv = res->max_abs;
  const int negBins = (bins - 1) / 2, posBins = bins / 2;
  QuantDev o;
  o.alg = WV_QUANT_UNIFORM;
  o.bin_count = bins;
  o.zero_bin = (bins - 1) >> 1;
  o.threshold = threshold;
  o.neg_scale = (float)(negBins - 1) / (maxv - threshold);
  o.pos_scale = (float)(posBins - 1) / (maxv - threshold);
  o.neg_offset = 1 - threshold * o.neg_scale;
  o.pos_offset = 1 - threshold * o.pos_scale;
  o.n_pos = o.n_neg = 0;
  o.table = nullptr;
  *q = o;
}

__global__ void __launch_bounds__(WV_BLOCK)
quant_hist_uniform_dev_kernel(const float *coef, int64_t n,
                              const QuantDev *qp, int32_t *bins,
                              uint32_t *__restrict__ counts) {
  __shared__ uint32_t h[WV_LDS_HIST];
  const QuantDev q = *qp;
#define QUANT_ONE(x) quant_one<WV_QUANT_UNIFORM>(q, nullptr, (x))
  const bool hist_in_lds = q.bin_count <= WV_LDS_HIST;
  if (hist_in_lds)
    for (int i = threadIdx.x; i < q.bin_count; i += WV_BLOCK) h[i] = 0;
  __syncthreads();
  unsigned zcount = 0;
  const int zb = q.zero_bin;
  // counts[bin_count] collects out-of-range bins (the log quantizer's bin
  // B above the max it was built with, quant.h:398-403)
  auto count = [&](int b) {
    if (b == zb) zcount++;
    else if (b >= q.bin_count) atomicAdd(&counts[q.bin_count], 1u);
    else if (hist_in_lds) atomicAdd(&h[b], 1u);
    else atomicAdd(&counts[b], 1u);
  };
  const bool aligned = (((uintptr_t)coef | (uintptr_t)bins) & 15) == 0;
  const int64_t nvec = aligned ? (n >> 2) : 0;
  const float4 *c4 = (const float4 *)coef;
  int4 *b4 = (int4 *)bins;
  // four independent 16-byte loads per thread in flight per iteration
  const int64_t stride = (int64_t)gridDim.x * WV_BLOCK;
  int64_t i0 = (int64_t)blockIdx.x * WV_BLOCK + threadIdx.x;
  for (; i0 + 3 * stride < nvec; i0 += 4 * stride) {
    float4 vv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) vv[u] = c4[i0 + u * stride];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const float4 v = vv[u];
      int4 r;
      r.x = QUANT_ONE(v.x);
      r.y = QUANT_ONE(v.y);
      r.z = QUANT_ONE(v.z);
      r.w = QUANT_ONE(v.w);
      b4[i0 + u * stride] = r;
      count(r.x); count(r.y); count(r.z); count(r.w);
    }
  }
  for (int64_t i = i0; i < nvec; i += stride) {
    const float4 v = c4[i];
    int4 r;
    r.x = quant_one<WV_QUANT_UNIFORM>(q, nullptr, v.x);
    r.y = quant_one<WV_QUANT_UNIFORM>(q, nullptr, v.y);
    r.z = quant_one<WV_QUANT_UNIFORM>(q, nullptr, v.z);
    r.w = quant_one<WV_QUANT_UNIFORM>(q, nullptr, v.w);
    b4[i] = r;
    count(r.x); count(r.y); count(r.z); count(r.w);
  }
  for (int64_t i = nvec * 4 + (int64_t)blockIdx.x * WV_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * WV_BLOCK) {
    const int r = quant_one<WV_QUANT_UNIFORM>(q, nullptr, coef[i]);
    bins[i] = r;
    count(r);
  }
  for (int off = 32; off > 0; off >>= 1)
    zcount += (unsigned)__shfl_xor((int)zcount, off);
  if ((threadIdx.x & 63) == 0 && zcount) {
    if (hist_in_lds) atomicAdd(&h[zb], zcount);
    else atomicAdd(&counts[zb], zcount);
  }
  __syncthreads();
  if (hist_in_lds)
    for (int i = threadIdx.x; i < q.bin_count; i += WV_BLOCK)
      if (h[i]) atomicAdd(&counts[i], h[i]);
}

hipError_t quantize_hist_uniform_dev(const float *coef, int64_t n,
                                     int bin_count,
                                     const wv_thresh_result *res,
                                     int32_t *bins, uint32_t *counts,
                                     hipStream_t s) {
  // the parameter block lives in the tail of the counts buffer region the
  // caller provides: counts has bin_count entries plus 64 spare words
  QuantDev *qp = (QuantDev *)(counts + ((bin_count + 63) & ~63));
  hipLaunchKernelGGL(zero_u32, dim3((bin_count + 256) / 256), dim3(256), 0, s,
                     counts, bin_count + 1);
  hipLaunchKernelGGL(uniform_params_kernel, dim3(1), dim3(1), 0, s, res,
                     bin_count, qp);
  hipLaunchKernelGGL(quant_hist_uniform_dev_kernel, dim3(grid_for(n)), dim3(WV_BLOCK),
                     0, s, coef, n, qp, bins, counts);
  return hipGetLastError();
}

// dequantization through a host-built table of the bin values
// Quantizer::dequantizeRow: out[i] = table[bins[i]] (0 for bins outside
// the table); 16-byte accesses, the table in LDS when it fits
#define DQ_LDS 8192
template <bool LDS>
__global__ void __launch_bounds__(WV_BLOCK)
dequant_kernel(const int32_t *__restrict__ bins, int64_t n, const float *__restrict__ table,
               int nb, float *__restrict__ out) {
  __shared__ float t[LDS ? DQ_LDS : 1];
  const float *tab = table;
  if (LDS) {
    for (int i = threadIdx.x; i < nb; i += WV_BLOCK) t[i] = table[i];
    __syncthreads();
    tab = t;
  }
  auto dq = [&](int b) { return (b >= 0 && b < nb) ? tab[b] : 0.f; };
  const int64_t stride = (int64_t)gridDim.x * WV_BLOCK;
  const bool aligned = (((uintptr_t)bins | (uintptr_t)out) & 15) == 0;
  const int64_t nvec = aligned ? n >> 2 : 0;
  const int4 *b4 = (const int4 *)bins;
  float4 *o4 = (float4 *)out;
  for (int64_t i = (int64_t)blockIdx.x * WV_BLOCK + threadIdx.x; i < nvec; i += stride) {
    const int4 b = b4[i];
    o4[i] = make_float4(dq(b.x), dq(b.y), dq(b.z), dq(b.w));
  }
  for (int64_t i = nvec * 4 + (int64_t)blockIdx.x * WV_BLOCK + threadIdx.x; i < n; i += stride)
    out[i] = dq(bins[i]);
}

hipError_t dequantize_table(const int32_t *bins, int64_t n, const float *table,
                            int bin_count, float *out, hipStream_t s) {
  int64_t g = (n / 4 + WV_BLOCK - 1) / WV_BLOCK;
  const int64_t cap = 8 * (int64_t)cu_count() * (2048 / WV_BLOCK);
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  if (bin_count <= DQ_LDS)
    hipLaunchKernelGGL(dequant_kernel<true>, dim3((unsigned)g), dim3(WV_BLOCK), 0, s, bins, n,
                       table, bin_count, out);
  else
    hipLaunchKernelGGL(dequant_kernel<false>, dim3((unsigned)g), dim3(WV_BLOCK), 0, s, bins, n,
                       table, bin_count, out);
  return hipGetLastError();
}

// ---- batches -------------------------------------------------------------
__device__ inline int find_qjob(const int *blk0, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// one block per job: counts[0 .. bin_count] = 0, and the uniform
// quantizer's parameters from the job's threshold result (the float
// expressions of QuantUniform::init, quant.h:227-239)
__global__ void bquant_prepare(const BQuantJob *J) {
  const BQuantJob &j = J[blockIdx.x];
  for (int i = threadIdx.x; i <= j.bin_count; i += blockDim.x) j.counts[i] = 0;
  if (threadIdx.x == 0 && j.alg == WV_QUANT_UNIFORM) {
    const wv_thresh_result *res = j.res;
    const float threshold = res->threshold, maxv = res->max_abs;
    const int bins = j.bin_count;
    const int negBins = (bins - 1) / 2, posBins = bins / 2;
    QuantDev o;
    o.alg = WV_QUANT_UNIFORM;
    o.bin_count = bins;
    o.zero_bin = (bins - 1) >> 1;
    o.threshold = threshold;
    o.neg_scale = (float)(negBins - 1) / (maxv - threshold);
    o.pos_scale = (float)(posBins - 1) / (maxv - threshold);
    o.neg_offset = 1 - threshold * o.neg_scale;
    o.pos_offset = 1 - threshold * o.pos_scale;
    o.n_pos = o.n_neg = 0;
    o.pos_clamp = 0.f;
    o.table = nullptr;
    *j.q = o;
  }
}

// LDS sized for the batch's largest bin count (dynamic: tab_cap floats of
// edges + slot tables, then hist_cap counters), so batches of small books
// keep more blocks resident than the 32 KB of the largest
__global__ void __launch_bounds__(WV_BLOCK)
bquant_hist(const BQuantJob *J, const int *blk0, int njobs, int tab_cap, int hist_cap) {
  const int jb = find_qjob(blk0, njobs, blockIdx.x);
  extern __shared__ uint32_t qlds[];
  float *tab = (float *)qlds;
  uint32_t *h = qlds + tab_cap;
  const BQuantJob &j = J[jb];
  const QuantDev q = *j.q;
  const int bidx = (int)blockIdx.x - j.blk0;
  if (q.alg == WV_QUANT_UNIFORM)
    quant_hist_body<WV_QUANT_UNIFORM>(j.coef, j.n, q, j.bins, j.counts, bidx, j.nblk, tab, h,
                                      tab_cap, hist_cap);
  else if (q.alg == WV_QUANT_LOG)
    quant_hist_body<WV_QUANT_LOG>(j.coef, j.n, q, j.bins, j.counts, bidx, j.nblk, tab, h,
                                  tab_cap, hist_cap);
  else
    quant_hist_body<WV_QUANT_LLOYD>(j.coef, j.n, q, j.bins, j.counts, bidx, j.nblk, tab, h,
                                    tab_cap, hist_cap);
}

// blocks of a batch job: 64 K samples each (a block's fixed costs -- table
// load, histogram clear and flush -- over more samples; one launch covers
// every job, so the grid stays large)
int quant_blocks_for(int64_t n) {
  const int64_t b = (n + 65535) / 65536;
  return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

hipError_t quantize_hist_batch(const BQuantJob *jobs_d, const int *blk0_d, int njobs,
                               int total_blocks, int max_bins, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  // edges (< max_bins) plus the log slot tables (2 x 1024), and the counters
  const int tab_cap = min(WV_LDS_TABLE, ((max_bins + 2 * 1024 + 63) & ~63));
  const int hist_cap = min(WV_LDS_HIST, ((max_bins + 63) & ~63));
  hipLaunchKernelGGL(bquant_prepare, dim3(njobs), dim3(256), 0, s, jobs_d);
  hipLaunchKernelGGL(bquant_hist, dim3(total_blocks), dim3(WV_BLOCK),
                     (size_t)(tab_cap + hist_cap) * 4, s, jobs_d, blk0_d, njobs, tab_cap,
                     hist_cap);
  return hipGetLastError();
}

}  // namespace wv
