// wv_select.hip -- threshold selection by radix select on the bit patterns
// of |x| (their order is the order of the values), replacing the full
// thrust::sort of the reference GPU path (test_compress_gpu.cu:632-662).
//
//   digit 1: bits 30..20 (2048 bins) -- histogram fused into the last DWT
//            pass (or wv_stats_compute);
//   scan 1 : one block finds the bucket holding rank idx = (int)(frac*n);
//   collect: one read of the coefficients appends the keys of that bucket
//            to a candidate list (wave-aggregated atomics) and histograms
//            digit 2 (bits 19..9) of them;
//   scan 2 / hist 3 (bits 8..0) / scan 3 over the candidates only.
// Everything stays on the device; the result struct is written by the last
// single-block kernel.
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

struct SelState {
  int32_t idx;          // target rank
  int32_t special;      // 1: frac >= 1 (threshold = max |x|)
  uint32_t b1, b2;      // chosen buckets
  uint32_t rank;        // rank inside the current bucket
  uint32_t ncand;       // candidates appended
  uint32_t pad[2];
  uint32_t hist2[WV_H2_BINS];
  uint32_t hist3[WV_H3_BINS];
};

int64_t threshold_ws_bytes(int64_t n) {
  return (int64_t)sizeof(SelState) + 256 + n * 4;
}

__global__ void stats_clear_kernel(WvStats *st) {
  for (int i = threadIdx.x; i < WV_H1_BINS; i += blockDim.x) st->hist[i] = 0;
  if (threadIdx.x == 0) {
    st->min_ord = 0xffffffffu;
    st->max_ord = 0u;
    st->max_key = 0u;
    st->count_lo = 0u;
  }
}

hipError_t stats_clear(void *stats, hipStream_t s) {
  hipLaunchKernelGGL(stats_clear_kernel, dim3(1), dim3(256), 0, s,
                     (WvStats *)stats);
  return hipGetLastError();
}

__global__ void __launch_bounds__(WV_BLOCK)
stats_kernel(const float *__restrict__ coef, int64_t n, WvStats *st) {
  __shared__ uint32_t h[WV_H1_BINS];
  for (int i = threadIdx.x; i < WV_H1_BINS; i += WV_BLOCK) h[i] = 0;
  __syncthreads();
  uint32_t mn = 0xffffffffu, mx = 0, mk = 0;
  for (int64_t i = (int64_t)blockIdx.x * WV_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * WV_BLOCK) {
    const float v = coef[i];
    const uint32_t o = wv_f2ord(v);
    mn = min(mn, o);
    mx = max(mx, o);
    const uint32_t key = __builtin_bit_cast(uint32_t, v) & 0x7fffffffu;
    mk = max(mk, key);
    atomicAdd(&h[key >> WV_H1_SHIFT], 1u);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    mk = max(mk, (uint32_t)__shfl_xor((int)mk, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&st->min_ord, mn);
    atomicMax(&st->max_ord, mx);
    atomicMax(&st->max_key, mk);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < WV_H1_BINS; i += WV_BLOCK)
    if (h[i]) atomicAdd(&st->hist[i], h[i]);
}

hipError_t stats_compute(const float *coef, int64_t n, void *stats,
                         hipStream_t s) {
  hipError_t e = stats_clear(stats, s);
  if (e != hipSuccess) return e;
  int64_t blocks = (n + WV_BLOCK * 16 - 1) / (WV_BLOCK * 16);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(stats_kernel, dim3((unsigned)blocks), dim3(WV_BLOCK), 0, s,
                     coef, n, (WvStats *)stats);
  return hipGetLastError();
}

// Block-wide search: given per-bin counts (nb bins) find the bin b with
// cum(b) <= rank < cum(b) + cnt(b).  1024 threads.
template <int NB>
__device__ void find_bucket(const uint32_t *cnt, uint32_t rank, uint32_t *out_b,
                            uint32_t *out_rank) {
  constexpr int PER = (NB + 1023) / 1024;
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  uint32_t local[PER];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int b = t * PER + k;
    local[k] = b < NB ? cnt[b] : 0;
    sum += local[k];
  }
  part[t] = sum;
  __syncthreads();
  // inclusive scan (Hillis-Steele)
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t before = t ? part[t - 1] : 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    if (rank >= before && rank < before + local[k]) {
      *out_b = (uint32_t)(t * PER + k);
      *out_rank = rank - before;
    }
    before += local[k];
  }
}

__global__ void __launch_bounds__(1024)
sel_scan1(const WvStats *st, int64_t n, float frac, SelState *S) {
  __shared__ uint32_t b, r;
  // (int)(compRatio * count) evaluated in float (thresh_cpu.cc:96)
  const float prod = frac * (float)n;
  int64_t idx = (int64_t)prod;
  if (threadIdx.x == 0) {
    S->special = (frac >= 1.0f || idx >= n) ? 1 : 0;
    S->idx = (int32_t)(idx >= n ? n : idx);
    S->ncand = 0;
    b = 0;
    r = 0;
  }
  for (int i = threadIdx.x; i < WV_H2_BINS; i += 1024) S->hist2[i] = 0;
  for (int i = threadIdx.x; i < WV_H3_BINS; i += 1024) S->hist3[i] = 0;
  __syncthreads();
  if (idx >= n) idx = n - 1;
  find_bucket<WV_H1_BINS>(st->hist, (uint32_t)idx, &b, &r);
  __syncthreads();
  if (threadIdx.x == 0) {
    S->b1 = b;
    S->rank = r;
  }
}

__global__ void __launch_bounds__(WV_BLOCK)
sel_collect(const float *__restrict__ coef, int64_t n, SelState *S,
            uint32_t *cand) {
  __shared__ uint32_t h[WV_H2_BINS];
  for (int i = threadIdx.x; i < WV_H2_BINS; i += WV_BLOCK) h[i] = 0;
  __syncthreads();
  const uint32_t b1 = S->b1;
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * WV_BLOCK;
  // whole waves iterate together so the ballot covers all lanes
  const int64_t nround = (n + stride - 1) / stride;
  for (int64_t r = 0; r < nround; r++) {
    const int64_t i = r * stride + (int64_t)blockIdx.x * WV_BLOCK + threadIdx.x;
    uint32_t key = 0;
    bool hit = false;
    if (i < n) {
      key = __builtin_bit_cast(uint32_t, coef[i]) & 0x7fffffffu;
      hit = (key >> WV_H1_SHIFT) == b1;
    }
    const uint64_t m = __ballot(hit);
    if (m) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&S->ncand, (uint32_t)__popcll(m));
      base = __shfl(base, 0);
      if (hit) {
        const uint32_t off = (uint32_t)__popcll(m & ((1ull << lane) - 1));
        cand[base + off] = key;
        atomicAdd(&h[(key >> WV_H2_SHIFT) & (WV_H2_BINS - 1)], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < WV_H2_BINS; i += WV_BLOCK)
    if (h[i]) atomicAdd(&S->hist2[i], h[i]);
}

__global__ void __launch_bounds__(1024) sel_scan2(SelState *S) {
  __shared__ uint32_t b, r;
  if (threadIdx.x == 0) { b = 0; r = 0; }
  __syncthreads();
  find_bucket<WV_H2_BINS>(S->hist2, S->rank, &b, &r);
  __syncthreads();
  if (threadIdx.x == 0) {
    S->b2 = b;
    S->rank = r;
  }
}

__global__ void __launch_bounds__(WV_BLOCK)
sel_hist3(const uint32_t *cand, SelState *S) {
  __shared__ uint32_t h[WV_H3_BINS];
  for (int i = threadIdx.x; i < WV_H3_BINS; i += WV_BLOCK) h[i] = 0;
  __syncthreads();
  const uint32_t nc = S->ncand, b2 = S->b2;
  for (uint32_t i = blockIdx.x * WV_BLOCK + threadIdx.x; i < nc;
       i += gridDim.x * WV_BLOCK) {
    const uint32_t key = cand[i];
    if (((key >> WV_H2_SHIFT) & (WV_H2_BINS - 1)) == b2)
      atomicAdd(&h[key & (WV_H3_BINS - 1)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < WV_H3_BINS; i += WV_BLOCK)
    if (h[i]) atomicAdd(&S->hist3[i], h[i]);
}

__global__ void __launch_bounds__(1024)
sel_scan3(const WvStats *st, int64_t n, SelState *S, wv_thresh_result *res) {
  __shared__ uint32_t b, r;
  if (threadIdx.x == 0) { b = 0; r = 0; }
  __syncthreads();
  find_bucket<WV_H3_BINS>(S->hist3, S->rank, &b, &r);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t key = (S->b1 << WV_H1_SHIFT) | (S->b2 << WV_H2_SHIFT) | b;
    float t = __builtin_bit_cast(float, key);
    wv_thresh_result o;
    // Traverse starts min at FLT_MAX and max at FLT_MIN (thresh_cpu.cc:67-68)
    float mn = wv_ord2f(st->min_ord), mx = wv_ord2f(st->max_ord);
    const float fmax = 3.40282347e+38f, fmin = 1.17549435e-38f;
    o.min_val = mn < fmax ? mn : fmax;
    o.max_val = mx > fmin ? mx : fmin;
    o.max_abs = __builtin_bit_cast(float, st->max_key);
    if (S->special) {
      // frac >= 1: the reference GPU driver takes max |x|
      // (test_compress_gpu.cu:299-301)
      o.threshold = o.max_abs;
      o.threshold_key = st->max_key;
      o.thresh_idx = (int32_t)n;
      o.nonzero_count = 0;
    } else {
      o.threshold_key = key;
      if (t == 0.f) t = 1e-16f;    // thresh_cpu.cc:98
      o.threshold = t;
      o.thresh_idx = S->idx;
      o.nonzero_count = (int32_t)(n - S->idx - 1);
    }
    o.n_candidates = (int32_t)S->ncand;
    *res = o;
  }
}

hipError_t threshold_select(const float *coef, int64_t n, float frac,
                            const void *stats, void *ws,
                            wv_thresh_result *res, hipStream_t s) {
  SelState *S = (SelState *)ws;
  uint32_t *cand = (uint32_t *)((char *)ws + ((sizeof(SelState) + 255) & ~(size_t)255));
  const WvStats *st = (const WvStats *)stats;
  hipLaunchKernelGGL(sel_scan1, dim3(1), dim3(1024), 0, s, st, n, frac, S);
  int64_t blocks = (n + WV_BLOCK * 8 - 1) / (WV_BLOCK * 8);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(sel_collect, dim3((unsigned)blocks), dim3(WV_BLOCK), 0, s,
                     coef, n, S, cand);
  hipLaunchKernelGGL(sel_scan2, dim3(1), dim3(1024), 0, s, S);
  int64_t b3 = (n / 64 + WV_BLOCK - 1) / WV_BLOCK;
  if (b3 > 1024) b3 = 1024;
  if (b3 < 1) b3 = 1;
  hipLaunchKernelGGL(sel_hist3, dim3((unsigned)b3), dim3(WV_BLOCK), 0, s, cand, S);
  hipLaunchKernelGGL(sel_scan3, dim3(1), dim3(1024), 0, s, st, n, S, res);
  return hipGetLastError();
}

}  // namespace wv
