// wv_select.hip -- threshold selection on the device, replacing the full
// thrust::sort of the reference GPU path (test_compress_gpu.cu:632-662) with
// the same result as thresh_cpu (thresh_cpu.cc:80-109): the rank
// idx = (int)(frac * (float)n) element of the ascending |x|.
//
// The bit patterns of |x| order like the values, so the select works on
// 31-bit keys:
//   sample   16K keys spread over the array (jittered strata: key i from a
//            hashed position inside the i-th 1/16K of the array -- the
//            32-bit hash of i times the stratum's width, shifted down 32
//            bits -- so the sample cannot alias with the sub-band layout of
//            the coefficients; gathered by 64 blocks so the scattered loads
//            are in flight from many CUs), and from their two-level
//            histogram one block picks a key window [lo, hi] around the
//            target rank with a margin of 1.25% of the ranks on each side
//            (the sampled rank's error is about 0.3% rms on cubelets of 4 M
//            keys);
//   collect  one streaming read of the coefficients by SEL_GRID blocks:
//            counts the keys below the window, appends the keys inside it
//            to a candidate list (buffered per wave in LDS, one global
//            atomic per flush; no block barrier in the streaming loop) and
//            writes per-block partials (count below, min, max, max |x|,
//            zeros).  Zero keys are only counted, never listed: exact zeros
//            are the one tie of any size in wavelet coefficients (flat
//            regions of integer data), and a window that starts at 0 would
//            otherwise list all of them;
//   radix    three rounds over the candidates, relative to the window:
//            each splits [base, base + 2^bits) into up to 2048 buckets (LDS
//            histograms merged with one atomic per bin).  Every block of a
//            round first derives the plan itself from the state, so no
//            launch needs a grid-wide synchronisation: round 0 reduces the
//            partials and verifies the window (the rank falls inside it when
//            below <= idx < below + candidates, else the exact fallback
//            takes every key from the full 31-bit range); a later round
//            takes the plan the round before it ran (stored in the state by
//            that round's block 0) and narrows it to the rank's bucket.  A
//            window of 20-24 bits takes two or three rounds, the fallback's
//            31 bits three; a round with no bits left returns at once;
//   finish   one block derives the final plan and writes the result record.
// Seven launches (gather, sample, collect, three rounds, finish);
// everything stays on the device.
#include <mutex>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define SEL_SAMPLE 16384
#define SEL_MARGIN 0.0125f
#define SEL_ITEMS 16
#define SEL_CHUNK (WV_BLOCK * SEL_ITEMS)
#define SEL_GRID 2048       // collect blocks (each streams n / SEL_GRID keys)
#define SEL_WAVE_CAP 512    // LDS candidate buffer per collect wave
#define SEL_RGRID 256       // blocks of a radix round
#define SEL_RBLOCK 1024     // threads of a radix round block

// The plan of a radix round (see sel_plan).
struct SelPlan {
  uint32_t below, min_ord, max_ord, max_key, zwin, crank, base, bits, ncand;
  int fallback;
};

struct SelState {
  int32_t idx;          // target rank
  int32_t special;      // 1: frac >= 1 (threshold = max |x|)
  uint32_t key_lo, key_hi;
  uint32_t below;       // keys < key_lo
  uint32_t ncand;
  int32_t fallback;
  uint32_t crank;       // rank inside the candidate list
  uint32_t base, bits;  // the key lies in [base, base + 2^bits) (radix rounds)
  uint32_t zwin;       // zero keys inside the window (counted, not listed)
  uint32_t min_ord, max_ord, max_key;
  uint32_t overflow;    // the candidate list outgrew its capacity
  uint32_t pad;
  uint32_t rhist[3][WV_H1_BINS];  // the radix rounds' digit histograms
  SelPlan splan[3];               // the plan each round ran (written by its block 0)
  uint32_t skeys[SEL_SAMPLE];     // the sampled keys (plain path: gathered by many blocks)
  // per collect block: below, min, max, max key, zeros, candidates
  uint32_t part[SEL_GRID][6];
};

int64_t threshold_ws_bytes(int64_t n) {
  return (int64_t)((sizeof(SelState) + 255) & ~(size_t)255) + n * 4 + 256;
}

// ---- statistics block (API: wv_stats_clear / wv_stats_compute) ----------
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

// full 2048-bin histogram of bits 30..20 of |x| plus extremes; grid-stride
__global__ void __launch_bounds__(WV_BLOCK)
hist1_kernel(const float *__restrict__ coef, int64_t n, uint32_t *hist,
             uint32_t *min_ord, uint32_t *max_ord, uint32_t *max_key,
             const int32_t *gate) {
  if (gate && !*gate) return;
  __shared__ uint32_t h[WV_H1_BINS];
  for (int i = threadIdx.x; i < WV_H1_BINS; i += WV_BLOCK) h[i] = 0;
  __syncthreads();
  uint32_t mn = 0xffffffffu, mx = 0, mk = 0;
  for (int64_t base = (int64_t)blockIdx.x * WV_BLOCK * 8; base < n;
       base += (int64_t)gridDim.x * WV_BLOCK * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int64_t i = base + (int64_t)u * WV_BLOCK + threadIdx.x;
      v[u] = i < n ? coef[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int64_t i = base + (int64_t)u * WV_BLOCK + threadIdx.x;
      const bool ok = i < n;
      const uint32_t key = __builtin_bit_cast(uint32_t, v[u]) & 0x7fffffffu;
      if (ok) {
        const uint32_t o = wv_f2ord(v[u]);
        mn = min(mn, o);
        mx = max(mx, o);
        mk = max(mk, key);
      }
      wv_wave_hist_add<WV_H1_BITS>(h, key >> WV_H1_SHIFT, ok);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    mk = max(mk, (uint32_t)__shfl_xor((int)mk, off));
  }
  if ((threadIdx.x & 63) == 0 && min_ord) {
    atomicMin(min_ord, mn);
    atomicMax(max_ord, mx);
    atomicMax(max_key, mk);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < WV_H1_BINS; i += WV_BLOCK)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

hipError_t stats_compute(const float *coef, int64_t n, void *stats,
                         hipStream_t s) {
  hipError_t e = stats_clear(stats, s);
  if (e != hipSuccess) return e;
  WvStats *st = (WvStats *)stats;
  int64_t blocks = (n + WV_BLOCK * 8 - 1) / (WV_BLOCK * 8);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(hist1_kernel, dim3((unsigned)blocks), dim3(WV_BLOCK), 0, s,
                     coef, n, st->hist, &st->min_ord, &st->max_ord, &st->max_key,
                     (const int32_t *)nullptr);
  return hipGetLastError();
}

// Block-wide search (1024 threads): the bin b of cnt[0..NB) with
// cum(b) <= rank < cum(b) + cnt(b).  Writes *out_b / *out_rank.
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
  for (int off = 1; off < 1024; off <<= 1) {   // inclusive scan
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
  __syncthreads();
}

// The same search for any block size (a multiple of 64, at most 1024).
template <int NB>
__device__ void find_bucket_blk(const uint32_t *cnt, uint32_t rank, uint32_t *out_b,
                                uint32_t *out_rank) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x, nt = blockDim.x;
  const int per = (NB + nt - 1) / nt;
  uint32_t sum = 0;
  for (int k = 0; k < per; k++) {
    const int b = t * per + k;
    sum += b < NB ? cnt[b] : 0u;
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {   // inclusive scan
    const uint32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t before = t ? part[t - 1] : 0;
  for (int k = 0; k < per; k++) {
    const int b = t * per + k;
    const uint32_t c = b < NB ? cnt[b] : 0u;
    if (rank >= before && rank < before + c) {
      *out_b = (uint32_t)b;
      *out_rank = rank - before;
    }
    before += c;
  }
  __syncthreads();
}

// ---- 1. sample -------------------------------------------------------------
// 32-bit integer mixer (xor-shift / multiply rounds with the "lowbias32"
// constants published by C. Wellons, hash-prospector)
__device__ inline uint32_t sample_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// position of sampled key i of n: one per stratum [i n / m, (i + 1) n / m)
// at a hashed offset, the hash scaled to the stratum's width (multiply-shift)
__device__ inline int64_t sample_pos(int i, int64_t n) {
  const int64_t lo = (int64_t)i * n / SEL_SAMPLE, hi = (int64_t)(i + 1) * n / SEL_SAMPLE;
  const uint64_t w = (uint64_t)(hi - lo);
  return lo + (int64_t)(((uint64_t)sample_hash((uint32_t)i) * w) >> 32);
}

// The plain path's gather of the sample (n > SEL_SAMPLE): one key per
// thread over SEL_SAMPLE / 256 blocks, so the scattered loads are in flight
// from many CUs at once rather than queued behind one.
__global__ void __launch_bounds__(WV_BLOCK)
sel_gather(const float *__restrict__ coef, int64_t n, SelState *S) {
  const int i = blockIdx.x * WV_BLOCK + threadIdx.x;
  if (i < SEL_SAMPLE) S->skeys[i] = __builtin_bit_cast(uint32_t, coef[sample_pos(i, n)]) & 0x7fffffffu;
}

// pre: the gathered sample (sel_gather), or null to gather it here
__device__ void sel_sample_dev(const float *__restrict__ coef, int64_t n, float frac,
                               SelState *S, const uint32_t *pre) {
  __shared__ uint32_t keys[SEL_SAMPLE];
  __shared__ uint32_t h[WV_H1_BINS];
  __shared__ uint32_t res_b[2], res_r[2];
  const int t = threadIdx.x;
  // (int)(compRatio * count) evaluated in float (thresh_cpu.cc:96)
  const float prod = frac * (float)n;
  int64_t idx = (int64_t)prod;
  if (idx < 0) idx = 0;
  const bool special = frac >= 1.0f || idx >= n;
  // sample: all keys when n is small, else one key per stratum
  // [i n / m, (i + 1) n / m) at a hashed offset
  const int m = n <= SEL_SAMPLE ? (int)n : SEL_SAMPLE;
  if (n <= SEL_SAMPLE) {
    for (int i = t; i < m; i += 1024)
      keys[i] = __builtin_bit_cast(uint32_t, coef[i]) & 0x7fffffffu;
  } else if (pre) {
    for (int i = t; i < SEL_SAMPLE; i += 1024) keys[i] = pre[i];
  } else {
    // all 16 loads of a thread issued before any is used
#pragma unroll
    for (int k = 0; k < SEL_SAMPLE / 1024; k++) {
      const int i = t + k * 1024;
      keys[i] = __builtin_bit_cast(uint32_t, coef[sample_pos(i, n)]) & 0x7fffffffu;
    }
  }
  for (int i = t; i < WV_H1_BINS; i += 1024) h[i] = 0;
  if (t < 2) { res_b[t] = 0; res_r[t] = 0; }
  __syncthreads();
  for (int i = t; i < m; i += 1024) atomicAdd(&h[keys[i] >> WV_H1_SHIFT], 1u);
  __syncthreads();
  // sample ranks of the window ends
  const float p = (float)idx / (float)n;
  const float plo = p - SEL_MARGIN, phi = p + SEL_MARGIN;
  const int r_lo = plo <= 0.f ? -1 : (int)(plo * m);
  const int r_hi = phi >= 1.f ? m : min(m - 1, (int)(phi * m) + 1);
  uint32_t lo_b = 0, lo_r = 0, hi_b = 0, hi_r = 0;
  if (r_lo >= 0) {
    find_bucket<WV_H1_BINS>(h, (uint32_t)r_lo, &res_b[0], &res_r[0]);
    lo_b = res_b[0]; lo_r = res_r[0];
  }
  if (r_hi < m) {
    find_bucket<WV_H1_BINS>(h, (uint32_t)r_hi, &res_b[1], &res_r[1]);
    hi_b = res_b[1]; hi_r = res_r[1];
  }
  // refine each end to a digit-2 sub-bucket with the samples of its bucket
  uint32_t key_lo = 0, key_hi = 0x7fffffffu;
  for (int end = 0; end < 2; end++) {
    const bool use = end == 0 ? r_lo >= 0 : r_hi < m;
    if (!use) continue;
    const uint32_t b = end == 0 ? lo_b : hi_b, r = end == 0 ? lo_r : hi_r;
    __syncthreads();
    for (int i = t; i < WV_H2_BINS; i += 1024) h[i] = 0;
    __syncthreads();
    for (int i = t; i < m; i += 1024)
      if ((keys[i] >> WV_H1_SHIFT) == b)
        atomicAdd(&h[(keys[i] >> WV_H2_SHIFT) & (WV_H2_BINS - 1)], 1u);
    __syncthreads();
    find_bucket<WV_H2_BINS>(h, r, &res_b[end], &res_r[end]);
    const uint32_t sub = res_b[end];
    if (end == 0) key_lo = (b << WV_H1_SHIFT) | (sub << WV_H2_SHIFT);
    else key_hi = (b << WV_H1_SHIFT) | (sub << WV_H2_SHIFT) | ((1u << WV_H2_SHIFT) - 1);
  }
  for (int i = t; i < 3 * WV_H1_BINS; i += 1024) (&S->rhist[0][0])[i] = 0;
  if (t == 0) {
    S->idx = (int32_t)(special ? n : idx);
    S->special = special;
    S->key_lo = key_lo;
    S->key_hi = key_hi;
    S->below = 0;
    S->ncand = 0;
    S->fallback = 0;
    S->crank = 0;
    S->zwin = 0;
    // the radix rounds start from the window: bits covers key_hi - key_lo
    S->base = key_lo;
    S->bits = key_hi > key_lo ? 32u - (uint32_t)__clz((int)(key_hi - key_lo)) : 0u;
    S->min_ord = 0xffffffffu;
    S->max_ord = 0u;
    S->max_key = 0u;
    S->overflow = 0;
  }
}

__global__ void __launch_bounds__(1024)
sel_sample(const float *__restrict__ coef, int64_t n, float frac, SelState *S) {
  sel_sample_dev(coef, n, frac, S, n > SEL_SAMPLE ? S->skeys : nullptr);
}

// ---- 2. collect ------------------------------------------------------------
// nblk blocks, each streaming a contiguous range in sub-chunks of SEL_CHUNK
// keys (16-byte loads, 16 keys per thread; the next chunk's loads are in
// flight while the current one is counted).  Each wave appends its
// in-window keys to its own LDS buffer and flushes it to the candidate
// list when it is nearly full, so the streaming loop has no block barrier.
// Per-block results go to S->part (no same-address global atomics per
// element).
__device__ inline void sel_load_full(const float *__restrict__ coef, int64_t c, int tid,
                                     float *v) {
  const WV_GLOBAL wv_f4 *c4 = (const WV_GLOBAL wv_f4 *)(gptr(coef) + c * SEL_CHUNK);
#pragma unroll
  for (int u = 0; u < SEL_ITEMS / 4; u++) {
    const wv_f4 q = c4[tid + u * WV_BLOCK];
    v[4 * u] = q.x; v[4 * u + 1] = q.y; v[4 * u + 2] = q.z; v[4 * u + 3] = q.w;
  }
}

// the index of element u of thread tid in a chunk (the order of the loads)
__device__ inline int64_t sel_elem_index(int64_t c, int tid, int u) {
  return c * SEL_CHUNK + (int64_t)(u >> 2) * 4 * WV_BLOCK + 4 * tid + (u & 3);
}

// A wave's buffered candidates to the candidate list (wave-uniform call).
// SEG (one select per launch): the block's own segment of the list, which
// starts at the block's first sample and cannot overflow (a block has no
// more candidates than samples); the offset comes from a block counter in
// LDS, so the collect takes no global atomic.  Otherwise (batches: many
// selects share the launch, each list holding cap keys) one global atomic
// per flush reserves the space.
template <bool SEG>
__device__ inline void sel_flush(uint32_t *buf, uint32_t &wcnt, SelState *S, uint32_t *cand,
                                 uint32_t cap, int lane, uint32_t *seg_cnt, int64_t seg0) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t base = 0;
  if (lane == 0) {
    if (SEG) {
      base = atomicAdd(seg_cnt, wcnt);
    } else {
      base = atomicAdd(&S->ncand, wcnt);
      if (base + wcnt > cap) S->overflow = 1;
    }
  }
  base = __shfl(base, 0);
  if (SEG)
    for (uint32_t i = lane; i < wcnt; i += 64) gptr(cand)[seg0 + base + i] = buf[i];
  else if (base + wcnt <= cap)
    for (uint32_t i = lane; i < wcnt; i += 64) gptr(cand)[base + i] = buf[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  wcnt = 0;
}

template <bool SEG>
__device__ __attribute__((always_inline)) void
sel_collect_dev(const float *__restrict__ coef, int64_t n, SelState *S, uint32_t *cand,
                uint32_t cap, int bidx, int nblk) {
  __shared__ uint32_t wbuf[WV_BLOCK / 64][SEL_WAVE_CAP];
  __shared__ uint32_t red[3][WV_BLOCK / 64];
  __shared__ uint32_t seg_cnt;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t *buf = wbuf[wave];
  uint32_t wcnt = 0;   // wave-uniform
  const uint32_t klo = S->key_lo, khi = S->key_hi;
  // One count serves both the keys below the window and the zeros: keys
  // below cut = max(klo, 1) are counted.  With klo > 0 that is the keys
  // below the window (zeros included); with klo = 0 nothing lies below and
  // the count is of the zeros, which the window holds but never lists.
  const uint32_t cut = klo > 1u ? klo : 1u;
  // listed: nonzero keys in the window, key - wlo <= span (unsigned)
  const uint32_t wlo = khi >= cut ? cut : 0x80000000u;   // empty: no 31-bit key
  const uint32_t span = khi >= cut ? khi - cut : 0u;
  const int64_t nchunk = (n + SEL_CHUNK - 1) / SEL_CHUNK, nfull = n / SEL_CHUNK;
  const int64_t c0 = nchunk * bidx / nblk;
  const int64_t c1 = nchunk * (bidx + 1) / nblk;
  const int64_t seg0 = c0 * SEL_CHUNK;
  if (SEG) {
    if (tid == 0) seg_cnt = 0;
    __syncthreads();
  }
  uint32_t cnt = 0, mn = 0xffffffffu, mx = 0;
  auto elem = [&](float x, bool valid) {
    const uint32_t bits = __builtin_bit_cast(uint32_t, x);
    const uint32_t key = bits & 0x7fffffffu;
    if (valid) {
      const uint32_t o = wv_f2ord(x);
      mn = min(mn, o);
      mx = max(mx, o);
      cnt += key < cut;
    }
    const bool hit = valid && key - wlo <= span;
    const uint64_t m = __ballot(hit);
    if (m) {
      if (hit) buf[wcnt + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = key;
      wcnt += (uint32_t)__popcll(m);
    }
  };
  // the buffer is checked once per four elements (at most 256 new keys),
  // which keeps the flush code out of the per-element path
  auto check = [&]() {
    if (wcnt > SEL_WAVE_CAP - 256) sel_flush<SEG>(buf, wcnt, S, cand, cap, lane, &seg_cnt, seg0);
  };
  // full chunks, two register sets taking turns; the prefetch index is
  // clamped to the block's last full chunk (a re-read of a chunk just read,
  // an L2 hit), so it needs no branch
  const int64_t cend = c1 < nfull ? c1 : nfull;
  if (c0 < cend) {
    float A[SEL_ITEMS], B[SEL_ITEMS];
    sel_load_full(coef, c0, tid, A);
    for (int64_t c = c0;;) {
      sel_load_full(coef, c + 1 < cend ? c + 1 : cend - 1, tid, B);
#pragma unroll
      for (int u = 0; u < SEL_ITEMS; u++) {
        elem(A[u], true);
        if ((u & 3) == 3) check();
      }
      if (++c >= cend) break;
      sel_load_full(coef, c + 1 < cend ? c + 1 : cend - 1, tid, A);
#pragma unroll
      for (int u = 0; u < SEL_ITEMS; u++) {
        elem(B[u], true);
        if ((u & 3) == 3) check();
      }
      if (++c >= cend) break;
    }
  }
  if (c1 > nfull) {
    // the partial last chunk
    float v[SEL_ITEMS];
#pragma unroll
    for (int u = 0; u < SEL_ITEMS; u++) {
      const int64_t i = sel_elem_index(nfull, tid, u);
      v[u] = i < n ? gptr(coef)[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < SEL_ITEMS; u++) {
      elem(v[u], sel_elem_index(nfull, tid, u) < n);
      if ((u & 3) == 3) check();
    }
  }
  if (wcnt) sel_flush<SEG>(buf, wcnt, S, cand, cap, lane, &seg_cnt, seg0);
  for (int off = 32; off > 0; off >>= 1) {
    cnt += (uint32_t)__shfl_xor((int)cnt, off);
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
  }
  if (lane == 0) { red[0][wave] = cnt; red[1][wave] = mn; red[2][wave] = mx; }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < WV_BLOCK / 64; w++) {
      cnt += red[0][w];
      mn = min(mn, red[1][w]);
      mx = max(mx, red[2][w]);
    }
    // the largest |x| is at one of the ends of the value order: the most
    // negative or the most positive value
    uint32_t mk = 0;
    if (mn <= mx) {
      mk = __builtin_bit_cast(uint32_t, wv_ord2f(mn)) & 0x7fffffffu;
      mk = max(mk, __builtin_bit_cast(uint32_t, wv_ord2f(mx)) & 0x7fffffffu);
    }
    S->part[bidx][0] = klo > 0 ? cnt : 0u;   // keys below the window
    S->part[bidx][1] = mn;
    S->part[bidx][2] = mx;
    S->part[bidx][3] = mk;
    S->part[bidx][4] = klo > 0 ? 0u : cnt;   // zeros inside the window
    S->part[bidx][5] = SEG ? seg_cnt : 0u;
  }
}

// ---- 3. verify -------------------------------------------------------------
__device__ void sel_verify_dev(SelState *S, int nparts) {
  __shared__ uint32_t r[5][16];
  const int t = threadIdx.x, nw = blockDim.x >> 6;
  uint32_t nb = 0, mn = 0xffffffffu, mx = 0, mk = 0, nz = 0;
  for (int i = t; i < nparts; i += blockDim.x) {
    nb += S->part[i][0];
    mn = min(mn, S->part[i][1]);
    mx = max(mx, S->part[i][2]);
    mk = max(mk, S->part[i][3]);
    nz += S->part[i][4];
  }
  for (int off = 32; off > 0; off >>= 1) {
    nb += (uint32_t)__shfl_xor((int)nb, off);
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    mk = max(mk, (uint32_t)__shfl_xor((int)mk, off));
    nz += (uint32_t)__shfl_xor((int)nz, off);
  }
  if ((t & 63) == 0) {
    r[0][t >> 6] = nb; r[1][t >> 6] = mn; r[2][t >> 6] = mx; r[3][t >> 6] = mk; r[4][t >> 6] = nz;
  }
  __syncthreads();
  if (t != 0) return;
  for (int w = 1; w < nw; w++) {
    nb += r[0][w];
    mn = min(mn, r[1][w]);
    mx = max(mx, r[2][w]);
    mk = max(mk, r[3][w]);
    nz += r[4][w];
  }
  S->below = nb;
  S->min_ord = mn;
  S->max_ord = mx;
  S->max_key = mk;
  // the zeros are in the window when it starts at 0 (else they are below)
  S->zwin = S->key_lo == 0 ? nz : 0u;
  if (S->special) { S->crank = 0; return; }
  const uint32_t idx = (uint32_t)S->idx;
  if (!S->overflow && S->below <= idx && idx < S->below + S->ncand + S->zwin) {
    S->crank = idx - S->below;
  } else {
    // exact fallback: the rank among all keys, from the full key range
    S->fallback = 1;
    S->ncand = 0;
    S->zwin = 0;
    S->crank = idx;
    S->base = 0;
    S->bits = 31;
  }
}

__global__ void __launch_bounds__(WV_BLOCK)
sel_collect(const float *__restrict__ coef, int64_t n, SelState *S, uint32_t *cand) {
  sel_collect_dev<true>(coef, n, S, cand, 0xffffffffu, blockIdx.x, gridDim.x);
}

// ---- 4. radix select -----------------------------------------------------------
// The key of rank crank lies in [base, base + 2^bits) among the candidates
// (or, after an exact fallback, among all keys, from base 0 and 31 bits).
// A round histograms the next d = min(bits, 11) bits of (key - base) over
// the keys in that range and narrows base / bits / crank to the rank's
// bucket.  The sampled window spans ~20 bits, so two rounds finish.
__device__ inline bool in_range(uint32_t key, uint32_t base, uint32_t bits) {
  return key >= base && (uint64_t)(key - base) < (1ull << bits);
}

__device__ inline uint32_t round_digit(uint32_t key, uint32_t base, uint32_t bits) {
  const uint32_t d = bits < 11u ? bits : 11u;
  return ((key - base) >> (bits - d)) & ((1u << d) - 1u);
}

__device__ void sel_result_dev(int64_t n, const SelState *S, wv_thresh_result *res) {
  wv_thresh_result o;
  // Traverse starts min at FLT_MAX and max at FLT_MIN (thresh_cpu.cc:67-68)
  const float mn = wv_ord2f(S->min_ord), mx = wv_ord2f(S->max_ord);
  const float fmax = 3.40282347e+38f, fmin = 1.17549435e-38f;
  o.min_val = mn < fmax ? mn : fmax;
  o.max_val = mx > fmin ? mx : fmin;
  o.max_abs = __builtin_bit_cast(float, S->max_key);
  if (S->special) {
    // frac >= 1: the reference GPU driver takes max |x|
    // (test_compress_gpu.cu:299-301)
    o.threshold = o.max_abs;
    o.threshold_key = S->max_key;
    o.thresh_idx = (int32_t)n;
    o.nonzero_count = 0;
  } else {
    const uint32_t key = S->base;
    float t = __builtin_bit_cast(float, key);
    o.threshold_key = key;
    if (t == 0.f) t = 1e-16f;    // thresh_cpu.cc:98
    o.threshold = t;
    o.thresh_idx = S->idx;
    o.nonzero_count = (int32_t)(n - S->idx - 1);
  }
  // candidates examined; the top bit records an exact-fallback run
  o.n_candidates = (int32_t)(S->ncand & 0x7fffffffu) |
                   (S->fallback ? (int32_t)0x80000000 : 0);
  *res = o;
}

// The plan every block of a round derives on its own, from what earlier
// launches left in the state, so no launch needs a grid-wide
// synchronisation (a device-scope release fence in every block costs an L2
// write-back).  Round 0 reduces the collect partials and checks the window;
// round r > 0 takes the plan round r - 1 ran (stored by that round's block
// 0) and narrows it to the rank's bucket in round r - 1's histogram.  The
// plan is deterministic, so every block agrees.  round = 3: the finish.
__device__ void sel_plan(SelState *S, int nparts, int round, SelPlan *pl) {
  __shared__ uint32_t r[6][16];
  __shared__ uint32_t rb, rr;
  const int t = threadIdx.x, nw = blockDim.x >> 6;
  if (round == 0) {
    uint32_t nb = 0, mn = 0xffffffffu, mx = 0, mk = 0, nz = 0, nc = 0;
    // every load issued before the first use
    constexpr int PER = SEL_GRID / SEL_RBLOCK;
    uint32_t v[PER][6];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = t + k * (int)blockDim.x;
#pragma unroll
      for (int f = 0; f < 6; f++) v[k][f] = i < nparts ? S->part[i][f] : 0u;
      if (i >= nparts) { v[k][1] = 0xffffffffu; }
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      nb += v[k][0];
      mn = min(mn, v[k][1]);
      mx = max(mx, v[k][2]);
      mk = max(mk, v[k][3]);
      nz += v[k][4];
      nc += v[k][5];
    }
    for (int i = t + PER * (int)blockDim.x; i < nparts; i += blockDim.x) {
      nb += S->part[i][0];
      mn = min(mn, S->part[i][1]);
      mx = max(mx, S->part[i][2]);
      mk = max(mk, S->part[i][3]);
      nz += S->part[i][4];
      nc += S->part[i][5];
    }
    for (int off = 32; off > 0; off >>= 1) {
      nb += (uint32_t)__shfl_xor((int)nb, off);
      mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
      mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
      mk = max(mk, (uint32_t)__shfl_xor((int)mk, off));
      nz += (uint32_t)__shfl_xor((int)nz, off);
      nc += (uint32_t)__shfl_xor((int)nc, off);
    }
    if ((t & 63) == 0) {
      r[0][t >> 6] = nb; r[1][t >> 6] = mn; r[2][t >> 6] = mx; r[3][t >> 6] = mk;
      r[4][t >> 6] = nz; r[5][t >> 6] = nc;
    }
    __syncthreads();
    if (t == 0) {
      for (int w = 1; w < nw; w++) {
        nb += r[0][w];
        mn = min(mn, r[1][w]);
        mx = max(mx, r[2][w]);
        mk = max(mk, r[3][w]);
        nz += r[4][w];
        nc += r[5][w];
      }
      pl->ncand = nc;
      pl->below = nb;
      pl->min_ord = mn;
      pl->max_ord = mx;
      pl->max_key = mk;
      // the zeros are in the window when it starts at 0 (else they are below)
      pl->zwin = S->key_lo == 0 ? nz : 0u;
      const uint32_t idx = (uint32_t)S->idx;
      if (S->special) {
        pl->fallback = 0;
        pl->crank = 0;
        pl->base = S->base;
        pl->bits = 0;
      } else if (nb <= idx && idx < nb + nc + pl->zwin) {
        pl->fallback = 0;
        pl->crank = idx - nb;
        pl->base = S->base;
        pl->bits = S->bits;
      } else {
        // exact fallback: the rank among all keys, from the full key range
        pl->fallback = 1;
        pl->zwin = 0;
        pl->crank = idx;
        pl->base = 0;
        pl->bits = 31;
      }
    }
    __syncthreads();
  } else {
    if (t == 0) *pl = S->splan[round - 1];
    __syncthreads();
    if (pl->bits != 0) {
      if (t == 0) { rb = 0; rr = 0; }
      __syncthreads();
      find_bucket_blk<WV_H1_BINS>(S->rhist[round - 1], pl->crank, &rb, &rr);
      if (t == 0) {
        const uint32_t bits = pl->bits;
        const uint32_t d = bits < 11u ? bits : 11u;
        pl->base += rb << (bits - d);
        pl->bits = bits - d;
        pl->crank = rr;
      }
      __syncthreads();
    }
  }
  if (round < 3 && blockIdx.x == 0 && t == 0) S->splan[round] = *pl;
}

// Round k over SEL_RGRID blocks: the next min(bits, 11) bits of (key - base)
// of the keys in [base, base + 2^bits) -- the candidates, or after an exact
// fallback all n coefficients -- histogrammed in LDS and merged into
// S->rhist[k] with one atomic per non-empty bin.  A round with no bits left
// returns at once.
// histogram of the keys src[a, e) in [base, base + 2^bits) into LDS h
__device__ inline void sel_hist_range(const uint32_t *src, int64_t a, int64_t e, uint32_t base,
                                      uint32_t bits, uint32_t *h) {
  const WV_GLOBAL uint32_t *g = gptr(src);
  for (int64_t i0 = a; i0 < e; i0 += (int64_t)blockDim.x * 4) {
    uint32_t kk[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + (int64_t)u * blockDim.x + threadIdx.x;
      kk[u] = i < e ? g[i] & 0x7fffffffu : 0xffffffffu;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t key = kk[u];
      if (key != 0xffffffffu && in_range(key, base, bits))
        atomicAdd(&h[round_digit(key, base, bits)], 1u);
    }
  }
}

// The histogram of the listed keys of a round block's segments (segments
// sg = blockIdx.x + k * gridDim.x), walked as one run of keys so that every
// thread has SEL_ITEMS loads in flight at a time across segment ends: the
// segment lengths are loaded first, then position v of the run maps to
// segment k with pre[k] <= v < pre[k + 1].
#define SEL_RSEG ((SEL_GRID + SEL_RGRID - 1) / SEL_RGRID)
__device__ inline void sel_hist_segments(const uint32_t *cand, const SelState *S, int64_t n,
                                         int nparts, uint32_t base, uint32_t bits,
                                         uint32_t *h) {
  const int64_t nchunk = (n + SEL_CHUNK - 1) / SEL_CHUNK;
  int64_t off[SEL_RSEG];    // segment start minus its position in the run
  uint32_t pre[SEL_RSEG + 1];
  uint32_t len[SEL_RSEG];
#pragma unroll
  for (int k = 0; k < SEL_RSEG; k++) {
    const int sg = blockIdx.x + k * gridDim.x;
    len[k] = sg < nparts ? S->part[sg][5] : 0u;
  }
  pre[0] = 0;
#pragma unroll
  for (int k = 0; k < SEL_RSEG; k++) {
    const int sg = blockIdx.x + k * gridDim.x;
    pre[k + 1] = pre[k] + len[k];
    off[k] = (sg < nparts ? nchunk * sg / nparts * SEL_CHUNK : 0) - (int64_t)pre[k];
  }
  const uint32_t total = pre[SEL_RSEG];
  const WV_GLOBAL uint32_t *g = gptr(cand);
  for (uint32_t i0 = 0; i0 < total; i0 += SEL_RBLOCK * SEL_ITEMS) {
    uint32_t kk[SEL_ITEMS];
#pragma unroll
    for (int u = 0; u < SEL_ITEMS; u++) {
      const uint32_t v = i0 + u * SEL_RBLOCK + threadIdx.x;
      int64_t o = off[0];
#pragma unroll
      for (int k = 1; k < SEL_RSEG; k++) o = v >= pre[k] ? off[k] : o;
      kk[u] = v < total ? g[o + v] & 0x7fffffffu : 0xffffffffu;
    }
#pragma unroll
    for (int u = 0; u < SEL_ITEMS; u++) {
      const uint32_t key = kk[u];
      if (key != 0xffffffffu && in_range(key, base, bits))
        atomicAdd(&h[round_digit(key, base, bits)], 1u);
    }
  }
}

__global__ void __launch_bounds__(SEL_RBLOCK)
sel_round(const uint32_t *cand, const float *coef, int64_t n, SelState *S, int nparts,
          int round) {
  __shared__ uint32_t h[WV_H1_BINS];
  __shared__ SelPlan pl;
  sel_plan(S, nparts, round, &pl);
  if (S->special || pl.bits == 0) return;
  for (int i = threadIdx.x; i < WV_H1_BINS; i += SEL_RBLOCK) h[i] = 0;
  __syncthreads();
  const uint32_t base = pl.base, bits = pl.bits;
  if (pl.fallback) {
    // every coefficient, split evenly over the blocks
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t a = (int64_t)blockIdx.x * per;
    sel_hist_range((const uint32_t *)coef, a, min(n, a + per), base, bits, h);
  } else {
    // the collect blocks' segments, a share of them per block
    sel_hist_segments(cand, S, n, nparts, base, bits, h);
  }
  // the listed keys exclude the window's zeros: block 0 adds them
  if (threadIdx.x == 0 && blockIdx.x == 0 && pl.zwin && in_range(0u, base, bits))
    atomicAdd(&h[round_digit(0u, base, bits)], pl.zwin);
  __syncthreads();
  for (int i = threadIdx.x; i < WV_H1_BINS; i += SEL_RBLOCK)
    if (h[i]) atomicAdd(&S->rhist[round][i], h[i]);
}

// one block: the plan after the three rounds -> the state and the result
__global__ void __launch_bounds__(WV_BLOCK)
sel_finish(int64_t n, SelState *S, int nparts, wv_thresh_result *res) {
  __shared__ SelPlan pl;
  sel_plan(S, nparts, 3, &pl);
  if (threadIdx.x == 0) {
    S->below = pl.below;
    S->min_ord = pl.min_ord;
    S->max_ord = pl.max_ord;
    S->max_key = pl.max_key;
    S->zwin = pl.zwin;
    S->fallback = pl.fallback;
    S->ncand = pl.fallback ? 0u : pl.ncand;
    S->crank = pl.crank;
    S->base = pl.base;
    S->bits = pl.bits;
    sel_result_dev(n, S, res);
  }
}

static_assert(SEL_RGRID * SEL_RSEG >= SEL_GRID, "a round must cover every collect segment");

hipError_t threshold_select(const float *coef, int64_t n, float frac,
                            const void *stats, void *ws,
                            wv_thresh_result *res, hipStream_t s) {
  (void)stats;
  SelState *S = (SelState *)ws;
  uint32_t *cand = (uint32_t *)((char *)ws + ((sizeof(SelState) + 255) & ~(size_t)255));
  if (n > SEL_SAMPLE)
    hipLaunchKernelGGL(sel_gather, dim3(SEL_SAMPLE / WV_BLOCK), dim3(WV_BLOCK), 0, s, coef, n, S);
  hipLaunchKernelGGL(sel_sample, dim3(1), dim3(1024), 0, s, coef, n, frac, S);
  const int64_t nchunk = (n + SEL_CHUNK - 1) / SEL_CHUNK;
  const int grid = (int)(nchunk < SEL_GRID ? nchunk : SEL_GRID);
  hipLaunchKernelGGL(sel_collect, dim3(grid), dim3(WV_BLOCK), 0, s, coef, n, S, cand);
  // three rounds cover any window (11 + 11 + 11 >= 31 bits)
  for (int r = 0; r < 3; r++)
    hipLaunchKernelGGL(sel_round, dim3(SEL_RGRID), dim3(SEL_RBLOCK), 0, s, cand, coef, n, S, grid,
                       r);
  hipLaunchKernelGGL(sel_finish, dim3(1), dim3(WV_BLOCK), 0, s, n, S, grid, res);
  return hipGetLastError();
}

// ---- batches: one select per cubelet, all cubelets in three launches -----
// Block b of the collect launch serves the last job j with blk0[j] <= b.
__device__ inline int find_job(const int *blk0, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(1024) bsel_sample(const BSelJob *J) {
  const BSelJob &j = J[blockIdx.x];
  sel_sample_dev(j.coef, j.n, j.frac, (SelState *)j.state, nullptr);
}

__global__ void __launch_bounds__(WV_BLOCK)
bsel_collect(const BSelJob *J, const int *blk0, int njobs) {
  const int jb = find_job(blk0, njobs, blockIdx.x);
  const BSelJob &j = J[jb];
  sel_collect_dev<false>(j.coef, j.n, (SelState *)j.state, j.cand, j.cap,
                  (int)blockIdx.x - j.blk0, j.nblk);
}

// Verify the window, then a three-digit radix select inside the block over
// the candidates -- or, when the window missed the rank or the candidate
// list overflowed, over every key of the cubelet -- and the result record.
__global__ void __launch_bounds__(1024) bsel_finish(const BSelJob *J) {
  __shared__ uint32_t h[WV_H1_BINS];
  __shared__ uint32_t rb, rr;
  const BSelJob &j = J[blockIdx.x];
  SelState *S = (SelState *)j.state;
  sel_verify_dev(S, j.nblk);
  __syncthreads();
  if (!S->special) {
    const bool fb = S->fallback != 0;
    const uint32_t *src = fb ? (const uint32_t *)j.coef : j.cand;
    const int64_t cnt = fb ? j.n : (int64_t)S->ncand;
    uint32_t rank = fb ? (uint32_t)S->idx : S->crank;
    uint32_t base = fb ? 0u : S->base, bits = fb ? 31u : S->bits;
    const uint32_t zwin = fb ? 0u : S->zwin;   // zeros of the window, not listed
    while (bits > 0) {
      for (int i = threadIdx.x; i < WV_H1_BINS; i += 1024) h[i] = 0;
      if (threadIdx.x == 0) { rb = 0; rr = 0; }
      __syncthreads();
      if (threadIdx.x == 0 && zwin && in_range(0u, base, bits))
        atomicAdd(&h[round_digit(0u, base, bits)], zwin);
      for (int64_t i = threadIdx.x; i < cnt; i += 1024) {
        const uint32_t key = src[i] & 0x7fffffffu;
        if (in_range(key, base, bits)) atomicAdd(&h[round_digit(key, base, bits)], 1u);
      }
      __syncthreads();
      find_bucket<WV_H1_BINS>(h, rank, &rb, &rr);
      const uint32_t d = bits < 11u ? bits : 11u;
      base += rb << (bits - d);
      bits -= d;
      rank = rr;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      S->base = base;
      S->bits = 0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) sel_result_dev(j.n, S, j.res);
}

// ---- the allocation-free threshold entry point ------------------------------
// Its workspace (the select state and a candidate list of up to
// WV_NOALLOC_MAX_N keys: the layout threshold_ws_bytes describes) is this
// file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes the array's name to a kernel: on the host it is a
// stand-in, not a device address.  Each device's address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
constexpr size_t kNoallocScratchBytes =
    ((sizeof(SelState) + 255) & ~(size_t)255) + (size_t)WV_NOALLOC_MAX_N * 4 + 256;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];   // per device; null until looked up
}  // namespace

int64_t sel_state_bytes() { return (int64_t)((sizeof(SelState) + 255) & ~(size_t)255); }
int sel_max_collect_blocks() { return SEL_GRID; }
int64_t sel_collect_chunk() { return SEL_CHUNK; }

hipError_t threshold_select_batch(const BSelJob *jobs_d, const int *blk0_d, int njobs,
                                  int total_blocks, hipStream_t s) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(bsel_sample, dim3(njobs), dim3(1024), 0, s, jobs_d);
  hipLaunchKernelGGL(bsel_collect, dim3(total_blocks), dim3(WV_BLOCK), 0, s, jobs_d, blk0_d,
                     njobs);
  hipLaunchKernelGGL(bsel_finish, dim3(njobs), dim3(1024), 0, s, jobs_d);
  return hipGetLastError();
}

// The current device's copy of the scratch array (looked up on the first
// call on each device).
static int noalloc_scratch(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("wv_threshold_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("wv_threshold_noalloc: device %d beyond %d", dev, kMaxDevices);
    return WV_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("wv_threshold_noalloc: scratch address: %s", hipGetErrorString(e));
      return WV_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return WV_OK;
}

}  // namespace wv

extern "C" int wv_threshold_noalloc(const float *coef_d, int64_t n, float frac,
                                    wv_thresh_result *result_d, void *stream) {
  using namespace wv;
  if (!coef_d || !result_d || n <= 0) {
    set_error("wv_threshold_noalloc: bad arguments");
    return WV_ERR_ARG;
  }
  if (n > WV_NOALLOC_MAX_N) {
    set_error("wv_threshold_noalloc: %lld coefficients exceed the static scratch's %d "
              "(use wv_threshold with a workspace)", (long long)n, (int)WV_NOALLOC_MAX_N);
    return WV_ERR_ARG;
  }
  void *ws = nullptr;
  const int rc = noalloc_scratch(&ws);
  if (rc != WV_OK) return rc;
  // the sample kernel initialises every state field the later kernels read,
  // so nothing is cleared from the host
  const hipError_t e = threshold_select(coef_d, n, frac, nullptr, ws, result_d,
                                        (hipStream_t)stream);
  if (e != hipSuccess) {
    set_error("wv_threshold_noalloc: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  return WV_OK;
}
