// wv_lloyd.hip -- Lloyd-Max codebook training on the device with the
// semantics of the reference CPU path (computeLloydQuantization,
// test_compress_cpu.cc:404-459; lloyd, lloyds.cpp:30-141):
//
//  * points: the |x| ranked above the threshold index, ascending, without
//    the largest (sorted[idx+1 .. n-2]).  They are gathered as the keys
//    > threshold plus the threshold's own duplicates ranked after idx, then
//    sorted with an LSD radix sort (4 x 8-bit digits, stable block ranks).
//  * every assignment (nearest codebook with fabsf and strict '<', the
//    'between' partition scan) is computed per point in parallel, exactly;
//  * every float sum the reference accumulates sequentially (per-cell sums
//    and the signed distortion) is accumulated sequentially here too, in
//    point order, by one wavefront per chain: the chains are independent,
//    so the cells and the distortion run concurrently.  When a cell's
//    members are not contiguous in point order (only possible with an
//    unsorted codebook) the chain scans all points and adds its own.
//  * the loop, its stop rule (reldist > 1e-6f) and the codebook / partition
//    updates run on the device.  The distortion only decides when the loop
//    stops, so iterations run in speculative windows of eight: the
//    trajectory runs ahead, the eight distortion chains of a window run
//    side by side, and the first iteration that meets the stop rule is the
//    result.  The host polls once per window.
#include "wv_device.h"
#include "wv_internal.h"
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace wv {

#define LL_SORT_ITEMS 16            // keys per thread per sort block
#define LL_SORT_CHUNK (WV_BLOCK * LL_SORT_ITEMS)
#define LL_MAX_CS 16383           // codebook held in LDS by the assignment
#define LL_HARD_CAP 100000        // uncapped runs stop with an error here

struct LloydState {
  int32_t psize, cs;
  int32_t iters, done, max_iters, nonmono_b, nonmono_n;
  float stop;
  float dist, reldist, pmin, pmax;
  uint32_t n_gt;                    // keys > threshold gathered
  int32_t final_slot;               // window slot of the last iteration run
  int32_t pad[2];
};

// ---------------------------------------------------------------- gather
// keys > threshold, block-buffered in LDS so the global counter takes one
// atomic per 4096 coefficients
#define LL_GITEMS 16
#define LL_GCHUNK (WV_BLOCK * LL_GITEMS)
__global__ void __launch_bounds__(WV_BLOCK)
ll_gather(const float *coef, int64_t n, uint32_t tkey, uint32_t *out,
          LloydState *S) {
  __shared__ uint32_t buf[LL_GCHUNK];
  __shared__ uint32_t cnt, base;
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) cnt = 0;
  __syncthreads();
  const int64_t start = (int64_t)blockIdx.x * LL_GCHUNK;
#pragma unroll
  for (int u = 0; u < LL_GITEMS; u++) {
    const int64_t i = start + (int64_t)u * WV_BLOCK + tid;
    uint32_t key = 0;
    bool hit = false;
    if (i < n) {
      key = __builtin_bit_cast(uint32_t, coef[i]) & 0x7fffffffu;
      hit = key > tkey;
    }
    const uint64_t m = __ballot(hit);
    if (m) {
      uint32_t off = 0;
      if (lane == 0) off = atomicAdd(&cnt, (uint32_t)__popcll(m));
      off = __shfl(off, 0);
      if (hit) buf[off + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = key;
    }
  }
  __syncthreads();
  const uint32_t c = cnt;
  if (tid == 0 && c) base = atomicAdd(&S->n_gt, c);
  __syncthreads();
  for (uint32_t i = tid; i < c; i += WV_BLOCK) out[base + i] = buf[i];
}

// ---------------------------------------------------------------- sort
// per-block digit histograms, block-major [block][256]
__global__ void __launch_bounds__(WV_BLOCK)
ll_sort_hist(const uint32_t *keys, const uint32_t *count, int shift,
             uint32_t *hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = *count;
  const uint32_t base = blockIdx.x * LL_SORT_CHUNK;
  for (int k = 0; k < LL_SORT_ITEMS; k++) {
    const uint32_t i = base + k * WV_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1u);
  }
  __syncthreads();
  hist[blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

// exclusive offsets: off[b][d] = sum_{d'<d} total[d'] + sum_{b'<b} hist[b'][d]
__global__ void __launch_bounds__(256)
ll_sort_scan(uint32_t *hist, int nblocks) {
  __shared__ uint32_t tot[256];
  const int d = threadIdx.x;
  uint32_t run = 0;
  for (int b = 0; b < nblocks; b++) run += hist[b * 256 + d];
  tot[d] = run;
  __syncthreads();
  // exclusive scan over digits
  if (d == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < 256; k++) { uint32_t t = tot[k]; tot[k] = acc; acc += t; }
  }
  __syncthreads();
  uint32_t acc = tot[d];
  int b = 0;
  for (; b + 8 <= nblocks; b += 8) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = hist[(b + k) * 256 + d];
#pragma unroll
    for (int k = 0; k < 8; k++) { hist[(b + k) * 256 + d] = acc; acc += v[k]; }
  }
  for (; b < nblocks; b++) {
    const uint32_t v = hist[b * 256 + d];
    hist[b * 256 + d] = acc;
    acc += v;
  }
}

// stable scatter: keys of a block are ranked in index order
__global__ void __launch_bounds__(WV_BLOCK)
ll_sort_scatter(const uint32_t *keys, uint32_t *out, const uint32_t *count,
                int shift, const uint32_t *offs) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wave_cnt[WV_BLOCK / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  base[tid] = offs[blockIdx.x * 256 + tid];
  const uint32_t n = *count;
  const uint32_t start = blockIdx.x * LL_SORT_CHUNK;
  __syncthreads();
  for (int k = 0; k < LL_SORT_ITEMS; k++) {
    const uint32_t i = start + k * WV_BLOCK + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? keys[i] : 0;
    const uint32_t dig = (key >> shift) & 255;
    // lanes of this wave holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
      const uint64_t b = __ballot((dig >> bit) & 1);
      peers &= ((dig >> bit) & 1) ? b : ~b;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1));
    for (int j = lane; j < 256; j += 64) wave_cnt[wave][j] = 0;
    __syncthreads();
    if (valid && rank == 0) wave_cnt[wave][dig] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (int w = 0; w < wave; w++) before += wave_cnt[w][dig];
      out[base[dig] + before + rank] = key;
    }
    __syncthreads();
    // advance the per-digit bases by this round's totals
    {
      uint32_t t = 0;
      for (int w = 0; w < WV_BLOCK / 64; w++) t += wave_cnt[w][tid];
      base[tid] += t;
    }
    __syncthreads();
  }
}

int64_t radix_sort_hist_bytes(int64_t max_n) {
  return ((max_n + LL_SORT_CHUNK - 1) / LL_SORT_CHUNK + 1) * 256 * 4;
}

// LSD radix sort of the first *count_d (at most max_n) 32-bit keys: four
// stable passes of 8 bits (per-block digit histograms, one scan, a scatter
// that ranks a block's keys in index order).  After the even number of
// passes the sorted keys are back in keys; tmp holds max_n keys, hist
// radix_sort_hist_bytes(max_n).
hipError_t radix_sort_u32(uint32_t *keys, uint32_t *tmp, uint32_t *hist,
                          const uint32_t *count_d, int64_t max_n, hipStream_t s) {
  const int nblocks = (int)((max_n + LL_SORT_CHUNK - 1) / LL_SORT_CHUNK);
  uint32_t *src = keys, *dst = tmp;
  for (int pass = 0; pass < 4 && nblocks > 0; pass++) {
    const int shift = pass * 8;
    hipLaunchKernelGGL(ll_sort_hist, dim3(nblocks), dim3(WV_BLOCK), 0, s, src, count_d, shift,
                       hist);
    hipLaunchKernelGGL(ll_sort_scan, dim3(1), dim3(256), 0, s, hist, nblocks);
    hipLaunchKernelGGL(ll_sort_scatter, dim3(nblocks), dim3(WV_BLOCK), 0, s, src, dst, count_d,
                       shift, hist);
    std::swap(src, dst);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------- Lloyd
// nearest codebook entry: first minimum of fabsf(p - c[j]) (lloyds.cpp:50-60)
__device__ inline int nearest(const float *cb, int cs, float p) {
  float mn = fabsf(p - cb[0]);
  int idx = 0;
  for (int j = 1; j < cs; j++) {
    const float d = fabsf(p - cb[j]);
    if (d < mn) { idx = j; mn = d; }
  }
  return idx;
}

// first j with c[j] <= p < c[j+1], or -1 (lloyds.cpp:90-99)
__device__ inline int between(const float *cb, int cs, float p) {
  for (int j = 0; j < cs - 1; j++)
    if (p >= cb[j] && p < cb[j + 1]) return j;
  return -1;
}

// ---- one run's device pointers ------------------------------------------
// Kernels read their run from a table (blockIdx.y), so one launch serves
// one run or many: in the serving loop the Lloyd runs of all admitted
// requests share every launch of a window.
struct LJob {
  LloydState *S;
  const float *pts;
  int32_t *grp;
  float *cb, *part, *sum;
  uint32_t *count;
  int32_t *first, *last, *nruns;
  float *dsum;
  uint16_t *ring;      // LL_WINDOW slots of ring_stride 16-bit groups
  float *cbring;       // LL_WINDOW codebooks of cs + 1 floats
  float *wdsum;        // LL_WINDOW distortion sums
  int64_t ring_stride;
  int32_t cs, psize;
  int32_t nw;          // iterations of the current window for this run
  int32_t pad;
};

// k: the iteration of the current window (-1 outside the windows); a run
// whose window holds fewer iterations (its iteration cap) skips the rest
__device__ inline bool ll_skip(const LJob &j, int k) {
  return j.S->done || (k >= 0 && k >= j.nw);
}

// Assignment of every point and its run boundaries in one pass (ring: a
// 16-bit copy of the nearest groups, kept for the distortion chain of
// window iteration k, which runs later).  A block takes 256 consecutive
// points at a time; a point starts a run of its group when the point
// before it is in another group (or it is the first point) and ends one
// when the point after it is: the neighbours' groups come from LDS, and at
// the chunk's two ends they are computed again from the points themselves.
// A group whose members are contiguous has exactly one run.  The run
// arrays must be cleared first (ll_clear_runs, or the update kernels).
#define LL_NOGROUP (-2147483647 - 1)
template <bool NEAREST>
__global__ void __launch_bounds__(WV_BLOCK)
ll_assign(const LJob *J, int k) {
  extern __shared__ float cb[];
  __shared__ int sg[WV_BLOCK + 2];
  const LJob j = J[blockIdx.y];
  if (ll_skip(j, k)) return;
  const int cs = j.cs;
  for (int t = threadIdx.x; t < cs; t += WV_BLOCK) cb[t] = j.cb[t];
  __syncthreads();
  uint16_t *ring = NEAREST && k >= 0 ? j.ring + (int64_t)k * j.ring_stride : nullptr;
  const int psize = j.psize;
  auto group = [&](int i) {
    const float p = j.pts[i];
    return NEAREST ? nearest(cb, cs, p) : between(cb, cs, p);
  };
  const int tid = threadIdx.x;
  for (int base = blockIdx.x * WV_BLOCK; base < psize; base += gridDim.x * WV_BLOCK) {
    const int i = base + tid;
    int g = LL_NOGROUP;
    if (i < psize) {
      g = group(i);
      j.grp[i] = g;
      if (ring) ring[i] = (uint16_t)g;
    }
    sg[tid + 1] = g;
    if (tid == 0) sg[0] = base > 0 ? group(base - 1) : LL_NOGROUP;
    if (tid == WV_BLOCK - 1) sg[WV_BLOCK + 1] = base + WV_BLOCK < psize ? group(base + WV_BLOCK)
                                                                     : LL_NOGROUP;
    __syncthreads();
    if (i < psize && g >= 0) {
      if (sg[tid] != g) { j.first[g] = i; atomicAdd(&j.nruns[g], 1); }
      if (sg[tid + 2] != g) j.last[g] = i;
    }
    __syncthreads();
  }
}

// the run arrays the next assignment fills
__device__ inline void clear_runs(const LJob &j) {
  for (int t = threadIdx.x; t < j.cs; t += blockDim.x) {
    j.nruns[t] = 0;
    j.first[t] = 0;
    j.last[t] = -1;
  }
}

__global__ void ll_clear_runs(const LJob *J, int k) {
  const LJob j = J[blockIdx.y];
  if (ll_skip(j, k)) return;
  clear_runs(j);
}

// one wavefront per chain: chains 0..cs-1 are the cells, chain cs the
// signed distortion (only when DIST).  The float sums must follow the
// points' order exactly (lloyds.cpp:56-64,124-132), so each chain is one
// dependent add per point.
//
// A batch of 64 consecutive points is loaded one per lane (a coalesced
// vector load), and the chain then takes the batch's lanes in order: each
// point is read into a scalar register (v_readlane) and added with one
// dependent v_add_f32.  The reads do not depend on the sum, so they issue
// ahead of the adds and the chain costs one dependent add per point (4.6 ns
// per point measured, against 6.0 for a sum carried across the lanes with
// DPP wave_ror adds and 7.2 for wave-uniform scalar loads of 16 points).
// Lanes outside the chain's range add -0.0f, the identity of IEEE
// addition.  Loads run LL_CU batches ahead of the adds.
#define LL_CU 8

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float chain64(float acc, float t) {
#pragma unroll
  for (int k0 = 0; k0 < 64; k0 += 16) {
    float v[16];   // read ahead, so no add waits on the read before it
#pragma unroll
    for (int k = 0; k < 16; k++)
      v[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), k0 + k));
#pragma unroll
    for (int k = 0; k < 16; k++) acc += v[k];
  }
  return acc;
}

// s = 0; s += pts[i] for i in [a, b), a < b.  Two register sets take
// turns: the loads of the next LL_CU batches are in flight while the
// current ones are added.  Load indices are clamped into [a, b), so every
// load is in bounds (also the one issued past the end, whose values are
// never used) and the prefetch needs no branch; the lanes outside the
// range add -0.0f.
__device__ __forceinline__ float run_chain(const float *__restrict__ pts, int a, int b) {
  const int lane = threadIdx.x;
  constexpr int G = 64 * LL_CU;
  float acc = 0.f;
  float pA[LL_CU], pB[LL_CU];
  auto load = [&](float *pp, int j) {
#pragma unroll
    for (int u = 0; u < LL_CU; u++) pp[u] = pts[max(a, min(j + u * 64 + lane, b - 1))];
  };
  auto run = [&](const float *pp, int j) {
#pragma unroll
    for (int u = 0; u < LL_CU; u++) {
      const int i = j + u * 64 + lane;
      if (j + u * 64 < b) acc = chain64(acc, (i >= a && i < b) ? pp[u] : -0.f);
    }
  };
  int j = a & ~63;
  load(pA, j);
  while (true) {
    load(pB, j + G);
    run(pA, j);
    j += G;
    if (j >= b) break;
    load(pA, j + G);
    run(pB, j);
    j += G;
    if (j >= b) break;
  }
  return acc;
}

// d = 0; d += cb[g[i]] - pts[i] for i in [0, psize), psize >= 1, with the
// codebook in LDS (lcb): its reads count on their own counter, so they do
// not wait for the next batches' loads in flight.
template <typename G>
__device__ __forceinline__ float dist_chain(const float *__restrict__ pts, const G *__restrict__ g,
                                            const float *lcb, int psize) {
  const int lane = threadIdx.x;
  constexpr int GR = 64 * LL_CU;
  float acc = 0.f;
  float pA[LL_CU], pB[LL_CU];
  int qA[LL_CU], qB[LL_CU];
  auto load = [&](float *pp, int *qq, int j) {
#pragma unroll
    for (int u = 0; u < LL_CU; u++) {
      const int i = min(j + u * 64 + lane, psize - 1);
      pp[u] = pts[i];
      qq[u] = (int)g[i];
    }
  };
  auto run = [&](const float *pp, const int *qq, int j) {
    float t[LL_CU];
#pragma unroll
    for (int u = 0; u < LL_CU; u++) t[u] = lcb[qq[u]] - pp[u];
#pragma unroll
    for (int u = 0; u < LL_CU; u++)
      if (j + u * 64 < psize) acc = chain64(acc, j + u * 64 + lane < psize ? t[u] : -0.f);
  };
  int j = 0;
  load(pA, qA, 0);
  while (true) {
    load(pB, qB, j + GR);
    run(pA, qA, j);
    j += GR;
    if (j >= psize) break;
    load(pA, qA, j + GR);
    run(pB, qB, j);
    j += GR;
    if (j >= psize) break;
  }
  return acc;
}

// the codebook (cs floats) into LDS, for a one-wave block
__device__ __forceinline__ void cb_to_lds(float *lcb, const float *cb, int cs) {
  for (int k = threadIdx.x; k < cs; k += 64) lcb[k] = cb[k];
  __syncthreads();
}

template <bool DIST>
__global__ void __launch_bounds__(64)
ll_chains(const LJob *J, int k) {
  const LJob j = J[blockIdx.y];
  const int chain = blockIdx.x;
  const int cs = uni(j.cs);
  if (chain > cs || (!DIST && chain == cs)) return;   // a larger run's grid
  if (ll_skip(j, k)) return;
  const int psize = uni(j.psize);
  if (chain < cs) {
    const int nr = uni(j.nruns[chain]);
    float s = 0.f;
    uint32_t c = 0;
    if (nr == 1) {
      const int a = uni(j.first[chain]), b = uni(j.last[chain] + 1);
      c = (uint32_t)(b - a);
      s = run_chain(j.pts, a, b);
    } else if (nr > 1) {
      // members scattered over several runs: scan every point in order
      // (wave-uniform scalar loads, one dependent add per member)
      const float *pts = j.pts;
      const int32_t *grp = j.grp;
      int i = 0;
      for (; i + 16 <= psize; i += 16) {
        float v[16];
        int g[16];
#pragma unroll
        for (int q = 0; q < 16; q++) { v[q] = pts[i + q]; g[q] = grp[i + q]; }
#pragma unroll
        for (int q = 0; q < 16; q++)
          if (g[q] == chain) { s += v[q]; c++; }
      }
      for (; i < psize; i++)
        if (grp[i] == chain) { s += pts[i]; c++; }
    }
    if (threadIdx.x == 0) { j.sum[chain] = s; j.count[chain] = c; }
  } else {
    // dist += codebook[idx] - p over all points in order (lloyds.cpp:64,132)
    extern __shared__ float lcb[];
    cb_to_lds(lcb, j.cb, cs);
    const float d = dist_chain(j.pts, j.grp, lcb, psize);
    if (threadIdx.x == 0) *j.dsum = d;
  }
}

// before the loop: partition from the initial codebook, dist, reldist
__global__ void ll_init_finish(const LJob *J) {
  const LJob j = J[blockIdx.y];
  LloydState *S = j.S;
  const int cs = j.cs;
  for (int i = threadIdx.x; i < cs - 1; i += blockDim.x)
    j.part[i] = (j.cb[i] + j.cb[i + 1]) / 2;
  if (threadIdx.x == 0) {
    // pmax / pmin by sequential comparisons equal the ends of the sorted
    // points (lloyds.cpp:40-49)
    S->pmin = j.pts[0];
    S->pmax = j.pts[S->psize - 1];
    float dist = *j.dsum / (float)(unsigned)S->psize;
    S->dist = dist;
    S->reldist = fabsf(dist);
    S->iters = 0;
    S->done = !(S->reldist > S->stop) || (S->max_iters >= 0 && S->max_iters == 0);
  }
}

// loop head + codebook update (lloyds.cpp:69-87), window iteration k >= 0;
// clears the runs for the partition pass that follows
__global__ void ll_update_codebook(const LJob *J, int k) {
  const LJob j = J[blockIdx.y];
  if (ll_skip(j, k)) return;
  clear_runs(j);
  const LloydState *S = j.S;
  const int cs = j.cs;
  // the iteration's codebook also goes to its window slot (nothing changes
  // it later in the iteration)
  float *snap = j.cbring + (int64_t)k * (cs + 1);
  for (int i = threadIdx.x; i < cs; i += blockDim.x) {
    float c;
    if (j.count[i] != 0) c = j.sum[i] / (float)j.count[i];
    else if (i == 0) c = (j.part[0] + S->pmin) / 2;
    else if (i == cs - 1) c = (j.part[i - 1] + S->pmax) / 2;
    else c = (j.part[i - 1] + j.part[i]) / 2;
    j.cb[i] = c;
    snap[i] = c;
  }
}

// partition update (lloyds.cpp:102-111); clears the runs for the nearest
// pass that follows (the cell sums it reads were taken already)
__global__ void ll_update_partition(const LJob *J, int k) {
  const LJob j = J[blockIdx.y];
  if (ll_skip(j, k)) return;
  clear_runs(j);
  const int cs = j.cs;
  for (int i = threadIdx.x; i < cs - 1; i += blockDim.x) {
    if (j.count[i] != 0) j.part[i] = j.sum[i] / (float)j.count[i];
    else j.part[i] = (j.cb[i] + j.cb[i + 1]) / 2;
  }
}

// ---- speculative windows ---------------------------------------------------
// The codebook trajectory depends only on the cell sums and counts
// (lloyds.cpp:73-111); the signed distortion decides only when the loop
// stops.  So a window runs its iterations of the trajectory back to back,
// each leaving its codebook and 16-bit nearest groups in a slot, then the
// distortion chains of the window run side by side (one wavefront each, the
// same sequential sum over the points in order as lloyds.cpp:124-133), and
// one thread applies the loop test to them in order: the first iteration
// that stops the loop is the result, the later ones of the window are
// dropped.

__global__ void __launch_bounds__(64)
ll_window_dist(const LJob *J) {
  const LJob j = J[blockIdx.y];
  const int w = blockIdx.x;
  if (j.S->done || w >= j.nw) return;
  extern __shared__ float lcb[];
  cb_to_lds(lcb, j.cbring + (int64_t)w * (j.cs + 1), uni(j.cs));
  const float d = dist_chain(j.pts, j.ring + (int64_t)w * j.ring_stride, lcb, uni(j.psize));
  if (threadIdx.x == 0) j.wdsum[w] = d;
}

// the loop tail of each iteration of the window, in order: reldist =
// |dist_old - dist_new|, then the loop test (lloyds.cpp:122-135, plus the
// iteration cap)
__global__ void ll_window_eval(const LJob *J) {
  const LJob j = J[blockIdx.y];
  LloydState *S = j.S;
  for (int w = 0; w < j.nw && !S->done; w++) {
    const float prev = S->dist;
    const float dist = j.wdsum[w] / (float)(unsigned)S->psize;
    S->dist = dist;
    S->reldist = fabsf(prev - dist);
    S->iters++;
    S->final_slot = w;
    S->done = !(S->reldist > S->stop) || (S->max_iters >= 0 && S->iters >= S->max_iters);
  }
}

// the done flags of a table of runs, for one host poll
__global__ void ll_poll(const LJob *J, int njobs, int32_t *done) {
  for (int t = threadIdx.x; t < njobs; t += blockDim.x) done[t] = J[t].S->done;
}

__global__ void ll_fill_dups(uint32_t *tail, uint32_t ndup, uint32_t tkey) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ndup;
       i += gridDim.x * blockDim.x)
    tail[i] = tkey;
}

__global__ void ll_state_init(LloydState *S, int psize, int cs, int max_iters) {
  S->psize = psize;
  S->cs = cs;
  S->iters = 0;
  S->done = 0;
  S->max_iters = max_iters;
  S->stop = (float)10e-7;    // DEFAULT_LLOYD_STOP_CRITERIA, lloyds.h:1
  S->n_gt = 0;
  S->final_slot = -1;
}

}  // namespace wv

using namespace wv;

namespace {
inline int64_t a256(int64_t v) { return (v + 255) & ~(int64_t)255; }
constexpr int LL_WINDOW = 8;   // iterations per speculative window
struct LWs {
  int64_t state, keys, keys2, hist, grp, cb, part, sum, count, first, last, nruns, dsum;
  int64_t ring, ring_stride, cbring, wdsum, jobtab, total;
};
LWs lloyd_layout(int64_t n, int bins) {
  LWs L{};
  const int cs = (bins - 1) / 2;
  const int64_t nblk = (n + LL_SORT_CHUNK - 1) / LL_SORT_CHUNK + 1;
  int64_t o = 0;
  L.state = o; o += a256(sizeof(LloydState));
  L.keys = o; o += a256(n * 4);
  L.keys2 = o; o += a256(n * 4);
  L.hist = o; o += a256(nblk * 256 * 4);
  L.grp = o; o += a256(n * 4);
  L.cb = o; o += a256((int64_t)cs * 4 + 4);
  L.part = o; o += a256((int64_t)cs * 4 + 4);
  L.sum = o; o += a256((int64_t)cs * 4 + 4);
  L.count = o; o += a256((int64_t)cs * 4 + 4);
  L.first = o; o += a256((int64_t)cs * 4 + 4);
  L.last = o; o += a256((int64_t)cs * 4 + 4);
  L.nruns = o; o += a256((int64_t)cs * 4 + 4);
  L.dsum = o; o += a256(16);
  L.ring_stride = a256(n * 2) / 2;   // uint16 elements per slot
  L.ring = o; o += L.ring_stride * 2 * LL_WINDOW;
  L.cbring = o; o += a256((int64_t)(cs + 1) * 4 * LL_WINDOW);
  L.wdsum = o; o += a256(4 * LL_WINDOW);
  L.jobtab = o; o += a256(sizeof(LJob));
  L.total = o;
  return L;
}
}  // namespace

extern "C" int64_t wv_lloyd_ws_bytes(int64_t n, int bin_count) {
  return lloyd_layout(n, bin_count).total;
}

#define LCHK(expr, what)                                          \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) {                                       \
      set_error("lloyd %s: %s", what, hipGetErrorString(_e));     \
      return WV_ERR_HIP;                                          \
    }                                                             \
  } while (0)

// initialLloydCodebook (test_compress_common.cc:749-779) with the host libm
static void initial_codebook(float *cb, int cs, float min_abs, float max_abs) {
  const float log_scale = logf(max_abs / min_abs) / cs;
  for (int i = 0; i < cs; i++) cb[i] = min_abs * expf(i * log_scale);
}

// setBinsFromCodebook (test_compress_common.cc:784-846)
static int bins_from_codebook(float *v, float *b, int bins, const float *c, int cs,
                              float thr, float mn, float mx) {
  int nv = 0, nb = 0;
  if ((bins & 1) == 0 && fabsf(mn) > mx) { v[nv++] = mn; b[nb++] = (mn + -c[cs - 1]) / 2; }
  v[nv++] = -c[cs - 1];
  for (int i = cs - 2; i >= 0; i--) { b[nb++] = (-c[i] + -c[i + 1]) / 2; v[nv++] = -c[i]; }
  b[nb++] = -thr;
  v[nv++] = 0;
  b[nb++] = thr;
  for (int i = 0; i < cs - 1; i++) { v[nv++] = c[i]; b[nb++] = (c[i] + c[i + 1]) / 2; }
  v[nv++] = c[cs - 1];
  if ((bins & 1) == 0 && mx >= fabsf(mn)) { v[nv++] = mx; b[nb++] = (mx + c[cs - 1]) / 2; }
  return (nv == bins && nb == bins - 1) ? WV_OK : WV_ERR_STATE;
}

namespace wv {

// One Lloyd run between its set-up and its result: the device pointers
// into its workspace, its one-entry run table there, and the window count.
struct LloydRun {
  LJob job;
  LJob *job_d;
  int ag, max_iters, enqueued;
  LloydState hs;
};

static int lloyd_setup_points(const float *nonzero, int64_t count, float threshold,
                              int bin_count, int max_iters, void *ws_d, int64_t n_ws,
                              LloydRun &R, hipStream_t s);

// points, initial codebook, initial assignment and distortion; the state
// comes back to R.hs (R.hs.done: the loop never runs)
static int lloyd_setup(const float *coef_d, int64_t n, const wv_thresh_result *res,
                       int bin_count, int max_iters, void *ws_d, LloydRun &R, hipStream_t s) {
  const int cs = (bin_count - 1) / 2;
  const int64_t nz = res->nonzero_count;
  const int64_t psize = nz - 1;
  if (cs < 2 || cs > LL_MAX_CS || psize < 2) {
    set_error("lloyd: %lld points / %d codebook entries unsupported", (long long)psize, cs);
    return WV_ERR_UNSUPPORTED;
  }
  const LWs L = lloyd_layout(n, bin_count);
  char *w = (char *)ws_d;
  LloydState *S = (LloydState *)(w + L.state);
  uint32_t *keys = (uint32_t *)(w + L.keys), *keys2 = (uint32_t *)(w + L.keys2);
  uint32_t *hist = (uint32_t *)(w + L.hist);

  hipLaunchKernelGGL(ll_state_init, dim3(1), dim3(1), 0, s, S, (int)psize, cs, max_iters);
  // 1. tail = threshold duplicates ranked after idx, then sorted keys > T
  const int64_t gb = (n + LL_GCHUNK - 1) / LL_GCHUNK;
  hipLaunchKernelGGL(ll_gather, dim3((unsigned)gb), dim3(WV_BLOCK), 0, s, coef_d, n,
                     res->threshold_key, keys, S);
  LCHK(hipMemcpyAsync(&R.hs, S, sizeof R.hs, hipMemcpyDeviceToHost, s), "D2H");
  LCHK(hipStreamSynchronize(s), "sync");
  const int64_t ngt = R.hs.n_gt;
  const int64_t ndup = nz - ngt;
  if (ndup < 0 || ndup > n) { set_error("lloyd: inconsistent threshold result"); return WV_ERR_STATE; }
  LCHK(radix_sort_u32(keys, keys2, hist, &S->n_gt, ngt, s), "sort");
  // the sorted keys are in keys; lay out the tail [dups..., sorted...] in
  // keys2
  uint32_t *tail = keys2;
  if (ndup > 0)
    hipLaunchKernelGGL(ll_fill_dups, dim3(64), dim3(256), 0, s, tail, (uint32_t)ndup,
                       res->threshold_key);
  if (ngt > 0)
    LCHK(hipMemcpyAsync(tail + ndup, keys, ngt * 4, hipMemcpyDeviceToDevice, s), "tail copy");
  return lloyd_setup_points((const float *)tail, nz, res->threshold, bin_count, max_iters,
                            ws_d, n, R, s);
}

// The set-up from sorted points: nonzero[0 .. count-1] ascending, the run
// over nonzero[0 .. count-2] (computeLloydQuantization,
// test_compress_cpu.cc:404-459: the largest value is left out), initial
// codebook log-spaced from the threshold to nonzero[count-2], initial
// assignment and distortion; the state comes back to R.hs.  ws_d is laid
// out for n_ws samples (n_ws >= count).
static int lloyd_setup_points(const float *nonzero, int64_t count, float threshold,
                              int bin_count, int max_iters, void *ws_d, int64_t n_ws,
                              LloydRun &R, hipStream_t s) {
  const int cs = (bin_count - 1) / 2;
  const int64_t psize = count - 1;
  if (cs < 2 || cs > LL_MAX_CS || psize < 2 || psize > 0x7fffffff) {
    set_error("lloyd: %lld points / %d codebook entries unsupported", (long long)psize, cs);
    return WV_ERR_UNSUPPORTED;
  }
  const LWs L = lloyd_layout(n_ws, bin_count);
  char *w = (char *)ws_d;
  LloydState *S = (LloydState *)(w + L.state);
  hipLaunchKernelGGL(ll_state_init, dim3(1), dim3(1), 0, s, S, (int)psize, cs, max_iters);
  const float *tail = nonzero;
  const int64_t nz = count;

  LJob &J = R.job;
  J = LJob{};
  J.S = S;
  J.pts = (const float *)tail;   // points = tail[0 .. nz-2]
  J.grp = (int32_t *)(w + L.grp);
  J.cb = (float *)(w + L.cb);
  J.part = (float *)(w + L.part);
  J.sum = (float *)(w + L.sum);
  J.count = (uint32_t *)(w + L.count);
  J.first = (int32_t *)(w + L.first);
  J.last = (int32_t *)(w + L.last);
  J.nruns = (int32_t *)(w + L.nruns);
  J.dsum = (float *)(w + L.dsum);
  J.ring = (uint16_t *)(w + L.ring);
  J.cbring = (float *)(w + L.cbring);
  J.wdsum = (float *)(w + L.wdsum);
  J.ring_stride = L.ring_stride;
  J.cs = cs;
  J.psize = (int32_t)psize;
  J.nw = 0;
  R.job_d = (LJob *)(w + L.jobtab);
  R.ag = (int)std::min<int64_t>((psize + WV_BLOCK - 1) / WV_BLOCK, 4096);
  R.max_iters = max_iters;
  R.enqueued = 0;

  // 2. initial codebook on the host: minAbs = threshold, maxAbs = tail[nz-2]
  float max2;
  LCHK(hipMemcpyAsync(&max2, tail + (nz - 2), 4, hipMemcpyDeviceToHost, s), "D2H max");
  LCHK(hipStreamSynchronize(s), "sync");
  std::vector<float> hcb(cs);
  initial_codebook(hcb.data(), cs, threshold, max2);
  LCHK(hipMemcpyAsync(J.cb, hcb.data(), cs * 4, hipMemcpyHostToDevice, s), "H2D codebook");
  LCHK(hipMemcpyAsync(R.job_d, &J, sizeof J, hipMemcpyHostToDevice, s), "H2D run table");

  // 3. initial assignment (lloyds.cpp:35-67): nearest groups, their cell
  // sums and the distortion
  const size_t cb_lds = (size_t)cs * 4;
  hipLaunchKernelGGL(ll_clear_runs, dim3(1), dim3(256), 0, s, R.job_d, -1);
  hipLaunchKernelGGL(ll_assign<true>, dim3(R.ag), dim3(WV_BLOCK), cb_lds, s, R.job_d, -1);
  hipLaunchKernelGGL(ll_chains<true>, dim3(cs + 1), dim3(64), cb_lds, s, R.job_d, -1);
  hipLaunchKernelGGL(ll_init_finish, dim3(1), dim3(256), 0, s, R.job_d);
  LCHK(hipMemcpyAsync(&R.hs, S, sizeof R.hs, hipMemcpyDeviceToHost, s), "D2H state");
  LCHK(hipStreamSynchronize(s), "sync");
  return WV_OK;
}

// the window's iterations nw for a run, from its iteration cap
static int window_iters(const LloydRun &R) {
  int nw = LL_WINDOW;
  if (R.max_iters >= 0 && R.max_iters - R.enqueued < nw) nw = R.max_iters - R.enqueued;
  return nw;
}

// one window of iterations for the njobs runs of the table J (run j takes
// iterations k < its nw; W = the largest nw); grids cover the largest run
static void enqueue_window(const LJob *J, int njobs, int max_cs, int max_ag, int W,
                           hipStream_t s) {
  const size_t cb_lds = (size_t)max_cs * 4;
  auto assign_and_chains = [&](bool nearest_pass, int k) {
    // the update kernel before each assignment has cleared its runs
    if (nearest_pass)
      hipLaunchKernelGGL(ll_assign<true>, dim3(max_ag, njobs), dim3(WV_BLOCK), cb_lds, s, J, k);
    else
      hipLaunchKernelGGL(ll_assign<false>, dim3(max_ag, njobs), dim3(WV_BLOCK), cb_lds, s, J, k);
    hipLaunchKernelGGL(ll_chains<false>, dim3(max_cs, njobs), dim3(64), 0, s, J, k);
  };
  for (int k = 0; k < W; k++) {
    hipLaunchKernelGGL(ll_update_codebook, dim3(1, njobs), dim3(256), 0, s, J, k);
    assign_and_chains(false, k);
    hipLaunchKernelGGL(ll_update_partition, dim3(1, njobs), dim3(256), 0, s, J, k);
    assign_and_chains(true, k);
  }
  hipLaunchKernelGGL(ll_window_dist, dim3(W, njobs), dim3(64), cb_lds, s, J);
  hipLaunchKernelGGL(ll_window_eval, dim3(1, njobs), dim3(1), 0, s, J);
}

// the windows of one run on its own stream, one host poll each
static int lloyd_loop(LloydRun &R, hipStream_t s) {
  while (!R.hs.done) {
    const int nw = window_iters(R);
    if (nw != R.job.nw) {
      R.job.nw = nw;
      LCHK(hipMemcpyAsync(R.job_d, &R.job, sizeof R.job, hipMemcpyHostToDevice, s),
           "H2D run table");
    }
    enqueue_window(R.job_d, 1, R.job.cs, R.ag, nw, s);
    R.enqueued += nw;
    LCHK(hipGetLastError(), "launch");
    LCHK(hipMemcpyAsync(&R.hs, R.job.S, sizeof R.hs, hipMemcpyDeviceToHost, s), "D2H state");
    LCHK(hipStreamSynchronize(s), "sync");
    if (!R.hs.done && R.enqueued >= LL_HARD_CAP) {
      set_error("lloyd: no convergence after %d iterations", R.enqueued);
      return WV_ERR_STATE;
    }
  }
  return WV_OK;
}

// ---- the Lloyd runs of concurrent requests, batched ----------------------
// A window batch: one thread owns a stream and a run table.  Runs join at
// window boundaries; each window is one set of launches for every active
// run (grid.y = run), then one poll of all their done flags.  A run leaves
// when its loop stops, and its worker continues.
struct LloydWindowBatch {
  struct Entry {
    LloydRun *run;
    int status = WV_OK;
    bool finished = false;
    char msg[160] = {0};   // the error, set in the worker's thread (messages are per thread)
  };
  static constexpr int kCap = 64;
  int device = 0;
  hipStream_t s = nullptr;
  LJob *tab_d = nullptr, *tab_h = nullptr;
  int32_t *flags_d = nullptr, *flags_h = nullptr;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Entry *> pending;
  bool stop = false;
  int64_t windows = 0, run_windows = 0;   // launches sets, and runs served by them
  std::thread th;

  int init(int dev) {
    device = dev;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&tab_d, kCap * sizeof(LJob)) != hipSuccess ||
        hipMalloc((void **)&flags_d, kCap * sizeof(int32_t)) != hipSuccess ||
        hipHostMalloc((void **)&tab_h, kCap * sizeof(LJob), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&flags_h, kCap * sizeof(int32_t), hipHostMallocDefault) !=
            hipSuccess) {
      set_error("lloyd batcher: device resources");
      return WV_ERR_HIP;
    }
    th = std::thread(&LloydWindowBatch::loop, this);
    return WV_OK;
  }

  ~LloydWindowBatch() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (th.joinable()) th.join();
    if (tab_d) (void)hipFree(tab_d);
    if (flags_d) (void)hipFree(flags_d);
    if (tab_h) (void)hipHostFree(tab_h);
    if (flags_h) (void)hipHostFree(flags_h);
    if (s) (void)hipStreamDestroy(s);
  }

  // runs R's windows with the other runs in flight; returns its status
  int run(LloydRun *R) {
    Entry e;
    e.run = R;
    std::unique_lock<std::mutex> lk(mu);
    pending.push_back(&e);
    cv.notify_all();
    cv_done.wait(lk, [&] { return e.finished; });
    if (e.status != WV_OK) set_error("%s", e.msg);
    return e.status;
  }

  void finish(std::vector<Entry *> &active, size_t i, int status, const char *msg) {
    std::lock_guard<std::mutex> lk(mu);
    active[i]->status = status;
    if (msg) snprintf(active[i]->msg, sizeof active[i]->msg, "%s", msg);
    active[i]->finished = true;
    cv_done.notify_all();
  }

  void loop() {
    (void)hipSetDevice(device);
    std::vector<Entry *> active;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
        if (stop && pending.empty() && active.empty()) break;
        while (!pending.empty() && (int)active.size() < kCap) {
          active.push_back(pending.front());
          pending.pop_front();
        }
      }
      const int nj = (int)active.size();
      int W = 0, max_cs = 0, max_ag = 0;
      for (int i = 0; i < nj; i++) {
        LloydRun &R = *active[i]->run;
        R.job.nw = window_iters(R);
        tab_h[i] = R.job;
        W = std::max(W, R.job.nw);
        max_cs = std::max(max_cs, R.job.cs);
        max_ag = std::max(max_ag, R.ag);
      }
      hipError_t e = hipMemcpyAsync(tab_d, tab_h, nj * sizeof(LJob), hipMemcpyHostToDevice, s);
      if (e == hipSuccess) {
        enqueue_window(tab_d, nj, max_cs, max_ag, W, s);
        hipLaunchKernelGGL(ll_poll, dim3(1), dim3(64), 0, s, tab_d, nj, flags_d);
        e = hipGetLastError();
      }
      if (e == hipSuccess)
        e = hipMemcpyAsync(flags_h, flags_d, nj * sizeof(int32_t), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      {
        std::lock_guard<std::mutex> lk(mu);
        windows++;
        run_windows += nj;
      }
      std::vector<Entry *> keep;
      for (int i = 0; i < nj; i++) {
        LloydRun &R = *active[i]->run;
        R.enqueued += R.job.nw;
        if (e != hipSuccess) {
          char m[160];
          snprintf(m, sizeof m, "lloyd batcher: %s", hipGetErrorString(e));
          finish(active, i, WV_ERR_HIP, m);
        } else if (flags_h[i]) {
          finish(active, i, WV_OK, nullptr);
        } else if (R.enqueued >= LL_HARD_CAP) {
          char m[160];
          snprintf(m, sizeof m, "lloyd: no convergence after %d iterations", R.enqueued);
          finish(active, i, WV_ERR_STATE, m);
        } else {
          keep.push_back(active[i]);
        }
      }
      active.swap(keep);
    }
  }
};

// A window lasts as long as its longest run, so runs are batched by size:
// one window batch per power-of-two class of the point count (a run waits
// for runs at most twice its size), each with its own thread and stream,
// created when a run of its class first arrives.  Runs of equal size share
// every launch; different sizes still run side by side.
struct LloydBatcher {
  int device = 0;
  std::mutex mu;
  LloydWindowBatch *cls[32] = {};

  ~LloydBatcher() {
    for (LloydWindowBatch *b : cls) delete b;
  }

  LloydWindowBatch *batch_for(int psize) {
    int c = 0;
    while (c < 31 && (2 << c) <= psize) c++;
    std::lock_guard<std::mutex> lk(mu);
    if (!cls[c]) {
      LloydWindowBatch *b = new LloydWindowBatch;
      if (b->init(device) != WV_OK) {
        delete b;
        return nullptr;
      }
      cls[c] = b;
    }
    return cls[c];
  }

  int run(LloydRun *R) {
    LloydWindowBatch *b = batch_for(R->job.psize);
    return b ? b->run(R) : WV_ERR_HIP;
  }
};

LloydBatcher *lloyd_batcher_create(int device) {
  LloydBatcher *b = new LloydBatcher;
  b->device = device;
  return b;
}

void lloyd_batcher_destroy(LloydBatcher *b) { delete b; }

void lloyd_batcher_stats(LloydBatcher *b, int64_t *windows, int64_t *run_windows) {
  *windows = *run_windows = 0;
  std::lock_guard<std::mutex> lk(b->mu);
  for (LloydWindowBatch *c : b->cls) {
    if (!c) continue;
    std::lock_guard<std::mutex> lc(c->mu);
    *windows += c->windows;
    *run_windows += c->run_windows;
  }
}

// the whole run: set-up, the windows (on this stream, or batched with the
// other runs in flight), the codebook of the iteration that stopped the
// loop, and the bins from it
int lloyd_run(const float *coef_d, int64_t n, const wv_thresh_result *res, int bin_count,
              int max_iters, void *ws_d, float *bin_boundaries, float *bin_values,
              int *iters_out, hipStream_t s, LloydBatcher *batcher) {
  LloydRun R;
  const int rc = lloyd_setup(coef_d, n, res, bin_count, max_iters, ws_d, R, s);
  if (rc != WV_OK) return rc;
  if (!R.hs.done) {
    const int lrc = batcher ? batcher->run(&R) : lloyd_loop(R, s);
    if (lrc != WV_OK) return lrc;
    if (batcher) {
      LCHK(hipMemcpyAsync(&R.hs, R.job.S, sizeof R.hs, hipMemcpyDeviceToHost, s), "D2H state");
      LCHK(hipStreamSynchronize(s), "sync");
    }
  }
  const int cs = R.job.cs;
  // the codebook of the iteration that stopped the loop (the initial one
  // when the loop never ran)
  const float *final_cb =
      R.hs.iters > 0 ? R.job.cbring + (int64_t)R.hs.final_slot * (cs + 1) : R.job.cb;
  std::vector<float> fin(cs);
  LCHK(hipMemcpyAsync(fin.data(), final_cb, cs * 4, hipMemcpyDeviceToHost, s), "D2H codebook");
  LCHK(hipStreamSynchronize(s), "sync");
  if (iters_out) *iters_out = R.hs.iters;
  return bins_from_codebook(bin_values, bin_boundaries, bin_count, fin.data(), cs,
                            res->threshold, res->min_val, res->max_val);
}

// Lloyd over points already sorted on the device: nonzero[0 .. count-1]
// ascending, the run over all but the largest (the optimiser's test points,
// which take nonzero = the sorted |x| from findSorted(threshold)).  ws_d:
// wv_lloyd_ws_bytes(n_ws, bin_count) with n_ws >= count.
int lloyd_run_sorted(const float *nonzero, int64_t count, float threshold, int bin_count,
                     int max_iters, void *ws_d, int64_t n_ws, float min_val, float max_val,
                     float *bin_boundaries, float *bin_values, int *iters_out, hipStream_t s) {
  LloydRun R;
  const int rc = lloyd_setup_points(nonzero, count, threshold, bin_count, max_iters, ws_d, n_ws,
                                    R, s);
  if (rc != WV_OK) return rc;
  if (!R.hs.done) {
    const int lrc = lloyd_loop(R, s);
    if (lrc != WV_OK) return lrc;
  }
  const int cs = R.job.cs;
  const float *final_cb =
      R.hs.iters > 0 ? R.job.cbring + (int64_t)R.hs.final_slot * (cs + 1) : R.job.cb;
  std::vector<float> fin(cs);
  LCHK(hipMemcpyAsync(fin.data(), final_cb, cs * 4, hipMemcpyDeviceToHost, s), "D2H codebook");
  LCHK(hipStreamSynchronize(s), "sync");
  if (iters_out) *iters_out = R.hs.iters;
  return bins_from_codebook(bin_values, bin_boundaries, bin_count, fin.data(), cs, threshold,
                            min_val, max_val);
}

}  // namespace wv

extern "C" int wv_lloyd(const float *coef_d, int64_t n, const wv_thresh_result *res,
                        int bin_count, int max_iters, void *ws_d,
                        float *bin_boundaries, float *bin_values, int *iters_out,
                        void *stream) {
  if (!coef_d || !res || !ws_d || !bin_boundaries || !bin_values || n <= 0) {
    set_error("bad arguments");
    return WV_ERR_ARG;
  }
  return lloyd_run(coef_d, n, res, bin_count, max_iters, ws_d, bin_boundaries, bin_values,
                   iters_out, (hipStream_t)stream, nullptr);
}
