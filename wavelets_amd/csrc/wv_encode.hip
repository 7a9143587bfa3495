// wv_encode.hip -- Huffman coding of the quantized bins on the device.
//
// The reference codes the bins on the host after copying all of them back
// (test_compress_gpu.cu:402, then writeQuantData: initHuffman +
// huffScanDupsOnly test_compress_common.cc:927-973, computeHuffmanCoding
// huffman.cc:42-97, encodeToStream / dupEncodeString huffman.h:241-296,
// BitStreamWriter bit_stream.h:87-162).  Here only the two run statistics
// and the finished payload cross PCIe; the code book (a tree over at most
// bin_count + 1 leaves) is still built on the host, by the host coder.
//
// The bins are cut into chunks of ENC_C, and every block of a launch owns
// a contiguous range of chunks, which it walks in order:
//   last     per range, the index of its last bin that is not the
//            duplicate value (read backwards from the range's end, so
//            usually only its last chunk)
//   count    each run of the duplicate value -- whose start is the last
//            other bin before it, found in the chunk or carried from the
//            chunks and ranges before -- splits into dup codes and single
//            codes (below); per range, the histogram of the other values
//            and the run sums; in total, the counts the code book needs
//   (host)   Huffman code book
//   bits     per range, its output bits: the histogram dotted with the
//            code lengths (or, for code books too large for LDS, a walk)
//   write    every block codes its range chunk by chunk at the running
//            bit offset into LDS words, then stores them (the first and
//            last word of a chunk, which neighbours share, by atomic OR
//            into the zeroed payload)
//
// A run of r copies of the duplicate value (dupEncodeString, huffman.h:
// 278-296) codes as q = r / 1024 dup codes with e = 7, then one dup code
// per set bit of r mod 1024 at positions 9..3 (e = position - 3, highest
// first), except that a remainder of exactly 8 stays as 8 single codes;
// the rest (at most 8) are single codes of the value.  A dup code is the
// dup key's code followed by e in 3 bits, LSB first.  A run's codes are
// written where its last bin is.  Only the first run that ends in a chunk
// can have started before it, so at most one run per chunk has q > 4: its
// q dup codes are a periodic bit string that the block writes word by
// word.
#include <climits>
#include <cstdlib>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define ENC_NT 256
#define ENC_EPT 16
#define ENC_C (ENC_NT * ENC_EPT)   // bins per chunk
#define ENC_NW (ENC_NT / 64)
#define ENC_TAB 4097               // code-book / histogram entries held in LDS
#define ENC_STAGE 2048             // payload words staged in LDS per chunk
#define ENC_SENT INT_MIN           // stands for bins past the end
#define ENC_LONG_Q 4               // runs with more dup codes are written
                                   // by the whole block
#define ENC_MAX_RANGES 2048        // blocks per launch (8 per CU)

// EncArgs (one coding job) is declared in wv_internal.h


// A launch over many jobs: block b serves the last job j with blk0[j] <= b
// (job tables on the device); a single job is passed by value with blk0 0.
__device__ inline int enc_job(const int *blk0, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// this block's range within its job
__device__ inline int range_id(const EncArgs &A) { return (int)blockIdx.x - A.blk0; }

__device__ inline int code_len(uint64_t c) { return (int)(c >> 56); }
__device__ inline uint64_t code_bits(uint64_t c) { return c & ((1ull << 56) - 1); }

struct RunSplit {
  int64_t q;          // dup codes with e = 7
  uint32_t hi;        // bit k set: one dup code with e = k (k = 6..0)
  uint32_t singles;   // single codes of the value
};

__device__ inline RunSplit split_run(int64_t r) {
  RunSplit s;
  if (r <= 8) {
    s.q = 0;
    s.hi = 0;
    s.singles = (uint32_t)r;
    return s;
  }
  s.q = r >> 10;
  const uint32_t r1 = (uint32_t)(r & 1023);
  s.hi = r1 >> 3;
  s.singles = r1 & 7u;
  if ((s.hi & 1u) && s.singles == 0) {   // remainder 8: eight single codes
    s.hi &= ~1u;
    s.singles = 8;
  }
  return s;
}

__device__ inline uint64_t shfl_xor64(uint64_t v, int d) {
  const uint32_t lo = __shfl_xor((uint32_t)v, d, 64);
  const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline int wave_incl_max(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(v, d, 64);
    if (lane >= d) v = v > y ? v : y;
  }
  return v;
}
__device__ inline uint32_t wave_incl_sum(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  return v;
}

// Exclusive max over the block's threads in order (-1 before thread 0);
// *total gets the block's max.  red: ENC_NW ints of LDS.  Ends with a
// barrier, so red can be reused.
__device__ inline int block_excl_max(int v, int *red, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_incl_max(v);
  if (lane == 63) red[w] = inc;
  __syncthreads();
  int before = -1, all = -1;
  for (int i = 0; i < ENC_NW; i++) {
    if (i < w) before = red[i] > before ? red[i] : before;
    all = red[i] > all ? red[i] : all;
  }
  int ex = __shfl_up(inc, 1, 64);
  if (lane == 0) ex = -1;
  *total = all;
  __syncthreads();
  return ex > before ? ex : before;
}

// Exclusive sum over the block's threads; *total gets the block's sum.
__device__ inline uint32_t block_excl_sum(uint32_t v, uint32_t *red, uint32_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_sum(v);
  if (lane == 63) red[w] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int i = 0; i < ENC_NW; i++) {
    if (i < w) before += red[i];
    all += red[i];
  }
  uint32_t ex = __shfl_up(inc, 1, 64);
  if (lane == 0) ex = 0;
  *total = all;
  __syncthreads();
  return before + ex;
}

// Sum (or signed max, identity -1) over the block of every thread's acc,
// for every thread.  red: ENC_NW entries of LDS.
template <bool MAX>
__device__ inline uint64_t block_all(uint64_t acc, uint64_t *red) {
  auto op = [](uint64_t x, uint64_t y) -> uint64_t {
    if (MAX) return (int64_t)x > (int64_t)y ? x : y;
    return x + y;
  };
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc = op(acc, shfl_xor64(acc, d));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  uint64_t r = MAX ? ~0ull : 0ull;
  for (int i = 0; i < ENC_NW; i++) r = op(r, red[i]);
  __syncthreads();
  return r;
}

// the same over a[0, m) of a global array
template <bool MAX>
__device__ inline uint64_t block_reduce(const uint64_t *a, int m, uint64_t *red) {
  uint64_t acc = MAX ? ~0ull : 0ull;
  for (int i = threadIdx.x; i < m; i += ENC_NT) {
    const uint64_t v = a[i];
    acc = MAX ? ((int64_t)v > (int64_t)acc ? v : acc) : acc + v;
  }
  return block_all<MAX>(acc, red);
}

__device__ inline void range_of(const EncArgs &A, int64_t &cb, int64_t &ce) {
  cb = (int64_t)range_id(A) * A.per;
  ce = cb + A.per < A.nchunks ? cb + A.per : A.nchunks;
}

// The thread's ENC_EPT bins of chunk c (ENC_SENT past the end) and whether
// the bin after its last one is the duplicate value.  first: ENC_NT ints
// of LDS; the caller has passed a barrier since their last use.
__device__ inline void load_bins(const EncArgs &A, int64_t c, int32_t e[ENC_EPT], int *first,
                                 bool &next_dup) {
  const int t = threadIdx.x;
  const int64_t i0 = c * ENC_C + (int64_t)ENC_EPT * t;
  if (A.vec && i0 + ENC_EPT <= A.n) {
    const WV_GLOBAL wv_i4 *p = (const WV_GLOBAL wv_i4 *)gptr(A.bins + i0);
#pragma unroll
    for (int j = 0; j < ENC_EPT / 4; j++) {
      const wv_i4 v = p[j];
      e[4 * j] = v.x;
      e[4 * j + 1] = v.y;
      e[4 * j + 2] = v.z;
      e[4 * j + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < ENC_EPT; k++) e[k] = i0 + k < A.n ? gptr(A.bins)[i0 + k] : ENC_SENT;
  }
  first[t] = e[0];
  __syncthreads();
  int nx;
  if (t + 1 < ENC_NT) {
    nx = first[t + 1];
  } else {
    const int64_t j = (c + 1) * ENC_C;
    nx = j < A.n ? gptr(A.bins)[j] : ENC_SENT;
  }
  next_dup = nx == A.zb;
}

// A chunk's bins in the thread's registers: its ENC_EPT bins (ENC_SENT
// past the end) and the bin after the chunk (every thread loads it; the
// last thread needs it).
struct Chunk {
  int32_t e[ENC_EPT];
  int32_t after;
};

// full chunk, 16-byte aligned bins: exactly five loads per thread
__device__ inline void fetch_fast(const EncArgs &A, int64_t c, Chunk &k) {
  const WV_GLOBAL wv_i4 *p =
      (const WV_GLOBAL wv_i4 *)gptr(A.bins + c * ENC_C + (int64_t)ENC_EPT * threadIdx.x);
#pragma unroll
  for (int j = 0; j < ENC_EPT / 4; j++) {
    const wv_i4 v = p[j];
    k.e[4 * j] = v.x;
    k.e[4 * j + 1] = v.y;
    k.e[4 * j + 2] = v.z;
    k.e[4 * j + 3] = v.w;
  }
  const int64_t j = (c + 1) * ENC_C;
  const int32_t a = gptr(A.bins)[j < A.n ? j : A.n - 1];
  k.after = j < A.n ? a : ENC_SENT;
}

__device__ inline void fetch_plain(const EncArgs &A, int64_t c, Chunk &k) {
  const int64_t i0 = c * ENC_C + (int64_t)ENC_EPT * threadIdx.x;
#pragma unroll
  for (int i = 0; i < ENC_EPT; i++) k.e[i] = i0 + i < A.n ? gptr(A.bins)[i0 + i] : ENC_SENT;
  const int64_t j = (c + 1) * ENC_C;
  k.after = j < A.n ? gptr(A.bins)[j] : ENC_SENT;
}

// Runs body(c, chunk, parity) over chunks [cb, ce) in order.  Full chunks
// of aligned bins are software-pipelined: the next chunk's loads are
// issued before the current one is processed, and stay in flight through
// it (the body's barriers wait on LDS only).  The last partial chunk, and
// unaligned bins, are loaded when they are processed.
template <bool PREFETCH, typename Body>
__device__ inline void for_chunks(const EncArgs &A, int64_t cb, int64_t ce, Body &&body) {
  int64_t c = cb;
  const int64_t nfull = A.n / ENC_C;
  const int64_t cfull = ce < nfull ? ce : nfull;
  int par = 0;
  if (PREFETCH && A.vec && c < cfull) {
    Chunk cur, nxt;
    fetch_fast(A, c, cur);
    for (; c < cfull; c++) {
      fetch_fast(A, c + 1 < nfull ? c + 1 : c, nxt);
      body(c, cur, par);
      cur = nxt;
      par ^= 1;
    }
  }
  for (; c < ce; c++) {
    Chunk k;
    if (A.vec && c < cfull) fetch_fast(A, c, k);
    else fetch_plain(A, c, k);
    body(c, k, par);
    par ^= 1;
  }
}

// Publishes the thread's first bin and its last non-duplicate index, and
// after one barrier returns the exclusive max (pz), the chunk's max (*tot)
// and whether the bin after the thread's last one is the duplicate value.
// first / red are the parity's buffers: a buffer is written again only
// two chunks later, after the next chunk's barrier.
__device__ inline int chunk_scan(const EncArgs &A, int64_t c, const Chunk &k, int *first,
                                 int *red, int *tot, bool *next_dup);

// chunk-local index of the thread's last bin that is not the duplicate, -1
__device__ inline int thread_last_nz(const EncArgs &A, int64_t c, const int32_t e[ENC_EPT]) {
  const int64_t i0 = c * ENC_C + (int64_t)ENC_EPT * threadIdx.x;
  int l = -1;
#pragma unroll
  for (int k = 0; k < ENC_EPT; k++)
    if (i0 + k < A.n && e[k] != A.zb) l = ENC_EPT * threadIdx.x + k;
  return l;
}

__device__ inline int chunk_scan(const EncArgs &A, int64_t c, const Chunk &k, int *first,
                                 int *red, int *tot, bool *next_dup) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  first[t] = k.e[0];
  const int inc = wave_incl_max(thread_last_nz(A, c, k.e));
  if (lane == 63) red[w] = inc;
  __syncthreads();
  int before = -1, all = -1;
  for (int i = 0; i < ENC_NW; i++) {
    if (i < w) before = red[i] > before ? red[i] : before;
    all = red[i] > all ? red[i] : all;
  }
  int ex = __shfl_up(inc, 1, 64);
  if (lane == 0) ex = -1;
  *tot = all;
  *next_dup = (t + 1 < ENC_NT ? first[t + 1] : k.after) == A.zb;
  return ex > before ? ex : before;
}

// Walks the thread's bins in order.  pz: chunk-local index of the last
// non-duplicate bin before them (-1: none in the chunk; the run then
// started after prev, the last one before the chunk, -1 for none).
// f(kind, value or run length): kind 0 = a non-duplicate bin, 1 = the last
// bin of a run.
template <typename F>
__device__ inline void walk_bins(const EncArgs &A, int64_t prev, int64_t c,
                                 const int32_t e[ENC_EPT], bool next_dup, int pz, F &&f) {
  const int64_t c0 = c * ENC_C;
  const int64_t i0 = c0 + (int64_t)ENC_EPT * threadIdx.x;
  const int64_t left = A.n - i0;
  const int nv = left >= ENC_EPT ? ENC_EPT : (left > 0 ? (int)left : 0);
  // a rolled loop over a shift register (the bins move down one place per
  // step, so every access has a constant index and stays in registers):
  // the body is large, and unrolled sixteen times it no longer fits the
  // instruction cache
  int32_t r[ENC_EPT];
#pragma unroll
  for (int j = 0; j < ENC_EPT; j++) r[j] = e[j];
#pragma unroll 1
  for (int k = 0; k < nv; k++) {
    const int v = r[0];
    if (v != A.zb) {
      f(0, (int64_t)v);
      pz = ENC_EPT * threadIdx.x + k;
    } else {
      const bool end = k + 1 < ENC_EPT ? r[1] != A.zb : !next_dup;
      if (end) f(1, i0 + k - (pz >= 0 ? c0 + pz : prev));
    }
#pragma unroll
    for (int j = 0; j + 1 < ENC_EPT; j++) r[j] = r[j + 1];
  }
}

__device__ inline int clamp_bin(const EncArgs &A, int64_t v) {
  // bins outside [0, nb1) do not occur (the host checks the histogram);
  // the clamp only keeps table accesses in bounds
  return v < 0 ? 0 : (v >= A.nb1 ? A.nb1 - 1 : (int)v);
}

// last: per range, its last non-duplicate bin, reading chunks backwards
template <bool BATCH>
__global__ void __launch_bounds__(ENC_NT) enc_range_last(EncArgs A0, const EncArgs *J, const int *JB, int nj) {
  const EncArgs A = BATCH ? J[enc_job(JB, nj, (int)blockIdx.x)] : A0;
  __shared__ int first[ENC_NT];
  __shared__ int red[ENC_NW];
  int64_t cb, ce;
  range_of(A, cb, ce);
  int64_t found = -1;
  for (int64_t c = ce - 1; c >= cb; c--) {
    int32_t e[ENC_EPT];
    bool nd;
    __syncthreads();
    load_bins(A, c, e, first, nd);
    int tot;
    block_excl_max(thread_last_nz(A, c, e), red, &tot);
    if (tot >= 0) {   // uniform
      found = c * ENC_C + tot;
      break;
    }
  }
  if (threadIdx.x == 0) A.range_last[range_id(A)] = found;
}

// count: run statistics per range and in total; with HIST also the
// range's histogram of the other values (dynamic LDS of nb1 entries).
// One barrier per chunk.
template <bool HIST, bool BATCH>
__global__ void __launch_bounds__(ENC_NT) enc_count(EncArgs A0, const EncArgs *J, const int *JB, int nj) {
  const EncArgs A = BATCH ? J[enc_job(JB, nj, (int)blockIdx.x)] : A0;
  extern __shared__ uint32_t hist[];
  __shared__ int first[2][ENC_NT];
  __shared__ int red[2][ENC_NW];
  __shared__ uint64_t red64[ENC_NW];
  if (HIST)
    for (int i = threadIdx.x; i < A.nb1; i += ENC_NT) hist[i] = 0;
  int64_t prev = A.zb >= 0 ? (int64_t)block_reduce<true>((const uint64_t *)A.range_last,
                                                           range_id(A), red64)
                           : -1;
  __syncthreads();
  uint64_t singles = 0, dups = 0;
  int64_t cb, ce;
  range_of(A, cb, ce);
  for_chunks<true>(A, cb, ce, [&](int64_t c, const Chunk &k, int par) {
    int tot;
    bool nd;
    const int pz = chunk_scan(A, c, k, first[par], red[par], &tot, &nd);
    walk_bins(A, prev, c, k.e, nd, pz, [&](int kind, int64_t x) {
      if (kind == 0) {
        if (HIST) atomicAdd(&hist[clamp_bin(A, x)], 1u);
      } else {
        const RunSplit s = split_run(x);
        singles += s.singles;
        dups += (uint64_t)s.q + (uint64_t)__popc(s.hi);
      }
    });
    if (tot >= 0) prev = c * ENC_C + tot;
  });
  __syncthreads();
  const uint64_t S = block_all<false>(singles, red64);
  const uint64_t D = block_all<false>(dups, red64);
  if (threadIdx.x == 0) {
    A.range_sd[2 * range_id(A)] = S;
    A.range_sd[2 * range_id(A) + 1] = D;
    if (S) atomicAdd(&A.sd[0], (unsigned long long)S);
    if (D) atomicAdd(&A.sd[1], (unsigned long long)D);
  }
  if (HIST)
    for (int i = threadIdx.x; i < A.nb1; i += ENC_NT)
      A.range_hist[(int64_t)range_id(A) * A.nb1 + i] = hist[i];
}

// Code book (LDS copy when it fits) and the two codes a run uses.
struct Book {
  const uint64_t *tab;
  uint64_t zcode, dcode;
  int zlen, dlen;
};

template <bool TAB_LDS>
__device__ inline Book load_book(const EncArgs &A, uint64_t *lds) {
  Book b;
  if (TAB_LDS) {
    for (int i = threadIdx.x; i < A.nb1; i += ENC_NT) lds[i] = A.book[i];
    __syncthreads();
    b.tab = lds;
  } else {
    b.tab = A.book;
  }
  if (A.zb >= 0) {
    const uint64_t z = A.book[A.zb], d = A.book[A.nb1 - 1];
    b.zcode = code_bits(z);
    b.zlen = code_len(z);
    b.dcode = code_bits(d);
    b.dlen = code_len(d);
  } else {
    b.zcode = b.dcode = 0;
    b.zlen = b.dlen = 0;
  }
  return b;
}

// The codes of a run of r < 1024 (its dup codes for the set bits of r
// above bit 2, then its single codes), as one bit string of up to 128
// bits, for every r: a run of any length r is q = r / 1024 dup codes with
// e = 7 followed by the string of r mod 1024.  *fast (block-uniform):
// every string fits.
struct RunCodes {
  uint64_t lo, hi;
};

__device__ inline bool build_run_codes(const Book &b, RunCodes *rc, uint16_t *rl) {
  bool ok = true;
  const int u = b.dlen + 3;
  for (int r = threadIdx.x; r < 1024; r += ENC_NT) {
    const RunSplit s = split_run(r);
    uint64_t lo = 0, hi = 0;
    int len = 0;
    auto app = [&](uint64_t v, int l) {
      // appends l (<= 64) bits; strings past 128 bits are only measured
      if (len + l <= 128) {
        if (len < 64) {
          lo |= v << len;
          if (len + l > 64) hi |= v >> (64 - len);
        } else {
          hi |= v << (len - 64);
        }
      }
      len += l;
    };
    for (int k = 6; k >= 0; k--)
      if (s.hi & (1u << k)) app(b.dcode | ((uint64_t)k << b.dlen), u);
    for (uint32_t i = 0; i < s.singles; i++) app(b.zcode, b.zlen);
    rc[r].lo = lo;
    rc[r].hi = hi;
    rl[r] = (uint16_t)(len < 65535 ? len : 65535);
    ok = ok && len <= 128;
  }
  return __syncthreads_and(ok) != 0;
}

__device__ inline uint32_t run_bits(const Book &b, int64_t r) {
  const RunSplit s = split_run(r);
  return (uint32_t)((uint64_t)(s.q + __popc(s.hi)) * (uint64_t)(b.dlen + 3) +
                    (uint64_t)s.singles * (uint64_t)b.zlen);
}

// Output bits of the thread's bins.
__device__ inline uint32_t thread_bits(const EncArgs &A, const Book &b, int64_t prev, int64_t c,
                                       const int32_t e[ENC_EPT], bool nd, int pz) {
  uint32_t bits = 0;
  walk_bins(A, prev, c, e, nd, pz, [&](int kind, int64_t x) {
    bits += kind == 0 ? (uint32_t)code_len(b.tab[clamp_bin(A, x)]) : run_bits(b, x);
  });
  return bits;
}

// bits per range from its histogram and run sums
template <bool BATCH>
__global__ void __launch_bounds__(ENC_NT) enc_range_bits_hist(EncArgs A0, const EncArgs *J, const int *JB, int nj) {
  const EncArgs A = BATCH ? J[enc_job(JB, nj, (int)blockIdx.x)] : A0;
  __shared__ uint64_t red64[ENC_NW];
  const int64_t b = range_id(A);
  uint64_t acc = 0;
  for (int v = threadIdx.x; v < A.nb1; v += ENC_NT)
    acc += (uint64_t)A.range_hist[b * A.nb1 + v] * (uint64_t)code_len(A.book[v]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += shfl_xor64(acc, d);
  if ((threadIdx.x & 63) == 0) red64[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int i = 0; i < ENC_NW; i++) t += red64[i];
    if (A.zb >= 0) {
      const uint64_t zl = (uint64_t)code_len(A.book[A.zb]);
      const uint64_t dl = (uint64_t)code_len(A.book[A.nb1 - 1]) + 3u;
      t += A.range_sd[2 * b] * zl + A.range_sd[2 * b + 1] * dl;
    }
    A.range_bits[b] = t;
  }
}

// bits per range by walking it (code books too large for LDS)
template <bool BATCH>
__global__ void __launch_bounds__(ENC_NT) enc_range_bits_walk(EncArgs A0, const EncArgs *J, const int *JB, int nj) {
  const EncArgs A = BATCH ? J[enc_job(JB, nj, (int)blockIdx.x)] : A0;
  __shared__ int first[ENC_NT];
  __shared__ int red[ENC_NW];
  __shared__ uint64_t red64[ENC_NW];
  const Book bk = load_book<false>(A, nullptr);
  int64_t prev = A.zb >= 0 ? (int64_t)block_reduce<true>((const uint64_t *)A.range_last,
                                                           range_id(A), red64)
                           : -1;
  uint64_t bits = 0;
  int64_t cb, ce;
  range_of(A, cb, ce);
  for (int64_t c = cb; c < ce; c++) {
    int32_t e[ENC_EPT];
    bool nd;
    __syncthreads();
    load_bins(A, c, e, first, nd);
    int tot;
    const int pz = block_excl_max(thread_last_nz(A, c, e), red, &tot);
    bits += thread_bits(A, bk, prev, c, e, nd, pz);
    if (tot >= 0) prev = c * ENC_C + tot;
  }
  const uint64_t t = block_all<false>(bits, red64);
  if (threadIdx.x == 0) A.range_bits[range_id(A)] = t;
}

// q copies of the u-bit string unit from bit p of dst, one word per thread
// at a time; inner words belong to the string alone and are stored, its
// first and last word are ORed.
__device__ inline void put_periodic(uint32_t *dst, uint64_t p, int64_t q, uint64_t unit, int u) {
  const uint64_t end = p + (uint64_t)q * (uint64_t)u;
  const uint64_t w0 = p >> 5, w1 = (end - 1) >> 5;
  for (uint64_t w = w0 + threadIdx.x; w <= w1; w += ENC_NT) {
    const uint64_t wb = w << 5;
    const uint64_t a = wb > p ? wb : p;
    const uint64_t b = wb + 32 < end ? wb + 32 : end;
    uint32_t o = (uint32_t)((a - p) % (uint64_t)u);
    uint32_t word = 0;
    for (uint64_t pos = a; pos < b;) {
      const uint64_t room = b - pos;
      const uint32_t take = (uint64_t)(u - o) < room ? (uint32_t)(u - o) : (uint32_t)room;
      const uint64_t piece = (unit >> o) & ((1ull << take) - 1ull);
      word |= (uint32_t)(piece << (pos - wb));
      pos += take;
      o = 0;
    }
    if (w == w0 || w == w1) atomicOr(&dst[w], word);
    else dst[w] = word;
  }
}

// Sequential bit packing over the thread's own output range: whole words
// are stored once; the first word it writes and the last (partial) one
// can hold a neighbour's bits and are ORed in.
struct BitWriter {
  uint32_t *dst;
  uint64_t w;       // word the pending bits start in
  uint64_t acc;     // pending bits, from bit 0 of word w
  int fill;         // pending bit count (< 32 between calls)
  bool first;
  __device__ __forceinline__ void start(uint32_t *d, uint64_t pos) {
    dst = d;
    w = pos >> 5;
    fill = (int)(pos & 31);
    acc = 0;
    first = true;
  }
  __device__ __forceinline__ void put32(uint32_t v, int len) {   // len <= 32, v < 2^len
    acc |= (uint64_t)v << fill;
    fill += len;
    if (fill >= 32) {
      const uint32_t word = (uint32_t)acc;
      if (first) {
        if (word) atomicOr(&dst[w], word);
        first = false;
      } else {
        dst[w] = word;
      }
      acc >>= 32;
      w++;
      fill -= 32;
    }
  }
  __device__ __forceinline__ void put(uint64_t v, int len) {   // len <= 64
    if (len > 32) {
      put32((uint32_t)v, 32);
      put32((uint32_t)(v >> 32), len - 32);
    } else if (len > 0) {
      put32((uint32_t)v, len);
    }
  }
  __device__ __forceinline__ uint64_t pos() const { return (w << 5) + (uint64_t)fill; }
  __device__ __forceinline__ void finish() {
    if (fill > 0 && (uint32_t)acc) atomicOr(&dst[w], (uint32_t)acc);
  }
  // leaves the next bits to another writer
  __device__ __forceinline__ void skip(uint64_t bits) {
    const uint64_t p = pos() + bits;
    finish();
    start(dst, p);
  }
};

// Codes the thread's bins from bit pos of dst.  A run with more than
// ENC_LONG_Q dup codes of e = 7 leaves them to the block: its position
// and count go to *lp / *lq.
// rc / rl: the run strings (null: runs are coded code by code).
__device__ inline void emit_bins(const EncArgs &A, const Book &b, int64_t prev, int64_t c,
                                 const int32_t e[ENC_EPT], bool nd, int pz, uint32_t *dst,
                                 uint64_t pos, uint64_t *lp, int64_t *lq, const RunCodes *rc,
                                 const uint16_t *rl) {
  BitWriter bw;
  bw.start(dst, pos);
  walk_bins(A, prev, c, e, nd, pz, [&](int kind, int64_t x) {
    if (kind == 0) {
      const uint64_t cd = b.tab[clamp_bin(A, x)];
      bw.put(code_bits(cd), code_len(cd));
      return;
    }
    const int u = b.dlen + 3;
    const int64_t q = x >> 10;
    if (q > ENC_LONG_Q) {
      *lp = bw.pos();
      *lq = q;
      bw.skip((uint64_t)q * (uint64_t)u);
    } else {
      for (int64_t i = 0; i < q; i++) bw.put(b.dcode | (7ull << b.dlen), u);
    }
    if (rc) {
      const int r1 = (int)(x & 1023);
      const RunCodes p = rc[r1];
      const int l = rl[r1];
      bw.put(p.lo, l < 64 ? l : 64);
      if (l > 64) bw.put(p.hi, l - 64);
      return;
    }
    const RunSplit s = split_run(x);
    for (uint32_t hi = s.hi; hi;) {
      const int k = 31 - __clz(hi);
      bw.put(b.dcode | ((uint64_t)k << b.dlen), u);
      hi &= ~(1u << k);
    }
    for (uint32_t i = 0; i < s.singles; i++) bw.put(b.zcode, b.zlen);
  });
  bw.finish();
}

// write: the block codes its range chunk by chunk from its bit offset
// (the sum of the ranges before it); TAB_LDS: the code book is copied to
// dynamic LDS.  Per chunk: one barrier for the run scan, one for the bit
// offsets, one after coding into the staged words (plus one when a long
// run's dup codes are written by the block); staged words are cleared as
// they are copied out, so the next chunk finds them zero.
template <bool TAB_LDS, bool PREFETCH, bool BATCH>
__global__ void __launch_bounds__(ENC_NT) enc_write(EncArgs A0, const EncArgs *J, const int *JB, int nj) {
  const EncArgs A = BATCH ? J[enc_job(JB, nj, (int)blockIdx.x)] : A0;
  extern __shared__ uint64_t tab[];
  __shared__ uint32_t stage[ENC_STAGE];
  __shared__ int first[2][ENC_NT];
  __shared__ int red[2][ENC_NW];
  __shared__ uint32_t red32[ENC_NW];
  __shared__ uint64_t red64[ENC_NW];
  __shared__ uint64_t long_p;
  __shared__ int64_t long_q;
  __shared__ RunCodes rcodes[1024];
  __shared__ uint16_t rlen[1024];
  const Book bk = load_book<TAB_LDS>(A, tab);
  const bool fast = A.zb >= 0 && build_run_codes(bk, rcodes, rlen);
  const RunCodes *rc = fast ? rcodes : nullptr;
  for (int i = threadIdx.x; i < ENC_STAGE; i += ENC_NT) stage[i] = 0;
  if (threadIdx.x == 0) long_q = 0;
  int64_t prev = A.zb >= 0 ? (int64_t)block_reduce<true>((const uint64_t *)A.range_last,
                                                           range_id(A), red64)
                           : -1;
  uint64_t off = block_reduce<false>(A.range_bits, range_id(A), red64);
  int64_t cb, ce;
  range_of(A, cb, ce);
  for_chunks<PREFETCH>(A, cb, ce, [&](int64_t c, const Chunk &k, int par) {
    int tot;
    bool nd;
    const int pz = chunk_scan(A, c, k, first[par], red[par], &tot, &nd);
    const uint32_t bits = thread_bits(A, bk, prev, c, k.e, nd, pz);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_sum(bits);
    if (lane == 63) red32[w] = inc;
    // every thread read long_q of the chunk before, before the last barrier
    if (threadIdx.x == 0) long_q = 0;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < ENC_NW; i++) {
      if (i < w) before += red32[i];
      total += red32[i];
    }
    before += inc - bits;
    const uint64_t base = off;
    const int64_t cprev = prev;
    off += total;
    if (tot >= 0) prev = c * ENC_C + tot;
    // uniform: nothing to write, or (with counts that do not belong to the
    // bins) a range past the payload buffer, which is left unwritten
    if (total == 0 || base + total > A.nbits_cap) return;
    const uint64_t wfirst = base >> 5;
    const uint64_t nw = ((base + total - 1) >> 5) - wfirst + 1;
    const bool staged = nw <= ENC_STAGE;
    uint32_t *dst = staged ? stage : A.words;
    const uint64_t pos = (staged ? (base & 31) : base) + before;
    uint64_t lp = 0;
    int64_t lq = 0;
    emit_bins(A, bk, cprev, c, k.e, nd, pz, dst, pos, &lp, &lq, rc, rlen);
    if (lq) {
      long_p = lp;
      long_q = lq;
    }
    __syncthreads();
    if (long_q) {
      put_periodic(dst, long_p, long_q, bk.dcode | (7ull << bk.dlen), bk.dlen + 3);
      __syncthreads();
    }
    if (staged) {
      for (int i = threadIdx.x; i < (int)nw; i += ENC_NT) {
        const uint32_t v = stage[i];
        stage[i] = 0;
        if (i == 0 || i == (int)nw - 1) {
          if (v) atomicOr(&A.words[wfirst + i], v);
        } else {
          gptr(A.words)[wfirst + i] = v;
        }
      }
    }
  });
}

// The payload writer for code books that fit LDS and run strings of at
// most 120 bits (the common case).  One table serves every coded event:
// entry v < nb1 is symbol v's code (bits | length << 56); from run0 (nb1
// rounded up to even) two entries per r < 1024 hold the codes of a run of
// r (lo, then hi | length << 56).  A non-duplicate bin and the last bin of
// a run then take the same path -- one paired LDS read and one put -- so
// a wave whose lanes hold both kinds runs it once.  Runs of 1024 or more
// add their q leading dup codes of e = 7.
struct Event {
  uint64_t lo, hi;
  int len;
  int64_t q;
};

template <typename F>
__device__ inline void walk_events(const EncArgs &A, const uint64_t *ctab, int run0, int64_t prev,
                                   int64_t c, const int32_t e[ENC_EPT], bool next_dup, int pz,
                                   F &&f) {
  const int64_t c0 = c * ENC_C;
  const int64_t i0 = c0 + (int64_t)ENC_EPT * threadIdx.x;
  const int64_t left = A.n - i0;
  const int nv = left >= ENC_EPT ? ENC_EPT : (left > 0 ? (int)left : 0);
  int32_t r[ENC_EPT];
#pragma unroll
  for (int j = 0; j < ENC_EPT; j++) r[j] = e[j];
#pragma unroll 1
  for (int k = 0; k < nv; k++) {
    const int v = r[0];
    const bool nz = v != A.zb;
    const bool rend = !nz && (k + 1 < ENC_EPT ? r[1] != A.zb : !next_dup);
    const int64_t rl = i0 + k - (pz >= 0 ? c0 + pz : prev);
    if (nz) pz = ENC_EPT * threadIdx.x + k;
    if (nz || rend) {
      const int idx = nz ? clamp_bin(A, v) : run0 + 2 * (int)(rl & 1023);
      const uint64_t x0 = ctab[idx], x1 = ctab[idx + 1];
      Event ev;
      ev.lo = nz ? code_bits(x0) : x0;
      ev.hi = nz ? 0 : code_bits(x1);
      ev.len = nz ? code_len(x0) : code_len(x1);
      ev.q = nz ? 0 : rl >> 10;
      f(ev);
    }
#pragma unroll
    for (int j = 0; j + 1 < ENC_EPT; j++) r[j] = r[j + 1];
  }
}

template <bool BATCH>
__global__ void __launch_bounds__(ENC_NT) enc_write_fast(EncArgs A0, const EncArgs *J, const int *JB, int nj) {
  const EncArgs A = BATCH ? J[enc_job(JB, nj, (int)blockIdx.x)] : A0;
  extern __shared__ uint64_t ctab[];
  __shared__ uint32_t stage[ENC_STAGE];
  __shared__ int first[2][ENC_NT];
  __shared__ int red[2][ENC_NW];
  __shared__ uint32_t red32[ENC_NW];
  __shared__ uint64_t red64[ENC_NW];
  __shared__ uint64_t long_p;
  __shared__ int64_t long_q;
  const int run0 = (A.nb1 + 1) & ~1;
  for (int i = threadIdx.x; i < A.nb1; i += ENC_NT) ctab[i] = A.book[i];
  if (threadIdx.x == 0 && run0 > A.nb1) ctab[A.nb1] = 0;
  const Book bk = load_book<false>(A, nullptr);
  const int u = bk.dlen + 3;
  const uint64_t unit7 = bk.dcode | (7ull << bk.dlen);
  if (A.zb >= 0) {
    for (int r = threadIdx.x; r < 1024; r += ENC_NT) {
      const RunSplit sp = split_run(r);
      uint64_t lo = 0, hi = 0;
      int len = 0;
      auto app = [&](uint64_t v, int l) {   // the host checked len <= 120
        if (len < 64) {
          lo |= v << len;
          if (len + l > 64) hi |= v >> (64 - len);
        } else {
          hi |= v << (len - 64);
        }
        len += l;
      };
      for (int k = 6; k >= 0; k--)
        if (sp.hi & (1u << k)) app(bk.dcode | ((uint64_t)k << bk.dlen), u);
      for (uint32_t i = 0; i < sp.singles; i++) app(bk.zcode, bk.zlen);
      ctab[run0 + 2 * r] = lo;
      ctab[run0 + 2 * r + 1] = (hi & ((1ull << 56) - 1)) | ((uint64_t)len << 56);
    }
  }
  for (int i = threadIdx.x; i < ENC_STAGE; i += ENC_NT) stage[i] = 0;
  if (threadIdx.x == 0) long_q = 0;
  int64_t prev = A.zb >= 0 ? (int64_t)block_reduce<true>((const uint64_t *)A.range_last,
                                                           range_id(A), red64)
                           : -1;
  uint64_t off = block_reduce<false>(A.range_bits, range_id(A), red64);
  int64_t cb, ce;
  range_of(A, cb, ce);
  for_chunks<false>(A, cb, ce, [&](int64_t c, const Chunk &k, int par) {
    int tot;
    bool nd;
    const int pz = chunk_scan(A, c, k, first[par], red[par], &tot, &nd);
    uint32_t bits = 0;
    walk_events(A, ctab, run0, prev, c, k.e, nd, pz, [&](const Event &ev) {
      bits += (uint32_t)ev.len + (uint32_t)ev.q * (uint32_t)u;
    });
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_sum(bits);
    if (lane == 63) red32[w] = inc;
    // every thread read long_q of the chunk before, before the last barrier
    if (threadIdx.x == 0) long_q = 0;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < ENC_NW; i++) {
      if (i < w) before += red32[i];
      total += red32[i];
    }
    before += inc - bits;
    const uint64_t base = off;
    const int64_t cprev = prev;
    off += total;
    if (tot >= 0) prev = c * ENC_C + tot;
    if (total == 0 || base + total > A.nbits_cap) return;   // uniform
    const uint64_t wfirst = base >> 5;
    const uint64_t nw = ((base + total - 1) >> 5) - wfirst + 1;
    const bool staged = nw <= ENC_STAGE;
    uint32_t *dst = staged ? stage : A.words;
    BitWriter bw;
    bw.start(dst, (staged ? (base & 31) : base) + before);
    walk_events(A, ctab, run0, cprev, c, k.e, nd, pz, [&](const Event &ev) {
      if (ev.q > ENC_LONG_Q) {
        long_p = bw.pos();
        long_q = ev.q;
        bw.skip((uint64_t)ev.q * (uint64_t)u);
      } else {
        for (int64_t i = 0; i < ev.q; i++) bw.put(unit7, u);
      }
      bw.put(ev.lo, ev.len < 64 ? ev.len : 64);
      if (ev.len > 64) bw.put(ev.hi, ev.len - 64);
    });
    bw.finish();
    __syncthreads();
    if (long_q) {
      put_periodic(dst, long_p, long_q, unit7, u);
      __syncthreads();
    }
    if (staged) {
      for (int i = threadIdx.x; i < (int)nw; i += ENC_NT) {
        const uint32_t v = stage[i];
        stage[i] = 0;
        if (i == 0 || i == (int)nw - 1) {
          if (v) atomicOr(&A.words[wfirst + i], v);
        } else {
          gptr(A.words)[wfirst + i] = v;
        }
      }
    }
  });
}

__global__ void enc_zero(uint32_t *p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

__global__ void enc_zero64(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// batches: the run-statistics of job j go to sd[2j..2j+1]; its histogram
// (counts[j], m entries) is packed into dst + j * m
__global__ void enc_zero_sd_batch(const EncArgs *J, int nj) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x) {
    J[j].sd[0] = 0;
    J[j].sd[1] = 0;
  }
}

__global__ void enc_gather_counts(const uint32_t *const *src, int m, uint32_t *dst) {
  const uint32_t *c = src[blockIdx.x];
  for (int i = threadIdx.x; i < m; i += blockDim.x) dst[(int64_t)blockIdx.x * m + i] = c[i];
}

// ---- host side ----------------------------------------------------------

namespace {

inline int64_t al256(int64_t v) { return (v + 255) & ~(int64_t)255; }

struct EncLayout {
  int64_t nchunks, per;
  int nranges;
  bool hist;
  int64_t last, bits, rsd, sd, book, hist_off, total;
};

EncLayout enc_layout(int64_t n, int bin_count) {
  EncLayout L;
  L.nchunks = (n + ENC_C - 1) / ENC_C;
  const int64_t g = L.nchunks < ENC_MAX_RANGES ? L.nchunks : ENC_MAX_RANGES;
  L.per = (L.nchunks + g - 1) / g;
  L.nranges = (int)((L.nchunks + L.per - 1) / L.per);
  const int64_t nb1 = (int64_t)bin_count + 1;
  L.hist = nb1 <= ENC_TAB;
  int64_t o = 0;
  L.last = o; o += al256(L.nranges * 8);
  L.bits = o; o += al256(L.nranges * 8);
  L.rsd = o; o += al256(L.nranges * 16);
  L.sd = o; o += 256;
  L.book = o; o += al256(nb1 * 8);
  L.hist_off = o; o += L.hist ? al256(L.nranges * nb1 * 4) : 0;
  L.total = o;
  return L;
}

EncArgs enc_args(const int32_t *bins, int64_t n, int zb, void *ws, int bin_count,
                 const EncLayout &L) {
  char *w = (char *)ws;
  EncArgs A{};
  A.bins = bins;
  A.n = n;
  A.nchunks = L.nchunks;
  A.zb = zb;
  A.nb1 = zb >= 0 ? bin_count + 1 : bin_count;
  A.vec = ((uintptr_t)bins & 15) == 0;
  A.nranges = L.nranges;
  A.per = L.per;
  A.range_last = (int64_t *)(w + L.last);
  A.range_bits = (uint64_t *)(w + L.bits);
  A.range_sd = (unsigned long long *)(w + L.rsd);
  A.sd = (unsigned long long *)(w + L.sd);
  A.book = (const uint64_t *)(w + L.book);
  A.range_hist = L.hist ? (uint32_t *)(w + L.hist_off) : nullptr;
  return A;
}

// WV_ENC_PREFETCH=1: the payload writer keeps the next chunk's loads in
// flight (more registers per thread, fewer resident blocks)
bool enc_prefetch() {
  static const bool v = [] {
    const char *e = getenv("WV_ENC_PREFETCH");
    return e && e[0] == '1';
  }();
  return v;
}

}  // namespace

int64_t enc_ws_bytes(int64_t n, int bin_count) { return enc_layout(n, bin_count).total; }

uint64_t *enc_book_ptr(void *ws, int64_t n, int bin_count) {
  return (uint64_t *)((char *)ws + enc_layout(n, bin_count).book);
}

hipError_t enc_count_runs(const int32_t *bins, int64_t n, int zb, int bin_count, void *ws,
                          unsigned long long *sd_host, hipStream_t s) {
  const EncLayout L = enc_layout(n, bin_count);
  const EncArgs A = enc_args(bins, n, zb, ws, bin_count, L);
  hipLaunchKernelGGL(enc_zero64, dim3(1), dim3(64), 0, s, A.sd, 2);
  if (zb >= 0)
    hipLaunchKernelGGL(enc_range_last<false>, dim3(L.nranges), dim3(ENC_NT), 0, s, A, nullptr, nullptr, 0);
  if (L.hist)
    hipLaunchKernelGGL((enc_count<true, false>), dim3(L.nranges), dim3(ENC_NT), A.nb1 * 4, s, A, nullptr, nullptr, 0);
  else if (zb >= 0)
    hipLaunchKernelGGL((enc_count<false, false>), dim3(L.nranges), dim3(ENC_NT), 0, s, A, nullptr, nullptr, 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(sd_host, A.sd, 16, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t enc_write_payload(const int32_t *bins, int64_t n, int zb, int bin_count, int nb1,
                             bool short_runs, void *ws, uint32_t *words, int64_t nwords,
                             hipStream_t s) {
  const EncLayout L = enc_layout(n, bin_count);
  EncArgs A = enc_args(bins, n, zb, ws, bin_count, L);
  if (A.nb1 != nb1) return hipErrorInvalidValue;
  A.words = words;
  A.nbits_cap = (uint64_t)nwords * 32u;
  int64_t zg = (nwords + 255) / 256;
  if (zg > 4096) zg = 4096;
  if (zg < 1) zg = 1;
  hipLaunchKernelGGL(enc_zero, dim3((unsigned)zg), dim3(256), 0, s, words, nwords);
  const bool pf = enc_prefetch();
  if (L.hist && short_runs && !pf) {
    hipLaunchKernelGGL(enc_range_bits_hist<false>, dim3(L.nranges), dim3(ENC_NT), 0, s, A, nullptr, nullptr, 0);
    const size_t lds = (size_t)(((nb1 + 1) & ~1) + 2048) * 8;
    hipLaunchKernelGGL(enc_write_fast<false>, dim3(L.nranges), dim3(ENC_NT), lds, s, A, nullptr, nullptr, 0);
  } else if (L.hist) {
    hipLaunchKernelGGL(enc_range_bits_hist<false>, dim3(L.nranges), dim3(ENC_NT), 0, s, A, nullptr, nullptr, 0);
    if (pf)
      hipLaunchKernelGGL((enc_write<true, true, false>), dim3(L.nranges), dim3(ENC_NT),
                         (size_t)nb1 * 8, s, A, nullptr, nullptr, 0);
    else
      hipLaunchKernelGGL((enc_write<true, false, false>), dim3(L.nranges), dim3(ENC_NT),
                         (size_t)nb1 * 8, s, A, nullptr, nullptr, 0);
  } else {
    hipLaunchKernelGGL(enc_range_bits_walk<false>, dim3(L.nranges), dim3(ENC_NT), 0, s, A, nullptr, nullptr, 0);
    hipLaunchKernelGGL((enc_write<false, false, false>), dim3(L.nranges), dim3(ENC_NT), 0, s, A, nullptr, nullptr, 0);
  }
  return hipGetLastError();
}

// ---- batches of jobs with one code-book size --------------------------
// ws: [job table][blk0][sd pairs][count pointers][packed counts][packed
// books][per-job workspaces].  Every phase is one launch over all jobs'
// ranges (block b serves the job with the last blk0 <= b).
EncBatch enc_batch_plan(const int32_t *const *bins, const int64_t *n, const int *zb, int nj,
                        int bin_count, void *ws) {
  EncBatch P;
  P.nj = nj;
  P.bin_count = bin_count;
  P.nb1 = nj > 0 && zb[0] >= 0 ? bin_count + 1 : bin_count;
  P.hist = P.nb1 <= ENC_TAB;
  int64_t o = 0;
  P.o_jobs = o; o += al256((int64_t)nj * sizeof(EncArgs));
  P.o_blk0 = o; o += al256((int64_t)nj * 4);
  P.o_sd = o; o += al256((int64_t)nj * 16);
  P.o_ptrs = o; o += al256((int64_t)nj * 8);
  P.o_counts = o; o += al256((int64_t)nj * (bin_count + 1) * 4);
  P.o_books = o; o += al256((int64_t)nj * P.nb1 * 8);
  P.jobs.resize((size_t)nj);
  P.blk0.resize((size_t)nj);
  P.ok = true;
  int blocks = 0;
  char *w = (char *)ws;
  for (int j = 0; j < nj; j++) {
    const EncLayout L = enc_layout(n[j], bin_count);
    EncArgs A = enc_args(bins[j], n[j], zb[j], w ? w + o : nullptr, bin_count, L);
    if (A.nb1 != P.nb1) P.ok = false;   // every job with or every job without a dup key
    A.sd = w ? (unsigned long long *)(w + P.o_sd) + 2 * j : nullptr;
    A.book = w ? (const uint64_t *)(w + P.o_books) + (int64_t)j * P.nb1 : nullptr;
    A.blk0 = blocks;
    P.blk0[j] = blocks;
    blocks += L.nranges;
    P.jobs[j] = A;
    o += al256(L.total);
  }
  P.blocks = blocks;
  P.total = o;
  return P;
}

hipError_t enc_batch_counts(EncBatch &P, const uint32_t *const *counts_d, void *ws,
                            uint32_t *counts_host, unsigned long long *sd_host, hipStream_t s) {
  if (!P.ok || P.nj <= 0) return hipErrorInvalidValue;
  char *w = (char *)ws;
  EncArgs *J = (EncArgs *)(w + P.o_jobs);
  const int *B = (const int *)(w + P.o_blk0);
  const int m = P.bin_count + 1;
  hipError_t e = hipMemcpyAsync(J, P.jobs.data(), P.jobs.size() * sizeof(EncArgs),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(w + P.o_blk0, P.blk0.data(), P.blk0.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(w + P.o_ptrs, counts_d, (size_t)P.nj * 8, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  const bool dup = P.jobs[0].zb >= 0;
  hipLaunchKernelGGL(enc_zero_sd_batch, dim3((P.nj + 255) / 256), dim3(256), 0, s, J, P.nj);
  if (dup)
    hipLaunchKernelGGL(enc_range_last<true>, dim3(P.blocks), dim3(ENC_NT), 0, s, EncArgs{}, J, B,
                       P.nj);
  if (P.hist)
    hipLaunchKernelGGL((enc_count<true, true>), dim3(P.blocks), dim3(ENC_NT), P.nb1 * 4, s,
                       EncArgs{}, J, B, P.nj);
  else if (dup)
    hipLaunchKernelGGL((enc_count<false, true>), dim3(P.blocks), dim3(ENC_NT), 0, s, EncArgs{},
                       J, B, P.nj);
  hipLaunchKernelGGL(enc_gather_counts, dim3(P.nj), dim3(256), 0, s,
                     (const uint32_t *const *)(w + P.o_ptrs), m, (uint32_t *)(w + P.o_counts));
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(counts_host, w + P.o_counts, (size_t)P.nj * m * 4, hipMemcpyDeviceToHost,
                       s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(sd_host, w + P.o_sd, (size_t)P.nj * 16, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

hipError_t enc_batch_payload(EncBatch &P, void *ws, const uint64_t *books_host, bool short_runs,
                             uint32_t *words, const int64_t *woff, const int64_t *nwords,
                             int64_t total_words, hipStream_t s) {
  if (!P.ok || P.nj <= 0) return hipErrorInvalidValue;
  char *w = (char *)ws;
  EncArgs *J = (EncArgs *)(w + P.o_jobs);
  const int *B = (const int *)(w + P.o_blk0);
  for (int j = 0; j < P.nj; j++) {
    P.jobs[j].words = words + woff[j];
    P.jobs[j].nbits_cap = (uint64_t)nwords[j] * 32u;
  }
  hipError_t e = hipMemcpyAsync(J, P.jobs.data(), P.jobs.size() * sizeof(EncArgs),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(w + P.o_books, books_host, (size_t)P.nj * P.nb1 * 8,
                       hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  int64_t zg = (total_words + 255) / 256;
  if (zg > 4096) zg = 4096;
  if (zg < 1) zg = 1;
  hipLaunchKernelGGL(enc_zero, dim3((unsigned)zg), dim3(256), 0, s, words, total_words);
  const int nb1 = P.nb1;
  if (P.hist && short_runs) {
    hipLaunchKernelGGL(enc_range_bits_hist<true>, dim3(P.blocks), dim3(ENC_NT), 0, s, EncArgs{},
                       J, B, P.nj);
    const size_t lds = (size_t)(((nb1 + 1) & ~1) + 2048) * 8;
    hipLaunchKernelGGL(enc_write_fast<true>, dim3(P.blocks), dim3(ENC_NT), lds, s, EncArgs{}, J,
                       B, P.nj);
  } else if (P.hist) {
    hipLaunchKernelGGL(enc_range_bits_hist<true>, dim3(P.blocks), dim3(ENC_NT), 0, s, EncArgs{},
                       J, B, P.nj);
    hipLaunchKernelGGL((enc_write<true, false, true>), dim3(P.blocks), dim3(ENC_NT),
                       (size_t)nb1 * 8, s, EncArgs{}, J, B, P.nj);
  } else {
    hipLaunchKernelGGL(enc_range_bits_walk<true>, dim3(P.blocks), dim3(ENC_NT), 0, s, EncArgs{},
                       J, B, P.nj);
    hipLaunchKernelGGL((enc_write<false, false, true>), dim3(P.blocks), dim3(ENC_NT), 0, s,
                       EncArgs{}, J, B, P.nj);
  }
  return hipGetLastError();
}

}  // namespace wv
