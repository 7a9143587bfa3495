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
//   last     per chunk of ENC_C bins, the index of its last bin that is
//            not the duplicate value; an exclusive prefix max over chunks
//            then gives every run of the duplicate value its start
//   count    each run's split into dup codes and single codes (below);
//            their sums replace the duplicate value's histogram count and
//            give the dup key's count
//   bits     bits of every chunk's output; exclusive prefix sum over chunks
//   write    every chunk codes its bins at its bit offset into LDS words,
//            then stores them (the first and last word, which neighbouring
//            chunks share, by atomic OR into the zeroed payload)
//
// A run of r copies of the duplicate value (dupEncodeString, huffman.h:
// 278-296) codes as q = r / 1024 dup codes with e = 7, then one dup code
// per set bit of r mod 1024 at positions 9..3 (e = position - 3, highest
// first), except that a remainder of exactly 8 stays as 8 single codes;
// the rest (at most 8) are single codes of the value.  A dup code is the
// dup key's code followed by e in 3 bits, LSB first.  Only the first run
// that ends in a chunk can have started before it, so at most one run per
// chunk has q > 4: its q dup codes are a periodic bit string that the
// whole block writes word by word.
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define ENC_NT 1024
#define ENC_EPT 4
#define ENC_C (ENC_NT * ENC_EPT)   // bins per chunk
#define ENC_NW (ENC_NT / 64)
#define ENC_TAB 4097               // code-book entries held in LDS
#define ENC_STAGE 8192             // payload words staged in LDS per chunk
#define ENC_SENT ((int32_t)0x80000000)   // stands for bins past the end
#define ENC_LONG_Q 4               // runs with more dup codes are written
                                   // by the whole block

struct EncArgs {
  const int32_t *bins;
  int64_t n;
  int64_t nchunks;
  int zb;                    // duplicate value; -1: no duplicate coding
  int nb1;                   // code-book entries (the dup key is the last)
  int vec;                   // bins are 16-byte aligned
  int64_t *last_nz;          // per chunk: its last non-duplicate index, -1
  int64_t *prev_nz;          // per chunk: the last one before it, -1
  const uint64_t *book;      // codes: bits | length << 56
  uint64_t *chunk_bits;
  const uint64_t *chunk_off;
  uint32_t *words;
  uint64_t nbits_cap;        // bits the payload buffer holds
  unsigned long long *sd;    // [0] single codes of zb, [1] dup codes
};

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

__device__ inline uint64_t shfl_up64(uint64_t v, int d) {
  const uint32_t lo = __shfl_up((uint32_t)v, d, 64);
  const uint32_t hi = __shfl_up((uint32_t)(v >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
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

// Exclusive max over the block's threads in order (-1 before thread 0).
// red: ENC_NW ints of LDS.  Ends with a barrier, so red can be reused.
__device__ inline int block_excl_max(int v, int *red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_incl_max(v);
  if (lane == 63) red[w] = inc;
  __syncthreads();
  int before = -1;
  for (int i = 0; i < w; i++) before = red[i] > before ? red[i] : before;
  int ex = __shfl_up(inc, 1, 64);
  if (lane == 0) ex = -1;
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

// The thread's ENC_EPT bins of chunk c (ENC_SENT past the end) and whether
// the bin after its last one is the duplicate value.  first: ENC_NT ints
// of LDS; the caller has passed a barrier since their last use.
__device__ inline void load_bins(const EncArgs &A, int64_t c0, int32_t e[ENC_EPT],
                                 int *first, bool &next_dup) {
  const int t = threadIdx.x;
  const int64_t i0 = c0 + (int64_t)ENC_EPT * t;
  if (A.vec && i0 + ENC_EPT <= A.n) {
    const int4 v = *(const int4 *)(A.bins + i0);
    e[0] = v.x; e[1] = v.y; e[2] = v.z; e[3] = v.w;
  } else {
#pragma unroll
    for (int k = 0; k < ENC_EPT; k++) e[k] = i0 + k < A.n ? A.bins[i0 + k] : ENC_SENT;
  }
  first[t] = e[0];
  __syncthreads();
  int nx;
  if (t + 1 < ENC_NT) {
    nx = first[t + 1];
  } else {
    const int64_t j = c0 + ENC_C;
    nx = j < A.n ? A.bins[j] : ENC_SENT;
  }
  next_dup = nx == A.zb;
}

// chunk-local index of the thread's last bin that is not the duplicate, -1
__device__ inline int thread_last_nz(const EncArgs &A, int64_t c0, const int32_t e[ENC_EPT]) {
  const int64_t i0 = c0 + (int64_t)ENC_EPT * threadIdx.x;
  int l = -1;
#pragma unroll
  for (int k = 0; k < ENC_EPT; k++)
    if (i0 + k < A.n && e[k] != A.zb) l = ENC_EPT * threadIdx.x + k;
  return l;
}

__global__ void __launch_bounds__(ENC_NT) enc_last_nz(EncArgs A) {
  __shared__ int first[ENC_NT];
  __shared__ int red[ENC_NW];
  for (int64_t c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
    const int64_t c0 = c * ENC_C;
    int32_t e[ENC_EPT];
    bool nd;
    __syncthreads();
    load_bins(A, c0, e, first, nd);
    const int m = wave_incl_max(thread_last_nz(A, c0, e));
    if ((threadIdx.x & 63) == 63) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      int b = -1;
      for (int i = 0; i < ENC_NW; i++) b = red[i] > b ? red[i] : b;
      A.last_nz[c] = b >= 0 ? c0 + b : -1;
    }
  }
}

// Exclusive prefix scan over chunks in one block of 1024 threads: MAX
// (identity -1, int64) or SUM (identity 0, uint64).
template <bool MAX>
__global__ void __launch_bounds__(1024) enc_chunk_scan(const uint64_t *in, int64_t nc,
                                                       uint64_t *out) {
  __shared__ uint64_t s[1024];
  const int t = threadIdx.x;
  const uint64_t ident = MAX ? ~0ull : 0ull;   // -1 as int64
  auto op = [](uint64_t a, uint64_t b) -> uint64_t {
    if (MAX) return (int64_t)a > (int64_t)b ? a : b;
    return a + b;
  };
  const int64_t per = (nc + 1023) / 1024;
  const int64_t b = (int64_t)t * per;
  const int64_t e = b + per < nc ? b + per : nc;
  uint64_t acc = ident;
  for (int64_t i = b; i < e; i++) acc = op(acc, in[i]);
  s[t] = acc;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const uint64_t y = t >= d ? s[t - d] : ident;
    __syncthreads();
    s[t] = op(s[t], y);
    __syncthreads();
  }
  uint64_t run = t > 0 ? s[t - 1] : ident;
  for (int64_t i = b; i < e; i++) {
    const uint64_t v = in[i];
    out[i] = run;
    run = op(run, v);
  }
}

// Walks the thread's bins in order.  pz: chunk-local index of the last
// non-duplicate bin before them (-1: none in the chunk, the run then
// started at prev_nz of the chunk + 1).  f(kind, k, value or run length):
// kind 0 = a non-duplicate bin, 1 = the last bin of a run.
template <typename F>
__device__ inline void walk_bins(const EncArgs &A, int64_t c, int64_t c0, const int32_t e[ENC_EPT],
                                 bool next_dup, int pz, F &&f) {
  const int64_t i0 = c0 + (int64_t)ENC_EPT * threadIdx.x;
#pragma unroll
  for (int k = 0; k < ENC_EPT; k++) {
    if (i0 + k >= A.n) break;
    const int v = e[k];
    if (v != A.zb) {
      f(0, k, (int64_t)v);
      pz = ENC_EPT * threadIdx.x + k;
    } else {
      const bool end = k + 1 < ENC_EPT ? e[k + 1] != A.zb : !next_dup;
      if (end) {
        const int64_t start_prev = pz >= 0 ? c0 + pz : A.prev_nz[c];
        f(1, k, i0 + k - start_prev);
      }
    }
  }
}

// count: single codes of the duplicate value and dup codes, over all runs
__global__ void __launch_bounds__(ENC_NT) enc_count(EncArgs A) {
  __shared__ int first[ENC_NT];
  __shared__ int red[ENC_NW];
  __shared__ unsigned long long red2[2 * ENC_NW];
  uint64_t singles = 0, dups = 0;
  for (int64_t c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
    const int64_t c0 = c * ENC_C;
    int32_t e[ENC_EPT];
    bool nd;
    __syncthreads();
    load_bins(A, c0, e, first, nd);
    const int pz = block_excl_max(thread_last_nz(A, c0, e), red);
    walk_bins(A, c, c0, e, nd, pz, [&](int kind, int, int64_t x) {
      if (kind == 1) {
        const RunSplit s = split_run(x);
        singles += s.singles;
        dups += (uint64_t)s.q + (uint64_t)__popc(s.hi);
      }
    });
  }
  for (int d = 32; d >= 1; d >>= 1) {
    singles += shfl_xor64(singles, d);
    dups += shfl_xor64(dups, d);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red2[2 * w] = singles;
    red2[2 * w + 1] = dups;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (int i = 0; i < ENC_NW; i++) {
      a += red2[2 * i];
      b += red2[2 * i + 1];
    }
    if (a) atomicAdd(&A.sd[0], a);
    if (b) atomicAdd(&A.sd[1], b);
  }
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

__device__ inline uint64_t book_at(const Book &b, const EncArgs &A, int64_t v) {
  // bins outside [0, nb1) do not occur (the host checks the histogram);
  // the clamp only keeps the lookup in bounds
  const int i = v < 0 ? 0 : (v >= A.nb1 ? A.nb1 - 1 : (int)v);
  return b.tab[i];
}

__device__ inline uint32_t run_bits(const Book &b, int64_t r) {
  const RunSplit s = split_run(r);
  return (uint32_t)((uint64_t)(s.q + __popc(s.hi)) * (uint64_t)(b.dlen + 3) +
                    (uint64_t)s.singles * (uint64_t)b.zlen);
}

// Output bits of the thread's bins.
__device__ inline uint32_t thread_bits(const EncArgs &A, const Book &b, int64_t c, int64_t c0,
                                       const int32_t e[ENC_EPT], bool nd, int pz) {
  uint32_t bits = 0;
  walk_bins(A, c, c0, e, nd, pz, [&](int kind, int, int64_t x) {
    bits += kind == 0 ? (uint32_t)code_len(book_at(b, A, x)) : run_bits(b, x);
  });
  return bits;
}

template <bool TAB_LDS>
__global__ void __launch_bounds__(ENC_NT) enc_bits(EncArgs A) {
  __shared__ uint64_t tab[TAB_LDS ? ENC_TAB : 1];
  __shared__ int first[ENC_NT];
  __shared__ int red[ENC_NW];
  __shared__ uint32_t red32[ENC_NW];
  const Book b = load_book<TAB_LDS>(A, tab);
  for (int64_t c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
    const int64_t c0 = c * ENC_C;
    int32_t e[ENC_EPT];
    bool nd;
    __syncthreads();
    load_bins(A, c0, e, first, nd);
    const int pz = block_excl_max(thread_last_nz(A, c0, e), red);
    uint32_t total;
    block_excl_sum(thread_bits(A, b, c, c0, e, nd, pz), red32, &total);
    if (threadIdx.x == 0) A.chunk_bits[c] = total;
  }
}

// ORs the low len (<= 64) bits of v into the bit string dst at bit p.
__device__ inline void put_bits(uint32_t *dst, uint64_t p, uint64_t v, int len) {
  if (len <= 0) return;
  const uint64_t w = p >> 5;
  const int sh = (int)(p & 31);
  const uint64_t lo = v << sh;
  atomicOr(&dst[w], (uint32_t)lo);
  if (sh + len > 32) atomicOr(&dst[w + 1], (uint32_t)(lo >> 32));
  if (sh + len > 64) atomicOr(&dst[w + 2], (uint32_t)(v >> (64 - sh)));
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

// Codes the thread's bins from bit pos of dst.  A run with more than
// ENC_LONG_Q dup codes of e = 7 leaves them to the block: its position
// and count go to *lp / *lq.
__device__ inline void emit_bins(const EncArgs &A, const Book &b, int64_t c, int64_t c0,
                                 const int32_t e[ENC_EPT], bool nd, int pz, uint32_t *dst,
                                 uint64_t pos, uint64_t *lp, int64_t *lq) {
  const uint64_t dunit = b.dcode;   // dup key code, then e in 3 bits
  walk_bins(A, c, c0, e, nd, pz, [&](int kind, int, int64_t x) {
    if (kind == 0) {
      const uint64_t cd = book_at(b, A, x);
      put_bits(dst, pos, code_bits(cd), code_len(cd));
      pos += (uint64_t)code_len(cd);
      return;
    }
    const RunSplit s = split_run(x);
    const int u = b.dlen + 3;
    if (s.q > ENC_LONG_Q) {
      *lp = pos;
      *lq = s.q;
      pos += (uint64_t)s.q * (uint64_t)u;
    } else {
      for (int64_t i = 0; i < s.q; i++) {
        put_bits(dst, pos, dunit | (7ull << b.dlen), u);
        pos += (uint64_t)u;
      }
    }
    for (int k = 6; k >= 0; k--) {
      if (s.hi & (1u << k)) {
        put_bits(dst, pos, dunit | ((uint64_t)k << b.dlen), u);
        pos += (uint64_t)u;
      }
    }
    for (uint32_t i = 0; i < s.singles; i++) {
      put_bits(dst, pos, b.zcode, b.zlen);
      pos += (uint64_t)b.zlen;
    }
  });
}

template <bool TAB_LDS>
__global__ void __launch_bounds__(ENC_NT) enc_write(EncArgs A) {
  __shared__ uint64_t tab[TAB_LDS ? ENC_TAB : 1];
  __shared__ uint32_t stage[ENC_STAGE];
  __shared__ int first[ENC_NT];
  __shared__ int red[ENC_NW];
  __shared__ uint32_t red32[ENC_NW];
  __shared__ uint64_t long_p;
  __shared__ int64_t long_q;
  const Book b = load_book<TAB_LDS>(A, tab);
  for (int64_t c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
    const int64_t c0 = c * ENC_C;
    int32_t e[ENC_EPT];
    bool nd;
    __syncthreads();
    if (threadIdx.x == 0) long_q = 0;
    load_bins(A, c0, e, first, nd);
    const int pz = block_excl_max(thread_last_nz(A, c0, e), red);
    uint32_t total;
    const uint32_t before = block_excl_sum(thread_bits(A, b, c, c0, e, nd, pz), red32, &total);
    const uint64_t base = A.chunk_off[c];
    // uniform: nothing to write, or (with counts that do not belong to the
    // bins) a range past the payload buffer, which is left unwritten
    if (total == 0 || base + total > A.nbits_cap) continue;
    const uint64_t wfirst = base >> 5;
    const uint64_t nw = ((base + total - 1) >> 5) - wfirst + 1;
    const bool staged = nw <= ENC_STAGE;
    uint32_t *dst;
    uint64_t pos;
    if (staged) {
      for (int i = threadIdx.x; i < (int)nw; i += ENC_NT) stage[i] = 0;
      __syncthreads();
      dst = stage;
      pos = (base & 31) + before;
    } else {
      dst = A.words;
      pos = base + before;
    }
    uint64_t lp = 0;
    int64_t lq = 0;
    emit_bins(A, b, c, c0, e, nd, pz, dst, pos, &lp, &lq);
    if (lq) {
      long_p = lp;
      long_q = lq;
    }
    __syncthreads();
    if (long_q) put_periodic(dst, long_p, long_q, b.dcode | (7ull << b.dlen), b.dlen + 3);
    if (staged) {
      __syncthreads();
      for (int i = threadIdx.x; i < (int)nw; i += ENC_NT) {
        const uint32_t v = stage[i];
        if (i == 0 || i == (int)nw - 1) {
          if (v) atomicOr(&A.words[wfirst + i], v);
        } else {
          A.words[wfirst + i] = v;
        }
      }
    }
  }
}

__global__ void enc_zero(uint32_t *p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

__global__ void enc_zero64(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---- host side ----------------------------------------------------------

namespace {

inline int64_t al256(int64_t v) { return (v + 255) & ~(int64_t)255; }

struct EncLayout {
  int64_t last, prev, cbits, coff, sd, book, total;
};

EncLayout enc_layout(int64_t n, int bin_count) {
  const int64_t nc = (n + ENC_C - 1) / ENC_C;
  EncLayout L;
  int64_t o = 0;
  L.last = o; o += al256(nc * 8);
  L.prev = o; o += al256(nc * 8);
  L.cbits = o; o += al256(nc * 8);
  L.coff = o; o += al256(nc * 8);
  L.sd = o; o += 256;
  L.book = o; o += al256(((int64_t)bin_count + 1) * 8);
  L.total = o;
  return L;
}

EncArgs enc_args(const int32_t *bins, int64_t n, int zb, void *ws, int bin_count) {
  const EncLayout L = enc_layout(n, bin_count);
  char *w = (char *)ws;
  EncArgs A{};
  A.bins = bins;
  A.n = n;
  A.nchunks = (n + ENC_C - 1) / ENC_C;
  A.zb = zb;
  A.vec = ((uintptr_t)bins & 15) == 0;
  A.last_nz = (int64_t *)(w + L.last);
  A.prev_nz = (int64_t *)(w + L.prev);
  A.chunk_bits = (uint64_t *)(w + L.cbits);
  A.chunk_off = (const uint64_t *)(w + L.coff);
  A.sd = (unsigned long long *)(w + L.sd);
  A.book = (const uint64_t *)(w + L.book);
  return A;
}

int enc_grid(int64_t nchunks) {
  // two 1024-thread blocks fit a CU (LDS ~70 KB each)
  const int64_t g = (int64_t)cu_count() * 2;
  return (int)(nchunks < g ? nchunks : g);
}

}  // namespace

int64_t enc_ws_bytes(int64_t n, int bin_count) { return enc_layout(n, bin_count).total; }

uint64_t *enc_book_ptr(void *ws, int64_t n, int bin_count) {
  return (uint64_t *)((char *)ws + enc_layout(n, bin_count).book);
}

hipError_t enc_count_runs(const int32_t *bins, int64_t n, int zb, int bin_count, void *ws,
                          unsigned long long *sd_host, hipStream_t s) {
  EncArgs A = enc_args(bins, n, zb, ws, bin_count);
  const int g = enc_grid(A.nchunks);
  hipLaunchKernelGGL(enc_zero64, dim3(1), dim3(64), 0, s, A.sd, 2);
  hipLaunchKernelGGL(enc_last_nz, dim3(g), dim3(ENC_NT), 0, s, A);
  hipLaunchKernelGGL(enc_chunk_scan<true>, dim3(1), dim3(1024), 0, s,
                     (const uint64_t *)A.last_nz, A.nchunks, (uint64_t *)A.prev_nz);
  hipLaunchKernelGGL(enc_count, dim3(g), dim3(ENC_NT), 0, s, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(sd_host, A.sd, 16, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t enc_write_payload(const int32_t *bins, int64_t n, int zb, int bin_count, int nb1,
                             void *ws, uint32_t *words, int64_t nwords, hipStream_t s) {
  EncArgs A = enc_args(bins, n, zb, ws, bin_count);
  A.nb1 = nb1;
  A.words = words;
  A.nbits_cap = (uint64_t)nwords * 32u;
  const int g = enc_grid(A.nchunks);
  int64_t zg = (nwords + 255) / 256;
  if (zg > 4096) zg = 4096;
  if (zg < 1) zg = 1;
  hipLaunchKernelGGL(enc_zero, dim3((unsigned)zg), dim3(256), 0, s, words, nwords);
  const bool lds = nb1 <= ENC_TAB;
  if (lds) hipLaunchKernelGGL(enc_bits<true>, dim3(g), dim3(ENC_NT), 0, s, A);
  else hipLaunchKernelGGL(enc_bits<false>, dim3(g), dim3(ENC_NT), 0, s, A);
  hipLaunchKernelGGL(enc_chunk_scan<false>, dim3(1), dim3(1024), 0, s,
                     (const uint64_t *)A.chunk_bits, A.nchunks, (uint64_t *)A.chunk_off);
  if (lds) hipLaunchKernelGGL(enc_write<true>, dim3(g), dim3(ENC_NT), 0, s, A);
  else hipLaunchKernelGGL(enc_write<false>, dim3(g), dim3(ENC_NT), 0, s, A);
  return hipGetLastError();
}

}  // namespace wv
