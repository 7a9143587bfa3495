// wv_decode.hip -- Huffman decoding on the device, for decompression.
//
// The reference decodes on the host, one bit at a time through the decode
// table (HuffmanDecoder::decodeFromStream, huffman.h:96-131: LSB-first
// words, bits past the end read as 0, a dup-key leaf reads 3 more bits e
// and stands for 2^(e+3) copies of the dup value, decoding stops at count
// values or when the stream is empty at the start of a value).  Here the
// stream is cut into subsequences of DEC_S bits, one per thread, and
// decoded in parallel using the self-synchronisation of Huffman codes:
//
//   sync     every thread decodes its subsequence from a start it assumes
//            (first i * DEC_S) up to the first code boundary at or past the
//            subsequence's end.  A thread whose start differs from where
//            its predecessor's decode ended decodes again from there; the
//            rounds repeat until no start changes (a decode from a wrong
//            position falls into step with the true code boundaries within
//            a few codes, so few rounds are needed).  The first
//            subsequence starts at bit 0, so by induction every start is
//            then a true code boundary.
//   scan     values per subsequence -> output offsets
//   write    every thread decodes its subsequence again and stores its
//            values.  When the stream has a dup key, the output is first
//            filled with the dup value, and only the other values are
//            stored: dup runs (most of a thresholded cube) cost nothing.
//
// Codes resolve through a DEC_K-bit lookup table built on the host (leaf:
// value and length; else the table node after DEC_K bits, from which the
// walk continues bit by bit).  The host checks that every child id of the
// decode table is larger than its parent's and inside the table, so every
// walk is bounded.
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define DEC_K 12
#define DEC_S 512
#define DEC_NT 256

struct DecArgs {
  const uint32_t *words;
  int64_t nwords, nbits;
  int64_t nsub;
  const uint32_t *lut;      // leaf << 31 | (value or node) << 4 | length
  const int32_t *table;     // decode table, pairs
  int n_table;
  int dup_key, dup_value;
  int64_t *start, *end_pos, *nstart;
  uint64_t *cnt, *off;
  int32_t *changed;         // per subsequence (sync rounds)
  unsigned long long *flag; // any start changed
  int32_t *out;
  int64_t count;
  int fill;                 // out was filled with dup_value
};

// The payload words of a block's subsequences (plus a few past them),
// staged in LDS: decoding reads its bits there, with no global-memory
// latency in the chain of dependent codes.  Reads past the staged words
// (long codes at the end of the last subsequence) go to global memory.
#define DEC_WPS (DEC_S / 32)
#define DEC_BW (DEC_NT * DEC_WPS + 4)
// word k sits at k + k / 16: the lanes of a wave read words 16 apart,
// which the spare word per 16 spreads over the LDS banks
#define DEC_SLOT(k) ((k) + ((k) >> 4))
#define DEC_BWP (DEC_SLOT(DEC_BW) + 1)

struct Bits {
  const uint32_t *words;
  int64_t nwords;
  const uint32_t *sw;   // staged words
  int64_t w0;           // index of sw[0]
  __device__ inline uint32_t word(int64_t w) const {
    const int64_t j = w - w0;
    if (j >= 0 && j < DEC_BW) return sw[DEC_SLOT(j)];
    return w < nwords ? words[w] : 0u;
  }
  __device__ inline uint32_t get32(int64_t p) const {
    const int64_t w = p >> 5;
    const int sh = (int)(p & 31);
    const uint32_t a = word(w), b = word(w + 1);
    return sh ? (a >> sh) | (b << (32 - sh)) : a;
  }
};

__device__ inline Bits stage_words(const DecArgs &A, uint32_t *sw) {
  const int64_t w0 = (int64_t)blockIdx.x * DEC_NT * DEC_WPS;
  for (int k = threadIdx.x; k < DEC_BW; k += DEC_NT) {
    const int64_t w = w0 + k;
    sw[DEC_SLOT(k)] = w < A.nwords ? A.words[w] : 0u;
  }
  __syncthreads();
  Bits b;
  b.words = A.words;
  b.nwords = A.nwords;
  b.sw = sw;
  b.w0 = w0;
  return b;
}

// Decodes the code at bit p: returns the value, advances p; a dup key also
// consumes its 3 length bits and sets *run to 2^(e+3) (else 1).
__device__ inline int decode_one(const DecArgs &A, const Bits &B, const uint32_t *lut,
                                 const int32_t *table, int64_t &p, int64_t &run) {
  const uint32_t w = B.get32(p);
  const uint32_t e = lut[w & ((1u << DEC_K) - 1)];
  const int ex = (int)((e >> 4) & 0x7ffffffu);
  int v;
  if (e >> 31) {
    v = ex;
    p += e & 15u;
  } else {
    // past DEC_K bits: the rest of the code one bit at a time, from the
    // bits already fetched while they last
    int pos = ex;
    uint32_t rest = w >> DEC_K;
    int left = 32 - DEC_K;
    p += DEC_K;
    for (;;) {
      if (left == 0) {
        rest = B.get32(p);
        left = 32;
      }
      const unsigned bit = rest & 1u;
      rest >>= 1;
      left--;
      p++;
      pos = table[pos + bit] * 2;
      if (table[pos] == 0) break;
    }
    v = table[pos + 1];
  }
  run = 1;
  if (v == A.dup_key) {
    const unsigned ebits = B.get32(p) & 7u;
    p += 3;
    run = (int64_t)1 << (ebits + 3);
  }
  return v;
}

// first bit position a subsequence's decode must reach
__device__ inline int64_t sub_limit(const DecArgs &A, int64_t i) {
  return i + 1 < A.nsub ? (i + 1) * DEC_S : A.nbits;
}

// the lookup table, and the decode table when it fits, into LDS
#define DEC_TAB_LDS 8192
__device__ inline const int32_t *load_tables(const DecArgs &A, uint32_t *lut, int32_t *tab) {
  for (int i = threadIdx.x; i < (1 << DEC_K); i += DEC_NT) lut[i] = A.lut[i];
  const bool in_lds = A.n_table <= DEC_TAB_LDS;
  if (in_lds)
    for (int i = threadIdx.x; i < A.n_table; i += DEC_NT) tab[i] = A.table[i];
  __syncthreads();
  return in_lds ? tab : A.table;
}

// decodes subsequence i from s: its end position and value count
__device__ inline void decode_sub(const DecArgs &A, const Bits &B, const uint32_t *lut,
                                  const int32_t *tab, int64_t i, int64_t s) {
  const int64_t lim = sub_limit(A, i);
  int64_t p = s;
  uint64_t n = 0;
  while (p < lim) {
    int64_t run;
    decode_one(A, B, lut, tab, p, run);
    n += (uint64_t)run;
  }
  A.end_pos[i] = p;
  A.cnt[i] = n;
}

__global__ void __launch_bounds__(DEC_NT) dec_first(DecArgs A) {
  __shared__ uint32_t lut[1 << DEC_K];
  __shared__ int32_t tab[DEC_TAB_LDS];
  __shared__ uint32_t sw[DEC_BWP];
  const int32_t *t = load_tables(A, lut, tab);
  const Bits B = stage_words(A, sw);
  const int64_t i = (int64_t)blockIdx.x * DEC_NT + threadIdx.x;
  if (i >= A.nsub) return;
  const int64_t s = i * DEC_S;
  A.start[i] = s;
  decode_sub(A, B, lut, t, i, s);
}

// a sync round, part 1: the start each subsequence must have
__global__ void dec_starts(DecArgs A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nsub) return;
  const int64_t s = i == 0 ? 0 : A.end_pos[i - 1];
  A.nstart[i] = s;
  A.changed[i] = s != A.start[i];
}

// part 2: subsequences whose start changed decode again
__global__ void __launch_bounds__(DEC_NT) dec_redo(DecArgs A) {
  __shared__ uint32_t lut[1 << DEC_K];
  __shared__ int32_t tab[DEC_TAB_LDS];
  __shared__ uint32_t sw[DEC_BWP];
  const int64_t i = (int64_t)blockIdx.x * DEC_NT + threadIdx.x;
  const bool mine = i < A.nsub && A.changed[i];
  if (!__syncthreads_or(mine)) return;   // uniform: nothing to redo here
  const int32_t *t = load_tables(A, lut, tab);
  const Bits B = stage_words(A, sw);
  if (mine) {
    const int64_t s = A.nstart[i];
    A.start[i] = s;
    decode_sub(A, B, lut, t, i, s);
  }
  if (threadIdx.x == 0) atomicAdd(A.flag, 1ull);
}

// values per block of DEC_NT subsequences
__global__ void __launch_bounds__(DEC_NT) dec_bsum(DecArgs A) {
  __shared__ uint64_t red[DEC_NT / 64];
  const int64_t i = (int64_t)blockIdx.x * DEC_NT + threadIdx.x;
  uint64_t v = i < A.nsub ? A.cnt[i] : 0;
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)v, d, 64), hi = __shfl_xor((uint32_t)(v >> 32), d, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int k = 0; k < DEC_NT / 64; k++) t += red[k];
    A.off[blockIdx.x] = t;
  }
}

// exclusive prefix over the block sums in place (one block of 1024, in
// coalesced chunks); the total goes to flag[1]
__global__ void __launch_bounds__(1024) dec_scan(DecArgs A, int64_t nblk) {
  __shared__ uint64_t s[1024];
  const int t = threadIdx.x;
  uint64_t carry = 0;
  for (int64_t c0 = 0; c0 < nblk; c0 += 1024) {
    const int64_t i = c0 + t;
    const uint64_t v = i < nblk ? A.off[i] : 0;
    s[t] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const uint64_t y = t >= d ? s[t - d] : 0;
      __syncthreads();
      s[t] += y;
      __syncthreads();
    }
    if (i < nblk) A.off[i] = carry + s[t] - v;
    carry += s[1023];
    __syncthreads();
  }
  if (t == 0) A.flag[1] = carry;   // values in the whole stream
}

__global__ void dec_fill(int32_t *out, int64_t n, int v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = v;
}

__global__ void __launch_bounds__(DEC_NT) dec_write(DecArgs A) {
  __shared__ uint32_t lut[1 << DEC_K];
  __shared__ int32_t tab[DEC_TAB_LDS];
  __shared__ uint32_t sw[DEC_BWP];
  __shared__ uint64_t sc[DEC_NT];
  const int32_t *t = load_tables(A, lut, tab);
  const Bits B = stage_words(A, sw);
  const int64_t i = (int64_t)blockIdx.x * DEC_NT + threadIdx.x;
  // the thread's output offset: the block's (scanned block sums) plus the
  // counts of the block's subsequences before it
  const uint64_t c = i < A.nsub ? A.cnt[i] : 0;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int d = 1; d < DEC_NT; d <<= 1) {
    const uint64_t y = (int)threadIdx.x >= d ? sc[threadIdx.x - d] : 0;
    __syncthreads();
    sc[threadIdx.x] += y;
    __syncthreads();
  }
  if (i >= A.nsub) return;
  int64_t o = (int64_t)(A.off[blockIdx.x] + sc[threadIdx.x] - c);
  if (o >= A.count) return;
  const int64_t lim = sub_limit(A, i);
  int64_t p = A.start[i];
  while (p < lim && o < A.count) {
    int64_t run;
    const int v = decode_one(A, B, lut, t, p, run);
    if (run > 1 || v == A.dup_key) {
      if (!A.fill) {
        const int64_t m = run < A.count - o ? run : A.count - o;
        for (int64_t k = 0; k < m; k++) A.out[o + k] = A.dup_value;
      }
    } else if (!A.fill || v != A.dup_value) {
      A.out[o] = v;
    }
    o += run;
  }
}

}  // namespace wv

using namespace wv;

namespace {

inline int64_t al256(int64_t v) { return (v + 255) & ~(int64_t)255; }

struct DecLayout {
  int64_t lut, table, start, end, nstart, cnt, off, changed, flag, total;
};

DecLayout dec_layout(int64_t nwords, int n_table) {
  const int64_t nbits = nwords * 32;
  const int64_t nsub = nbits > 0 ? (nbits + DEC_S - 1) / DEC_S : 1;
  DecLayout L;
  int64_t o = 0;
  L.lut = o; o += al256(((int64_t)1 << DEC_K) * 4);
  L.table = o; o += al256((int64_t)n_table * 4);
  L.start = o; o += al256(nsub * 8);
  L.end = o; o += al256(nsub * 8);
  L.nstart = o; o += al256(nsub * 8);
  L.cnt = o; o += al256(nsub * 8);
  L.off = o; o += al256(nsub * 8);
  L.changed = o; o += al256(nsub * 4);
  L.flag = o; o += 256;
  L.total = o;
  return L;
}

// the decode table must be a tree: every internal pair's child ids point
// further into the table and inside it (BFS numbering), so walks end; leaf
// values are bins or the dup key, at most WV_MAX_BINS (the lookup entries
// keep 27 bits of them)
bool table_ok(const int32_t *t, int n) {
  if (n < 2 || (n & 1)) return false;
  const int pairs = n / 2;
  for (int h = 0; h < pairs; h++) {
    if (t[2 * h] == 0) {   // leaf
      if (t[2 * h + 1] < 0 || t[2 * h + 1] > WV_MAX_BINS) return false;
      continue;
    }
    for (int k = 0; k < 2; k++) {
      const int c = t[2 * h + k];
      if (c <= h || c >= pairs) return false;
    }
  }
  return t[0] != 0;   // a one-leaf tree has zero-length codes: nothing to decode
}

}  // namespace

extern "C" int64_t wv_huff_decode_ws_bytes(int64_t nwords, int n_table) {
  return nwords < 0 || n_table < 0 ? WV_ERR_ARG : dec_layout(nwords, n_table).total;
}

extern "C" int wv_huff_decode_device(const int32_t *table, int n_table, int dup_key,
                                     int dup_value, const uint32_t *words_d, int64_t nwords,
                                     int32_t *out_d, int64_t count, void *ws_d, int *rounds_out,
                                     void *stream) {
  if (!table || !out_d || !ws_d || count <= 0 || nwords < 0 || (nwords > 0 && !words_d)) {
    set_error("wv_huff_decode_device: bad argument");
    return WV_ERR_ARG;
  }
  if (n_table == 2 && table[0] == 0) {
    // one leaf: its code has no bits, and the reference decoder reads no
    // value from the (empty) stream
    set_error("read only 0 of %lld values", (long long)count);
    return WV_ERR_ARG;
  }
  if (!table_ok(table, n_table)) {
    set_error("wv_huff_decode_device: malformed decode table");
    return WV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const DecLayout L = dec_layout(nwords, n_table);
  char *w = (char *)ws_d;
  // lookup table (the host decoder's rule): leaf within DEC_K bits, or the
  // node reached after DEC_K bits
  std::vector<uint32_t> lut((size_t)1 << DEC_K);
  for (uint32_t p = 0; p < (1u << DEC_K); p++) {
    int pos = 0;
    uint32_t e = 0;
    bool leaf = false;
    for (int b = 0; b < DEC_K; b++) {
      pos = table[pos + ((p >> b) & 1u)] * 2;
      if (table[pos] == 0) {
        e = (1u << 31) | ((uint32_t)table[pos + 1] << 4) | (uint32_t)(b + 1);
        leaf = true;
        break;
      }
    }
    if (!leaf) e = ((uint32_t)pos << 4) | DEC_K;
    lut[p] = e;
  }
  // the walk past DEC_K bits reads the table on the device
  const int64_t tab_bytes = (int64_t)n_table * 4;
  void *tab_d = w + L.table;
  DecArgs A{};
  A.words = words_d;
  A.nwords = nwords;
  A.nbits = nwords * 32;
  A.nsub = A.nbits > 0 ? (A.nbits + DEC_S - 1) / DEC_S : 1;
  A.lut = (const uint32_t *)(w + L.lut);
  A.table = (const int32_t *)tab_d;
  A.n_table = n_table;
  A.dup_key = dup_key;
  A.dup_value = dup_value;
  A.start = (int64_t *)(w + L.start);
  A.end_pos = (int64_t *)(w + L.end);
  A.nstart = (int64_t *)(w + L.nstart);
  A.cnt = (uint64_t *)(w + L.cnt);
  A.off = (uint64_t *)(w + L.off);
  A.changed = (int32_t *)(w + L.changed);
  A.flag = (unsigned long long *)(w + L.flag);
  A.out = out_d;
  A.count = count;
  A.fill = dup_key >= 0;
  int rc = WV_OK;
  int rounds = 0;
  unsigned long long hflag[2] = {0, 0};
  const unsigned gsub = (unsigned)((A.nsub + DEC_NT - 1) / DEC_NT);
  hipError_t e = hipMemcpyAsync(w + L.lut, lut.data(), lut.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tab_d, table, tab_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(dec_first, dim3(gsub), dim3(DEC_NT), 0, s, A);
    e = hipGetLastError();
  }
  // sync rounds until no start changes (each round fixes at least the first
  // wrong start, so nsub rounds always suffice)
  while (e == hipSuccess && rounds < A.nsub) {
    rounds++;
    hflag[0] = 0;
    e = hipMemsetAsync(A.flag, 0, 8, s);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(dec_starts, dim3(gsub), dim3(DEC_NT), 0, s, A);
    hipLaunchKernelGGL(dec_redo, dim3(gsub), dim3(DEC_NT), 0, s, A);
    e = hipMemcpyAsync(hflag, A.flag, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess || hflag[0] == 0) break;
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(dec_bsum, dim3(gsub), dim3(DEC_NT), 0, s, A);
    hipLaunchKernelGGL(dec_scan, dim3(1), dim3(1024), 0, s, A, (int64_t)gsub);
    e = hipMemcpyAsync(hflag, A.flag, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e == hipSuccess && (int64_t)hflag[1] < count) {
    set_error("read only %llu of %lld values", (unsigned long long)hflag[1], (long long)count);
    rc = WV_ERR_ARG;
  }
  if (e == hipSuccess && rc == WV_OK) {
    if (A.fill) {
      int64_t g = (count + 255) / 256;
      if (g > 8192) g = 8192;
      hipLaunchKernelGGL(dec_fill, dim3((unsigned)g), dim3(256), 0, s, out_d, count, dup_value);
    }
    hipLaunchKernelGGL(dec_write, dim3(gsub), dim3(DEC_NT), 0, s, A);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);   // lut / table are host locals
  if (e != hipSuccess) {
    set_error("wv_huff_decode_device: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  if (rounds_out) *rounds_out = rounds;
  return rc;
}
