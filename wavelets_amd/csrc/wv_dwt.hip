// wv_dwt.hip -- fused axis passes of the 3-D Haar / CDF 9/7 transform for
// gfx950.
//
// One kernel, instantiated per (wavelet, direction, source dtype, clamping,
// tile width), processes *line tiles*: T lines of one axis, each of length
// L (the padded axis length), held in LDS as tile[row][line] with row
// stride T+1 (line-contiguous and position-contiguous accesses are both
// bank-conflict free) and 4 spare rows before and after the line.
//
// The grid is persistent: block b handles tiles b, b+G, b+2G, ...  For
// each tile the block
//   1. moves the elements it prefetched into registers to LDS (u8/i16/i32
//      converted to float; REPEAT padding by clamping coordinates into the
//      unpadded source), then issues the loads of the NEXT tile into the
//      same registers -- they stay in flight through steps 2 and 3, since
//      barriers only wait for LDS traffic,
//   2. runs every level of the 1-D transform in LDS: each thread owns up to
//      EPT/2 consecutive (low, high) output pairs of one line and walks a
//      9-value register window over them; all threads write back after a
//      barrier,
//   3. stores the tile back in x,y,z layout or, for the z pass, in the
//      reference's rotated z,x,y output layout.
// Global accesses are either whole lines (consecutive lanes on consecutive
// positions) or T-float segments across lines, so a pass moves 8 B per
// element (the input's own size plus 4 B on the first pass).  Elements past
// the line end and lines past the volume edge are clamped onto the last
// valid ones, whose values they then duplicate: every load and store is
// unconditional and the memory phases carry no branches.
//
// Forward CDF 9/7 levels use a gap layout: the high band of a level is
// written 4 rows after the end of its low band, so rows [n, n+3) past a
// level's input are free, and the threads at the two line ends write the
// mirrored values there (and at rows -4..-1) before reading them.  Every
// chunk then reads plain rows -- no per-element mirroring -- and the final
// coefficients sit at row p (p < final low length) or p + 4.
//
// Small planes (at most 32 x 32 samples) take a plane pass instead of the x
// and y passes: dwt_plane_kernel below, x and y levels of whole xy-planes in
// LDS.
//
// Arithmetic follows the reference CPU code operation by operation with
// contraction disabled (-ffp-contract=off), so coefficients are
// bit-identical to cdf97 / haar (dwt_cpu.cc:139-203,683-799).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

// dwt_cpu.cc:15-43 float initialisers
constexpr float F0 = .85269867900940f, F1 = .377402855612650f,
                F2 = -.110624404418420f, F3 = -.02384946501938f,
                F4 = .037828455506995f;
constexpr float G0 = -.788485616405660f, G1 = .418092273222210f,
                G2 = .040689417609558f, G3 = -.064538882628938f;
constexpr float S0 = .788485616405660f, S1 = .377402855612650f,
                S2 = -.040689417609558f, S3 = -.02384946501938f;
constexpr float H0 = -.85269867900940f, H1 = .418092273222210f,
                H2 = .110624404418420f, H3 = -.064538882628938f,
                H4 = -.037828455506995f;
constexpr float INV_SQRT2 = 0.70710678118654752440f;   // dwt_cpu.cc:11

#define WV_EPT (2 * WV_MAXP)   // most tile elements per thread (EPT): L*T <= NT*EPT
#define WV_HALO 4              // spare rows before and after a line

template <typename S> __device__ inline float to_f(S v) { return (float)v; }

// An int behind an empty asm statement.  Values derived from it are
// recomputed where they are used instead of being hoisted out of the tile
// loop and held in registers (or spilled) for its whole length.
__device__ inline int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ inline int opaque_tid() { return opaque((int)threadIdx.x); }

// ---- tiles ----------------------------------------------------------------
// A pass's lines form a grid of Nu neighbouring lines (u) by Nv rows (v); a
// tile is T consecutive u of one row.
struct Tile {
  int nvalid;      // valid lines in this tile (>= 1)
  int a, b;        // first u, row v
  const void *src; // this tile's cubelet
  float *dst;
};

// Global tile index -> cubelet and tile.  Divisors pass through opaque()
// so no reciprocal is hoisted out of the tile loop.
__device__ inline Tile tile_at(const DwtPass &P, int T, int tile) {
  Tile t;
  if (P.src_tab) {
    const int n1 = opaque(P.ntiles1);
    const int c = tile / n1;
    tile -= c * n1;
    t.src = P.src_tab[c];
    t.dst = P.dst_tab[c];
  } else {
    t.src = P.src;
    t.dst = P.dst;
  }
  const int ntu = opaque(P.ntu);
  const int v = tile / ntu;
  t.a = (tile - v * ntu) * T;
  t.b = v;
  t.nvalid = min(T, P.Nu - t.a);
  return t;
}

__device__ inline bool line_major_load(const DwtPass &P) { return P.ld_line != 0; }
__device__ inline bool line_major_store(const DwtPass &P) { return P.st_line != 0; }

// global offset of (line l, position 0); lines past the edge map onto the
// last valid line, coordinates past the source onto its last sample
__device__ inline size_t line_base(const LineMap &m, const Tile &t, int l) {
  const int u = min(t.a + min(l, t.nvalid - 1), m.ucap);
  const int v = min(t.b, m.vcap);
  return (size_t)((int64_t)u * m.su + (int64_t)v * m.sv);
}

// Position-major access (neighbouring lines adjacent in memory, su == 1):
// element (line l, position p) at rowbase + min(p, pcap) * pstride +
// min(a + l, xcap).
struct PosMap {
  size_t rowbase, pstride;
  int pcap, xcap;
};

__device__ inline PosMap pos_map(const LineMap &m, const Tile &t) {
  PosMap r;
  r.rowbase = (size_t)((int64_t)min(t.b, m.vcap) * m.sv);
  r.pstride = (size_t)m.sp;
  r.pcap = m.pcap;
  r.xcap = min(t.a + t.nvalid - 1, m.ucap);
  return r;
}

// ---- element <-> thread mappings ------------------------------------------
// General mappings (any tile):
//   line-major, compile-time TT >= 8 with L % 64 == 0: wave w owns lines
//     w*TT/8 .. w*TT/8 + TT/8 - 1; element u = li*JN + j of a lane is line
//     w*TT/8 + li at position j*64 + lane (JN = EPT / (TT/8));
//   line-major, generic: element u is flat index tid + u*NT of the
//     line-major tile (division by L through magic_L);
//   position-major: element u is line tid % T at position
//     tid / T + u * (NT / T).
// Exact tiles (compile-time TT, L*T == NT*EPT, no padding, all T
// lines valid) use the same element order with no clamping, and address
// global memory as a wave-uniform base plus a 32-bit lane offset.
// compile-time line-major mapping: whole 64-position segments per wave
// instruction, at most EPT lines per wave
template <int TT, int EPT, int NT> constexpr bool lm_fast() {
  return TT >= NT / 64 && TT / (NT / 64) <= EPT;
}

template <int TT, int EPT, int NT>
__device__ inline void lm_index(const DwtPass &P, int tid, int u, int &l, int &p) {
  if (lm_fast<TT, EPT, NT>()) {
    constexpr int LPW = lm_fast<TT, EPT, NT>() ? TT / (NT / 64) : 1;
    constexpr int JN = EPT / LPW;
    l = (tid >> 6) * LPW + u / JN;
    p = (u % JN) * 64 + (tid & 63);
  } else {
    const int i = min(tid + u * NT, P.T * P.L - 1);
    l = (int)__umulhi((unsigned)i, P.magic_L);
    p = i - l * P.L;
  }
}

template <int TT, int EPT, int NT>
__device__ inline void pm_index(const DwtPass &P, int tid, int u, int &l, int &p) {
  if (TT > 0) {
    l = tid % TT;
    p = tid / TT + u * (NT / TT);
  } else {
    l = tid & (P.T - 1);
    p = (tid >> P.log2T) + (u << (P.log2dblock_T));
  }
}

template <int TT, int EPT, int NT>
__device__ inline bool exact_tile(const DwtPass &P, const Tile &t) {
  return lm_fast<TT, EPT, NT>() && P.exact && t.nvalid == TT;
}

// Global accesses of full tiles are non-temporal (WV_NT_LD / WV_NT_ST in
// wv_device.h): a pass reads and writes every sample once, so neither side
// is reused from the caches (measured on the 512^3 float cubelet: 58.6% ->
// 60.0% of 8 TB/s over the three forward passes).
//
// The cubelet pointers reach the kernel through the pass struct and its
// tables, so the compiler cannot tell they address global memory and would
// emit flat instructions.  Flat accesses also count on LGKM_CNT, so the
// lgkmcnt(0) wait in front of every LDS barrier of the levels would wait
// for the next tile's loads as well and undo the software pipelining.
// Every access therefore goes through a global (address space 1) pointer
// (gptr, wv_device.h).

template <typename T>
__device__ inline T ld_tile(const T *p) {
  if (WV_NT_LD) return __builtin_nontemporal_load(gptr(p));
  return *gptr(p);
}
template <typename T>
__device__ inline T ld_g(const T *p) { return *gptr(p); }
__device__ inline void st_tile(float *p, float v) {
  if (WV_NT_ST) __builtin_nontemporal_store(v, gptr(p));
  else *gptr(p) = v;
}
__device__ inline void st_g(float *p, float v) { *gptr(p) = v; }
typedef float wv_vec4 __attribute__((ext_vector_type(4)));
__device__ inline void st_tile4(wv_vec4 *p, wv_vec4 v) {
  if (WV_NT_ST) __builtin_nontemporal_store(v, gptr(p));
  else *gptr(p) = v;
}

// 64-bit value that is the same in every lane, moved to scalar registers
__device__ inline size_t uniform64(size_t v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((size_t)hi << 32) | lo;
}

// Issues the loads of this thread's elements of tile t into v[] without
// waiting for them.  Every path issues exactly EPT loads.
template <typename SRC, bool CLAMP, int TT, int EPT, int NT>
__device__ inline void load_regs(const DwtPass &P, const Tile &t, const size_t *lb,
                                 float (&v)[EPT]) {
  const int tid = opaque_tid();
  const SRC *src = (const SRC *)t.src;
  if (exact_tile<TT, EPT, NT>(P, t)) {
    if (line_major_load(P)) {
      constexpr int LPW = lm_fast<TT, EPT, NT>() ? TT / (NT / 64) : 1;
      constexpr int JN = EPT / LPW;
      const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
      const int lane = tid & 63;
#pragma unroll
      for (int li = 0; li < LPW; li++) {
        const SRC *base = src + uniform64(lb[w * LPW + li]);
#pragma unroll
        for (int j = 0; j < JN; j++) v[li * JN + j] = to_f(ld_tile(base + j * 64 + lane));
      }
    } else {
      constexpr int PSTEP = NT / (TT > 0 ? TT : 1);
      const PosMap m = pos_map(P.ld, t);
      const char *base = (const char *)(src + m.rowbase + t.a);
      constexpr int TW = TT > 0 ? TT : 1;
      const unsigned voff = (unsigned)(((tid / TW) * (unsigned)m.pstride + tid % TW) * sizeof(SRC));
      const size_t ustep = (size_t)PSTEP * m.pstride * sizeof(SRC);
#pragma unroll
      for (int u = 0; u < EPT; u++)
        v[u] = to_f(ld_tile((const SRC *)(base + u * ustep + voff)));
    }
    return;
  }
  if (TT == 0 && P.rt_exact && t.nvalid == P.T && !line_major_load(P)) {
    // the runtime-T form of the exact position-major loads: no clamping,
    // a uniform base plus a 32-bit lane offset
    const PosMap m = pos_map(P.ld, t);
    const char *base = (const char *)(src + m.rowbase + t.a);
    const int lg = P.log2T;
    const unsigned voff =
        (((unsigned)tid >> lg) * (unsigned)m.pstride + ((unsigned)tid & (unsigned)(P.T - 1))) *
        (unsigned)sizeof(SRC);
    const size_t ustep = ((size_t)NT >> lg) * m.pstride * sizeof(SRC);
#pragma unroll
    for (int u = 0; u < EPT; u++)
      v[u] = to_f(ld_tile((const SRC *)(base + u * ustep + voff)));
    return;
  }
  if (line_major_load(P)) {
    const int pmax = P.ld.pcap;
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      int l, p;
      lm_index<TT, EPT, NT>(P, tid, u, l, p);
      v[u] = to_f(ld_g(src + lb[l] + min(p, pmax)));
    }
  } else {
    const PosMap m = pos_map(P.ld, t);
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      int l, p;
      pm_index<TT, EPT, NT>(P, tid, u, l, p);
      v[u] = to_f(ld_g(src + m.rowbase + (size_t)min(p, m.pcap) * m.pstride +
                       min(t.a + l, m.xcap)));
    }
  }
}

template <int TT, int EPT, int NT>
__device__ inline void regs_to_lds(const DwtPass &P, const Tile &t, float *tile,
                                   const float (&v)[EPT]) {
  const int LS = (TT > 0 ? TT : P.T) + 1;
  const int tid = opaque_tid();
  if (exact_tile<TT, EPT, NT>(P, t)) {
    if (line_major_load(P)) {
      constexpr int LPW = lm_fast<TT, EPT, NT>() ? TT / (NT / 64) : 1;
      constexpr int JN = EPT / LPW;
      float *b = tile + (tid & 63) * LS + (tid >> 6) * LPW;
#pragma unroll
      for (int li = 0; li < LPW; li++)
#pragma unroll
        for (int j = 0; j < JN; j++) b[j * 64 * LS + li] = v[li * JN + j];
    } else {
      constexpr int PSTEP = NT / (TT > 0 ? TT : 1);
      constexpr int TW = TT > 0 ? TT : 1;
      float *b = tile + (tid / TW) * LS + tid % TW;
#pragma unroll
      for (int u = 0; u < EPT; u++) b[u * PSTEP * LS] = v[u];
    }
    return;
  }
  const int plast = P.L - 1;
  if (line_major_load(P)) {
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      int l, p;
      lm_index<TT, EPT, NT>(P, tid, u, l, p);
      tile[min(p, plast) * LS + l] = v[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      int l, p;
      pm_index<TT, EPT, NT>(P, tid, u, l, p);
      tile[min(p, plast) * LS + l] = v[u];
    }
  }
}

// Stores tile t from LDS; tout is the row of output position 0 (forward
// CDF 9/7 leaves its coefficients WV_HALO rows down).
template <int TT, int EPT, int NT>
__device__ inline void store_tile(const DwtPass &P, const Tile &t, const float *tout,
                                  const size_t *lb) {
  const int LS = (TT > 0 ? TT : P.T) + 1;
  const int tid = opaque_tid();
  if (exact_tile<TT, EPT, NT>(P, t)) {
    if (line_major_store(P)) {
      constexpr int LPW = lm_fast<TT, EPT, NT>() ? TT / (NT / 64) : 1;
      constexpr int JN = EPT / LPW;
      const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
      const int lane = tid & 63;
      const float *b = tout + lane * LS + w * LPW;
#pragma unroll
      for (int li = 0; li < LPW; li++) {
        float *base = t.dst + uniform64(lb[w * LPW + li]);
#pragma unroll
        for (int j = 0; j < JN; j++) st_tile(base + j * 64 + lane, b[j * 64 * LS + li]);
      }
    } else if (TT >= 32 && TT % 32 == 0 && P.vec_st) {
      // 16-byte vectors of 4 neighbouring lines: a half-wave covers 4 rows
      // of 32 lines (8 lanes per 128-byte row piece), so its LDS reads
      // (row stride TT + 1) fall on 32 distinct banks
      constexpr int CB = TT > 0 ? TT / 32 : 1;   // 32-line blocks per row
      const PosMap m = pos_map(P.st, t);
      float *base = t.dst + m.rowbase + t.a;
      const size_t ps = (size_t)m.pstride;
      const int cq = tid & 7, rr = (tid >> 3) & 3;
#pragma unroll
      for (int j = 0; j < EPT / 4; j++) {
        const int rest = j * (NT / 32) + (tid >> 5);
        const int p = (rest / CB) * 4 + rr, l = (rest % CB) * 32 + cq * 4;
        const float *r = tout + p * LS + l;
        st_tile4((wv_vec4 *)(base + (size_t)p * ps + l), wv_vec4{r[0], r[1], r[2], r[3]});
      }
    } else {
      constexpr int PSTEP = NT / (TT > 0 ? TT : 1);
      const PosMap m = pos_map(P.st, t);
      char *base = (char *)(t.dst + m.rowbase + t.a);
      constexpr int TW = TT > 0 ? TT : 1;
      const unsigned voff = (unsigned)(((tid / TW) * (unsigned)m.pstride + tid % TW) * 4u);
      const size_t ustep = (size_t)PSTEP * m.pstride * 4;
      const float *b = tout + (tid / TW) * LS + tid % TW;
#pragma unroll
      for (int u = 0; u < EPT; u++)
        st_tile((float *)(base + u * ustep + voff), b[u * PSTEP * LS]);
    }
    return;
  }
  if (TT == 0 && P.st_block && t.nvalid == P.T) {
    // the tile's lines are adjacent in the output: one contiguous block of
    // T*L floats (line l, position p at l*L + p), stored as 16-byte vectors
    // of 4 positions of one line (L % 4 == 0)
    float *base = t.dst + lb[0];
#pragma unroll
    for (int j = 0; j < EPT / 4; j++) {
      const int e = (tid + j * NT) * 4;
      const int l = (int)__umulhi((unsigned)e, P.magic_L);
      const int p = e - l * P.L;
      const float *r = tout + p * LS + l;
      st_tile4((wv_vec4 *)(base + e), wv_vec4{r[0], r[LS], r[2 * LS], r[3 * LS]});
    }
    return;
  }
  const int plast = P.L - 1;
  if (line_major_store(P)) {
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      int l, p;
      lm_index<TT, EPT, NT>(P, tid, u, l, p);
      p = min(p, plast);
      st_g(t.dst + lb[l] + p, tout[p * LS + l]);
    }
  } else {
    const PosMap m = pos_map(P.st, t);
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      int l, p;
      pm_index<TT, EPT, NT>(P, tid, u, l, p);
      p = min(p, plast);
      st_g(t.dst + m.rowbase + (size_t)p * m.pstride + min(t.a + l, m.xcap), tout[p * LS + l]);
    }
  }
}

// ---- forward levels ---------------------------------------------------------
// CDF 9/7 with the gap layout.  Level input: rows [0, n); the window of the
// pair j spans rows 2j-4 .. 2j+4, i.e. rows -4 .. n+2 over the line.
template <int TT, int EPT, int NT>
__device__ void forward97_levels(float *tile, const DwtPass &P) {
  const int T = TT > 0 ? TT : P.T;
  const int LS = T + 1;
  const int tid = opaque_tid();
  const int line = tid % T;
  const int k = tid / T;
  const int tpl = NT / T;
  float *col = tile + line;
  int n = P.L;
  for (int s = 0; s < P.levels; s++) {
    const int pairs = n >> 1;
    const int cpt = (pairs + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, pairs - j0);
    float lo[(EPT / 2)], hi[(EPT / 2)];
    if (cnt > 0) {
      const int c = 2 * j0;
      // mirrored halo rows this chunk reads (MirroredArray, dwt_cpu.h:162-176);
      // threads of neighbouring chunks may write the same values
      if (c < 4) {
#pragma unroll
        for (int h = 1; h <= 4; h++) col[-h * LS] = col[wv_mirror(-h, n) * LS];
      }
      if (c + 2 * cnt + 2 >= n) {
#pragma unroll
        for (int h = 0; h < 3; h++) col[(n + h) * LS] = col[wv_mirror(n + h, n) * LS];
      }
      const float *cc = col + c * LS;
      float w0 = cc[-4 * LS], w1 = cc[-3 * LS], w2 = cc[-2 * LS], w3 = cc[-LS],
            w4 = cc[0], w5 = cc[LS], w6 = cc[2 * LS], w7 = cc[3 * LS], w8 = cc[4 * LS];
#pragma unroll
      for (int q = 0; q < (EPT / 2); q++) {
        if (q < cnt) {
          if (q > 0) {
            w0 = w2; w1 = w3; w2 = w4; w3 = w5; w4 = w6; w5 = w7; w6 = w8;
            w7 = cc[(2 * q + 3) * LS];
            w8 = cc[(2 * q + 4) * LS];
          }
          // cdf97_tmpl dwt_cpu.cc:709-723: centre tap first, then
          // F[k]*(a[i+k]+a[i-k]) for k = 1..4, no contraction
          float sl = F0 * w4;
          sl += F1 * (w5 + w3);
          sl += F2 * (w6 + w2);
          sl += F3 * (w7 + w1);
          sl += F4 * (w8 + w0);
          float sh = G0 * w5;
          sh += G1 * (w6 + w4);
          sh += G2 * (w7 + w3);
          sh += G3 * (w8 + w2);
          lo[q] = sl;
          hi[q] = sh;
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
      // the last level also moves its low band down WV_HALO rows, so every
      // final coefficient p sits at row p + WV_HALO
      const int lshift = s == P.levels - 1 ? WV_HALO : 0;
      float *lw = col + (j0 + lshift) * LS, *hw = col + (pairs + WV_HALO + j0) * LS;
#pragma unroll
      for (int q = 0; q < (EPT / 2); q++) {
        if (q < cnt) {
          lw[q * LS] = lo[q];
          hw[q * LS] = hi[q];
        }
      }
    }
    __syncthreads();
    n = pairs;
  }
}

template <int TT, int EPT, int NT>
__device__ void forward_haar_levels(float *tile, const DwtPass &P) {
  const int T = TT > 0 ? TT : P.T;
  const int LS = T + 1;
  const int tid = opaque_tid();
  const int line = tid % T;
  const int k = tid / T;
  const int tpl = NT / T;
  float *col = tile + line;
  int n = P.L;
  for (int s = 0; s < P.levels; s++) {
    const int pairs = n >> 1;
    const int cpt = (pairs + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, pairs - j0);
    float lo[(EPT / 2)], hi[(EPT / 2)];
    if (cnt > 0) {
      const float *cc = col + 2 * j0 * LS;
#pragma unroll
      for (int q = 0; q < (EPT / 2); q++) {
        if (q < cnt) {
          const float a = cc[(2 * q) * LS], b = cc[(2 * q + 1) * LS];
          // haar_internal dwt_cpu.cc:159-162
          hi[q] = (a - b) * INV_SQRT2;
          lo[q] = (a + b) * INV_SQRT2;
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
      float *lw = col + j0 * LS, *hw = col + (pairs + j0) * LS;
#pragma unroll
      for (int q = 0; q < (EPT / 2); q++) {
        if (q < cnt) {
          lw[q * LS] = lo[q];
          hw[q * LS] = hi[q];
        }
      }
    }
    __syncthreads();
    n = pairs;
  }
}

// ---- inverse levels -------------------------------------------------------
// interleaved view of the current level: t(i) = low[i/2] for even i,
// high[i/2] for odd i, mirrored on the interleaved length n
__device__ inline float ldi(const float *col, int i, int n, int half, int LS) {
  if ((unsigned)i >= (unsigned)n) i = wv_mirror(i, n);
  return (i & 1) ? col[(half + (i >> 1)) * LS] : col[(i >> 1) * LS];
}

template <int WAVE, int TT, int EPT, int NT>
__device__ void inverse_levels(float *tile, const DwtPass &P) {
  const int T = TT > 0 ? TT : P.T;
  const int LS = T + 1;
  const int tid = opaque_tid();
  const int line = tid % T;
  const int k = tid / T;
  const int tpl = NT / T;
  float *col = tile + line;
  int n = P.L >> (P.levels - 1);
  for (int s = 0; s < P.levels; s++) {
    const int half = n >> 1;
    const int cpt = (half + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, half - j0);
    float ev[(EPT / 2)], od[(EPT / 2)];
    if (cnt > 0) {
      if (WAVE == WV_WAVELET_CDF97) {
        const int c = 2 * j0;
        float u0 = ldi(col, c - 3, n, half, LS), u1 = ldi(col, c - 2, n, half, LS),
              u2 = ldi(col, c - 1, n, half, LS), u3 = ldi(col, c, n, half, LS),
              u4 = ldi(col, c + 1, n, half, LS), u5 = ldi(col, c + 2, n, half, LS),
              u6 = ldi(col, c + 3, n, half, LS), u7 = ldi(col, c + 4, n, half, LS),
              u8 = ldi(col, c + 5, n, half, LS);
#pragma unroll
        for (int q = 0; q < (EPT / 2); q++) {
          if (q < cnt) {
            if (q > 0) {
              const int nc = c + 2 * q;
              u0 = u2; u1 = u3; u2 = u4; u3 = u5; u4 = u6; u5 = u7; u6 = u8;
              u7 = ldi(col, nc + 4, n, half, LS);
              u8 = ldi(col, nc + 5, n, half, LS);
            }
            // cdf97_inverse_tmpl dwt_cpu.cc:784-795
            float e = u3 * S0;
            e += (u4 + u2) * S1;
            e += (u5 + u1) * S2;
            e += (u6 + u0) * S3;
            float o = u4 * H0;
            o += (u5 + u3) * H1;
            o += (u6 + u2) * H2;
            o += (u7 + u1) * H3;
            o += (u8 + u0) * H4;
            ev[q] = e;
            od[q] = o;
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < (EPT / 2); q++) {
          if (q < cnt) {
            const int j = j0 + q;
            const float sv = col[j * LS], dv = col[(half + j) * LS];
            // haar_inv_internal dwt_cpu.cc:193-196
            ev[q] = INV_SQRT2 * (sv + dv);
            od[q] = INV_SQRT2 * (sv - dv);
          }
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
      float *w = col + 2 * j0 * LS;
#pragma unroll
      for (int q = 0; q < (EPT / 2); q++) {
        if (q < cnt) {
          w[(2 * q) * LS] = ev[q];
          w[(2 * q + 1) * LS] = od[q];
        }
      }
    }
    __syncthreads();
    n <<= 1;
  }
}

// ---- the pass kernel ------------------------------------------------------
template <int WAVE, bool INV, int TT, int EPT, int NT>
__device__ inline void run_levels(float *tile, const DwtPass &P) {
  if (P.levels <= 0) return;
  if (INV) inverse_levels<WAVE, TT, EPT, NT>(tile, P);
  else if (WAVE == WV_WAVELET_CDF97) forward97_levels<TT, EPT, NT>(tile, P);
  else forward_haar_levels<TT, EPT, NT>(tile, P);
}

template <bool CLAMP, int NT>
__device__ inline void set_bases(const DwtPass &P, const Tile &t, size_t *lb_ld,
                                 const Tile *ts, size_t *lb_st) {
  const int T = P.T;
  for (int l = threadIdx.x; l < T; l += NT) {
    if (lb_ld) lb_ld[l] = line_base(P.ld, t, l);
    if (lb_st) lb_st[l] = line_base(P.st, *ts, l);
  }
}

// One steady-state step: tile t goes from registers to LDS, the loads of
// tile tn are issued, t is transformed and stored.
template <int WAVE, bool INV, typename SRC, bool CLAMP, int TT, int EPT, int NT>
__device__ __attribute__((always_inline)) inline void
pipe_step(const DwtPass &P, float *tile, const float *tout, const Tile &t, const Tile &tn,
          size_t *lb_next, size_t *lb_st, float (&v)[EPT]) {
  regs_to_lds<TT, EPT, NT>(P, t, tile, v);
  set_bases<CLAMP, NT>(P, tn, line_major_load(P) ? lb_next : nullptr, &t,
                   line_major_store(P) ? lb_st : nullptr);
  __syncthreads();
  load_regs<SRC, CLAMP, TT, EPT, NT>(P, tn, lb_next, v);
  run_levels<WAVE, INV, TT, EPT, NT>(tile, P);
  store_tile<TT, EPT, NT>(P, t, tout, lb_st);
  __syncthreads();
}

// The first steady-state step is peeled off the loop so the loop is entered
// with the same memory operations in flight as at its back edge (the
// previous tile's stores behind the next tile's loads); waits on the loads
// then never cover those stores.
template <int WAVE, bool INV, typename SRC, bool CLAMP, int TT, int EPT, int NT>
__global__ void __launch_bounds__(NT, EPT == 8 ? WV_DWAVES8
                                              : (EPT == 16 && TT > 0 ? WV_DWAVES16 : WV_DWAVES))
dwt_pass_kernel(DwtPass P, int ntiles) {
  constexpr int GAP = (WAVE == WV_WAVELET_CDF97 && !INV) ? WV_HALO : 0;
  extern __shared__ float smem[];
  __shared__ size_t lb_ld[2][WV_MAXT];   // line bases of the tile being loaded
  __shared__ size_t lb_st[WV_MAXT];      // line bases of the tile being stored
  const int T = TT > 0 ? TT : P.T;
  float *tile = smem + WV_HALO * (T + 1);
  const float *tout = tile + (P.levels > 0 ? GAP : 0) * (T + 1);
  // block b takes tiles b, b + G, ...  (an XCD-grouped order -- each of
  // the 8 XCDs owning an eighth of the tiles -- measured slower)
  int g = blockIdx.x;
  const int G = gridDim.x;
  const int gend = ntiles;
  if (g >= gend) return;
  Tile t = tile_at(P, T, g);
  set_bases<CLAMP, NT>(P, t, line_major_load(P) ? lb_ld[0] : nullptr, &t, nullptr);
  __syncthreads();
  float v[EPT];
  load_regs<SRC, CLAMP, TT, EPT, NT>(P, t, lb_ld[0], v);
  int buf = 0;
  g += G;
  if (g < gend) {
    Tile tn = tile_at(P, T, g);
    pipe_step<WAVE, INV, SRC, CLAMP, TT, EPT, NT>(P, tile, tout, t, tn, lb_ld[1], lb_st, v);
    t = tn;
    buf = 1;
    g += G;
    while (g < gend) {
      tn = tile_at(P, T, g);
      pipe_step<WAVE, INV, SRC, CLAMP, TT, EPT, NT>(P, tile, tout, t, tn, lb_ld[buf ^ 1], lb_st, v);
      t = tn;
      buf ^= 1;
      g += G;
    }
  }
  // last tile: nothing left to prefetch
  regs_to_lds<TT, EPT, NT>(P, t, tile, v);
  set_bases<CLAMP, NT>(P, t, nullptr, &t, line_major_store(P) ? lb_st : nullptr);
  __syncthreads();
  run_levels<WAVE, INV, TT, EPT, NT>(tile, P);
  store_tile<TT, EPT, NT>(P, t, tout, lb_st);
}

// ---- the plane pass ---------------------------------------------------------
// x and y levels of PP consecutive xy-planes in one pass.  The kernel takes
// planes of up to 128 x 128; forward_specs uses it for planes of at most
// WV_DWT_PLANE_MAX samples (32 x 32 by default: on larger planes the three
// axis passes are faster, since the plane tile's two sets of levels make
// the pass latency-bound).  The planes are loaded into LDS (converted from
// the source type, coordinates clamped into the unpadded source), every row
// is transformed along x, then every column along y, and the planes are
// stored back in x,y,z (the z pass then writes z,x,y): two HBM round trips
// for the three axes instead of three.  Plane p's row y sits at LDS row
// p*(Y+8) + 4 + y, its column x at 4 + x, with four spare rows above and
// below each plane and four spare columns left and right of each row; the
// row stride X + 9 is odd, so a wave's accesses along a row (y levels:
// neighbouring lines are neighbouring columns) and across rows (x levels:
// neighbouring lines are neighbouring rows) are both bank-conflict free.
// The levels are forward97_levels / forward_haar_levels with a line base
// and a position stride in place of the tile's column and row stride: the
// same arithmetic, the same gap layout of the 9/7 bands.
#define WV_PL_EPT 32   // tile elements per thread: PP*X*Y <= 512*32

template <typename Base>
__device__ void plane97_levels(float *lds, int nl, Base base, int ps, int n0, int levels) {
  const int tid = opaque_tid();
  const int line = tid % nl;
  const int k = tid / nl;
  const int tpl = WV_DBLOCK / nl;
  float *col = lds + base(line);
  int n = n0;
  for (int s = 0; s < levels; s++) {
    const int pairs = n >> 1;
    const int cpt = (pairs + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, pairs - j0);
    float lo[WV_PL_EPT / 2], hi[WV_PL_EPT / 2];
    if (cnt > 0) {
      const int c = 2 * j0;
      // mirrored halo positions this chunk reads (MirroredArray,
      // dwt_cpu.h:162-176)
      if (c < 4) {
#pragma unroll
        for (int h = 1; h <= 4; h++) col[-h * ps] = col[wv_mirror(-h, n) * ps];
      }
      if (c + 2 * cnt + 2 >= n) {
#pragma unroll
        for (int h = 0; h < 3; h++) col[(n + h) * ps] = col[wv_mirror(n + h, n) * ps];
      }
      const float *cc = col + c * ps;
      float w0 = cc[-4 * ps], w1 = cc[-3 * ps], w2 = cc[-2 * ps], w3 = cc[-ps],
            w4 = cc[0], w5 = cc[ps], w6 = cc[2 * ps], w7 = cc[3 * ps], w8 = cc[4 * ps];
#pragma unroll
      for (int q = 0; q < (WV_PL_EPT / 2); q++) {
        if (q < cnt) {
          if (q > 0) {
            w0 = w2; w1 = w3; w2 = w4; w3 = w5; w4 = w6; w5 = w7; w6 = w8;
            w7 = cc[(2 * q + 3) * ps];
            w8 = cc[(2 * q + 4) * ps];
          }
          // cdf97_tmpl dwt_cpu.cc:709-723
          float sl = F0 * w4;
          sl += F1 * (w5 + w3);
          sl += F2 * (w6 + w2);
          sl += F3 * (w7 + w1);
          sl += F4 * (w8 + w0);
          float sh = G0 * w5;
          sh += G1 * (w6 + w4);
          sh += G2 * (w7 + w3);
          sh += G3 * (w8 + w2);
          lo[q] = sl;
          hi[q] = sh;
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
      const int lshift = s == levels - 1 ? WV_HALO : 0;
      float *lw = col + (j0 + lshift) * ps, *hw = col + (pairs + WV_HALO + j0) * ps;
#pragma unroll
      for (int q = 0; q < (WV_PL_EPT / 2); q++) {
        if (q < cnt) {
          lw[q * ps] = lo[q];
          hw[q * ps] = hi[q];
        }
      }
    }
    __syncthreads();
    n = pairs;
  }
}

template <typename Base>
__device__ void plane_haar_levels(float *lds, int nl, Base base, int ps, int n0, int levels) {
  const int tid = opaque_tid();
  const int line = tid % nl;
  const int k = tid / nl;
  const int tpl = WV_DBLOCK / nl;
  float *col = lds + base(line);
  int n = n0;
  for (int s = 0; s < levels; s++) {
    const int pairs = n >> 1;
    const int cpt = (pairs + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, pairs - j0);
    float lo[WV_PL_EPT / 2], hi[WV_PL_EPT / 2];
    if (cnt > 0) {
      const float *cc = col + 2 * j0 * ps;
#pragma unroll
      for (int q = 0; q < (WV_PL_EPT / 2); q++) {
        if (q < cnt) {
          const float a = cc[(2 * q) * ps], b = cc[(2 * q + 1) * ps];
          // haar_internal dwt_cpu.cc:159-162
          hi[q] = (a - b) * INV_SQRT2;
          lo[q] = (a + b) * INV_SQRT2;
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
      float *lw = col + j0 * ps, *hw = col + (pairs + j0) * ps;
#pragma unroll
      for (int q = 0; q < (WV_PL_EPT / 2); q++) {
        if (q < cnt) {
          lw[q * ps] = lo[q];
          hw[q * ps] = hi[q];
        }
      }
    }
    __syncthreads();
    n = pairs;
  }
}

// tile -> (cubelet source, destination, first plane)
struct PlTile {
  const void *src;
  float *dst;
  int z0, np;   // first plane, planes in this tile
};

__device__ inline PlTile plane_tile(const DwtPass &P, int tile) {
  PlTile t;
  if (P.src_tab) {
    const int n1 = opaque(P.ntiles1);
    const int c = tile / n1;
    tile -= c * n1;
    t.src = P.src_tab[c];
    t.dst = P.dst_tab[c];
  } else {
    t.src = P.src;
    t.dst = P.dst;
  }
  t.z0 = tile * P.T;
  t.np = min(P.T, P.Z - t.z0);
  return t;
}

// element i of thread tid: flat index tid + i*NT of the tile's planes
// (x fastest); division by X and X*Y through exact reciprocals
__device__ inline void plane_xyz(const DwtPass &P, int idx, int &p, int &y, int &x) {
  p = (int)__umulhi((unsigned)idx, P.magic_XY);
  const int r = idx - p * opaque(P.X * P.Y);
  y = (int)__umulhi((unsigned)r, P.magic_L);
  x = r - y * opaque(P.X);
}

template <typename SRC>
__device__ inline void plane_load(const DwtPass &P, const PlTile &t, float (&v)[WV_PL_EPT]) {
  const WV_GLOBAL SRC *src = gptr((const SRC *)t.src);
  const int64_t sy = P.src_sy > 0 ? P.src_sy : (int64_t)P.ox;
  const int64_t sz = P.src_sz > 0 ? P.src_sz : (int64_t)P.ox * P.oy;
  const int ne = t.np * P.X * P.Y;
  const int tid = opaque_tid();
  // unconditional loads (elements past the tile re-read its last one): a
  // load under a branch would make the branch join wait for it, and the
  // loads would no longer be in flight together
#pragma unroll
  for (int i = 0; i < WV_PL_EPT; i++) {
    const int idx = min(tid + i * WV_DBLOCK, ne - 1);
    int p, y, x;
    plane_xyz(P, idx, p, y, x);
    const int zc = min(t.z0 + p, P.oz - 1), yc = min(y, P.oy - 1), xc = min(x, P.ox - 1);
    v[i] = to_f(src[zc * sz + yc * sy + xc]);
  }
}

__device__ inline void plane_to_lds(const DwtPass &P, const PlTile &t, const float (&v)[WV_PL_EPT],
                                    float *lds) {
  const int ne = t.np * P.X * P.Y;
  const int RS = P.X + 2 * WV_HALO + 1, PR = P.Y + 2 * WV_HALO;
  const int tid = opaque_tid();
#pragma unroll
  for (int i = 0; i < WV_PL_EPT; i++) {
    const int idx = tid + i * WV_DBLOCK;
    if (idx < ne) {
      int p, y, x;
      plane_xyz(P, idx, p, y, x);
      lds[(p * PR + WV_HALO + y) * RS + WV_HALO + x] = v[i];
    }
  }
}

__device__ inline void plane_store(const DwtPass &P, const PlTile &t, const float *lds, int xs,
                                   int ys) {
  WV_GLOBAL float *dst = gptr(t.dst);
  const int XY = P.X * P.Y;
  const int ne = t.np * XY;
  const int RS = P.X + 2 * WV_HALO + 1, PR = P.Y + 2 * WV_HALO;
  const int tid = opaque_tid();
  const int64_t base = (int64_t)t.z0 * XY;
  // every LDS read first (unconditional, so they issue together), then the
  // stores of the elements inside the tile
  float o[WV_PL_EPT];
#pragma unroll
  for (int i = 0; i < WV_PL_EPT; i++) {
    const int idx = min(tid + i * WV_DBLOCK, ne - 1);
    int p, y, x;
    plane_xyz(P, idx, p, y, x);
    o[i] = lds[(p * PR + WV_HALO + y + ys) * RS + WV_HALO + x + xs];
  }
#pragma unroll
  for (int i = 0; i < WV_PL_EPT; i++) {
    const int idx = tid + i * WV_DBLOCK;
    if (idx < ne) dst[base + idx] = o[i];
  }
}

// P.T planes per tile, P.levels x levels, P.levels2 y levels; persistent,
// the next tile's loads in flight during this tile's levels and stores.
template <int WAVE, typename SRC>
__global__ void __launch_bounds__(WV_DBLOCK, WV_DWAVES)
dwt_plane_kernel(DwtPass P, int ntiles) {
  extern __shared__ float smem[];
  const int X = P.X, Y = P.Y;
  const int RS = X + 2 * WV_HALO + 1, PR = Y + 2 * WV_HALO;
  const int xs = WAVE == WV_WAVELET_CDF97 && P.levels > 0 ? WV_HALO : 0;
  const int ys = WAVE == WV_WAVELET_CDF97 && P.levels2 > 0 ? WV_HALO : 0;
  // x levels: line l is row l % Y of plane l / Y; y levels: line l is
  // column l % X of plane l / X (after the x levels' shift)
  auto xbase = [&](int l) {
    const int p = l / Y;
    return (p * PR + WV_HALO + (l - p * Y)) * RS + WV_HALO;
  };
  auto ybase = [&](int l) {
    const int p = l / X;
    return (p * PR + WV_HALO) * RS + WV_HALO + xs + (l - p * X);
  };
  int g = blockIdx.x;
  const int G = gridDim.x;
  if (g >= ntiles) return;
  PlTile t = plane_tile(P, g);
  float v[WV_PL_EPT];
  plane_load<SRC>(P, t, v);
  for (;;) {
    plane_to_lds(P, t, v, smem);
    __syncthreads();
    const int gn = g + G;
    PlTile tn = t;
    if (gn < ntiles) {
      tn = plane_tile(P, gn);
      plane_load<SRC>(P, tn, v);
    }
    if (WAVE == WV_WAVELET_CDF97) {
      plane97_levels(smem, P.T * Y, xbase, 1, X, P.levels);
      plane97_levels(smem, P.T * X, ybase, RS, Y, P.levels2);
    } else {
      plane_haar_levels(smem, P.T * Y, xbase, 1, X, P.levels);
      plane_haar_levels(smem, P.T * X, ybase, RS, Y, P.levels2);
    }
    plane_store(P, t, smem, xs, ys);
    __syncthreads();
    if (gn >= ntiles) break;
    g = gn;
    t = tn;
  }
}

// ---- host side ------------------------------------------------------------
// WV_DWT_VEC_ST=0 keeps the segment stores scalar (for comparison)
static bool vec_store_off() {
  static const bool v = [] {
    const char *e = getenv("WV_DWT_VEC_ST");
    return e && e[0] == '0';
  }();
  return v;
}

// resident blocks per CU override (autotuning hook): WV_DWT_BPC
static int bpc_override() {
  static const int v = [] {
    const char *e = getenv("WV_DWT_BPC");
    const int x = e ? atoi(e) : 0;
    return (x < 0 || x > 16) ? 0 : x;
  }();
  return v;
}

int cu_count() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (!n) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
      n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

static size_t tile_lds_bytes(const DwtPass &P) {
  return (size_t)(P.L + 2 * WV_HALO) * (P.T + 1) * 4;
}

template <int WAVE, bool INV, typename SRC, bool CLAMP, int TT, int EPT, int NT>
static hipError_t launch_one(const DwtPass &P, int ntiles, hipStream_t s) {
  auto k = dwt_pass_kernel<WAVE, INV, SRC, CLAMP, TT, EPT, NT>;
  const size_t lds = tile_lds_bytes(P);
  // 160 KB of LDS per workgroup on gfx950, minus the kernel's static LDS;
  // set once per instantiation (thread-safe)
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 152 * 1024);
  if (attr != hipSuccess) return attr;
  // persistent grid: as many blocks as stay resident at once
  int per_cu = P.bpc > 0 ? P.bpc : bpc_override();
  int occ = 0;
  {
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, NT, lds);
    if (e != hipSuccess) return e;
  }
  if (per_cu > occ) per_cu = occ;
  if (!per_cu) per_cu = occ > 0 ? occ : 1;
  const long cap = (long)per_cu * cu_count();
  const int grid = (int)(ntiles < cap ? ntiles : cap);
  hipLaunchKernelGGL(k, dim3(grid), dim3(NT), lds, s, P, ntiles);
  return hipGetLastError();
}

// Compile-time tile widths need L % 64 == 0 (whole 64-position line
// segments per wave instruction); anything else runs the generic mapping.
template <int WAVE, bool INV, typename SRC, bool CLAMP>
static hipError_t launch_tile(const DwtPass &P, int ntiles, hipStream_t s) {
  if (P.L % 64 == 0 && P.nt == 512 && P.ept == 16) {
    switch (P.T) {
      case 128: return launch_one<WAVE, INV, SRC, CLAMP, 128, 16, 512>(P, ntiles, s);
      case 64: return launch_one<WAVE, INV, SRC, CLAMP, 64, 16, 512>(P, ntiles, s);
      case 32: return launch_one<WAVE, INV, SRC, CLAMP, 32, 16, 512>(P, ntiles, s);
      case 16: return launch_one<WAVE, INV, SRC, CLAMP, 16, 16, 512>(P, ntiles, s);
      default: break;
    }
  }
  if (P.L % 64 == 0 && P.nt == 512 && P.ept == 32) {
    switch (P.T) {
      case 256: return launch_one<WAVE, INV, SRC, CLAMP, 256, 32, 512>(P, ntiles, s);
      case 128: return launch_one<WAVE, INV, SRC, CLAMP, 128, 32, 512>(P, ntiles, s);
      case 64: return launch_one<WAVE, INV, SRC, CLAMP, 64, 32, 512>(P, ntiles, s);
      case 32: return launch_one<WAVE, INV, SRC, CLAMP, 32, 32, 512>(P, ntiles, s);
      case 16: return launch_one<WAVE, INV, SRC, CLAMP, 16, 32, 512>(P, ntiles, s);
      case 8: return launch_one<WAVE, INV, SRC, CLAMP, 8, 32, 512>(P, ntiles, s);
      default: break;
    }
  }
  if (P.L % 64 == 0 && P.nt == 512 && P.ept == 8) {
    switch (P.T) {
      case 64: return launch_one<WAVE, INV, SRC, CLAMP, 64, 8, 512>(P, ntiles, s);
      case 32: return launch_one<WAVE, INV, SRC, CLAMP, 32, 8, 512>(P, ntiles, s);
      case 16: return launch_one<WAVE, INV, SRC, CLAMP, 16, 8, 512>(P, ntiles, s);
      case 8: return launch_one<WAVE, INV, SRC, CLAMP, 8, 8, 512>(P, ntiles, s);
      default: break;
    }
  }
  // generic mapping (holds any tile with L*T <= 512*EPT)
  if (P.ept == 8) return launch_one<WAVE, INV, SRC, CLAMP, 0, 8, 512>(P, ntiles, s);
  if (P.ept == 16) return launch_one<WAVE, INV, SRC, CLAMP, 0, 16, 512>(P, ntiles, s);
  return launch_one<WAVE, INV, SRC, CLAMP, 0, 32, 512>(P, ntiles, s);
}

template <int WAVE, bool INV>
static hipError_t launch_wave(const DwtPass &P, int src_dtype, bool clamp, int ntiles,
                              hipStream_t s) {
  if (INV) return launch_tile<WAVE, true, float, false>(P, ntiles, s);
  if (clamp) {
    if (src_dtype == WV_DTYPE_UINT8) return launch_tile<WAVE, false, uint8_t, true>(P, ntiles, s);
    if (src_dtype == WV_DTYPE_INT32) return launch_tile<WAVE, false, int32_t, true>(P, ntiles, s);
    if (src_dtype == WV_DTYPE_INT16) return launch_tile<WAVE, false, int16_t, true>(P, ntiles, s);
    return launch_tile<WAVE, false, float, true>(P, ntiles, s);
  }
  return launch_tile<WAVE, false, float, false>(P, ntiles, s);
}

static int ilog2(int v) { int r = 0; while ((1 << (r + 1)) <= v) r++; return r; }

// tile width cap (autotuning hook): WV_DWT_TILE in the environment
static int tile_cap() {
  static const int cap = [] {
    const char *e = getenv("WV_DWT_TILE");
    const int c = e ? atoi(e) : WV_MAXT;
    return (c < 1 || c > WV_MAXT || (c & (c - 1))) ? WV_MAXT : c;
  }();
  return cap;
}

// elements per thread for x-line passes (autotuning hook): WV_DWT_EPT0
static int ept_axis0() {
  static const int v = [] {
    const char *e = getenv("WV_DWT_EPT0");
    const int x = e ? atoi(e) : 16;
    return (x != 16 && x != 32) ? 16 : x;
  }();
  return v;
}

// element strides of the x, y and z axes in a layout
static void layout_strides(int layout, int64_t X, int64_t Y, int64_t Z, int64_t st[3]) {
  if (layout == WV_LAYOUT_YXZ) { st[0] = Y; st[1] = 1; st[2] = X * Y; }
  else if (layout == WV_LAYOUT_ZXY) { st[0] = Z; st[1] = X * Z; st[2] = 1; }
  else { st[0] = 1; st[1] = X; st[2] = X * Y; }
}

// WV_DWT_PLANE=0 turns the plane pass off (for comparison)
static bool plane_off() {
  static const bool v = [] {
    const char *e = getenv("WV_DWT_PLANE");
    return e && e[0] == '0';
  }();
  return v;
}

// autotuning hooks: WV_DWT_PLANE_MAX (largest X*Y taking the plane pass),
// WV_DWT_PLANE_PP (planes per tile, clamped to what fits)
static int env_int(const char *name, int dflt) {
  const char *e = getenv(name);
  return e ? atoi(e) : dflt;
}

// Planes per tile of the plane pass for X x Y planes, or 0 when the planes
// do not take it: at most WV_PL_EPT elements per thread, at most one thread
// per line and at most 16 pairs per thread at the first level in either
// direction, and the tile within one block's LDS.
static int plane_count(int X, int Y, int Z) {
  static const int max_xy = env_int("WV_DWT_PLANE_MAX", 1024);
  static const int want = env_int("WV_DWT_PLANE_PP", 0);
  if (X < 8 || Y < 8 || X > 128 || Y > 128 || X * Y > max_xy) return 0;
  const long plane_lds = (long)(Y + 2 * WV_HALO) * (X + 2 * WV_HALO + 1) * 4;
  int pp = (int)std::min<long>({(long)WV_DBLOCK * WV_PL_EPT / (X * Y), (long)WV_DBLOCK / std::max(X, Y),
                                (want > 0 ? 150 : 76) * 1024 / plane_lds, (long)Z});
  if (want > 0 && want < pp) pp = want;
  for (; pp >= 1; pp--) {
    const int tx = WV_DBLOCK / (pp * Y), ty = WV_DBLOCK / (pp * X);   // threads per line
    if (tx >= 1 && ty >= 1 && (X / 2 + tx - 1) / tx <= WV_PL_EPT / 2 &&
        (Y / 2 + ty - 1) / ty <= WV_PL_EPT / 2)
      return pp;
  }
  return 0;
}

int forward_specs(const int padded[3], const int steps[3], PassSpec s[3]) {
  const int *S = steps;
  const int Z = padded[2];
  int n = 0;
  if (!S[0] && !S[1] && !S[2]) {            // no transform, no rotation
    s[n++] = {0, 0, WV_LAYOUT_XYZ, WV_LAYOUT_XYZ};
    return n;
  }
  if (S[0] && S[1] && !plane_off() && plane_count(padded[0], padded[1], Z) > 0) {
    s[n++] = {WV_PASS_PLANE, S[0], WV_LAYOUT_XYZ, WV_LAYOUT_XYZ, S[1]};
    if (S[2] || Z > 1) s[n++] = {2, S[2], WV_LAYOUT_XYZ, WV_LAYOUT_ZXY};
    return n;
  }
  if (S[0] && Z >= 32) {
    if (S[1]) {
      s[n++] = {0, S[0], WV_LAYOUT_XYZ, WV_LAYOUT_YXZ};
      s[n++] = {1, S[1], WV_LAYOUT_YXZ, WV_LAYOUT_ZXY};
    } else {
      s[n++] = {0, S[0], WV_LAYOUT_XYZ, WV_LAYOUT_ZXY};
    }
    if (S[2]) s[n++] = {2, S[2], WV_LAYOUT_ZXY, WV_LAYOUT_ZXY};
    return n;
  }
  if (S[0]) s[n++] = {0, S[0], WV_LAYOUT_XYZ, WV_LAYOUT_XYZ};
  if (S[1]) s[n++] = {1, S[1], WV_LAYOUT_XYZ, WV_LAYOUT_XYZ};
  if (S[2] || Z > 1) s[n++] = {2, S[2], WV_LAYOUT_XYZ, WV_LAYOUT_ZXY};
  return n;
}

int inverse_specs(const int padded[3], const int steps[3], PassSpec s[3]) {
  const int *S = steps;
  int n = 0;
  if (!S[0] && !S[1] && !S[2]) return 0;
  if (S[2] || padded[2] > 1) s[n++] = {2, S[2], WV_LAYOUT_ZXY, WV_LAYOUT_XYZ};
  if (S[1]) s[n++] = {1, S[1], WV_LAYOUT_XYZ, WV_LAYOUT_XYZ};
  if (S[0]) s[n++] = {0, S[0], WV_LAYOUT_XYZ, WV_LAYOUT_XYZ};
  return n;
}

void pass_buffers(const PassSpec *s, int n, int dst_out[3]) {
  if (n <= 0) return;
  dst_out[n - 1] = 1;
  for (int i = n - 2; i >= 0; i--) {
    const PassSpec &next = s[i + 1];
    dst_out[i] = next.lin == next.lout ? dst_out[i + 1] : !dst_out[i + 1];
  }
}

// Fills the tile geometry of a pass; returns the number of tiles or -1.
// Lines run along s.axis; u is the axis of neighbouring lines -- the fastest
// axis of whichever side is not line-contiguous, so its accesses are T-float
// segments across lines -- and v the remaining axis.  When both sides are
// line-contiguous and u, v are contiguous in both, they merge into one run
// of lines.  Each thread carries EPT tile elements (and owns at most EPT/2
// pairs per level); T is the largest power of two with L*T <= WV_DBLOCK*EPT,
// reduced when u has fewer lines.  Passes that are line-contiguous on both
// sides with 64..512 samples use EPT = 16 (half-size tiles, more resident
// blocks per CU); the others keep EPT = 32, whose segments across lines
// are twice as long.
int plan_pass(DwtPass &P, const PassSpec &s, bool clamp) {
  if (s.axis == WV_PASS_PLANE) {
    const int pp = plane_count(P.X, P.Y, P.Z);
    if (pp <= 0) return -1;
    P.plane = 1;
    P.levels = s.levels;
    P.levels2 = s.levels2;
    P.L = P.X;
    P.T = pp;
    P.nt = WV_DBLOCK;
    P.ept = WV_PL_EPT;
    P.magic_L = (unsigned)(((1ull << 32) + (unsigned)P.X - 1) / (unsigned)P.X);
    P.magic_XY = (unsigned)(((1ull << 32) + (unsigned)(P.X * P.Y) - 1) / (unsigned)(P.X * P.Y));
    P.ntiles1 = (P.Z + pp - 1) / pp;
    (void)clamp;
    const long long total = (long long)P.ntiles1 * (P.src_tab ? (P.ncub > 0 ? P.ncub : 0) : 1);
    return total > 0x7fffffff ? -1 : (int)total;
  }
  const int dims[3] = {P.X, P.Y, P.Z};
  const int odims[3] = {P.ox, P.oy, P.oz};
  const int a = s.axis;
  P.L = dims[a];
  P.levels = s.levels;
  if (P.L > WV_MAX_LINE) return -1;
  int64_t si[3], so[3];
  if (clamp) layout_strides(s.lin, P.ox, P.oy, P.oz, si);
  else layout_strides(s.lin, P.X, P.Y, P.Z, si);
  if (clamp && P.src_sy > 0 && s.lin == WV_LAYOUT_XYZ) {   // a box of a larger volume
    si[1] = P.src_sy;
    si[2] = P.src_sz;
  }
  layout_strides(s.lout, P.X, P.Y, P.Z, so);
  P.ld_line = si[a] == 1;
  P.st_line = so[a] == 1;
  int b = -1, c = -1;
  for (int k = 0; k < 3; k++)
    if (k != a) { if (b < 0) b = k; else c = k; }
  int u;
  if (!P.ld_line) u = si[b] == 1 ? b : c;
  else if (!P.st_line) u = so[b] == 1 ? b : c;
  else u = si[b] <= si[c] ? b : c;
  const int v = u == b ? c : b;
  if (!P.ld_line && so[u] != 1 && !P.st_line) return -1;   // no common fast axis
  const int *cap = clamp ? odims : dims;
  P.Nu = dims[u];
  P.Nv = dims[v];
  P.ld = {si[u], si[v], si[a], cap[u] - 1, cap[v] - 1, cap[a] - 1};
  P.st = {so[u], so[v], so[a], dims[u] - 1, dims[v] - 1, dims[a] - 1};
  const bool unpadded = !clamp || (P.ox == P.X && P.oy == P.Y && P.oz == P.Z);
  // rows that continue each other on both sides (line (u, v) at
  // (u + v * Nu) * su) form one run of lines -- for line-contiguous sides
  // and for position-contiguous ones (su == 1, whole rows of the fast axis)
  if (unpadded && P.ld.su * P.Nu == P.ld.sv && P.st.su * P.Nu == P.st.sv) {
    P.Nu *= P.Nv;
    P.Nv = 1;
    P.ld.ucap = P.st.ucap = P.Nu - 1;
    P.ld.vcap = P.st.vcap = 0;
  }
  P.ept = WV_EPT;
  P.nt = WV_DBLOCK;
  if (P.ld_line && P.st_line && P.L % 64 == 0 && P.L >= 64 && P.L <= 512) {
    P.ept = P.ept0 == 16 || P.ept0 == 32 ? P.ept0 : ept_axis0();
  }
  int T = P.tcap > 0 ? P.tcap : tile_cap();
  if (T > WV_MAXT) T = WV_MAXT;
  while (T & (T - 1)) T &= T - 1;
  while (T > 1 && (long)P.L * T > (long)P.nt * P.ept) T >>= 1;
  while (T > 1 && T / 2 >= P.Nu) T >>= 1;
  // the fewest elements per thread (32, 16 or 8) that hold the tile: short
  // lines or few lines leave most of a 32-element tile empty, and smaller
  // tiles keep more blocks resident
  while (P.ept > 8 && (long)P.L * T * 2 <= (long)P.nt * P.ept) P.ept >>= 1;
  P.T = T;
  P.log2T = ilog2(T);
  P.log2dblock_T = ilog2(P.nt / T);
  P.magic_L = (unsigned)(((1ull << 32) + (unsigned)P.L - 1) / (unsigned)P.L);
  P.ntu = (P.Nu + T - 1) / T;
  const long long n1 = (long long)P.ntu * P.Nv;
  if (n1 > 0x7fffffff) return -1;
  P.ntiles1 = (int)n1;
  const long long total = n1 * (P.src_tab ? (P.ncub > 0 ? P.ncub : 0) : 1);
  return total > 0x7fffffff ? -1 : (int)total;
}

template <int WAVE, typename SRC>
static hipError_t launch_plane_one(const DwtPass &P, int ntiles, hipStream_t s) {
  auto k = dwt_plane_kernel<WAVE, SRC>;
  const size_t lds = (size_t)P.T * (P.Y + 2 * WV_HALO) * (P.X + 2 * WV_HALO + 1) * 4;
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 152 * 1024);
  if (attr != hipSuccess) return attr;
  int occ = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, WV_DBLOCK, lds);
  if (e != hipSuccess) return e;
  const long cap = (long)(occ > 0 ? occ : 1) * cu_count();
  const int grid = (int)(ntiles < cap ? ntiles : cap);
  hipLaunchKernelGGL(k, dim3(grid), dim3(WV_DBLOCK), lds, s, P, ntiles);
  return hipGetLastError();
}

template <int WAVE>
static hipError_t launch_plane(const DwtPass &P, int src_dtype, int ntiles, hipStream_t s) {
  if (src_dtype == WV_DTYPE_UINT8) return launch_plane_one<WAVE, uint8_t>(P, ntiles, s);
  if (src_dtype == WV_DTYPE_INT16) return launch_plane_one<WAVE, int16_t>(P, ntiles, s);
  if (src_dtype == WV_DTYPE_INT32) return launch_plane_one<WAVE, int32_t>(P, ntiles, s);
  return launch_plane_one<WAVE, float>(P, ntiles, s);
}

hipError_t launch_pass(const DwtPass &P, int ntiles, int wavelet, bool inverse,
                       int src_dtype, bool clamp, bool stats, hipStream_t s) {
  (void)stats;
  if (ntiles <= 0) return hipSuccess;
  if (P.plane) {
    if (inverse) return hipErrorInvalidValue;
    return wavelet == WV_WAVELET_CDF97 ? launch_plane<WV_WAVELET_CDF97>(P, src_dtype, ntiles, s)
                                       : launch_plane<WV_WAVELET_HAAR>(P, src_dtype, ntiles, s);
  }
  DwtPass Q = P;
  // exact tiles: a tile holds exactly nt*ept elements, the source
  // needs no clamping, and lane offsets (< nt/T rows of the widest
  // stride, plus T) fit 32 bits
  const bool unpadded = !clamp || (P.ox == P.X && P.oy == P.Y && P.oz == P.Z);
  long long stride = 0;     // widest position stride of a position-major side
  if (!P.ld_line) stride = P.ld.sp;
  if (!P.st_line && P.st.sp > stride) stride = P.st.sp;
  Q.exact = (long long)P.L * P.T == (long long)P.nt * P.ept && P.T >= 8 &&
            P.L % 64 == 0 && unpadded &&
            ((long long)(P.nt / P.T) * stride + P.T) * 4 < (1ll << 31);
  // segments of 32+ lines as 16-byte vector stores: rows, positions and the
  // output base 16-byte aligned
  const bool dst_aligned = P.dst_tab ? P.dst_tab_aligned != 0
                                     : ((uintptr_t)P.dst & 15) == 0;
  Q.vec_st = Q.exact && !P.st_line && P.T % 32 == 0 && P.st.sp % 4 == 0 &&
             P.st.sv % 4 == 0 && dst_aligned && !vec_store_off();
  Q.rt_exact = (long long)P.L * P.T == (long long)P.nt * P.ept && P.T >= 8 && unpadded &&
               ((long long)(P.nt / P.T) * stride + P.T) * 4 < (1ll << 31);
  Q.st_block = Q.rt_exact && P.st_line && P.st.su == P.L && P.L % 4 == 0 &&
               P.st.sv % 4 == 0 && dst_aligned && !vec_store_off();
  if (wavelet == WV_WAVELET_CDF97)
    return inverse ? launch_wave<WV_WAVELET_CDF97, true>(Q, src_dtype, clamp, ntiles, s)
                   : launch_wave<WV_WAVELET_CDF97, false>(Q, src_dtype, clamp, ntiles, s);
  return inverse ? launch_wave<WV_WAVELET_HAAR, true>(Q, src_dtype, clamp, ntiles, s)
                 : launch_wave<WV_WAVELET_HAAR, false>(Q, src_dtype, clamp, ntiles, s);
}

}  // namespace wv
