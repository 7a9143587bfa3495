// wv_dwt.hip -- fused axis passes of the 3-D Haar / CDF 9/7 transform for
// gfx950.
//
// One kernel, instantiated per (wavelet, direction, source dtype, clamping,
// tile width), processes a *line tile*: T lines of one axis, each of length
// L (the padded axis length).  The block
//   1. loads the tile into LDS as tile[pos][line] (row stride T+1 so both
//      line-contiguous and position-contiguous accesses are bank-conflict
//      free), converting u8/i16/i32 to float and REPEAT-padding by clamping
//      coordinates into the unpadded source; every thread keeps WV_UNR
//      independent loads in flight,
//   2. runs every level of the 1-D transform in LDS: each thread owns up to
//      WV_MAXP consecutive (low, high) output pairs of one line and walks a
//      9-value register window over its inputs (mirroring only in the
//      chunks that touch a line end), then all threads write back after a
//      barrier,
//   3. stores the tile either back in x,y,z layout or, for the z pass, in the
//      reference's rotated z,x,y output layout.
// Every global access of a tile is either a contiguous block of the tile or
// T-float segments (128 B at T = 32), so each pass moves 8 B per element
// (the input's own size plus 4 B on the first pass).
//
// With the tile width a compile-time constant, every LDS address inside a
// line chunk is an immediate offset from one base register.
//
// Arithmetic follows the reference CPU code operation by operation with
// contraction disabled (-ffp-contract=off), so coefficients are
// bit-identical to cdf97 / haar (dwt_cpu.cc:139-203,683-799).
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

#define WV_UNR 32  // independent loads / stores in flight per thread

template <typename S> __device__ inline float to_f(S v) { return (float)v; }

struct Tile {
  int nvalid;      // valid lines in this tile
  int a, b;        // axis0: first row; axis1: x0, z; axis2: x0, y
};

__device__ inline Tile decode_tile(const DwtPass &P, int T) {
  Tile t;
  const int tile = blockIdx.x;
  if (P.axis == 0) {
    const int rows = P.Y * P.Z;
    t.a = tile * T;
    t.b = 0;
    t.nvalid = min(T, rows - t.a);
  } else {
    const int xt = tile % P.ntiles_x;
    t.a = xt * T;
    t.b = tile / P.ntiles_x;
    t.nvalid = min(T, P.X - t.a);
  }
  return t;
}

// ---- tile addressing -------------------------------------------------------
// A load/store walks the tile either line-major (consecutive threads step
// along a line: x lines, and z lines of the rotated z,x,y layout) or
// position-major (consecutive threads step across lines: y and z lines of
// the x,y,z layout, where T neighbouring lines are T neighbouring x).
__device__ inline bool line_major_load(const DwtPass &P) {
  return P.axis == 0 || (P.axis == 2 && P.load_zxy);
}
__device__ inline bool line_major_store(const DwtPass &P) {
  return P.axis == 0 || (P.axis == 2 && P.store_zxy);
}

// global offset of (line l, position 0) for a line-major access
template <bool CLAMP>
__device__ inline size_t line_base_load(const DwtPass &P, const Tile &t, int l) {
  if (P.axis == 0) {
    const int r = t.a + l;
    int y = r % P.Y, z = r / P.Y;
    if (CLAMP) {
      y = min(y, P.oy - 1);
      z = min(z, P.oz - 1);
      return ((size_t)z * P.oy + y) * P.ox;
    }
    return ((size_t)z * P.Y + y) * P.X;
  }
  return ((size_t)t.b * P.X + (t.a + l)) * P.Z;     // z,x,y layout
}
__device__ inline size_t line_base_store(const DwtPass &P, const Tile &t, int l) {
  if (P.axis == 0) return (size_t)(t.a + l) * P.X;
  return ((size_t)t.b * P.X + (t.a + l)) * P.Z;
}

// global offset of (line 0, position p) for a position-major access; the
// line index is added as +x (clamped for the unpadded source)
template <bool CLAMP>
__device__ inline size_t pos_base_load(const DwtPass &P, const Tile &t, int p) {
  if (CLAMP) {
    if (P.axis == 1) {
      const int z = min(t.b, P.oz - 1), y = min(p, P.oy - 1);
      return ((size_t)z * P.oy + y) * P.ox;
    }
    const int z = min(p, P.oz - 1), y = min(t.b, P.oy - 1);
    return ((size_t)z * P.oy + y) * P.ox;
  }
  if (P.axis == 1) return ((size_t)t.b * P.Y + p) * P.X;
  return ((size_t)p * P.Y + t.b) * P.X;
}
__device__ inline size_t pos_base_store(const DwtPass &P, const Tile &t, int p) {
  if (P.axis == 1) return ((size_t)t.b * P.Y + p) * P.X;
  return ((size_t)p * P.Y + t.b) * P.X;
}

// ---- forward levels -------------------------------------------------------
template <bool MIRROR>
__device__ inline float ldm(const float *col, int p, int n, int LS) {
  if (MIRROR && (unsigned)p >= (unsigned)n) p = wv_mirror(p, n);
  return col[p * LS];
}

// one thread's chunk of pairs [j0, j0+cnt) of one line at level length n
template <bool MIRROR>
__device__ inline void cdf97_chunk(const float *col, int LS, int n, int j0, int cnt,
                                   float *lo, float *hi) {
  const int c = 2 * j0;
  const float *cc = col + c * LS;       // positions relative to c
  float w0, w1, w2, w3, w4, w5, w6, w7, w8;
  if (MIRROR) {
    w0 = ldm<true>(col, c - 4, n, LS); w1 = ldm<true>(col, c - 3, n, LS);
    w2 = ldm<true>(col, c - 2, n, LS); w3 = ldm<true>(col, c - 1, n, LS);
    w4 = ldm<true>(col, c, n, LS);     w5 = ldm<true>(col, c + 1, n, LS);
    w6 = ldm<true>(col, c + 2, n, LS); w7 = ldm<true>(col, c + 3, n, LS);
    w8 = ldm<true>(col, c + 4, n, LS);
  } else {
    w0 = cc[-4 * LS]; w1 = cc[-3 * LS]; w2 = cc[-2 * LS]; w3 = cc[-1 * LS];
    w4 = cc[0];       w5 = cc[LS];      w6 = cc[2 * LS];  w7 = cc[3 * LS];
    w8 = cc[4 * LS];
  }
#pragma unroll
  for (int q = 0; q < WV_MAXP; q++) {
    if (q < cnt) {
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
      w0 = w2; w1 = w3; w2 = w4; w3 = w5; w4 = w6; w5 = w7; w6 = w8;
      if (MIRROR) {
        const int nc = c + 2 * q + 2;
        w7 = ldm<true>(col, nc + 3, n, LS);
        w8 = ldm<true>(col, nc + 4, n, LS);
      } else {
        w7 = cc[(2 * q + 5) * LS];
        w8 = cc[(2 * q + 6) * LS];
      }
    }
  }
}

template <int WAVE, int TT>
__device__ void forward_levels(float *tile, const DwtPass &P) {
  const int T = TT > 0 ? TT : P.T;
  const int LS = T + 1;
  const int tid = threadIdx.x;
  const int line = tid % T;
  const int k = tid / T;
  const int tpl = WV_DBLOCK / T;
  float *col = tile + line;
  int n = P.L;
  for (int s = 0; s < P.levels; s++) {
    const int pairs = n >> 1;
    const int cpt = (pairs + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, pairs - j0);
    float lo[WV_MAXP], hi[WV_MAXP];
    if (cnt > 0) {
      if (WAVE == WV_WAVELET_CDF97) {
        // the window reaches 4 positions beyond the chunk on both sides
        if (2 * j0 - 4 >= 0 && 2 * (j0 + cnt) + 4 < n)
          cdf97_chunk<false>(col, LS, n, j0, cnt, lo, hi);
        else
          cdf97_chunk<true>(col, LS, n, j0, cnt, lo, hi);
      } else {
        const float *cc = col + 2 * j0 * LS;
#pragma unroll
        for (int q = 0; q < WV_MAXP; q++) {
          if (q < cnt) {
            const float a = cc[(2 * q) * LS], b = cc[(2 * q + 1) * LS];
            // haar_internal dwt_cpu.cc:159-162
            hi[q] = (a - b) * INV_SQRT2;
            lo[q] = (a + b) * INV_SQRT2;
          }
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
      float *lw = col + j0 * LS, *hw = col + (pairs + j0) * LS;
#pragma unroll
      for (int q = 0; q < WV_MAXP; q++) {
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

template <int WAVE, int TT>
__device__ void inverse_levels(float *tile, const DwtPass &P) {
  const int T = TT > 0 ? TT : P.T;
  const int LS = T + 1;
  const int tid = threadIdx.x;
  const int line = tid % T;
  const int k = tid / T;
  const int tpl = WV_DBLOCK / T;
  float *col = tile + line;
  int n = P.L >> (P.levels - 1);
  for (int s = 0; s < P.levels; s++) {
    const int half = n >> 1;
    const int cpt = (half + tpl - 1) / tpl;
    const int j0 = k * cpt;
    const int cnt = min(cpt, half - j0);
    float ev[WV_MAXP], od[WV_MAXP];
    if (cnt > 0) {
      if (WAVE == WV_WAVELET_CDF97) {
        const int c = 2 * j0;
        float u0 = ldi(col, c - 3, n, half, LS), u1 = ldi(col, c - 2, n, half, LS),
              u2 = ldi(col, c - 1, n, half, LS), u3 = ldi(col, c, n, half, LS),
              u4 = ldi(col, c + 1, n, half, LS), u5 = ldi(col, c + 2, n, half, LS),
              u6 = ldi(col, c + 3, n, half, LS), u7 = ldi(col, c + 4, n, half, LS),
              u8 = ldi(col, c + 5, n, half, LS);
#pragma unroll
        for (int q = 0; q < WV_MAXP; q++) {
          if (q < cnt) {
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
            const int nc = c + 2 * q + 2;
            u0 = u2; u1 = u3; u2 = u4; u3 = u5; u4 = u6; u5 = u7; u6 = u8;
            u7 = ldi(col, nc + 4, n, half, LS);
            u8 = ldi(col, nc + 5, n, half, LS);
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < WV_MAXP; q++) {
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
      for (int q = 0; q < WV_MAXP; q++) {
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
template <int WAVE, bool INV, typename SRC, bool CLAMP, int TT>
__global__ void __launch_bounds__(WV_DBLOCK, 2)
dwt_pass_kernel(DwtPass P) {
  extern __shared__ float smem[];
  __shared__ size_t lbase[32];
  float *tile = smem;
  const int T = TT > 0 ? TT : P.T;
  const int LS = T + 1;
  const int tid = threadIdx.x;
  const Tile t = decode_tile(P, T);
  const int total = T * P.L;
  const SRC *src = (const SRC *)P.src;

  // 1. load
  if (line_major_load(P)) {
    if (tid < T) lbase[tid] = tid < t.nvalid ? line_base_load<CLAMP>(P, t, tid) : 0;
    __syncthreads();
    const int pmax = (CLAMP && P.axis == 0) ? P.ox - 1 : P.L - 1;
    for (int i0 = tid; i0 < total; i0 += WV_DBLOCK * WV_UNR) {
      float v[WV_UNR];
#pragma unroll
      for (int u = 0; u < WV_UNR; u++) {
        const int i = i0 + u * WV_DBLOCK;
        v[u] = 0.f;
        if (i < total) {
          const int l = (int)__umulhi((unsigned)i, P.magic_L);
          const int p = i - l * P.L;
          if (l < t.nvalid) v[u] = to_f(src[lbase[l] + min(p, pmax)]);
        }
      }
#pragma unroll
      for (int u = 0; u < WV_UNR; u++) {
        const int i = i0 + u * WV_DBLOCK;
        if (i < total) {
          const int l = (int)__umulhi((unsigned)i, P.magic_L);
          const int p = i - l * P.L;
          tile[p * LS + l] = v[u];
        }
      }
    }
  } else {
    const int l = tid % T;
    const int pstep = WV_DBLOCK / T;
    const bool lvalid = l < t.nvalid;
    const int x = CLAMP ? min(t.a + l, P.ox - 1) : t.a + l;
    for (int p0 = tid / T; p0 < P.L; p0 += pstep * WV_UNR) {
      float v[WV_UNR];
#pragma unroll
      for (int u = 0; u < WV_UNR; u++) {
        const int p = p0 + u * pstep;
        v[u] = 0.f;
        if (p < P.L && lvalid) v[u] = to_f(src[pos_base_load<CLAMP>(P, t, p) + x]);
      }
#pragma unroll
      for (int u = 0; u < WV_UNR; u++) {
        const int p = p0 + u * pstep;
        if (p < P.L) tile[p * LS + l] = v[u];
      }
    }
  }
  __syncthreads();

  // 2. levels
  if (P.levels > 0) {
    if (INV) inverse_levels<WAVE, TT>(tile, P);
    else forward_levels<WAVE, TT>(tile, P);
  }

  // 3. store
  if (line_major_store(P)) {
    if (tid < T) lbase[tid] = tid < t.nvalid ? line_base_store(P, t, tid) : 0;
    __syncthreads();
    for (int i0 = tid; i0 < total; i0 += WV_DBLOCK * WV_UNR) {
#pragma unroll
      for (int u = 0; u < WV_UNR; u++) {
        const int i = i0 + u * WV_DBLOCK;
        if (i < total) {
          const int l = (int)__umulhi((unsigned)i, P.magic_L);
          const int p = i - l * P.L;
          if (l < t.nvalid) P.dst[lbase[l] + p] = tile[p * LS + l];
        }
      }
    }
  } else {
    const int l = tid % T;
    const int pstep = WV_DBLOCK / T;
    if (l < t.nvalid) {
      for (int p0 = tid / T; p0 < P.L; p0 += pstep * WV_UNR) {
#pragma unroll
        for (int u = 0; u < WV_UNR; u++) {
          const int p = p0 + u * pstep;
          if (p < P.L) P.dst[pos_base_store(P, t, p) + t.a + l] = tile[p * LS + l];
        }
      }
    }
  }
}

// ---- host side ------------------------------------------------------------
template <int WAVE, bool INV, typename SRC, bool CLAMP, int TT>
static hipError_t launch_one(const DwtPass &P, int ntiles, hipStream_t s) {
  auto k = dwt_pass_kernel<WAVE, INV, SRC, CLAMP, TT>;
  const size_t lds = (size_t)P.L * (P.T + 1) * 4;
  static bool attr_set = false;
  if (!attr_set) {
    // 160 KB of LDS per workgroup on gfx950, minus the kernel's static LDS
    hipError_t e = hipFuncSetAttribute((const void *)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       159 * 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL(k, dim3(ntiles), dim3(WV_DBLOCK), lds, s, P);
  return hipGetLastError();
}

template <int WAVE, bool INV, typename SRC, bool CLAMP>
static hipError_t launch_tile(const DwtPass &P, int ntiles, hipStream_t s) {
  switch (P.T) {
    case 32: return launch_one<WAVE, INV, SRC, CLAMP, 32>(P, ntiles, s);
    case 16: return launch_one<WAVE, INV, SRC, CLAMP, 16>(P, ntiles, s);
    case 8: return launch_one<WAVE, INV, SRC, CLAMP, 8>(P, ntiles, s);
    default: return launch_one<WAVE, INV, SRC, CLAMP, 0>(P, ntiles, s);
  }
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
  static int cap = -1;
  if (cap < 0) {
    const char *e = getenv("WV_DWT_TILE");
    cap = e ? atoi(e) : 32;
    if (cap != 8 && cap != 16 && cap != 32) cap = 32;
  }
  return cap;
}

// Fills the tile geometry of a pass; returns the number of tiles or -1 if
// the line is longer than one LDS tile can hold.
int plan_pass(DwtPass &P, int axis) {
  const int dims[3] = {P.X, P.Y, P.Z};
  P.axis = axis;
  P.L = dims[axis];
  if (P.L > WV_MAX_LINE) return -1;
  int T = tile_cap();
  while (T > 1 && (long)P.L * T > 2L * WV_DBLOCK * WV_MAXP) T >>= 1;
  if (axis != 0) {
    while (T > 8 && T / 2 >= P.X) T >>= 1;   // narrow cubelets
  }
  P.T = T;
  P.log2T = ilog2(T);
  P.magic_L = (unsigned)(((1ull << 32) + (unsigned)P.L - 1) / (unsigned)P.L);
  if (axis == 0) {
    const int rows = P.Y * P.Z;
    return (rows + T - 1) / T;
  }
  P.ntiles_x = (P.X + T - 1) / T;
  return P.ntiles_x * (axis == 1 ? P.Z : P.Y);
}

hipError_t launch_pass(const DwtPass &P, int ntiles, int wavelet, bool inverse,
                       int src_dtype, bool clamp, bool stats, hipStream_t s) {
  (void)stats;
  if (ntiles <= 0) return hipSuccess;
  if (wavelet == WV_WAVELET_CDF97)
    return inverse ? launch_wave<WV_WAVELET_CDF97, true>(P, src_dtype, clamp, ntiles, s)
                   : launch_wave<WV_WAVELET_CDF97, false>(P, src_dtype, clamp, ntiles, s);
  return inverse ? launch_wave<WV_WAVELET_HAAR, true>(P, src_dtype, clamp, ntiles, s)
                 : launch_wave<WV_WAVELET_HAAR, false>(P, src_dtype, clamp, ntiles, s);
}

}  // namespace wv
