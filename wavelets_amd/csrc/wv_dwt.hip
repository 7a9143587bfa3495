// wv_dwt.hip -- fused axis passes of the 3-D Haar / CDF 9/7 transform for
// gfx950.
//
// One kernel, instantiated per (wavelet, direction, source dtype, clamping,
// fused statistics), processes a *line tile*: T lines of one axis, each of
// length L (the padded axis length).  The block
//   1. loads the tile into LDS as tile[pos][line] (row stride T+1 so both
//      line-contiguous and position-contiguous accesses are bank-conflict
//      free), converting u8/i32 to float and REPEAT-padding by clamping
//      coordinates into the unpadded source,
//   2. runs every level of the 1-D transform in LDS: each thread owns up to
//      WV_MAXP consecutive (low, high) output pairs of one line and walks a
//      9-value register window over its inputs, then all threads write back
//      after a barrier,
//   3. stores the tile either back in x,y,z layout or, for the z pass, in the
//      reference's rotated z,x,y output layout, optionally folding every
//      stored coefficient into the threshold statistics (a 2048-bin LDS
//      histogram of |x| bits 30..20, min, max, max |x|).
// Every global access of a tile is either a contiguous block of the tile or
// T-float (128 B at T = 32) segments, so each pass moves 8 B per element
// (plus the input's own size on the first pass).
//
// Arithmetic follows the reference CPU code operation by operation with
// contraction disabled (-ffp-contract=off), so coefficients are
// bit-identical to cdf97 / haar (dwt_cpu.cc:139-203,683-799).
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

template <typename S> __device__ inline float to_f(S v) { return (float)v; }

struct Tile {
  int nvalid;      // valid lines in this tile
  int a, b;        // fixed coordinates: axis0 -> first row; axis1 -> x0,z;
                   // axis2 -> x0,y
};

__device__ inline Tile decode_tile(const DwtPass &P) {
  Tile t;
  const int tile = blockIdx.x;
  if (P.axis == 0) {
    const int rows = P.Y * P.Z;
    t.a = tile * P.T;
    t.b = 0;
    t.nvalid = min(P.T, rows - t.a);
  } else {
    const int xt = tile % P.ntiles_x;
    t.a = xt * P.T;
    t.b = tile / P.ntiles_x;
    t.nvalid = min(P.T, P.X - t.a);
  }
  return t;
}

// source index of tile element (line l, position p)
template <bool CLAMP>
__device__ inline size_t src_index(const DwtPass &P, const Tile &t, int l,
                                   int p) {
  int x, y, z;
  if (P.axis == 0) {
    const int r = t.a + l;
    y = r % P.Y;
    z = r / P.Y;
    x = p;
  } else if (P.axis == 1) {
    x = t.a + l; y = p; z = t.b;
  } else {
    x = t.a + l; y = t.b; z = p;
    if (P.load_zxy) return ((size_t)y * P.X + x) * P.Z + z;
  }
  if (CLAMP) {
    x = min(x, P.ox - 1);
    y = min(y, P.oy - 1);
    z = min(z, P.oz - 1);
    return ((size_t)z * P.oy + y) * P.ox + x;
  }
  return ((size_t)z * P.Y + y) * P.X + x;
}

__device__ inline size_t dst_index(const DwtPass &P, const Tile &t, int l,
                                   int p) {
  if (P.axis == 0) return (size_t)(t.a + l) * P.X + p;
  if (P.axis == 1) return ((size_t)t.b * P.Y + p) * P.X + (t.a + l);
  if (P.store_zxy) return ((size_t)t.b * P.X + (t.a + l)) * P.Z + p;
  return ((size_t)p * P.Y + t.b) * P.X + (t.a + l);
}

// line-major traversal (consecutive i walk along a line) or position-major
// (consecutive i walk across lines at one position)
__device__ inline bool line_major_load(const DwtPass &P) {
  return P.axis == 0 || (P.axis == 2 && P.load_zxy);
}
__device__ inline bool line_major_store(const DwtPass &P) {
  return P.axis == 0 || (P.axis == 2 && P.store_zxy);
}

__device__ inline void split(const DwtPass &P, bool line_major, int i, int &l,
                             int &p) {
  if (line_major) {
    l = (int)__umulhi((unsigned)i, P.magic_L);
    p = i - l * P.L;
  } else {
    l = i & (P.T - 1);
    p = i >> P.log2T;
  }
}

// ---- forward levels -------------------------------------------------------
__device__ inline float ldm(const float *col, int p, int n, int LS) {
  if ((unsigned)p >= (unsigned)n) p = wv_mirror(p, n);
  return col[p * LS];
}

template <int WAVE>
__device__ void forward_levels(float *tile, const DwtPass &P) {
  const int tid = threadIdx.x;
  const int LS = P.T + 1;
  const int line = tid & (P.T - 1);
  const int k = tid >> P.log2T;
  const int tpl = WV_BLOCK >> P.log2T;
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
        const int c = 2 * j0;
        float w0 = ldm(col, c - 4, n, LS), w1 = ldm(col, c - 3, n, LS),
              w2 = ldm(col, c - 2, n, LS), w3 = ldm(col, c - 1, n, LS),
              w4 = ldm(col, c, n, LS), w5 = ldm(col, c + 1, n, LS),
              w6 = ldm(col, c + 2, n, LS), w7 = ldm(col, c + 3, n, LS),
              w8 = ldm(col, c + 4, n, LS);
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
            const int nc = c + 2 * q + 2;
            w0 = w2; w1 = w3; w2 = w4; w3 = w5; w4 = w6; w5 = w7; w6 = w8;
            w7 = ldm(col, nc + 3, n, LS);
            w8 = ldm(col, nc + 4, n, LS);
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < WV_MAXP; q++) {
          if (q < cnt) {
            const int j = j0 + q;
            const float a = col[(2 * j) * LS], b = col[(2 * j + 1) * LS];
            // haar_internal dwt_cpu.cc:159-162
            hi[q] = (a - b) * INV_SQRT2;
            lo[q] = (a + b) * INV_SQRT2;
          }
        }
      }
    }
    __syncthreads();
    if (cnt > 0) {
#pragma unroll
      for (int q = 0; q < WV_MAXP; q++) {
        if (q < cnt) {
          col[(j0 + q) * LS] = lo[q];
          col[(pairs + j0 + q) * LS] = hi[q];
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

template <int WAVE>
__device__ void inverse_levels(float *tile, const DwtPass &P) {
  const int tid = threadIdx.x;
  const int LS = P.T + 1;
  const int line = tid & (P.T - 1);
  const int k = tid >> P.log2T;
  const int tpl = WV_BLOCK >> P.log2T;
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
#pragma unroll
      for (int q = 0; q < WV_MAXP; q++) {
        if (q < cnt) {
          col[(2 * (j0 + q)) * LS] = ev[q];
          col[(2 * (j0 + q) + 1) * LS] = od[q];
        }
      }
    }
    __syncthreads();
    n <<= 1;
  }
}

// ---- the pass kernel ------------------------------------------------------
template <int WAVE, bool INV, typename SRC, bool CLAMP, bool STATS>
__global__ void __launch_bounds__(WV_BLOCK)
dwt_pass_kernel(DwtPass P) {
  extern __shared__ float smem[];
  float *tile = smem;
  const int LS = P.T + 1;
  const int tid = threadIdx.x;
  const Tile t = decode_tile(P);
  const int total = P.T * P.L;
  const SRC *src = (const SRC *)P.src;

  uint32_t *sh_hist = nullptr;
  if (STATS) {
    sh_hist = (uint32_t *)(smem + P.L * LS);
    for (int i = tid; i < WV_H1_BINS; i += WV_BLOCK) sh_hist[i] = 0;
  }

  // 1. load
  {
    const bool lm = line_major_load(P);
    for (int i = tid; i < total; i += WV_BLOCK) {
      int l, p;
      split(P, lm, i, l, p);
      float v = 0.f;
      if (l < t.nvalid) v = to_f(src[src_index<CLAMP>(P, t, l, p)]);
      tile[p * LS + l] = v;
    }
  }
  __syncthreads();

  // 2. levels
  if (P.levels > 0) {
    if (INV) inverse_levels<WAVE>(tile, P);
    else forward_levels<WAVE>(tile, P);
  }

  // 3. store (+ statistics)
  uint32_t mn = 0xffffffffu, mx = 0u, mk = 0u;
  unsigned zero_run = 0;   // |x| keys landing in bin 0 (counted in a register)
  {
    const bool lm = line_major_store(P);
    for (int i = tid; i < total; i += WV_BLOCK) {
      int l, p;
      split(P, lm, i, l, p);
      if (l < t.nvalid) {
        const float v = tile[p * LS + l];
        P.dst[dst_index(P, t, l, p)] = v;
        if (STATS) {
          const uint32_t o = wv_f2ord(v);
          mn = min(mn, o);
          mx = max(mx, o);
          const uint32_t key = __builtin_bit_cast(uint32_t, v) & 0x7fffffffu;
          mk = max(mk, key);
          const uint32_t bin = key >> WV_H1_SHIFT;
          if (bin == 0) zero_run++;
          else atomicAdd(&sh_hist[bin], 1u);
        }
      }
    }
  }
  if (STATS) {
    WvStats *st = P.stats;
    // wave reductions, then one global atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
      mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
      mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
      mk = max(mk, (uint32_t)__shfl_xor((int)mk, off));
      zero_run += (unsigned)__shfl_xor((int)zero_run, off);
    }
    if ((tid & 63) == 0) {
      if (zero_run) atomicAdd(&sh_hist[0], zero_run);
      atomicMin(&st->min_ord, mn);
      atomicMax(&st->max_ord, mx);
      atomicMax(&st->max_key, mk);
    }
    __syncthreads();
    for (int i = tid; i < WV_H1_BINS; i += WV_BLOCK) {
      const uint32_t c = sh_hist[i];
      if (c) atomicAdd(&st->hist[i], c);
    }
  }
}

// ---- host side ------------------------------------------------------------
template <int WAVE, bool INV, typename SRC, bool CLAMP, bool STATS>
static hipError_t launch_one(const DwtPass &P, int ntiles, hipStream_t s) {
  auto k = dwt_pass_kernel<WAVE, INV, SRC, CLAMP, STATS>;
  const size_t lds = ((size_t)P.L * (P.T + 1) + (STATS ? WV_H1_BINS : 0)) * 4;
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void *)k,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL(k, dim3(ntiles), dim3(WV_BLOCK), lds, s, P);
  return hipGetLastError();
}

template <int WAVE, bool INV>
static hipError_t launch_wave(const DwtPass &P, int src_dtype, bool clamp,
                              bool stats, int ntiles, hipStream_t s) {
  if (INV) {   // inverse passes read float coefficients, never clamp
    return launch_one<WAVE, true, float, false, false>(P, ntiles, s);
  }
  if (clamp) {
    if (src_dtype == WV_DTYPE_UINT8)
      return stats ? launch_one<WAVE, false, uint8_t, true, true>(P, ntiles, s)
                   : launch_one<WAVE, false, uint8_t, true, false>(P, ntiles, s);
    if (src_dtype == WV_DTYPE_INT32)
      return stats ? launch_one<WAVE, false, int32_t, true, true>(P, ntiles, s)
                   : launch_one<WAVE, false, int32_t, true, false>(P, ntiles, s);
    return stats ? launch_one<WAVE, false, float, true, true>(P, ntiles, s)
                 : launch_one<WAVE, false, float, true, false>(P, ntiles, s);
  }
  return stats ? launch_one<WAVE, false, float, false, true>(P, ntiles, s)
               : launch_one<WAVE, false, float, false, false>(P, ntiles, s);
}

static int ilog2(int v) { int r = 0; while ((1 << (r + 1)) <= v) r++; return r; }

// Fills the tile geometry of a pass; returns the number of tiles or -1 if
// the line is longer than one LDS tile can hold.
int plan_pass(DwtPass &P, int axis) {
  const int dims[3] = {P.X, P.Y, P.Z};
  P.axis = axis;
  P.L = dims[axis];
  if (P.L > WV_MAX_LINE) return -1;
  int T = 32;
  while (T > 1 && (long)P.L * T > 2L * WV_BLOCK * WV_MAXP) T >>= 1;
  if (axis != 0) {
    while (T > 1 && T / 2 >= P.X) T >>= 1;   // narrow cubelets
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
  if (ntiles <= 0) return hipSuccess;
  if (wavelet == WV_WAVELET_CDF97)
    return inverse ? launch_wave<WV_WAVELET_CDF97, true>(P, src_dtype, clamp, stats, ntiles, s)
                   : launch_wave<WV_WAVELET_CDF97, false>(P, src_dtype, clamp, stats, ntiles, s);
  return inverse ? launch_wave<WV_WAVELET_HAAR, true>(P, src_dtype, clamp, stats, ntiles, s)
                 : launch_wave<WV_WAVELET_HAAR, false>(P, src_dtype, clamp, stats, ntiles, s);
}

}  // namespace wv
