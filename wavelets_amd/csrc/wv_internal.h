// wv_internal.h -- declarations shared between the gfx950 kernel files and
// the C ABI implementation.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/wv_hip.h"

#define WV_MAX_LINE 16384   // longest axis one LDS tile holds
#define WV_MAXT 256         // most lines in one LDS tile

namespace wv {

// One fused axis pass over the whole (padded) volume.
struct DwtPass {
  const void *src;
  float *dst;
  struct WvStats *stats;
  int X, Y, Z;        // padded volume dims (x, y, z)
  int ox, oy, oz;     // dims of the clamped source (the unpadded input)
  int axis;           // 0: x lines, 1: y lines, 2: z lines
  int load_zxy;       // source is in the rotated z,x,y layout (axis 2)
  int store_zxy;      // destination is in the z,x,y layout (axis 2)
  int levels;
  int T, log2T;       // lines per tile (a power of two <= WV_MAXT)
  int log2dblock_T;   // log2(threads per block / T)
  int L;              // line length
  unsigned magic_L;   // ceil(2^32 / L) for exact division of i < 2^16
  int ntiles_x;
  int ept;            // tile elements per thread (16 or 32)
  int exact;          // L*T fills a tile exactly, no padding, 32-bit lane offsets
};

int plan_pass(DwtPass &P, int axis);
hipError_t launch_pass(const DwtPass &P, int ntiles, int wavelet, bool inverse,
                       int src_dtype, bool clamp, bool stats, hipStream_t s);

// statistics / threshold select
hipError_t stats_clear(void *stats, hipStream_t s);
hipError_t stats_compute(const float *coef, int64_t n, void *stats,
                         hipStream_t s);
hipError_t threshold_select(const float *coef, int64_t n, float frac,
                            const void *stats, void *ws,
                            wv_thresh_result *res, hipStream_t s);
int64_t threshold_ws_bytes(int64_t n);

// quantize + histogram
struct QuantDev {
  int alg, bin_count, zero_bin;
  float threshold;
  float neg_scale, neg_offset, pos_scale, pos_offset;
  int n_pos, n_neg;   // log edges / codebook boundaries
  float pos_clamp;    // log: |x| from which the top bin is clamped to B-1
  const float *table; // device: log: pos edges then neg edges;
                      //         codebook: boundaries
};
hipError_t quantize_hist(const float *coef, int64_t n, const QuantDev &q,
                         int32_t *bins, uint32_t *counts, hipStream_t s);
hipError_t quantize_hist_uniform_dev(const float *coef, int64_t n,
                                     int bin_count,
                                     const wv_thresh_result *res,
                                     int32_t *bins, uint32_t *counts,
                                     hipStream_t s);
hipError_t dequantize_table(const int32_t *bins, int64_t n, const float *table,
                            int bin_count, float *out, hipStream_t s);

void set_error(const char *fmt, ...);

}  // namespace wv
