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
  // batches: ncub cubelets of this same geometry, cubelet c reading
  // src_tab[c] and writing dst_tab[c] (src / dst unused then); ntiles1 tiles
  // per cubelet
  const void *const *src_tab;
  float *const *dst_tab;
  int ncub, ntiles1;
  // launch overrides (0: defaults): tile width cap, elements per thread of
  // x-line passes, resident blocks per CU
  int tcap, ept0, bpc;
};

// Pass sequence of the forward transform.  The reference visits x, rotates,
// y, rotates, z (cdf97_3d dwt_cpu.cc:943-965); here every intermediate
// stays in x,y,z layout and only the last pass writes the rotated z,x,y
// layout.  With depth 1 the z,x,y and x,y,z memory orders coincide, so no
// z pass is needed unless it has levels.
struct PassPlan {
  int axis[3];
  int levels[3];
  int n;
};

inline PassPlan forward_plan(const wv_geom *g) {
  PassPlan p{};
  const int *S = g->steps;
  const bool any = S[0] || S[1] || S[2];
  if (!any) {            // no transform, no rotation: conversion only
    p.axis[0] = 0; p.levels[0] = 0; p.n = 1;
    return p;
  }
  if (S[0]) { p.axis[p.n] = 0; p.levels[p.n++] = S[0]; }
  if (S[1]) { p.axis[p.n] = 1; p.levels[p.n++] = S[1]; }
  if (S[2] || g->padded[2] > 1) { p.axis[p.n] = 2; p.levels[p.n++] = S[2]; }
  return p;
}

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

// batched threshold select: one job per cubelet; the collect launch gives
// job j blocks [blk0, blk0 + nblk) (nblk <= sel_max_collect_blocks())
struct BSelJob {
  const float *coef;
  int64_t n;
  float frac;
  void *state;            // sel_state_bytes() of device memory
  uint32_t *cand;         // candidate keys, capacity cap
  uint32_t cap;
  wv_thresh_result *res;  // device
  int blk0, nblk;
};
int64_t sel_state_bytes();
int sel_max_collect_blocks();
int64_t sel_collect_chunk();
hipError_t threshold_select_batch(const BSelJob *jobs_d, const int *blk0_d, int njobs,
                                  int total_blocks, hipStream_t s);

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
// batched quantize + histogram: job j gets blocks [blk0, blk0 + nblk) of
// one launch; q points at the job's parameter block in device memory
// (filled by the host for log / codebook, by bquant_prepare for uniform)
struct BQuantJob {
  const float *coef;
  int64_t n;
  int32_t *bins;
  uint32_t *counts;       // bin_count + 1 entries
  QuantDev *q;
  const wv_thresh_result *res;
  int alg, bin_count;
  int blk0, nblk;
};
int quant_blocks_for(int64_t n);
QuantDev quant_dev_of(const wv_quant_params *p);   // table = p->table_d
hipError_t quantize_hist_batch(const BQuantJob *jobs_d, const int *blk0_d, int njobs,
                               int total_blocks, hipStream_t s);
hipError_t dequantize_table(const int32_t *bins, int64_t n, const float *table,
                            int bin_count, float *out, hipStream_t s);

void set_error(const char *fmt, ...);

}  // namespace wv
