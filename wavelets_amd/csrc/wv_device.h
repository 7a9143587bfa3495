// wv_device.h -- device-side helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WV_BLOCK 256        // threads per block of the streaming kernels
#define WV_DBLOCK 512       // threads per block of the DWT tile kernel
#ifndef WV_MAXP
#define WV_MAXP 16          // max (low,high) pairs a thread owns per level
#endif
// waves per SIMD the DWT kernel is compiled for (the launch-bounds second
// argument is the waves-per-EU hint on AMDGPU): 4 caps it at 128 VGPRs;
// the half-size tile variant (16 elements per thread) asks for 6 (80 VGPRs,
// three 512-thread blocks per CU), which measured faster than 8
#ifndef WV_DWAVES
#define WV_DWAVES 4
#endif
#ifndef WV_DWAVES16
#define WV_DWAVES16 6
#endif
// 8 elements per thread (short lines, few lines): 8 waves per EU, 64 VGPRs
#ifndef WV_DWAVES8
#define WV_DWAVES8 8
#endif
#define WV_H1_BITS 11       // threshold select digit 1: bits 30..20 of |x|
#define WV_H1_BINS 2048
#define WV_H1_SHIFT 20
#define WV_H2_BINS 2048     // digit 2: bits 19..9
#define WV_H2_SHIFT 9
#define WV_H3_BINS 512      // digit 3: bits 8..0

// The DWT pass marks its full-tile global accesses non-temporal (each
// sample is read and written once per pass).  WV_NT_LD / WV_NT_ST = 0 turn
// the hints off.  The quantize pass measured slower with them, the select
// unchanged, so those kernels use plain accesses.
#ifndef WV_NT_LD
#define WV_NT_LD 1
#endif
#ifndef WV_NT_ST
#define WV_NT_ST 1
#endif

// Global (address space 1) views of generic pointers.  Pointers that reach a
// kernel through a struct or a table are generic to the compiler, which
// then emits flat instructions; flat accesses also count on LGKM_CNT, so
// every wait for an LDS result would wait for the loads in flight as well.
#define WV_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline const WV_GLOBAL T *gptr(const T *p) { return (const WV_GLOBAL T *)p; }
template <typename T>
__device__ inline WV_GLOBAL T *gptr(T *p) { return (WV_GLOBAL T *)p; }
// 16-byte vectors for those views (HIP's float4 / int4 classes cannot be
// copied out of address space 1)
typedef float wv_f4 __attribute__((ext_vector_type(4)));
typedef int wv_i4 __attribute__((ext_vector_type(4)));

// Fused statistics block (device).  Layout is part of the ABI size
// WV_STATS_BYTES.
struct WvStats {
  uint32_t hist[WV_H1_BINS];
  uint32_t min_ord, max_ord, max_key, count_lo;
  uint32_t pad[12];
};

// Total order on floats through their bit patterns (no NaNs occur).
__device__ __host__ inline uint32_t wv_f2ord(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __host__ inline float wv_ord2f(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __builtin_bit_cast(float, u);
}

// MirroredArray::getOffset (dwt_cpu.h:162-176): whole-sample symmetric
// extension, folded as many times as needed (lines of length 2 and 4).
__device__ inline int wv_mirror(int o, int n) {
  if (o < 0) o = -o;
  while (o >= n) {
    o = n * 2 - o - 2;
    if (o < 0) o = -o;
  }
  return o;
}

// Wave-aggregated histogram increment: lanes holding the same bin are found
// with one ballot per bin bit and the lowest of them adds the group's
// count, so an LDS bin takes one atomic per wave instead of one per lane.
template <int BITS>
__device__ inline void wv_wave_hist_add(uint32_t *h, uint32_t bin, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < BITS; b++) {
    const bool bit = (bin >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  const int lane = threadIdx.x & 63;
  if (valid && (peers & ((1ull << lane) - 1)) == 0)
    atomicAdd(&h[bin], (uint32_t)__popcll(peers));
}
