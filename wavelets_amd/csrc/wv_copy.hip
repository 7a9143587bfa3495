// wv_copy.hip -- a device-to-device streaming copy with 128-bit accesses:
// the measured HBM ceiling the kernels' roofline fractions are read
// against (the published 8 TB/s is a peak no access pattern reaches).
// No counterpart in the reference; a measurement helper only.
//
// Each block copies contiguous chunks of 256 x U 16-byte vectors: every
// thread issues its U loads before any store (U = 4, 8 or 16: 4-16 KiB in
// flight per wave), and a grid of blocks_per_cu blocks per CU strides over
// the chunks.
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

typedef float cp_vec __attribute__((ext_vector_type(4)));   // 16 bytes

// U vectors per thread per chunk; a block's chunk is WV_BLOCK * U vectors
template <bool NT, int U>
__global__ void __launch_bounds__(WV_BLOCK)
copy_kernel(const cp_vec *__restrict__ src, cp_vec *__restrict__ dst, int64_t n4) {
  constexpr int CH = WV_BLOCK * U;
  const int t = threadIdx.x;
  const int64_t nfull = n4 / CH;
  for (int64_t c = blockIdx.x; c < nfull; c += gridDim.x) {
    const cp_vec *s = src + c * CH + t;
    cp_vec *d = dst + c * CH + t;
    cp_vec v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
      v[u] = NT ? __builtin_nontemporal_load(s + u * WV_BLOCK) : s[u * WV_BLOCK];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (NT) __builtin_nontemporal_store(v[u], d + u * WV_BLOCK);
      else d[u * WV_BLOCK] = v[u];
    }
  }
  // the tail (less than one chunk): block 0
  if (blockIdx.x == 0)
    for (int64_t i = nfull * CH + t; i < n4; i += WV_BLOCK) dst[i] = src[i];
}

template <int U>
hipError_t launch_copy(const void *src, void *dst, int64_t n4, bool nt, int bpc, hipStream_t s) {
  const int64_t chunks = n4 / (WV_BLOCK * U);
  int64_t grid = (int64_t)cu_count() * bpc;
  if (grid > chunks) grid = chunks > 0 ? chunks : 1;
  if (nt)
    hipLaunchKernelGGL((copy_kernel<true, U>), dim3((unsigned)grid), dim3(WV_BLOCK), 0, s,
                       (const cp_vec *)src, (cp_vec *)dst, n4);
  else
    hipLaunchKernelGGL((copy_kernel<false, U>), dim3((unsigned)grid), dim3(WV_BLOCK), 0, s,
                       (const cp_vec *)src, (cp_vec *)dst, n4);
  return hipGetLastError();
}

}  // namespace wv

extern "C" int wv_copy_stream(const void *src_d, void *dst_d, int64_t bytes, int nontemporal,
                              int blocks_per_cu, int vec_per_thread, void *stream) {
  using namespace wv;
  if (!src_d || !dst_d || bytes <= 0 || (bytes & 15) || blocks_per_cu < 1 ||
      blocks_per_cu > 64 || ((uintptr_t)src_d & 15) || ((uintptr_t)dst_d & 15) ||
      (vec_per_thread != 4 && vec_per_thread != 8 && vec_per_thread != 16)) {
    set_error("wv_copy_stream: bad arguments (16-byte multiples and alignment; 4, 8 or 16 "
              "vectors per thread)");
    return WV_ERR_ARG;
  }
  const int64_t n4 = bytes / 16;
  hipStream_t s = (hipStream_t)stream;
  const bool nt = nontemporal != 0;
  const hipError_t e =
      vec_per_thread == 4   ? launch_copy<4>(src_d, dst_d, n4, nt, blocks_per_cu, s)
      : vec_per_thread == 8 ? launch_copy<8>(src_d, dst_d, n4, nt, blocks_per_cu, s)
                            : launch_copy<16>(src_d, dst_d, n4, nt, blocks_per_cu, s);
  if (e != hipSuccess) {
    set_error("wv_copy_stream: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  return WV_OK;
}
