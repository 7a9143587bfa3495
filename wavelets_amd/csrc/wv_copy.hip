// wv_copy.hip -- a device-to-device streaming copy with 128-bit accesses:
// the measured HBM ceiling the kernels' roofline fractions are read
// against (the published 8 TB/s is a peak no access pattern reaches).
// No counterpart in the reference; a measurement helper only.
//
// Each block copies contiguous 64 KiB chunks: every thread issues its
// sixteen 16-byte loads before any store, so a wave keeps 16 KiB in flight,
// and a grid of blocks_per_cu blocks per CU strides over the chunks.
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define CP_UNROLL 16                                // vectors per thread per chunk
#define CP_CHUNK (WV_BLOCK * CP_UNROLL)             // vectors per block chunk

typedef float cp_vec __attribute__((ext_vector_type(4)));   // 16 bytes

template <bool NT>
__global__ void __launch_bounds__(WV_BLOCK)
copy_kernel(const cp_vec *__restrict__ src, cp_vec *__restrict__ dst, int64_t n4) {
  const int t = threadIdx.x;
  const int64_t nfull = n4 / CP_CHUNK;
  for (int64_t c = blockIdx.x; c < nfull; c += gridDim.x) {
    const cp_vec *s = src + c * CP_CHUNK + t;
    cp_vec *d = dst + c * CP_CHUNK + t;
    cp_vec v[CP_UNROLL];
#pragma unroll
    for (int u = 0; u < CP_UNROLL; u++)
      v[u] = NT ? __builtin_nontemporal_load(s + u * WV_BLOCK) : s[u * WV_BLOCK];
#pragma unroll
    for (int u = 0; u < CP_UNROLL; u++) {
      if (NT) __builtin_nontemporal_store(v[u], d + u * WV_BLOCK);
      else d[u * WV_BLOCK] = v[u];
    }
  }
  // the tail (less than one chunk): block 0
  if (blockIdx.x == 0)
    for (int64_t i = nfull * CP_CHUNK + t; i < n4; i += WV_BLOCK) dst[i] = src[i];
}

}  // namespace wv

extern "C" int wv_copy_stream(const void *src_d, void *dst_d, int64_t bytes, int nontemporal,
                              int blocks_per_cu, void *stream) {
  using namespace wv;
  if (!src_d || !dst_d || bytes <= 0 || (bytes & 15) || blocks_per_cu < 1 ||
      blocks_per_cu > 64 || ((uintptr_t)src_d & 15) || ((uintptr_t)dst_d & 15)) {
    set_error("wv_copy_stream: bad arguments (16-byte multiples and alignment)");
    return WV_ERR_ARG;
  }
  const int64_t n4 = bytes / 16;
  const int64_t chunks = n4 / CP_CHUNK;
  int64_t grid = (int64_t)cu_count() * blocks_per_cu;
  if (grid > chunks) grid = chunks > 0 ? chunks : 1;
  hipStream_t s = (hipStream_t)stream;
  if (nontemporal)
    hipLaunchKernelGGL(copy_kernel<true>, dim3((unsigned)grid), dim3(WV_BLOCK), 0, s,
                       (const cp_vec *)src_d, (cp_vec *)dst_d, n4);
  else
    hipLaunchKernelGGL(copy_kernel<false>, dim3((unsigned)grid), dim3(WV_BLOCK), 0, s,
                       (const cp_vec *)src_d, (cp_vec *)dst_d, n4);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("wv_copy_stream: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  return WV_OK;
}
