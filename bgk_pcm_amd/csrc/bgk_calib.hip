// bgk_calib.hip -- reference copy kernels for HBM counter calibration.
//
// Not part of the time-step ABI (libbgk_calib.so, loaded only by bench.py
// --calibrate-copy).  rocprofv3's FETCH_SIZE derived counter does not read
// the same fraction of the true bytes for every access type on gfx950, so
// the fused kernels' counter readings are scaled by those of a copy of known
// size made with the same accesses: 8-byte loads and stores, plain or
// non-temporal.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef float f32x2_t __attribute__((ext_vector_type(2)));

template <bool NT>
__global__ void k_copy8(const f32x2_t* __restrict__ a, f32x2_t* __restrict__ b, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    if constexpr (NT) {
      __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
    } else {
      b[i] = a[i];
    }
  }
}

}  // namespace

extern "C" int bgk_calib_copy(const void* src, void* dst, size_t bytes, int nontemporal,
                              void* stream) {
  if (!src || !dst || bytes % 8 != 0) return -1;
  const size_t n = bytes / 8;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (nontemporal)
    hipLaunchKernelGGL(k_copy8<true>, dim3(65536), dim3(256), 0, s,
                       static_cast<const f32x2_t*>(src), static_cast<f32x2_t*>(dst), n);
  else
    hipLaunchKernelGGL(k_copy8<false>, dim3(65536), dim3(256), 0, s,
                       static_cast<const f32x2_t*>(src), static_cast<f32x2_t*>(dst), n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
