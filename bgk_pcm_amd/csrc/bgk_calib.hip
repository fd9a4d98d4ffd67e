// bgk_calib.hip -- reference copy kernels for HBM counter calibration.
//
// Not part of the time-step ABI (libbgk_calib.so, loaded only by bench.py
// --calibrate-copy and --layout-probe).  rocprofv3's FETCH_SIZE derived counter does not read
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

// Streams the bytes of the FP32 moment step, 56 per node in and out (a
// 16-byte, b 16-byte, c 16-byte part and an 8-byte phi), non-temporal, one
// node per thread.  TILED = false: the parts are three planes (a, b, c at
// 0, 16n and 32n bytes); true: 64-node tiles [64 a][64 b][64 c], so that a
// wave's access to a part is one contiguous kilobyte.  phi at 48n.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
template <bool TILED>
__global__ void k_stream56(const char* __restrict__ in, char* __restrict__ out, int64_t n) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const f32x4_t* a = reinterpret_cast<const f32x4_t*>(in);
  f32x4_t* o = reinterpret_cast<f32x4_t*>(out);
  int64_t ia, ib, ic;
  if constexpr (TILED) {
    ia = (q >> 6) * 192 + (q & 63);
    ib = ia + 64;
    ic = ia + 128;
  } else {
    ia = q;
    ib = n + q;
    ic = 2 * n + q;
  }
  f32x4_t x = __builtin_nontemporal_load(a + ia), y = __builtin_nontemporal_load(a + ib),
          z = __builtin_nontemporal_load(a + ic);
  const double* ph = reinterpret_cast<const double*>(in + 48 * n);
  const double f = __builtin_nontemporal_load(ph + q);
  x.x += 1.0f;
  y.y += 1.0f;
  z.z += 1.0f;
  __builtin_nontemporal_store(x, o + ia);
  __builtin_nontemporal_store(y, o + ib);
  __builtin_nontemporal_store(z, o + ic);
  __builtin_nontemporal_store(f + 1.0, reinterpret_cast<double*>(out + 48 * n) + q);
}

}  // namespace

// in and out hold 56 * nodes bytes each; nodes a multiple of 64.
extern "C" int bgk_calib_stream56(const void* in, void* out, int64_t nodes, int tiled,
                                  void* stream) {
  if (!in || !out || nodes <= 0 || nodes % 64 != 0) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)((nodes + 255) / 256));
  if (tiled)
    hipLaunchKernelGGL(k_stream56<true>, grid, dim3(256), 0, s, static_cast<const char*>(in),
                       static_cast<char*>(out), nodes);
  else
    hipLaunchKernelGGL(k_stream56<false>, grid, dim3(256), 0, s, static_cast<const char*>(in),
                       static_cast<char*>(out), nodes);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

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
