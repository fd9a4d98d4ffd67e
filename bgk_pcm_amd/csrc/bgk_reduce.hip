// bgk_reduce.hip -- interior reductions and gathers for the diagnostics.
//
// field_check, mass_check and diag (STEP_7/field_check.f90:11-21,
// mass_check.f90:12-29, diag.f90:63-104) sum over the interior serially.
// Here every block owns a fixed set of rows, every thread a fixed set of
// columns, and each accumulates a compensated (Neumaier) FP64 sum; the
// (sum, error) pairs are then combined by error-free TwoSum in a fixed tree
// (cross-lane shuffles within each wavefront, then the wavefronts in order)
// and a second one-block pass.  The result is deterministic and within one
// rounding of the exact sum, so it differs from the serial sum only by that
// sum's own rounding error.  Counts are exact integers.
#include <mutex>

#include "bgk_internal.h"

namespace bgk {
namespace {

constexpr int RB = 256;        // threads per block
constexpr int MAX_PARTS = 1024;  // first-pass blocks

struct CSum {
  double s, c;
};

__device__ __forceinline__ void cadd(CSum& a, double x) {
  // Neumaier compensated addition
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x))
    a.c += (a.s - t) + x;
  else
    a.c += (x - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ CSum cmerge(CSum a, CSum b) {
  // TwoSum of the leading parts, errors added
  const double s = a.s + b.s;
  const double bb = s - a.s;
  const double err = (a.s - (s - bb)) + (b.s - bb);
  return CSum{s, a.c + b.c + err};
}

// Workgroup reduction of N compensated sums: a butterfly within each
// 64-wide wavefront (cross-lane shuffles, no LDS), then the RB/64 wavefront
// results combined in wave order by the first thread.  Fixed order:
// deterministic.  The result is valid in thread 0.
template <int N>
__device__ __forceinline__ void block_reduce(CSum (&v)[N], CSum (*lds)[RB / 64]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      CSum o;
      o.s = __shfl_down(v[k].s, off, 64);
      o.c = __shfl_down(v[k].c, off, 64);
      v[k] = cmerge(v[k], o);
    }
  if ((t & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) lds[k][t >> 6] = v[k];
  __syncthreads();
  if (t == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) {
      CSum a = lds[k][0];
      for (int w = 1; w < RB / 64; ++w) a = cmerge(a, lds[k][w]);
      v[k] = a;
    }
}

// Sums of three planes over the interior (rho, T, phi).
__global__ __launch_bounds__(RB) void k_sum3(const double* __restrict__ a,
                                             const double* __restrict__ b,
                                             const double* __restrict__ c3, Geom gm,
                                             CSum* __restrict__ part) {
  __shared__ CSum lds[3][RB / 64];
  CSum acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  for (int j = blockIdx.x + 1; j <= gm.ny; j += gridDim.x) {
    const int64_t r = (int64_t)j * gm.pitch;
    for (int i = threadIdx.x + 1; i <= gm.nx; i += RB) {
      cadd(acc[0], a[r + i]);
      cadd(acc[1], b[r + i]);
      cadd(acc[2], c3[r + i]);
    }
  }
  block_reduce<3>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[blockIdx.x * 4 + k] = acc[k];
}

// Terms of STEP_7/diag.f90:63-104 per row/node:
//  k=0: Ny*((Tin - T(2,j))/2)/Delta_T                  (column 2)
//  k=1: Ny*|(-3T(1,j)+4T(2,j)-T(3,j))/2|/Delta_T        (columns 1..3)
//  k=2: count of phi > 0;  k=3: sum of i where phi == 0
__global__ __launch_bounds__(RB) void k_diag(const double* __restrict__ T,
                                             const double* __restrict__ phi, Geom gm,
                                             double tin, double dT, CSum* __restrict__ part) {
  __shared__ CSum lds[4][RB / 64];
  CSum acc[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  long long cnt = 0, isum = 0;
  const double dny = (double)gm.ny;
  for (int j = blockIdx.x + 1; j <= gm.ny; j += gridDim.x) {
    const int64_t r = (int64_t)j * gm.pitch;
    if (threadIdx.x == 0) {
      const double t1 = T[r + 1], t2 = T[r + 2], t3 = T[r + 3];
      cadd(acc[0], dny * ((tin - t2) / 2.0) / dT);
      const double gt = fabs((-(3.0 * t1) + 4.0 * t2 - t3) / 2.0);
      cadd(acc[1], dny * gt / dT);
    }
    for (int i = threadIdx.x + 1; i <= gm.nx; i += RB) {
      const double p = phi[r + i];
      if (p > 0.0) cnt += 1;
      if (p == 0.0) isum += i;
    }
  }
  acc[2].s = (double)cnt;
  acc[3].s = (double)isum;
  block_reduce<4>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) part[blockIdx.x * 4 + k] = acc[k];
}

// Second pass: combine nparts partials of nk sums (stride 4).
__global__ __launch_bounds__(RB) void k_final(const CSum* __restrict__ part, int nparts, int nk,
                                              double* __restrict__ out) {
  __shared__ CSum lds[4][RB / 64];
  CSum acc[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  for (int p = threadIdx.x; p < nparts; p += RB)
    for (int k = 0; k < 4; ++k) acc[k] = cmerge(acc[k], part[p * 4 + k]);
  block_reduce<4>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < nk; ++k) out[k] = acc[k].s + acc[k].c;
}

__global__ void k_gather(const double* __restrict__ plane, int64_t pitch, int n,
                         const int32_t* __restrict__ ij, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = plane[(int64_t)ij[2 * t + 1] * pitch + ij[2 * t]];
}

inline int nparts(const Geom& g) { return g.ny < MAX_PARTS ? g.ny : MAX_PARTS; }

// Scratch of the allocation-free entry point (field_sums_planes): the first
// pass's partial sums, at most MAX_PARTS workgroups x 4, then the 4 results
// of the second pass.  64 KiB per GPU, held until the process ends; every
// call on a GPU shares it.  The first pass writes every partial the second
// reads, so nothing needs clearing between calls.
constexpr int kScratchSums = MAX_PARTS * 4 + 4;
__device__ __align__(256) CSum scratch[kScratchSums];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
CSum* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (each GPU holds its own
// copy of the array), looked up once per device.  The array's name is only a
// host-side stand-in: host code passes this address, never the name.
hipError_t scratch_of_device(int dev, CSum** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<CSum*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

}  // namespace

int max_scratch_devices() { return kMaxDevices; }

hipError_t field_sums_planes(const double* a, const double* b, const double* c3, const Geom& g,
                             hipStream_t stream, double out_host[3]) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  CSum* part = nullptr;
  if ((e = scratch_of_device(dev, &part)) != hipSuccess) return e;
  double* res = reinterpret_cast<double*>(part + MAX_PARTS * 4);
  const int np = nparts(g);
  hipLaunchKernelGGL(k_sum3, dim3(np), dim3(RB), 0, stream, a, b, c3, g, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, stream, part, np, 3, res);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(out_host, res, 3 * sizeof(double), hipMemcpyDeviceToHost, stream)) !=
      hipSuccess)
    return e;
  return hipStreamSynchronize(stream);
}

int reduce_partial_count(const bgk_handle*) { return MAX_PARTS * 4 * 2; }

hipError_t launch_field_sums(bgk_handle* h, double* out_dev) {
  const int np = nparts(h->geom);
  CSum* part = reinterpret_cast<CSum*>(h->partial);
  hipLaunchKernelGGL(k_sum3, dim3(np), dim3(RB), 0, h->stream, h->rho, h->T, h->phi[h->pc],
                     h->geom, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, h->stream, part, np, 3, out_dev);
  return hipGetLastError();
}

hipError_t launch_diag_sums(bgk_handle* h, double* out_dev) {
  const int np = nparts(h->geom);
  CSum* part = reinterpret_cast<CSum*>(h->partial);
  hipLaunchKernelGGL(k_diag, dim3(np), dim3(RB), 0, h->stream, h->T, h->phi[h->pc], h->geom,
                     h->c.tin, h->c.delta_t, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, h->stream, part, np, 4, out_dev);
  return hipGetLastError();
}

hipError_t launch_gather(bgk_handle* h, const double* plane, int n, const int32_t* ij_dev,
                         double* out_dev) {
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     plane, h->geom.pitch, n, ij_dev, out_dev);
  return hipGetLastError();
}

}  // namespace bgk
