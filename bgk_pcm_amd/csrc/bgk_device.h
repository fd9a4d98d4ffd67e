// bgk_device.h -- device-side building blocks of the BGK_PCM step.
//
// The arithmetic here reproduces the reference STEP_7 expressions operand by
// operand (Fortran evaluates a*b*c as (a*b)*c and a+b-c as (a+b)-c, and so
// does C++), so that translation units compiled with -ffp-contract=off give
// the reference's FP64 results bit for bit.  The FP32-storage kernel reuses
// the lattice constants only.
#ifndef BGK_DEVICE_H
#define BGK_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bgk_tanh.h"

namespace bgk {

// D2Q9 velocities, STEP_7/initialization.f90:52-70: 0 rest, 1 E, 2 N, 3 W,
// 4 S, 5 NE, 6 NW, 7 SW, 8 SE; OPP[k] is the reversed direction.
constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Lattice constants as literals (STEP_7/setup.f90:79-99, the same doubles
// set_constants computes), so that kernels need no scalar registers for
// them: w_k, cs2 and coef_k = (0.5 w_k)/(cs2 cs2).
constexpr double kCs2 = 1.0 / 3.0;
constexpr double kW[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr double coef_of(int k) { return (0.5 * kW[k]) / (kCs2 * kCs2); }

// Lattice geometry of a handle: plane pointers address node (0,0); node
// (i,j) is at [j*pitch + i].
struct Geom {
  int nx, ny;
  int64_t pitch;
};

// Populations are stored as (f_k, g_k) pairs: plane k of a buffer holds,
// for node q, f_k at element 2q and g_k at element 2q+1.  Both are pulled
// from the same neighbour, so one 2-wide load serves both, and a step
// streams 9 population planes in and 9 out instead of 18 and 18.
// Strided2 is the view of one member of the pair (base = plane for f,
// plane + 1 for g).
template <typename T>
struct Strided2 {
  T* base;
  __host__ __device__ __forceinline__ T& operator[](int64_t q) const { return base[2 * q]; }
};

struct Pops64 {
  Strided2<double> p[9];
};
struct Pops32 {
  Strided2<float> p[9];
};
// Streaming loads and stores of population pairs.  Every pair is read once
// and written once per step, far more than the caches hold, so they go
// through the non-temporal path (measured: +6..9% step throughput).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 stream_load(const float2* p) {
  const f32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 stream_load(const double2* p) {
  const f64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f64x2_t*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void stream_store(float2* p, float x, float y) {
  f32x2_t v;
  v.x = x;
  v.y = y;
  __builtin_nontemporal_store(v, reinterpret_cast<f32x2_t*>(p));
}
__device__ __forceinline__ void stream_store(double2* p, double x, double y) {
  f64x2_t v;
  v.x = x;
  v.y = y;
  __builtin_nontemporal_store(v, reinterpret_cast<f64x2_t*>(p));
}
__device__ __forceinline__ void stream_store(double* p, double x) {
  __builtin_nontemporal_store(x, p);
}

// Pair views of the same planes (used by the fused kernels).
struct Pairs64 {
  double2* p[9];
};
struct Pairs32 {
  float2* p[9];
};

// Scalars of module `shared` the step reads (STEP_7/setup.f90:79-105,
// initialization.f90:5-12, forces.f90:12-13, phase_field.f90:15-19).
struct Consts {
  double w[9];
  double coef[9];  // (0.5*w_k)/(cs2*cs2), collision.f90:94-130
  double cs2;
  double omega_f, omega_g;
  double rho0, rho1;
  double t0, tin, t_in_sol;
  double u0, v0;
  double delta_t;   // T0 - T_in_sol
  double dt_st;     // delta_T / stefan_number
  double beta;      // (1 - rho1/rho0)/delta_T, FP32-rounded under STEP7 rounding
  double grav;
  double t_width;
  double g_ref;     // FP32 storage: g is stored as g_k - w_k*g_ref (g_ref = 273)
  // f6(Nx+1,0) and f7(Nx+1,Ny+1): ghosts no BC_f branch writes, frozen at
  // their initial equilibria (moment storage keeps them as constants)
  double frozen6, frozen7;
  double inv_omega_f;  // 1/omega_f, for kernels that need not round as STEP_7
  double two_over_tw;  // 2/T_width, same
  // extensions (bgk_ext; STEP_7: t_liquid0 = T0, k_ratio = 1)
  double t_liquid0;    // initial temperature of the liquid strip
  double k_ratio;      // solid/liquid conductivity ratio
  double alpha_l;      // liquid thermal diffusivity cs2 (tau_g - 1/2)
  double inv_stefan;   // 1/stefan_number (adjoint)
  int east_cold;       // east wall at t_cold instead of the averaged T(Nx, j)
  double t_cold;
  double noise_amp;    // initial T noise amplitude in kelvin (0: none)
  uint64_t noise_seed;
  double crit_shift;   // 273 - t_crit of the phase field (0: STEP_7)
  double t_crit;       // critical temperature of the phase field (273: STEP_7)
  double t_ref;        // buoyancy reference temperature (0: STEP_7's absolute T)
};

// Thermal relaxation rate of a node whose phase field (after phase_field)
// is phi: c.omega_g, or with the conductivity-ratio extension alpha(phi) =
// alpha_l (fl + k_ratio (1 - fl)), fl = (1 + phi)/2, tau_g = 1/2 +
// alpha/cs2, in the oracle's order of operations so that the FP64 kernels
// stay bit for bit equal to it.
__device__ __forceinline__ double omega_g_at(const Consts& c, double phi) {
  if (c.k_ratio == 1.0) return c.omega_g;
  const double fl = 0.5 * (1.0 + phi);
  const double alpha = c.alpha_l * (fl + c.k_ratio * (1.0 - fl));
  return 1.0 / (0.5 + alpha / kCs2);
}

// Uniform [0,1) number of node q for the initial noise: SplitMix64 output
// for the state seed + (q+1)*0x9E3779B97F4A7C15 (Steele, Lea and Flood's
// generator with Vigna's constants), so that every node's value is
// independent of the traversal order and equal on host and device.
__host__ __device__ inline double node_uniform(uint64_t seed, int64_t q) {
  uint64_t z = seed + (uint64_t)(q + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Buoyancy of forces.f90:20, rho*grav*beta*T*(1 + phi/2), with T - t_ref
// in place of T (t_ref = 0 subtracts an exact zero: STEP_7's bits).
__device__ __forceinline__ double buoyancy(const Consts& c, double rho, double T, double phi) {
  return rho * c.grav * c.beta * (T - c.t_ref) * (1.0 + phi / 2.0);
}

// Temperature of the east ghost column for row j (BC_g.f90:21-31: the
// averaged T(Nx, j) of the previous step, or the cold-wall extension).
__device__ __forceinline__ double east_T(const Consts& c, const double* te, int j) {
  return c.east_cold ? c.t_cold : te[j];
}

// XCD-aware tiling of the fused kernels.  The lattice interior is cut into
// tiles of one row j and TILE_W consecutive columns.  gfx950 dispatches
// workgroup b to XCD b % 8, and each XCD has its own L2; a pull reads one
// cache line past each end of its row segment, which the neighbouring
// segment of the same row also reads.  So all segments of a row go to one
// XCD (workgroup b serves row (b/8 / nseg)*8 + b%8 + 1, segment b/8 % nseg):
// the shared lines are L2 hits, and at any moment the eight XCDs stream
// eight neighbouring rows.  Workgroups past the last row return at once.
constexpr int NXCD = 8;
constexpr int TILE_W = 256;

struct TileGrid {
  int nseg;        // tiles per row
  int64_t ntiles;  // nseg * ny
  int64_t per;     // workgroups per XCD: ceil(ny/8) * nseg
};

inline TileGrid make_tile_grid(const Geom& g) {
  TileGrid t;
  t.nseg = (g.nx + TILE_W - 1) / TILE_W;
  t.ntiles = (int64_t)t.nseg * g.ny;
  t.per = (int64_t)((g.ny + NXCD - 1) / NXCD) * t.nseg;
  return t;
}

// Node of this thread; false for threads of padding workgroups or past Nx.
__device__ __forceinline__ bool tile_node(const TileGrid& t, const Geom& g, int* i, int* j) {
  const int64_t b = blockIdx.x;
  const int64_t k = b / NXCD;
  *j = (int)((k / t.nseg) * NXCD + (b % NXCD)) + 1;
  if (*j > g.ny) return false;
  *i = (int)(k % t.nseg) * TILE_W + (int)threadIdx.x + 1;
  return *i <= g.nx;
}

// Second-order equilibria of the nine directions for moment m (rho or T),
// STEP_7/collision.f90:17-41.
__device__ __forceinline__ void equilibria(const Consts& c, double m, double u, double v,
                                           double eq[9]) {
  const double uu = u * u;
  const double vv = v * v;
  const double uv = 1.5 * (uu + vv);
  const double upv = u + v;
  const double umv = u - v;
  eq[0] = c.w[0] * m * (1.0 - uv);
  eq[1] = c.w[1] * m * (1.0 + 3.0 * u + 4.5 * uu - uv);
  eq[2] = c.w[2] * m * (1.0 + 3.0 * v + 4.5 * vv - uv);
  eq[3] = c.w[3] * m * (1.0 - 3.0 * u + 4.5 * uu - uv);
  eq[4] = c.w[4] * m * (1.0 - 3.0 * v + 4.5 * vv - uv);
  eq[5] = c.w[5] * m * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
  eq[6] = c.w[6] * m * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
  eq[7] = c.w[7] * m * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
  eq[8] = c.w[8] * m * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
}

// One equilibrium direction, same operand order as equilibria().
__device__ __forceinline__ double equilibrium_k(const Consts& c, int k, double m, double u,
                                                double v) {
  const double uu = u * u;
  const double vv = v * v;
  const double uv = 1.5 * (uu + vv);
  const double upv = u + v;
  const double umv = u - v;
  switch (k) {
    case 0: return c.w[0] * m * (1.0 - uv);
    case 1: return c.w[1] * m * (1.0 + 3.0 * u + 4.5 * uu - uv);
    case 2: return c.w[2] * m * (1.0 + 3.0 * v + 4.5 * vv - uv);
    case 3: return c.w[3] * m * (1.0 - 3.0 * u + 4.5 * uu - uv);
    case 4: return c.w[4] * m * (1.0 - 3.0 * v + 4.5 * vv - uv);
    case 5: return c.w[5] * m * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
    case 6: return c.w[6] * m * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
    case 7: return c.w[7] * m * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
    default: return c.w[8] * m * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
  }
}

template <bool STEP7>
__device__ __forceinline__ double round_tmp(double x) {
  // STEP7: the value is stored into an implicitly typed default-REAL
  // variable (ft0..ft8, gt0..gt8), i.e. rounded to FP32 to nearest.
  if constexpr (STEP7) return (double)(float)x;
  return x;
}

// Regularised BGK collision of one population set, STEP_7/collision.f90:44-150.
// m = rho (flow) or T (temperature); pop = fp or gp; out = f or g.
template <bool STEP7>
__device__ __forceinline__ void regularised_bgk(const Consts& c, double omega, double m,
                                                double u, double v, const double pop[9],
                                                double out[9]) {
  double eq[9];
  equilibria(c, m, u, v, eq);
  const double a = -c.cs2;
  const double b = 1.0 - c.cs2;
  double d[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) d[k] = pop[k] - eq[k];
  const double pxx = a * d[0] + b * d[1] + a * d[2] + b * d[3] + a * d[4] + b * d[5] +
                     b * d[6] + b * d[7] + b * d[8];
  const double pyy = a * d[0] + a * d[1] + b * d[2] + a * d[3] + b * d[4] + b * d[5] +
                     b * d[6] + b * d[7] + b * d[8];
  const double pxy = d[5] - d[6] + d[7] - d[8];
  double t[9];
  t[0] = round_tmp<STEP7>(eq[0] + c.coef[0] * (a * pxx + a * pyy));
  t[1] = round_tmp<STEP7>(eq[1] + c.coef[1] * (b * pxx + a * pyy));
  t[2] = round_tmp<STEP7>(eq[2] + c.coef[2] * (a * pxx + b * pyy));
  t[3] = round_tmp<STEP7>(eq[3] + c.coef[3] * (b * pxx + a * pyy));
  t[4] = round_tmp<STEP7>(eq[4] + c.coef[4] * (a * pxx + b * pyy));
  t[5] = round_tmp<STEP7>(eq[5] + c.coef[5] * (b * pxx + b * pyy + 2.0 * pxy));
  t[6] = round_tmp<STEP7>(eq[6] + c.coef[6] * (b * pxx + b * pyy - 2.0 * pxy));
  t[7] = round_tmp<STEP7>(eq[7] + c.coef[7] * (b * pxx + b * pyy + 2.0 * pxy));
  t[8] = round_tmp<STEP7>(eq[8] + c.coef[8] * (b * pxx + b * pyy - 2.0 * pxy));
#pragma unroll
  for (int k = 0; k < 9; ++k) out[k] = t[k] - omega * (t[k] - eq[k]);
}

// Initial macroscopic state of node (i,j), halo included, in the order
// STEP_7/initialization.f90:15-152 assigns it (later assignments win).
struct NodeInit {
  double T, rho, u, v, phi;
};

__device__ __forceinline__ NodeInit initial_node(const Consts& c, const Geom& gm, int i, int j,
                                                 int interf, int west_bc) {
  NodeInit s;
  if (i <= interf) {
    s.T = c.t_liquid0;
    s.rho = c.rho1;
  } else {
    s.T = c.t_in_sol;
    s.rho = c.rho0;
  }
  s.u = c.u0;
  s.v = c.v0;
  if (i == 0) s.T = c.tin;
  // extension: seeded noise on the interior temperature (0: STEP_7)
  if (c.noise_amp > 0.0 && i >= 1 && i <= gm.nx && j >= 1 && j <= gm.ny)
    s.T += c.noise_amp * (2.0 * node_uniform(c.noise_seed, (int64_t)j * (gm.nx + 2) + i) - 1.0);
  if (j == 0 || j == gm.ny + 1) {
    s.u = 0.0;
    s.v = 0.0;
  }
  if (west_bc == 3) {
    s.u = 0.0;
    s.v = 0.0;
    if (i == 0) s.rho = c.rho1;
  }
  if (i == 0 && j >= 1 && j <= gm.ny) {
    if (west_bc == 1) {
      s.u = c.u0;
      s.v = c.v0;
    } else if (west_bc == 2) {
      const double yj = (double)j;
      const double np1 = (double)(gm.ny + 1);
      s.u = -(6.0 * c.u0 / (np1 * np1)) * (yj * yj - np1 * yj);
      s.v = c.v0;
    }
  }
  s.phi = (i <= interf) ? 1.0 : -1.0;
  return s;
}

// Phase-field update of one node, STEP_7/phase_field.f90:24-52.  tap[] are
// phi_prec at C, E, N, W, S, NE, NW, SW, SE; returns the new phi.
__device__ __forceinline__ double phase_update(const Consts& c, double T, const double tap[9]) {
  double a = 0.5 * (1.0 + tap[0]);
  double kp = a * a;
  double b = 0.5 * (1.0 - tap[0]);
  double km = b * b;
#pragma unroll
  for (int m = 1; m < 9; ++m) {
    a = 0.5 * (1.0 + tap[m]);
    kp = kp + a * a;
  }
#pragma unroll
  for (int m = 1; m < 9; ++m) {
    b = 0.5 * (1.0 - tap[m]);
    km = km + b * b;
  }
  const double pc = tap[0];
  // Inside one phase (all nine taps -1, or all +1) each reaction term is
  // multiplied by an exact zero (kp = 0 or 1 + pc = 0, km = 0 or 1 - pc = 0)
  // and the update returns pc bit for bit; the tanh is needed only near the
  // front.  (With a non-finite T the reference would give NaN here too; T
  // itself carries the NaN either way.)
  if ((kp == 0.0 || 1.0 - pc == 0.0) && (km == 0.0 || 1.0 + pc == 0.0)) return pc;
  // T_act = 0 and T_crit = 273 (phase_field.f90:15-16; c.t_crit is 273
  // unless the extension moves it): both reaction rates take tanh of the
  // same argument.
  const double th = tanh_fdlibm(((T - c.t_crit) - 0.0) / c.t_width);
  const double react_plus = 0.5 * (1.0 + th);
  const double react_minus = 0.5 * (1.0 - th);
  kp = (kp / 8.0) * react_plus;
  km = (km / 8.0) * react_minus;
  return pc + (0.5 * kp * (1.0 - pc) - 0.5 * km * (1.0 + pc));
}

}  // namespace bgk

#endif  // BGK_DEVICE_H
