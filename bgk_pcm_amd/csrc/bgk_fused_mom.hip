// bgk_fused_mom.hip -- the fused time step with FP32 moment storage.
//
// After the regularised collision (STEP_7/collision.f90:44-150, clean
// rounding) every population of a node is a function of ten numbers:
//   f_k = w_k rho (1 + P_k(u,v)) + coef_k H_k : Pf,   Pf = (1 - omega_f) Pi_f^neq
//   g_k = w_k T   (1 + P_k(u,v)) + coef_k H_k : Pg,   Pg = (1 - omega_g) Pi_g^neq
// (P_k the equilibrium polynomial minus its leading 1, H_k = c_k c_k - cs2 I,
// coef_k = w_k / (2 cs2^2)).  So a node stores those ten numbers instead of
// eighteen populations, and the pull rebuilds population k of the upstream
// node from its moments.  Layout per node, 48 bytes in three 16-byte planes:
//   A double2 (rho, T - T_crit)           FP64
//   B float4  (u, v, Pf_xx, Pf_yy)        FP32
//   C float4  (Pf_xy, Pg_xx, Pg_yy, Pg_xy) FP32
// rho and T stay FP64: the rounding of rho (not of u) is what drives the
// velocity error of FP32 storage, since a neighbour's density error enters
// the momentum balance through sum_k c_k w_k rho(x - c_k).  u, v and the
// non-equilibrium moments carry their own relative precision in FP32.
//
// Tiling: a workgroup owns W columns and a strip of consecutive rows, and
// walks the strip row by row.  Row j+2 is fetched from HBM while row j is
// computed; rows j-1..j+2 live in a four-row LDS ring, so every node's
// moments and phi are read from HBM once per step (plus the strip's two
// boundary rows and the two halo columns).  phi_prec's 9-point stencil comes
// from the same ring.  BC_f (no-slip) and BC_g are folded in as in
// bgk_fused.hip: ghost rows of g and phi are the clamped rows, own opposite
// populations for bounce-back, the two frozen corner ghosts as constants.
//
// Arithmetic is FP64 throughout.  STEP_7's FP32 temporaries (ft, gt) cannot
// be represented by moments, so this path follows the clean rounding policy;
// beta follows bgk_params.rounding as everywhere.
#include <algorithm>

#include "bgk_internal.h"

namespace bgk {
namespace {

constexpr int TX = 64, TY = 4;  // block of the initialization kernel

// Lattice constants as literals (STEP_7/setup.f90:79-99), so that they
// need no scalar registers: w_k, cs2 and coef_k = (0.5 w_k)/(cs2 cs2).
constexpr double kCs2 = 1.0 / 3.0;
constexpr double kW[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr double coef_of(int k) { return (0.5 * kW[k]) / (kCs2 * kCs2); }
constexpr double kTcrit = 273.0;  // T_crit, STEP_7/phase_field.f90:16

typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 stream_load4(const float4* p) {
  const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stream_store4(float4* p, float x, float y, float z, float w) {
  f32x4_t v;
  v.x = x;
  v.y = y;
  v.z = z;
  v.w = w;
  __builtin_nontemporal_store(v, reinterpret_cast<f32x4_t*>(p));
}

struct Macro {
  double *rho, *u, *v, *T;
};

// Element of the third plane.  With omega_f = 1 (tau_f = 1, the reference
// inputs) the regularised flow collision returns the equilibrium, Pf = 0,
// and a node needs 36 bytes: B = (u, v, Pg_xx, Pg_yy), C = Pg_xy.
template <bool EQF>
struct CPlane {
  using T = float4;  // Pf_xy, Pg_xx, Pg_yy, Pg_xy
};
template <>
struct CPlane<true> {
  using T = float;  // Pg_xy
};

// Moments of one node as held in the ring.
template <bool EQF>
struct NodeM {
  double2 a;  // rho, T - T_crit
  float4 b;   // u, v, Pf_xx, Pf_yy  (EQF: u, v, Pg_xx, Pg_yy)
  typename CPlane<EQF>::T c;
  double phi;
};

// LDS ring of NS rows of W + 2 nodes (W columns and the two halo columns).
template <int W, int NS, bool EQF>
struct Ring {
  double2 a[NS][W + 2];
  float4 b[NS][W + 2];
  typename CPlane<EQF>::T c[NS][W + 2];
  double phi[NS][W + 2];
};

// P_k(u,v): the equilibrium polynomial of direction k minus its leading 1.
__device__ __forceinline__ double eq_poly(int k, double u, double v) {
  const double uv = 1.5 * (u * u + v * v);
  switch (k) {
    case 0: return -uv;
    case 1: return 3.0 * u + 4.5 * (u * u) - uv;
    case 2: return 3.0 * v + 4.5 * (v * v) - uv;
    case 3: return -3.0 * u + 4.5 * (u * u) - uv;
    case 4: return -3.0 * v + 4.5 * (v * v) - uv;
    case 5: { const double s = u + v; return 3.0 * s + 4.5 * (s * s) - uv; }
    case 6: { const double s = u - v; return -3.0 * s + 4.5 * (s * s) - uv; }
    case 7: { const double s = u + v; return -3.0 * s + 4.5 * (s * s) - uv; }
    default: { const double s = u - v; return 3.0 * s + 4.5 * (s * s) - uv; }
  }
}

// H_k : P for P = (pxx, pyy, pxy), H_k = c_k c_k - cs2 I, evaluated in FP32: its operands are FP32 already and the result
// keeps their relative precision.
__device__ __forceinline__ float h_k32(int k, float pxx, float pyy, float pxy) {
  constexpr float A = -1.0f / 3.0f, B = 2.0f / 3.0f;
  switch (k) {
    case 0: return A * (pxx + pyy);
    case 1:
    case 3: return B * pxx + A * pyy;
    case 2:
    case 4: return A * pxx + B * pyy;
    case 5:
    case 7: return B * (pxx + pyy) + 2.0f * pxy;
    default: return B * (pxx + pyy) - 2.0f * pxy;
  }
}

// Population k of a node from its moments: f_k and g_k - w_k*T_crit.
__device__ __forceinline__ void rebuild(int k, const double2 a, const float4 b, const float4 cc,
                                        double* fk, double* gdk) {
  const double p = eq_poly(k, (double)b.x, (double)b.y);
  const double hf = (double)h_k32(k, b.z, b.w, cc.x);
  const double hg = (double)h_k32(k, cc.y, cc.z, cc.w);
  *fk = kW[k] * a.x * (1.0 + p) + coef_of(k) * hf;
  *gdk = kW[k] * (a.y + (kTcrit + a.y) * p) + coef_of(k) * hg;
}
__device__ __forceinline__ void rebuild(int k, const double2 a, const float4 b, const float cc,
                                        double* fk, double* gdk) {
  const double p = eq_poly(k, (double)b.x, (double)b.y);
  const double hg = (double)h_k32(k, b.z, b.w, cc);
  *fk = kW[k] * a.x * (1.0 + p);
  *gdk = kW[k] * (a.y + (kTcrit + a.y) * p) + coef_of(k) * hg;
}

__device__ __forceinline__ double rebuild_f(int k, const double2 a, const float4 b,
                                            const float4 cc) {
  const double p = eq_poly(k, (double)b.x, (double)b.y);
  const double hf = (double)h_k32(k, b.z, b.w, cc.x);
  return kW[k] * a.x * (1.0 + p) + coef_of(k) * hf;
}
__device__ __forceinline__ double rebuild_f(int k, const double2 a, const float4 b, float) {
  return kW[k] * a.x * (1.0 + eq_poly(k, (double)b.x, (double)b.y));
}

// Equilibrium deviation of a boundary ghost, g_k - w_k*T_crit.
__device__ __forceinline__ double ghost_gdev(int k, double T, double u, double v) {
  return kW[k] * (T - kTcrit) + kW[k] * T * eq_poly(k, u, v);
}

struct MomIn {
  const double2* a;
  const float4* b;
  const void* c;  // float4 or float (CPlane)
};
struct MomOut {
  double2* a;
  float4* b;
  void* c;
};

template <bool EQF>
__device__ __forceinline__ typename CPlane<EQF>::T c_at(const void* c, int64_t q) {
  return static_cast<const typename CPlane<EQF>::T*>(c)[q];
}
template <bool EQF>
__device__ __forceinline__ typename CPlane<EQF>::T c_stream(const void* c, int64_t q) {
  if constexpr (EQF)
    return __builtin_nontemporal_load(static_cast<const float*>(c) + q);
  else
    return stream_load4(static_cast<const float4*>(c) + q);
}

// Row strips: workgroup b runs on XCD b % 8; all column segments of a strip
// go to one XCD (its halo columns are L2 hits) and the eight XCDs work on
// eight neighbouring strips.
struct StripGrid {
  int nseg;    // column segments of W
  int rows;    // rows per strip
  int nstrip;  // ceil(ny / rows)
  int nmain;   // workgroups of the strips (padded to a multiple of 8 strips)
};

template <bool EQF>
__device__ __forceinline__ NodeM<EQF> load_node(const MomIn& in, const double* __restrict__ pa,
                                                const Geom& gm, int ic, int jr) {
  NodeM<EQF> n;
  n.a = make_double2(0.0, 0.0);
  n.b = make_float4(0.f, 0.f, 0.f, 0.f);
  n.c = typename CPlane<EQF>::T{};
  n.phi = 0.0;
  const int64_t q = (int64_t)jr * gm.pitch + ic;
  if (ic >= 1 && ic <= gm.nx) {
    n.a = stream_load(in.a + q);
    n.b = stream_load4(in.b + q);
    n.c = c_stream<EQF>(in.c, q);
  }
  if (ic >= 0 && ic <= gm.nx + 1) n.phi = pa[q];
  return n;
}

template <int W, int NS, bool EQF>
__device__ __forceinline__ void put(Ring<W, NS, EQF>& L, int r, int s, const NodeM<EQF>& n) {
  L.a[r][s] = n.a;
  L.b[r][s] = n.b;
  L.c[r][s] = n.c;
  L.phi[r][s] = n.phi;
}

// Phase-field update of STEP_7/phase_field.f90:24-52 with the reaction rates
// in closed form: 0.5 (1 -/+ tanh x) = 1 / (1 + exp(+/-2x)), x = (T - 273) /
// T_width.  One exp and one division instead of the branchy glibc-exact tanh
// of the FP64 kernels (which this path, rounding as clean FP64, does not
// need to reproduce bit for bit); the absolute error of the rates is a few
// ulp of 1, like that of 0.5 (1 -/+ tanh x).  Evaluated at every node
// without a branch: inside one phase the reaction terms are multiplied by
// exact zeros anyway, and a branch here splits the node's code and costs
// ~70 registers.
__device__ __forceinline__ double phase_update_exp(const Consts& c, double T,
                                                   const double tap[9]) {
  double a = 0.5 * (1.0 + tap[0]);
  double kp = a * a;
  double b = 0.5 * (1.0 - tap[0]);
  double km = b * b;
#pragma unroll
  for (int m = 1; m < 9; ++m) {
    a = 0.5 * (1.0 + tap[m]);
    kp = kp + a * a;
    b = 0.5 * (1.0 - tap[m]);
    km = km + b * b;
  }
  const double pc = tap[0];
  const double react_minus = 1.0 / (1.0 + exp((T - kTcrit) * c.two_over_tw));
  const double react_plus = 1.0 - react_minus;
  kp = (kp * 0.125) * react_plus;
  km = (km * 0.125) * react_minus;
  return pc + (0.5 * kp * (1.0 - pc) - 0.5 * km * (1.0 + pc));
}

// Moments of the pulled populations of a node: zeroth, first and second
// moments of f, and of gd = g - w*T_crit (its first moment is not needed).
struct PullSums {
  double rho, jx, jy, fxx, fyy, fxy;
  double dT, gxx, gyy, gxy;
};

__device__ __forceinline__ void pull_zero(PullSums& S) {
  S.rho = S.jx = S.jy = S.fxx = S.fyy = S.fxy = 0.0;
  S.dT = S.gxx = S.gyy = S.gxy = 0.0;
}

// Add population k (f_k, gd_k) to the sums.
__device__ __forceinline__ void pull_add(PullSums& S, int k, double fk, double gk) {
  S.rho += fk;
  S.dT += gk;
  if (CX[k] != 0) {
    S.jx += CX[k] > 0 ? fk : -fk;
    S.fxx += fk;
    S.gxx += gk;
  }
  if (CY[k] != 0) {
    S.jy += CY[k] > 0 ? fk : -fk;
    S.fyy += fk;
    S.gyy += gk;
  }
  if (CX[k] * CY[k] > 0) {
    S.fxy += fk;
    S.gxy += gk;
  } else if (CX[k] * CY[k] < 0) {
    S.fxy -= fk;
    S.gxy -= gk;
  }
}

// momentsPre, phase field, forces and collision of node q from the moments
// of its pulled populations and its phi_prec taps (C, E, N, W, S, NE, NW,
// SW, SE); stores the node's moments and phi.  Returns what the averaged
// moments of STEP_7/moments.f90:11-25 need: rho, the averaged T (Tav), u, v.
struct NodeOut {
  double rho, Tav, uav, vav;
};

template <bool EQF>
__device__ __forceinline__ NodeOut update_node(const Consts& c, const PullSums& S,
                                               const double tap[9], const MomOut& out,
                                               double* __restrict__ pb, int64_t q) {
  // ---- momentsPre (STEP_7/momentsPre.f90:8-21)
  const double rho = S.rho;
  double dT = S.dT;
  const double irho = 1.0 / rho;
  double u = S.jx * irho;
  double v = S.jy * irho;
  const double dT2 = dT, u2 = u, v2 = v;

  // ---- phase field (STEP_7/phase_field.f90:24-52)
  const double phi = phase_update_exp(c, kTcrit + dT, tap);
  const double pc = tap[0];

  // ---- latent heat and forces (STEP_7/forces.f90:18-26)
  dT = dT - (c.dt_st * (phi - pc) * 0.5);
  const double T = kTcrit + dT;
  {
    double fy = rho * c.grav * c.beta * T * (1.0 + phi * 0.5);
    const double fx = -(u * (1.0 - phi) * 0.5);
    fy = fy - v * (1.0 - phi) * 0.5;
    const double s = irho * c.inv_omega_f;
    u = u + fx * s;
    v = v + fy * s;
  }

  // ---- collision (STEP_7/collision.f90:44-92) in moment form: the
  // non-equilibrium second moments of the pulled populations about the
  // equilibria at (rho, u, v) and (T, u, v); for the second-order D2Q9
  // equilibrium sum_k H_k feq_k = rho u u.
  const double omg1 = 1.0 - c.omega_g;
  const double pgxx = omg1 * (S.gxx - kCs2 * dT2 - T * u * u);
  const double pgyy = omg1 * (S.gyy - kCs2 * dT2 - T * v * v);
  const double pgxy = omg1 * (S.gxy - T * u * v);
  const float us = (float)u, vs = (float)v;
  stream_store(out.a + q, rho, dT);
  if constexpr (EQF) {
    stream_store4(out.b + q, us, vs, (float)pgxx, (float)pgyy);
    __builtin_nontemporal_store((float)pgxy, static_cast<float*>(out.c) + q);
  } else {
    const double omf1 = 1.0 - c.omega_f;
    const double pfxx = omf1 * (S.fxx - kCs2 * rho - rho * u * u);
    const double pfyy = omf1 * (S.fyy - kCs2 * rho - rho * v * v);
    const double pfxy = omf1 * (S.fxy - rho * u * v);
    stream_store4(out.b + q, us, vs, (float)pfxx, (float)pfyy);
    stream_store4(static_cast<float4*>(out.c) + q, (float)pfxy, (float)pgxx, (float)pgyy,
                  (float)pgxy);
  }
  stream_store(&pb[q], phi);

  // ---- moments of the stored state (STEP_7/moments.f90:11-25): sum_k f_k =
  // rho, sum_k g_k = T and sum_k c_k f_k / rho = u exactly for the rebuilt
  // populations
  NodeOut o;
  o.rho = rho;
  o.Tav = kTcrit + (dT + dT2) * 0.5;
  o.uav = ((double)us + u2) * 0.5;
  o.vav = ((double)vs + v2) * 0.5;
  return o;
}

// Interior nodes (1 < i < Nx, 1 < j < Ny) by row strips through the LDS
// ring.  W columns per workgroup, NS rows in the ring (4: one barrier per
// row; 3: two, for less LDS).  Workgroups from sg.nmain on serve the
// boundary-adjacent nodes (mom_edge_nodes).
template <int W, int NS, bool EQF>
__device__ __forceinline__ void mom_interior(const MomIn& in, const MomOut& out,
                                             const double* __restrict__ pa,
                                             double* __restrict__ pb, const Macro& m,
                                             const Consts& c, const Geom& gm, const StripGrid& sg,
                                             int write_all) {
  __shared__ Ring<W, NS, EQF> L;
  const int nx = gm.nx, ny = gm.ny;
  const int64_t P = gm.pitch;
  const int blk = (int)blockIdx.x;
  const int kq = blk / NXCD;
  const int strip = (kq / sg.nseg) * NXCD + blk % NXCD;
  if (strip >= sg.nstrip) return;  // whole workgroup
  const int seg = kq % sg.nseg;
  const int j0 = strip * sg.rows + 1;
  const int j1 = min(j0 + sg.rows - 1, ny);
  const int t = (int)threadIdx.x;
  const int i0 = seg * W + 1;  // column of slot 1
  const int i = i0 + t;        // this thread's column, slot t + 1
  // threads 0 and 1 also fetch the halo columns i0 - 1 (slot 0) and i0 + W (slot W + 1)
  const int hcol = (t == 0) ? i0 - 1 : i0 + W;
  const int hslot = (t == 0) ? 0 : W + 1;
  auto clampj = [ny](int jj) { return jj < 1 ? 1 : (jj > ny ? ny : jj); };
  auto slot = [j0](int jj) { return (jj - j0 + 1) % NS; };
  const bool col_in = (i > 1 && i < nx);

  for (int jj = j0 - 1; jj <= j0 + 1; ++jj) {
    put(L, slot(jj), t + 1, load_node<EQF>(in, pa, gm, i, clampj(jj)));
    if (t < 2) put(L, slot(jj), hslot, load_node<EQF>(in, pa, gm, hcol, clampj(jj)));
  }
  __syncthreads();

  for (int j = j0; j <= j1; ++j) {
    const bool pre = (j + 2 <= j1 + 1);
    NodeM<EQF> nn, nh;
    if (pre) {
      nn = load_node<EQF>(in, pa, gm, i, clampj(j + 2));
      if (t < 2) nh = load_node<EQF>(in, pa, gm, hcol, clampj(j + 2));
    }
    if (col_in && j > 1 && j < ny) {
      const int rS = slot(j - 1), rC = slot(j), rN = slot(j + 1);
      const int s = t + 1;
      const int64_t q = (int64_t)j * P + i;
      // ---- streaming: rebuild population k of node (i - cx_k, j - cy_k)
      PullSums S;
      pull_zero(S);
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const int r = CY[k] > 0 ? rS : (CY[k] < 0 ? rN : rC);
        const int ss = s - CX[k];
        double fk, gk;
        rebuild(k, L.a[r][ss], L.b[r][ss], L.c[r][ss], &fk, &gk);
        pull_add(S, k, fk, gk);
      }
      const double tap[9] = {L.phi[rC][s],     L.phi[rC][s + 1], L.phi[rN][s],
                             L.phi[rC][s - 1], L.phi[rS][s],     L.phi[rN][s + 1],
                             L.phi[rN][s - 1], L.phi[rS][s - 1], L.phi[rS][s + 1]};
      const NodeOut o = update_node<EQF>(c, S, tap, out, pb, q);
      if (write_all) {
        m.rho[q] = o.rho;
        m.T[q] = o.Tav;
        m.u[q] = o.uav;
        m.v[q] = o.vav;
      }
    }
    if constexpr (NS == 3) __syncthreads();  // row j+2 overwrites row j-1
    if (pre) {
      put(L, slot(j + 2), t + 1, nn);
      if (t < 2) put(L, slot(j + 2), hslot, nh);
    }
    __syncthreads();
  }
}

// Node (i, j) of the boundary-adjacent ring i in {1, Nx} or j in {1, Ny}
// for edge index e: row 1, row Ny, then columns 1 and Nx of rows 2..Ny-1.
__device__ __forceinline__ bool edge_node(int e, int nx, int ny, int* i, int* j) {
  const int nrow = ny > 1 ? 2 : 1;
  if (e < nrow * nx) {
    *j = (e < nx) ? 1 : ny;
    *i = e % nx + 1;
    return true;
  }
  e -= nrow * nx;
  const int ncol = nx > 1 ? 2 : 1;
  if (ny <= 2 || e >= ncol * (ny - 2)) return false;
  *j = e / ncol + 2;
  *i = (e % ncol == 0) ? 1 : nx;
  return true;
}

// The boundary-adjacent nodes, one per thread, with BC_f type 4 and BC_g
// (STEP_7/BC_f.f90:62-68,101-107,125-136,181-192; BC_g.f90:9-64) folded into
// the pulls: bounce-back of the node's own opposite populations, the two
// frozen corner ghosts, W/E thermal ghosts at Tin and at the previous
// step's averaged T(Nx, j), N/S ghost rows of g and phi equal to rows Ny / 1;
// then the halo bookkeeping of the macroscopic fields.
template <bool EQF>
__device__ __forceinline__ void mom_edge_nodes(const MomIn& in, const MomOut& out,
                                               const double* __restrict__ pa,
                                               double* __restrict__ pb,
                                               const double* __restrict__ te_in,
                                               double* __restrict__ te_out, const Macro& m,
                                               const Consts& c, const Geom& gm, int e) {
  const int nx = gm.nx, ny = gm.ny;
  const int64_t P = gm.pitch;
  int i, j;
  if (!edge_node(e, nx, ny, &i, &j)) return;
  const int64_t q = (int64_t)j * P + i;
  const double2 oa = in.a[q];
  const float4 ob = in.b[q];
  const typename CPlane<EQF>::T oc = c_at<EQF>(in.c, q);
  double f[9], gd[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int si = i - CX[k], sj = j - CY[k];
    const int sjc = sj < 1 ? 1 : (sj > ny ? ny : sj);
    if (si >= 1 && si <= nx) {
      // interior source column: rebuild from the node (si, sjc); for a ghost
      // row this is the N/S copy of g, and f is replaced below
      const int64_t h = (int64_t)sjc * P + si;
      rebuild(k, in.a[h], in.b[h], c_at<EQF>(in.c, h), &f[k], &gd[k]);
    } else {
      const int64_t h = (int64_t)sjc * P + si;
      gd[k] = ghost_gdev(k, si == 0 ? c.tin : te_in[sjc], m.u[h], m.v[h]);
    }
    if (si < 1 || si > nx || sj < 1 || sj > ny) {
      if (k == 6 && si == nx + 1 && sj == 0)
        f[k] = c.frozen6;
      else if (k == 7 && si == nx + 1 && sj == ny + 1)
        f[k] = c.frozen7;
      else
        f[k] = rebuild_f(OPP[k], oa, ob, oc);
    }
  }
  const int64_t rn = (int64_t)(j < ny ? j + 1 : ny) * P;
  const int64_t rs = (int64_t)(j > 1 ? j - 1 : 1) * P;
  const int64_t r0 = (int64_t)j * P;
  const double tap[9] = {pa[q],          pa[r0 + i + 1], pa[rn + i],     pa[r0 + i - 1], pa[rs + i],
                         pa[rn + i + 1], pa[rn + i - 1], pa[rs + i - 1], pa[rs + i + 1]};
  PullSums S;
  pull_zero(S);
#pragma unroll
  for (int k = 0; k < 9; ++k) pull_add(S, k, f[k], gd[k]);
  const NodeOut o = update_node<EQF>(c, S, tap, out, pb, q);

  // ---- halo bookkeeping: BC_f N/S rows (rho copied, u = v = 0), BC_g
  // T and phi ghosts, the east column's averaged T for the next step
  const bool w_edge = (i == 1), e_edge = (i == nx), s_edge = (j == 1), n_edge = (j == ny);
  if (s_edge) {
    m.rho[i] = m.rho[q];
    m.T[i] = m.T[q];
    m.u[i] = 0.0;
    m.v[i] = 0.0;
    pb[i] = pa[q];
  }
  if (n_edge) {
    const int64_t h = (int64_t)(ny + 1) * P + i;
    m.rho[h] = m.rho[q];
    m.T[h] = m.T[q];
    m.u[h] = 0.0;
    m.v[h] = 0.0;
    pb[h] = pa[q];
  }
  if (w_edge) {
    m.T[r0] = c.tin;
    pb[r0] = pa[r0];
    if (s_edge) {
      m.T[0] = c.tin;
      pb[0] = pa[P];
    }
    if (n_edge) {
      const int64_t h = (int64_t)(ny + 1) * P;
      m.T[h] = c.tin;
      pb[h] = pa[(int64_t)ny * P];
    }
  }
  if (e_edge) {
    const int64_t h = r0 + nx + 1;
    m.T[h] = te_in[j];
    pb[h] = pa[h];
    te_out[j] = o.Tav;
    if (s_edge) {
      m.T[nx + 1] = te_in[1];
      pb[nx + 1] = pa[P + nx + 1];
    }
    if (n_edge) {
      const int64_t hn = (int64_t)(ny + 1) * P + nx + 1;
      m.T[hn] = te_in[ny];
      pb[hn] = pa[(int64_t)ny * P + nx + 1];
    }
  }
  m.rho[q] = o.rho;
  m.T[q] = o.Tav;
  m.u[q] = o.uav;
  m.v[q] = o.vav;
}

template <int W, int NS, bool EQF>
__global__ __launch_bounds__(W) void k_fused_mom(
    MomIn in, MomOut out, const double* __restrict__ pa, double* __restrict__ pb,
    const double* __restrict__ te_in, double* __restrict__ te_out, Macro m, Consts c, Geom gm,
    StripGrid sg, int write_all) {
  if ((int)blockIdx.x >= sg.nmain) {
    const int e = ((int)blockIdx.x - sg.nmain) * W + (int)threadIdx.x;
    mom_edge_nodes<EQF>(in, out, pa, pb, te_in, te_out, m, c, gm, e);
    return;
  }
  mom_interior<W, NS, EQF>(in, out, pa, pb, m, c, gm, sg, write_all);
}

// STEP_7/initialization.f90:15-152 for a moment handle: the initial
// populations are equilibria, i.e. moments with zero non-equilibrium part.
template <bool EQF>
__global__ void k_initialization_mom(MomOut st, Macro m, double* phi, double* phip, Consts c,
                                     Geom gm, int interf, int west_bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const NodeInit s0 = initial_node(c, gm, i, j, interf, west_bc);
  st.a[q] = make_double2(s0.rho, s0.T - kTcrit);
  st.b[q] = make_float4((float)s0.u, (float)s0.v, 0.f, 0.f);
  static_cast<typename CPlane<EQF>::T*>(st.c)[q] = typename CPlane<EQF>::T{};
  m.rho[q] = s0.rho;
  m.u[q] = s0.u;
  m.v[q] = s0.v;
  m.T[q] = s0.T;
  phi[q] = s0.phi;
  phip[q] = s0.phi;
}

// Row strips of about `target` rows.  The strip count is chosen, where one
// exists within a factor two, as a multiple of 8 (the XCDs) whose workgroups
// fill whole rounds of the device (slots = CUs x resident workgroups per
// CU): measured at 4096^2, 64-row strips (512 workgroups) run at 57 GLUPS
// and 96-row strips (43 strips, 1.5 rounds) at 43.
StripGrid make_strip_grid(const bgk_handle* h, int w, int target, int slots) {
  StripGrid s;
  const int ny = h->ny;
  s.nseg = (h->nx + w - 1) / w;
  if (h->strip_rows > 0) {
    s.rows = h->strip_rows;
  } else {
    const int base = std::max(1, (ny + target - 1) / target);
    int pick = base;
    for (int n = base; n <= 2 * base && n <= ny; ++n)
      if (n % NXCD == 0 && slots > 0 && ((int64_t)n * s.nseg) % slots == 0) {
        pick = n;
        break;
      }
    s.rows = (ny + pick - 1) / pick;
  }
  s.nstrip = (ny + s.rows - 1) / s.rows;
  s.nmain = ((s.nstrip + NXCD - 1) / NXCD) * NXCD * s.nseg;
  return s;
}

template <int W, int NS, bool EQF>
void launch_mom_t(bgk_handle* h, MomIn in, MomOut out, const double* te_in, double* te_out,
                  bool write_all) {
  // resident workgroups per CU of this instantiation (LDS and registers)
  static const int per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_mom<W, NS, EQF>, W, 0) !=
        hipSuccess)
      n = 0;
    return n;
  }();
  const StripGrid sg = make_strip_grid(h, W, EQF ? 64 : 32, per_cu * h->n_cu);
  const int nx = h->nx, ny = h->ny;
  const int nedge = nx * (ny > 1 ? 2 : 1) + (ny > 2 ? (ny - 2) * (nx > 1 ? 2 : 1) : 0);
  const int64_t nwg = (int64_t)sg.nmain + (nedge + W - 1) / W;
  Macro m{h->rho, h->u, h->v, h->T};
  hipLaunchKernelGGL((k_fused_mom<W, NS, EQF>), dim3((unsigned)nwg), dim3(W), 0, h->stream, in,
                     out, h->phi[h->pc], h->phi[1 - h->pc], te_in, te_out, m, h->c, h->geom, sg,
                     write_all ? 1 : 0);
}

template <int W, int NS>
void launch_mom(bgk_handle* h, MomIn in, MomOut out, const double* te_in, double* te_out,
                bool write_all) {
  if (h->mom_eqf)
    launch_mom_t<W, NS, true>(h, in, out, te_in, te_out, write_all);
  else
    launch_mom_t<W, NS, false>(h, in, out, te_in, te_out, write_all);
}

}  // namespace

hipError_t launch_fused_step_mom(bgk_handle* h, int step_parity, bool write_all) {
  const double* te_in = h->teast[step_parity & 1];
  double* te_out = h->teast[(step_parity + 1) & 1];
  const int a = h->cur, b = 1 - h->cur;
  MomIn in{h->ma[a], h->mb[a], h->mc[a]};
  MomOut out{h->ma[b], h->mb[b], h->mc[b]};
  // Tiling (workgroup width W, ring rows NS).  Measured at 4096^2 on MI355X:
  // omega_f = 1 (36-byte nodes, 64-row strips) W = 512 / NS = 4 57.2 GLUPS,
  // 512/3 53.1, 256/4 53.9 (32-row strips), 128/4 45.1, 64/4 45.8; general
  // layout (48 bytes, 32-row strips) 512/3 45.9, 512/4 45.6, 256/3 42.0.
  // Wider segments stream longer contiguous runs of each row.
  int v = h->mom_variant;
  if (v == 0)
    v = h->nx >= 2048 ? (h->mom_eqf ? 7 : 8) : h->nx >= 512 ? 5 : h->nx >= 128 ? 3 : 1;
  switch (v) {
    case 1: launch_mom<64, 4>(h, in, out, te_in, te_out, write_all); break;
    case 2: launch_mom<64, 3>(h, in, out, te_in, te_out, write_all); break;
    case 3: launch_mom<128, 4>(h, in, out, te_in, te_out, write_all); break;
    case 4: launch_mom<128, 3>(h, in, out, te_in, te_out, write_all); break;
    case 5: launch_mom<256, 4>(h, in, out, te_in, te_out, write_all); break;
    case 6: launch_mom<256, 3>(h, in, out, te_in, te_out, write_all); break;
    case 7: launch_mom<512, 4>(h, in, out, te_in, te_out, write_all); break;
    default: launch_mom<512, 3>(h, in, out, te_in, te_out, write_all); break;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    h->cur = b;
    h->pc = 1 - h->pc;
  }
  return e;
}

hipError_t launch_initialization_mom(bgk_handle* h) {
  Macro m{h->rho, h->u, h->v, h->T};
  const Geom& g = h->geom;
  MomOut st{h->ma[h->cur], h->mb[h->cur], h->mc[h->cur]};
  dim3 grd((unsigned)((g.nx + 2 + TX - 1) / TX), (unsigned)((g.ny + 2 + TY - 1) / TY));
  if (h->mom_eqf)
    hipLaunchKernelGGL(k_initialization_mom<true>, grd, dim3(TX, TY), 0, h->stream, st, m,
                       h->phi[h->pc], h->phi[1 - h->pc], h->c, g, g.nx / 40, h->prm.west_bc);
  else
    hipLaunchKernelGGL(k_initialization_mom<false>, grd, dim3(TX, TY), 0, h->stream, st, m,
                       h->phi[h->pc], h->phi[1 - h->pc], h->c, g, g.nx / 40, h->prm.west_bc);
  return hipGetLastError();
}

}  // namespace bgk
