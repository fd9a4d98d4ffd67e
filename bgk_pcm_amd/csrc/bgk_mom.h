// bgk_mom.h -- moment form of the BGK_PCM step, shared by the moment-storage
// forward kernels (bgk_fused_mom.hip) and the adjoint (bgk_adjoint.hip).
//
// After the regularised collision (STEP_7/collision.f90:44-150, clean
// rounding) every population of a node is a function of ten numbers:
//   f_k = w_k rho (1 + P_k(u,v)) + coef_k H_k : Pf,   Pf = (1 - omega_f) Pi_f^neq
//   g_k = w_k T   (1 + P_k(u,v)) + coef_k H_k : Pg,   Pg = (1 - omega_g) Pi_g^neq
// with P_k the equilibrium polynomial minus its leading 1, H_k = c_k c_k -
// cs2 I and coef_k = w_k / (2 cs2^2).  A node stores
//   A = (rho, T - T_crit)                 double2
//   B = (u, v, Pf_xx, Pf_yy)              R4 (float4 or double4)
//   C = (Pf_xy, Pg_xx, Pg_yy, Pg_xy)      R4
// and, when omega_f = 1 (Pf = 0, "EQF"): B = (u, v, Pg_xx, Pg_yy), C = Pg_xy.
// The node update depends on its pulled populations only through ten
// moments (PullSums), which is what makes the adjoint compact.
#ifndef BGK_MOM_H
#define BGK_MOM_H

#include "bgk_internal.h"

namespace bgk {
namespace mom {

constexpr double kTcrit = 273.0;  // T_crit, STEP_7/phase_field.f90:16

// ---- storage traits: R = float (FP32 moments) or double (FP64 moments)
template <typename R>
struct Vec;
template <>
struct Vec<float> {
  using V4 = float4;
  typedef float nt4 __attribute__((ext_vector_type(4)));
};
template <>
struct Vec<double> {
  using V4 = double4;
  typedef double nt4 __attribute__((ext_vector_type(4)));
};
template <typename R, bool EQF>
struct CType {
  using T = typename Vec<R>::V4;
};
template <typename R>
struct CType<R, true> {
  using T = R;
};

template <typename R>
__device__ __forceinline__ typename Vec<R>::V4 make4(R x, R y, R z, R w) {
  typename Vec<R>::V4 v;
  v.x = x;
  v.y = y;
  v.z = z;
  v.w = w;
  return v;
}
template <typename R>
__device__ __forceinline__ typename Vec<R>::V4 nt_load4(const typename Vec<R>::V4* p) {
  const typename Vec<R>::nt4 v =
      __builtin_nontemporal_load(reinterpret_cast<const typename Vec<R>::nt4*>(p));
  return make4<R>(v.x, v.y, v.z, v.w);
}
template <typename R>
__device__ __forceinline__ void nt_store4(typename Vec<R>::V4* p, R x, R y, R z, R w) {
  typename Vec<R>::nt4 v;
  v.x = x;
  v.y = y;
  v.z = z;
  v.w = w;
  __builtin_nontemporal_store(v, reinterpret_cast<typename Vec<R>::nt4*>(p));
}

// Planes of one moment state (typed by the kernels through R, EQF).
struct MomIn {
  const double2* a;
  const void* b;
  const void* c;
};
struct MomOut {
  double2* a;
  void* b;
  void* c;
};

template <typename R, bool EQF>
struct Node {
  double2 a;                       // rho, T - T_crit
  typename Vec<R>::V4 b;           // u, v, Pf_xx, Pf_yy  (EQF: u, v, Pg_xx, Pg_yy)
  typename CType<R, EQF>::T c;     // Pf_xy, Pg_xx, Pg_yy, Pg_xy  (EQF: Pg_xy)
  double phi;
};

template <typename R, bool EQF>
__device__ __forceinline__ typename Vec<R>::V4 b_at(const void* b, int64_t q) {
  return static_cast<const typename Vec<R>::V4*>(b)[q];
}
template <typename R, bool EQF>
__device__ __forceinline__ typename CType<R, EQF>::T c_at(const void* c, int64_t q) {
  return static_cast<const typename CType<R, EQF>::T*>(c)[q];
}
template <typename R, bool EQF>
__device__ __forceinline__ typename CType<R, EQF>::T c_stream(const void* c, int64_t q) {
  if constexpr (EQF)
    return __builtin_nontemporal_load(static_cast<const R*>(c) + q);
  else
    return nt_load4<R>(static_cast<const typename Vec<R>::V4*>(c) + q);
}

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

// H_k : P for P = (pxx, pyy, pxy), H_k = c_k c_k - cs2 I, in the storage
// precision (FP32 operands keep their relative precision in FP32).
template <typename R>
__device__ __forceinline__ R h_k(int k, R pxx, R pyy, R pxy) {
  constexpr R A = R(-1) / R(3), B = R(2) / R(3);
  switch (k) {
    case 0: return A * (pxx + pyy);
    case 1:
    case 3: return B * pxx + A * pyy;
    case 2:
    case 4: return A * pxx + B * pyy;
    case 5:
    case 7: return B * (pxx + pyy) + R(2) * pxy;
    default: return B * (pxx + pyy) - R(2) * pxy;
  }
}

// Population k of a node from its moments: f_k and gd_k = g_k - w_k*T_crit.
template <typename R, bool EQF>
__device__ __forceinline__ void rebuild(int k, const double2 a, const typename Vec<R>::V4 b,
                                        const typename CType<R, EQF>::T cc, double* fk,
                                        double* gdk) {
  const double p = eq_poly(k, (double)b.x, (double)b.y);
  if constexpr (EQF) {
    const double hg = (double)h_k<R>(k, b.z, b.w, cc);
    *fk = kW[k] * a.x * (1.0 + p);
    *gdk = kW[k] * (a.y + (kTcrit + a.y) * p) + coef_of(k) * hg;
  } else {
    const double hf = (double)h_k<R>(k, b.z, b.w, cc.x);
    const double hg = (double)h_k<R>(k, cc.y, cc.z, cc.w);
    *fk = kW[k] * a.x * (1.0 + p) + coef_of(k) * hf;
    *gdk = kW[k] * (a.y + (kTcrit + a.y) * p) + coef_of(k) * hg;
  }
}

template <typename R, bool EQF>
__device__ __forceinline__ double rebuild_f(int k, const double2 a, const typename Vec<R>::V4 b,
                                            const typename CType<R, EQF>::T cc) {
  const double p = eq_poly(k, (double)b.x, (double)b.y);
  if constexpr (EQF) {
    return kW[k] * a.x * (1.0 + p);
  } else {
    const double hf = (double)h_k<R>(k, b.z, b.w, cc.x);
    return kW[k] * a.x * (1.0 + p) + coef_of(k) * hf;
  }
}

// Equilibrium deviation of a boundary ghost, g_k - w_k*T_crit.
__device__ __forceinline__ double ghost_gdev(int k, double T, double u, double v) {
  return kW[k] * (T - kTcrit) + kW[k] * T * eq_poly(k, u, v);
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

// The node update in moment form: momentsPre (STEP_7/momentsPre.f90:8-21),
// phase field (phase_field.f90:24-52, reaction rates in closed form),
// latent heat and forces (forces.f90:18-26), and the collision's
// non-equilibrium moments (collision.f90:44-92: for the second-order D2Q9
// equilibrium sum_k H_k feq_k = rho u u).  Everything the adjoint needs is
// kept in the struct.
//
// Phase field: 0.5 (1 -/+ tanh x) = 1 / (1 + exp(+/-2x)), x = (T - 273) /
// T_width: one exp and one division instead of the branchy glibc-exact
// tanh of the FP64 population kernels, which this clean-rounding path need
// not reproduce bit for bit (the rates' absolute error stays a few ulp of
// 1).  No branch inside one phase: there the reaction terms are multiplied
// by exact zeros anyway, and a branch splits the node's code (~70 more
// registers).
//
// Extensions: the phase field's T_crit (273 + crit_shift off) and a
// buoyancy reference t_ref (0: STEP_7's absolute T; both exact when off).
// Off when k_ratio == 1: solid/liquid conductivity ratio through
// a phase-dependent thermal relaxation, alpha(phi) = alpha_l (fl + k_ratio
// (1 - fl)), fl = (1 + phi)/2 the liquid fraction, tau_g = 1/2 + alpha/cs2.
struct Core {
  double rho, irho, u2, v2, dT2;     // momentsPre
  double kp_sum, km_sum, rm, phi, pc;  // phase field
  double dT, T;                        // after latent heat
  double u, v;                         // after forces
  double omega_g;                      // thermal relaxation rate used
  double qfxx, qfyy, qfxy;             // Pi_f^neq
  double qgxx, qgyy, qgxy;             // Pi_g^neq
};

// 1/d as the hardware estimate refined by two Newton steps (r <- r + r (1 -
// d r), each doubling the correct bits): within an ulp or two of the IEEE
// quotient and about ten instructions cheaper.  The moment path uses it for
// its per-node reciprocals (the two divisions of the conductivity ratio's
// rate cost 1.2% of the moment step at 4096^2); the FP64 population kernels
// keep IEEE division to stay bit-exact.
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

__device__ __forceinline__ double omega_g_of(const Consts& c, double phi) {
  if (c.k_ratio == 1.0) return c.omega_g;
  const double fl = 0.5 * (1.0 + phi);
  const double alpha = c.alpha_l * (fl + c.k_ratio * (1.0 - fl));
  return rcp_nr(0.5 + alpha * 3.0);  // 1/(1/2 + alpha/cs2)
}

__device__ __forceinline__ Core core_forward(const Consts& c, const PullSums& S,
                                             const double tap[9]) {
  Core o;
  o.rho = S.rho;
  o.dT2 = S.dT;
  o.irho = rcp_nr(o.rho);
  o.u2 = S.jx * o.irho;
  o.v2 = S.jy * o.irho;

  // sum_m ((1 +/- t_m)/2)^2 = (9 +/- 2 sum t + sum t^2) / 4
  double st = tap[0], st2 = tap[0] * tap[0];
#pragma unroll
  for (int m = 1; m < 9; ++m) {
    st = st + tap[m];
    st2 = fma(tap[m], tap[m], st2);
  }
  const double kp = 0.25 * ((9.0 + st2) + 2.0 * st);
  const double km = 0.25 * ((9.0 + st2) - 2.0 * st);
  o.kp_sum = kp;
  o.km_sum = km;
  o.pc = tap[0];
  // Both reaction terms vanish exactly when kp = 0 (all taps -1) or pc = 1,
  // and km = 0 (all taps +1) or pc = -1: phi = pc whatever the rates, which
  // are then not evaluated (rm = 1/2 keeps every product below finite; the
  // adjoint multiplies its rate terms by the same exact zeros).
  const bool react = !((kp == 0.0 || o.pc == 1.0) && (km == 0.0 || o.pc == -1.0));
  o.rm = 0.5;
  if (react) o.rm = 1.0 / (1.0 + exp((o.dT2 + c.crit_shift) * c.two_over_tw));
  const double rp = 1.0 - o.rm;
  const double kpr = (kp * 0.125) * rp;
  const double kmr = (km * 0.125) * o.rm;
  o.phi = o.pc + (0.5 * kpr * (1.0 - o.pc) - 0.5 * kmr * (1.0 + o.pc));

  o.dT = o.dT2 - (c.dt_st * (o.phi - o.pc) * 0.5);
  o.T = kTcrit + o.dT;
  double fy = o.rho * c.grav * c.beta * (o.T - c.t_ref) * (1.0 + o.phi * 0.5);
  const double fx = -(o.u2 * (1.0 - o.phi) * 0.5);
  fy = fy - o.v2 * (1.0 - o.phi) * 0.5;
  const double s = o.irho * c.inv_omega_f;
  o.u = o.u2 + fx * s;
  o.v = o.v2 + fy * s;

  o.omega_g = omega_g_of(c, o.phi);
  o.qfxx = S.fxx - kCs2 * o.rho - o.rho * o.u * o.u;
  o.qfyy = S.fyy - kCs2 * o.rho - o.rho * o.v * o.v;
  o.qfxy = S.fxy - o.rho * o.u * o.v;
  o.qgxx = S.gxx - kCs2 * o.dT2 - o.T * o.u * o.u;
  o.qgyy = S.gyy - kCs2 * o.dT2 - o.T * o.v * o.v;
  o.qgxy = S.gxy - o.T * o.u * o.v;
  return o;
}

// Averaged moments of STEP_7/moments.f90:11-25 from the stored state: sum_k
// f_k = rho, sum_k g_k = T and sum_k c_k f_k / rho = u exactly for the
// rebuilt populations.
struct NodeOut {
  double rho, Tav, uav, vav;
};

// One launch of the moment-storage step: state and phi in and out, the
// east-wall temperature column of the previous and of this step, an
// optional probe record (T averaged as moments.f90 at the nodes whose
// probe_idx is >= 0, written to probe_out[probe_idx]; probe_idx in column 0
// of a row is >= 0 when the row holds a probe) and optional
// macroscopic planes rho, u, v, T to write instead of the handle's (their
// halo columns of u, v must hold the ghost velocities).
struct MomStep {
  double* macro[4];
  MomIn in;
  MomOut out;
  const double* pa;
  double* pb;
  const double* te_in;
  double* te_out;
  bool write_all;
  const int* probe_idx;
  double* probe_out;
};

}  // namespace mom

hipError_t launch_mom_step(bgk_handle* h, const mom::MomStep& st);
}  // namespace bgk

#endif  // BGK_MOM_H
