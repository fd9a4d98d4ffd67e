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

// The helpers below take the arithmetic type A: double for the forward
// kernels (and the FP64-moment adjoint), float for the adjoint of the FP32
// moment step, which evaluates its Jacobians in FP32.

// P_k(u,v): the equilibrium polynomial of direction k minus its leading 1.
template <typename A>
__device__ __forceinline__ A eq_poly_t(int k, A u, A v) {
  const A uv = A(1.5) * (u * u + v * v);
  switch (k) {
    case 0: return -uv;
    case 1: return A(3) * u + A(4.5) * (u * u) - uv;
    case 2: return A(3) * v + A(4.5) * (v * v) - uv;
    case 3: return -A(3) * u + A(4.5) * (u * u) - uv;
    case 4: return -A(3) * v + A(4.5) * (v * v) - uv;
    case 5: { const A s = u + v; return A(3) * s + A(4.5) * (s * s) - uv; }
    case 6: { const A s = u - v; return -A(3) * s + A(4.5) * (s * s) - uv; }
    case 7: { const A s = u + v; return -A(3) * s + A(4.5) * (s * s) - uv; }
    default: { const A s = u - v; return A(3) * s + A(4.5) * (s * s) - uv; }
  }
}
__device__ __forceinline__ double eq_poly(int k, double u, double v) {
  return eq_poly_t<double>(k, u, v);
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
template <typename R, bool EQF, typename A = double, typename A2 = double2>
__device__ __forceinline__ void rebuild(int k, const A2 a, const typename Vec<R>::V4 b,
                                        const typename CType<R, EQF>::T cc, A* fk, A* gdk) {
  const A p = eq_poly_t<A>(k, (A)b.x, (A)b.y);
  const A ax = (A)a.x, ay = (A)a.y;
  if constexpr (EQF) {
    const A hg = (A)h_k<R>(k, b.z, b.w, cc);
    *fk = A(kW[k]) * ax * (A(1) + p);
    *gdk = A(kW[k]) * (ay + (A(kTcrit) + ay) * p) + A(coef_of(k)) * hg;
  } else {
    const A hf = (A)h_k<R>(k, b.z, b.w, cc.x);
    const A hg = (A)h_k<R>(k, cc.y, cc.z, cc.w);
    *fk = A(kW[k]) * ax * (A(1) + p) + A(coef_of(k)) * hf;
    *gdk = A(kW[k]) * (ay + (A(kTcrit) + ay) * p) + A(coef_of(k)) * hg;
  }
}

template <typename R, bool EQF, typename A = double, typename A2 = double2>
__device__ __forceinline__ A rebuild_f(int k, const A2 a, const typename Vec<R>::V4 b,
                                       const typename CType<R, EQF>::T cc) {
  const A p = eq_poly_t<A>(k, (A)b.x, (A)b.y);
  if constexpr (EQF) {
    return A(kW[k]) * (A)a.x * (A(1) + p);
  } else {
    const A hf = (A)h_k<R>(k, b.z, b.w, cc.x);
    return A(kW[k]) * (A)a.x * (A(1) + p) + A(coef_of(k)) * hf;
  }
}

// Equilibrium deviation of a boundary ghost, g_k - w_k*T_crit.
template <typename A = double>
__device__ __forceinline__ A ghost_gdev(int k, A T, A u, A v) {
  return A(kW[k]) * (T - A(kTcrit)) + A(kW[k]) * T * eq_poly_t<A>(k, u, v);
}

// Moments of the pulled populations of a node: zeroth, first and second
// moments of f, and of gd = g - w*T_crit (its first moment is not needed).
template <typename A>
struct PullSumsT {
  A rho, jx, jy, fxx, fyy, fxy;
  A dT, gxx, gyy, gxy;
};
using PullSums = PullSumsT<double>;

template <typename A>
__device__ __forceinline__ void pull_zero(PullSumsT<A>& S) {
  S.rho = S.jx = S.jy = S.fxx = S.fyy = S.fxy = A(0);
  S.dT = S.gxx = S.gyy = S.gxy = A(0);
}

// Add population k (f_k, gd_k) to the sums.
template <typename A>
__device__ __forceinline__ void pull_add(PullSumsT<A>& S, int k, A fk, A gk) {
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
template <typename A>
struct CoreT {
  A rho, irho, u2, v2, dT2;     // momentsPre
  A kp_sum, km_sum, rm, phi, pc;  // phase field
  A dT, T;                        // after latent heat
  A u, v;                         // after forces
  A omega_g;                      // thermal relaxation rate used
  A qfxx, qfyy, qfxy;             // Pi_f^neq
  A qgxx, qgyy, qgxy;             // Pi_g^neq
};
using Core = CoreT<double>;

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
// FP32: the hardware estimate and one Newton step
__device__ __forceinline__ float rcp_nr(float d) {
  float r = __builtin_amdgcn_rcpf(d);
  return fmaf(r, fmaf(-d, r, 1.0f), r);
}

// FP32 copies of the constants the node update and its adjoint read, for
// the FP32-arithmetic adjoint (filled from Consts on the host, so that no
// kernel converts them).
struct ConstsF {
  float crit_shift, two_over_tw, dt_st, grav, beta, t_ref, inv_omega_f, omega_f, omega_g,
      alpha_l, k_ratio, inv_stefan;
};
inline ConstsF consts_f(const Consts& c) {
  return ConstsF{(float)c.crit_shift, (float)c.two_over_tw, (float)c.dt_st,   (float)c.grav,
                 (float)c.beta,       (float)c.t_ref,       (float)c.inv_omega_f,
                 (float)c.omega_f,    (float)c.omega_g,     (float)c.alpha_l,
                 (float)c.k_ratio,    (float)c.inv_stefan};
}

// C: Consts, or ConstsF (the same fields in FP32) for FP32 arithmetic.
template <typename A, typename C>
__device__ __forceinline__ A omega_g_of(const C& c, A phi) {
  if (c.k_ratio == 1.0) return (A)c.omega_g;
  const A fl = A(0.5) * (A(1) + phi);
  const A alpha = (A)c.alpha_l * (fl + (A)c.k_ratio * (A(1) - fl));
  return rcp_nr(A(0.5) + alpha * A(3));  // 1/(1/2 + alpha/cs2)
}

template <typename A, typename C>
__device__ __forceinline__ CoreT<A> core_forward_t(const C& c, const PullSumsT<A>& S,
                                                   const A tap[9]) {
  CoreT<A> o;
  o.rho = S.rho;
  o.dT2 = S.dT;
  o.irho = rcp_nr(o.rho);
  o.u2 = S.jx * o.irho;
  o.v2 = S.jy * o.irho;

  // sum_m ((1 +/- t_m)/2)^2 = (9 +/- 2 sum t + sum t^2) / 4
  A st = tap[0], st2 = tap[0] * tap[0];
#pragma unroll
  for (int m = 1; m < 9; ++m) {
    st = st + tap[m];
    st2 = fma(tap[m], tap[m], st2);
  }
  const A kp = A(0.25) * ((A(9) + st2) + A(2) * st);
  const A km = A(0.25) * ((A(9) + st2) - A(2) * st);
  o.kp_sum = kp;
  o.km_sum = km;
  o.pc = tap[0];
  // Both reaction terms vanish exactly when kp = 0 (all taps -1) or pc = 1,
  // and km = 0 (all taps +1) or pc = -1: phi = pc whatever the rates, which
  // are then not evaluated (rm = 1/2 keeps every product below finite; the
  // adjoint multiplies its rate terms by the same exact zeros).
  const bool react = !((kp == A(0) || o.pc == A(1)) && (km == A(0) || o.pc == A(-1)));
  o.rm = A(0.5);
  if (react) o.rm = A(1) / (A(1) + exp((o.dT2 + (A)c.crit_shift) * (A)c.two_over_tw));
  const A rp = A(1) - o.rm;
  const A kpr = (kp * A(0.125)) * rp;
  const A kmr = (km * A(0.125)) * o.rm;
  o.phi = o.pc + (A(0.5) * kpr * (A(1) - o.pc) - A(0.5) * kmr * (A(1) + o.pc));

  o.dT = o.dT2 - ((A)c.dt_st * (o.phi - o.pc) * A(0.5));
  o.T = A(kTcrit) + o.dT;
  A fy = o.rho * (A)c.grav * (A)c.beta * (o.T - (A)c.t_ref) * (A(1) + o.phi * A(0.5));
  const A fx = -(o.u2 * (A(1) - o.phi) * A(0.5));
  fy = fy - o.v2 * (A(1) - o.phi) * A(0.5);
  const A s = o.irho * (A)c.inv_omega_f;
  o.u = o.u2 + fx * s;
  o.v = o.v2 + fy * s;

  o.omega_g = omega_g_of<A, C>(c, o.phi);
  o.qfxx = S.fxx - A(kCs2) * o.rho - o.rho * o.u * o.u;
  o.qfyy = S.fyy - A(kCs2) * o.rho - o.rho * o.v * o.v;
  o.qfxy = S.fxy - o.rho * o.u * o.v;
  o.qgxx = S.gxx - A(kCs2) * o.dT2 - o.T * o.u * o.u;
  o.qgyy = S.gyy - A(kCs2) * o.dT2 - o.T * o.v * o.v;
  o.qgxy = S.gxy - o.T * o.u * o.v;
  return o;
}

__device__ __forceinline__ Core core_forward(const Consts& c, const PullSums& S,
                                             const double tap[9]) {
  return core_forward_t<double, Consts>(c, S, tap);
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
