// bgk_adjoint.hip -- discrete adjoint of the moment-storage time step.
//
// Forward model: the moment-storage step (bgk_fused_mom.hip, clean rounding,
// no-slip cavity, extensions t_hot and k_ratio) with FP64 moments, or FP32
// moments (R = float: the adjoint of the step evaluated at the stored
// states, the storage rounding taken as the identity).  Loss: any
// function of the averaged temperature T (STEP_7/moments.f90:11-25) at a
// set of probe nodes after every step; the caller supplies dL/dT(step,
// probe) and receives dL/d(stefan_number, k_ratio, t_hot).
//
// Adjoint state per node (step n): the adjoints of the ten pulled moments
// the node update depends on (PullSums of bgk_mom.h: rho, j, the second
// moments of f; T and the second moments of g) and of the two tap sums and
// the centre tap of the phase field: 13 values, 10 with omega_f = 1, stored
// and computed in the state's precision.  The pull fp_k(x) =
// f_k(x - c_k) is reversed as a gather, so no atomics: node y collects,
// for every direction k, the adjoint of the population k that node y + c_k
// pulled from it, and turns it into the adjoint of its own stored moments
// through the derivative of the population rebuild.  The node update is
// reversed by hand (core_backward below), on values recomputed with the
// same core_forward the forward kernels run: the adjoint of step n reads
// only the state before step n.
//
// Checkpointing, binomial (bgk_adj_setup_binomial): the optimal schedule for
// a number of state slots (see rev_cost below).  Levels (bgk_adj_setup,
// bgk_adj_setup_levels): bgk_adj_forward keeps the state every seg_len steps;
// bgk_adj_backward recomputes one segment at a time from its checkpoint,
// keeping every state of the segment, then runs the adjoint kernels over
// the segment backwards.  Memory: (nsteps/seg_len + seg_len + 1) states.
// With two levels the forward run keeps every seg_outer-th state, and each
// outer segment is first recomputed to checkpoints every seg_inner steps:
// nsteps/seg_outer + seg_outer/seg_inner + seg_inner + 1 states.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "bgk_mom.h"

namespace bgk {
namespace {

using namespace mom;

// one stored state: moment planes, phi, east-wall column
struct StatePtrs {
  double2* a;
  void* b;
  void* c;
  double* phi;
  double* te;
};

// adjoint planes of one step (structure of arrays, node q at [q]), stored
// in the state's precision R (the arithmetic is FP64 either way)
struct AdjPlanes {
  void* v[13];  // rho jx jy fxx fyy fxy dT gxx gyy gxy kp km pc
};
template <typename R>
__device__ __forceinline__ double plane_at(const AdjPlanes& L, int p, int64_t q) {
  return (double)static_cast<const R*>(L.v[p])[q];
}
template <typename R>
__device__ __forceinline__ void plane_set(const AdjPlanes& L, int p, int64_t q, double x) {
  static_cast<R*>(L.v[p])[q] = (R)x;
}
enum { L_RHO, L_JX, L_JY, L_FXX, L_FYY, L_FXY, L_DT, L_GXX, L_GYY, L_GXY, L_KP, L_KM, L_PC };

struct AdjArgs {
  // the state before step n (the inputs of step n); the state after it, whose
  // moments the gather of step n+1's pulls differentiates at, is recomputed
  // from it in the kernel (core = 1) or is this state itself (core = 0)
  StatePtrs sprev;
  const double* uh;  // macroscopic u, v planes: their halo columns are the ghost velocities
  const double* vh;
  AdjPlanes next;    // adjoint of step n+1 (read)
  AdjPlanes cur;     // adjoint of step n (written)
  int has_next;      // 0 at the last step: no step n+1
  int core;          // 0: only the gather (the initial state)
  const int* probe_idx;
  const double* seeds;  // dL/dT at the probes after step n (null: none)
  double* partial;      // 3 per workgroup: dL/dSt, dL/dk_ratio, dL/dt_hot (accumulated)
  Consts c;
  ConstsF cf;  // FP32 copies for the FP32 kernel
  Geom gm;
  int interf;
  int ntx, nty;  // tiles across and down
};

// The workgroup's window in LDS: a tile of BX x BY nodes and a one-node
// halo, with the adjoint planes of step n+1 (those read at neighbours: not
// L_PC, and with omega_f = 1 (EQF) not the second moments of f, whose
// adjoints are (1 - omega_f) times finite values, exact zeros, and are
// neither stored nor read) and the state before step n (moments and phi).
// Every read of a node's update lies in its 3 x 3 neighbourhood, so each
// value comes from HBM once per tile instead of from L2 for every neighbour.
// Tiles of 64 x 8 nodes (one workgroup of 8 waves) with FP32 moments, 64 x 4
// (4 waves) with FP64 moments.  The window holds 68 B per node with FP32
// moments and omega_f = 1 (45 KB at 64 x 8: three workgroups per CU, 6
// waves per SIMD, 72 VGPRs), 92 B with the general FP32 layout (61 KB: 4
// waves per SIMD), 136 B and 168 B with FP64 moments (64 x 4: 3 and 2 waves
// per SIMD).  Measured at 8192^2 (FP32, omega_f = 1, XCD bands of 8-tile
// groups) per launch: 64 x 8 2.05 ms, 64 x 12 (two workgroups of 12 waves)
// 2.07 ms, 64 x 4 (six of 4 waves) 2.19 ms.  Per wave: 765 VALU
// instructions, and HBM reads of 1.17x the algorithmic bytes (the halo
// rows and columns are L2 hits, see tile_of).
constexpr int BX = 64, WX = BX + 2;
template <typename R, bool EQF>
constexpr int tile_rows() {
  return sizeof(R) == sizeof(float) ? 8 : 4;
}

// Window types of the state's FP64 members (rho, T - T_crit and phi): the
// FP32 kernel evaluates in FP32, so its window holds them rounded once, in
// 12 bytes instead of 24.
template <typename R>
struct WinT {
  using A2 = double2;
  using P = double;
  static __device__ __forceinline__ A2 a2(const double2 v) { return v; }
};
template <>
struct WinT<float> {
  using A2 = float2;
  using P = float;
  static __device__ __forceinline__ A2 a2(const double2 v) {
    return make_float2((float)v.x, (float)v.y);
  }
};

template <bool EQF>
struct LSlots {
  static constexpr int N = EQF ? 9 : 12;
  // plane staged in LDS slot s
  static constexpr int plane[12] = {L_RHO, L_JX, L_JY, L_DT, L_GXX, L_GYY,
                                    L_GXY, L_KP, L_KM, L_FXX, L_FYY, L_FXY};
  enum { RHO, JX, JY, DT, GXX, GYY, GXY, KP, KM, FXX, FYY, FXY };
};

template <typename R, bool EQF>
struct Win {
  static constexpr int WY = tile_rows<R, EQF>() + 2;
  R lam[LSlots<EQF>::N][WY][WX];  // as stored: every read widens exactly
  typename WinT<R>::A2 a[WY][WX];
  typename Vec<R>::V4 b[WY][WX];
  typename CType<R, EQF>::T c[WY][WX];
  typename WinT<R>::P phi[WY][WX];
};

// Adjoint of population k as pulled by the node at window (x, y), from that
// node's adjoint of its pulled moments (PullSums is linear in the
// populations).
template <typename R, bool EQF>
__device__ __forceinline__ R lam_f(const Win<R, EQF>& W, int k, int x, int y) {
  using S = LSlots<EQF>;
  R r = W.lam[S::RHO][y][x];
  if constexpr (EQF) {
    if (CX[k] != 0) r += CX[k] > 0 ? W.lam[S::JX][y][x] : -W.lam[S::JX][y][x];
    if (CY[k] != 0) r += CY[k] > 0 ? W.lam[S::JY][y][x] : -W.lam[S::JY][y][x];
  } else {
    if (CX[k] != 0)
      r += (CX[k] > 0 ? W.lam[S::JX][y][x] : -W.lam[S::JX][y][x]) + W.lam[S::FXX][y][x];
    if (CY[k] != 0)
      r += (CY[k] > 0 ? W.lam[S::JY][y][x] : -W.lam[S::JY][y][x]) + W.lam[S::FYY][y][x];
    if (CX[k] * CY[k] != 0)
      r += CX[k] * CY[k] > 0 ? W.lam[S::FXY][y][x] : -W.lam[S::FXY][y][x];
  }
  return r;
}
template <typename R, bool EQF>
__device__ __forceinline__ R lam_g(const Win<R, EQF>& W, int k, int x, int y) {
  using S = LSlots<EQF>;
  R r = W.lam[S::DT][y][x];
  if (CX[k] != 0) r += W.lam[S::GXX][y][x];
  if (CY[k] != 0) r += W.lam[S::GYY][y][x];
  if (CX[k] * CY[k] != 0) r += CX[k] * CY[k] > 0 ? W.lam[S::GXY][y][x] : -W.lam[S::GXY][y][x];
  return r;
}

// Stored moments of a node (EQF: Pf = 0), or their adjoints, in the
// arithmetic type A.
template <typename A>
struct Mom {
  A rho, th, u, v, pf[3], pg[3];
};

// The moments a forward step stores for a node (bgk_fused_mom.hip's
// update_node), widened back to FP64: rho and T - T_crit as computed, u, v
// and the scaled non-equilibrium moments rounded to the storage type R.
template <typename R, bool EQF, typename A, typename C>
__device__ __forceinline__ Mom<A> stored_mom(const C& c, const CoreT<A>& o) {
  Mom<A> m;
  m.rho = o.rho;
  m.th = o.dT;
  m.u = (A)(R)o.u;
  m.v = (A)(R)o.v;
  const A omg1 = A(1) - o.omega_g;
  m.pg[0] = (A)(R)(omg1 * o.qgxx);
  m.pg[1] = (A)(R)(omg1 * o.qgyy);
  m.pg[2] = (A)(R)(omg1 * o.qgxy);
  if constexpr (EQF) {
    m.pf[0] = m.pf[1] = m.pf[2] = A(0);
  } else {
    const A omf1 = A(1) - (A)c.omega_f;
    m.pf[0] = (A)(R)(omf1 * o.qfxx);
    m.pf[1] = (A)(R)(omf1 * o.qfyy);
    m.pf[2] = (A)(R)(omf1 * o.qfxy);
  }
  return m;
}

// Moments of a node held in the window (the state before the step).
template <typename R, bool EQF, typename A, typename A2>
__device__ __forceinline__ Mom<A> win_mom(const A2& a, const typename Vec<R>::V4& b,
                                          const typename CType<R, EQF>::T& cc) {
  Mom<A> m;
  m.rho = (A)a.x;
  m.th = (A)a.y;
  m.u = (A)b.x;
  m.v = (A)b.y;
  if constexpr (EQF) {
    m.pf[0] = m.pf[1] = m.pf[2] = A(0);
    m.pg[0] = (A)b.z;
    m.pg[1] = (A)b.w;
    m.pg[2] = (A)cc;
  } else {
    m.pf[0] = (A)b.z;
    m.pf[1] = (A)b.w;
    m.pf[2] = (A)cc.x;
    m.pg[0] = (A)cc.y;
    m.pg[1] = (A)cc.z;
    m.pg[2] = (A)cc.w;
  }
  return m;
}

// Weighted moments sum_k w_k c_k^n a_k (n = 0, 1, 2) of a vector over the
// nine directions.
template <typename A>
struct WMom {
  A m0, mx, my, mxx, myy, mxy;
};
template <typename A>
__device__ __forceinline__ WMom<A> wmoments(const A a[9]) {
  const A s13 = a[1] + a[3], s24 = a[2] + a[4];
  const A sd = (a[5] + a[6]) + (a[7] + a[8]);
  WMom<A> m;
  m.m0 = A(kW[0]) * a[0] + A(kW[1]) * (s13 + s24) + A(kW[5]) * sd;
  m.mx = A(kW[1]) * (a[1] - a[3]) + A(kW[5]) * ((a[5] + a[8]) - (a[6] + a[7]));
  m.my = A(kW[1]) * (a[2] - a[4]) + A(kW[5]) * ((a[5] + a[6]) - (a[7] + a[8]));
  m.mxx = A(kW[1]) * s13 + A(kW[5]) * sd;
  m.myy = A(kW[1]) * s24 + A(kW[5]) * sd;
  m.mxy = A(kW[5]) * ((a[5] + a[7]) - (a[6] + a[8]));
  return m;
}

// Accumulate into lm the adjoint of the node's moments from the adjoints
// lf[k], lg[k] of its nine populations f_k and gd_k (the transposed rebuild
// Jacobian).  Population k is w_k rho (1 + p_k) + c_k H_k : Pf for f and
// w_k (dT + T p_k) + c_k H_k : Pg for gd, with p_k = 3 cu + 9/2 cu^2 -
// 3/2 |u|^2, cu = c_k . u, c_k = w_k / (2 cs^4) and H_k = c_k c_k - cs^2 I;
// its derivatives are polynomials in c_k, so the sums over k reduce to the
// weighted moments of lf and lg.
template <typename A>
__device__ __forceinline__ void rebuild_adjoint(const Mom<A>& m, const A lf[9], const A lg[9],
                                                Mom<A>* lm) {
  const WMom<A> F = wmoments<A>(lf), G = wmoments<A>(lg);
  const A u = m.u, v = m.v, uu = u * u, vv = v * v, uv2 = A(2) * u * v;
  const A c0 = A(1) - A(1.5) * (uu + vv);
  // sum_k w_k a_k (1 + p_k)
  lm->rho += F.m0 * c0 + A(3) * (u * F.mx + v * F.my) +
             A(4.5) * (uu * F.mxx + uv2 * F.mxy + vv * F.myy);
  lm->th += G.m0 * c0 + A(3) * (u * G.mx + v * G.my) +
            A(4.5) * (uu * G.mxx + uv2 * G.mxy + vv * G.myy);
  // sum_k w_k L_k dp_k/du, L_k = rho lf_k + T lg_k (and dv)
  const A rho = m.rho, tt = A(kTcrit) + m.th;
  const A L0 = rho * F.m0 + tt * G.m0;
  const A Lx = rho * F.mx + tt * G.mx, Ly = rho * F.my + tt * G.my;
  const A Lxx = rho * F.mxx + tt * G.mxx, Lyy = rho * F.myy + tt * G.myy;
  const A Lxy = rho * F.mxy + tt * G.mxy;
  lm->u += A(3) * (Lx - u * L0) + A(9) * (u * Lxx + v * Lxy);
  lm->v += A(3) * (Ly - v * L0) + A(9) * (u * Lxy + v * Lyy);
  // sum_k c_k H_k a_k = (9/2) (Mxx - cs2 M0, Myy - cs2 M0, 2 Mxy)
  lm->pf[0] += A(4.5) * (F.mxx - A(kCs2) * F.m0);
  lm->pf[1] += A(4.5) * (F.myy - A(kCs2) * F.m0);
  lm->pf[2] += A(9) * F.mxy;
  lm->pg[0] += A(4.5) * (G.mxx - A(kCs2) * G.m0);
  lm->pg[1] += A(4.5) * (G.myy - A(kCs2) * G.m0);
  lm->pg[2] += A(9) * G.mxy;
}

// d ghost_gdev / dT for direction k at velocity (u, v)
template <typename A>
__device__ __forceinline__ A ghost_dT(int k, A u, A v) {
  return A(kW[k]) * (A(1) + eq_poly_t<A>(k, u, v));
}

// Pulled sums and taps of step n at node (i, j), window (ox, oy), from the
// state before the step, with the boundary folding of the forward kernels
// (mom_edge_nodes).  EDGE = false: the node and its eight neighbours are
// interior nodes (no boundary tests).
template <typename R, bool EQF, typename A, bool EDGE>
__device__ __forceinline__ void pull_step(const AdjArgs& a, const Win<R, EQF>& W, int i, int j,
                                          int ox, int oy, PullSumsT<A>* S, A tap[9]) {
  const int nx = a.gm.nx, ny = a.gm.ny;
  const int64_t P = a.gm.pitch;
  pull_zero(*S);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int si = i - CX[k], sj = j - CY[k];
    const int sjc = !EDGE ? sj : (sj < 1 ? 1 : (sj > ny ? ny : sj));
    const int wx = ox - CX[k], wy = oy + (sjc - j);
    A fk = A(0), gk;
    if (!EDGE || (si >= 1 && si <= nx)) {
      rebuild<R, EQF, A, typename WinT<R>::A2>(k, W.a[wy][wx], W.b[wy][wx], W.c[wy][wx], &fk,
                                               &gk);
    } else {
      const int64_t h = (int64_t)sjc * P + si;
      gk = ghost_gdev<A>(k, (A)(si == 0 ? a.c.tin : east_T(a.c, a.sprev.te, sjc)),
                         (A)a.uh[h], (A)a.vh[h]);
    }
    if (EDGE && (si < 1 || si > nx || sj < 1 || sj > ny)) {
      if (k == 6 && si == nx + 1 && sj == 0)
        fk = (A)a.c.frozen6;
      else if (k == 7 && si == nx + 1 && sj == ny + 1)
        fk = (A)a.c.frozen7;
      else
        fk = rebuild_f<R, EQF, A, typename WinT<R>::A2>(OPP[k], W.a[oy][ox], W.b[oy][ox],
                                                        W.c[oy][ox]);
    }
    pull_add(*S, k, fk, gk);
  }
  const int yn = oy + (!EDGE || j < ny ? 1 : 0), ys = oy - (!EDGE || j > 1 ? 1 : 0);
  tap[0] = (A)W.phi[oy][ox];
  tap[1] = (A)W.phi[oy][ox + 1];
  tap[2] = (A)W.phi[yn][ox];
  tap[3] = (A)W.phi[oy][ox - 1];
  tap[4] = (A)W.phi[ys][ox];
  tap[5] = (A)W.phi[yn][ox + 1];
  tap[6] = (A)W.phi[yn][ox - 1];
  tap[7] = (A)W.phi[ys][ox - 1];
  tap[8] = (A)W.phi[ys][ox + 1];
}

// Reverse of core_forward (bgk_mom.h) and of the state it stores.  lo: the
// adjoint of the node's stored moments after the step (rho, T - T_crit, u,
// v, Pf, Pg) and of its phi; ldT2 an extra adjoint on the pulled T (the
// averaged T of moments.f90 seeds half of its adjoint here).  Returns the
// adjoints of the pulled sums in lS, of the tap sums in lkp, lkm and of the
// centre tap (direct terms) in lpc; adds parameter adjoints to g[3].
template <typename A, typename C>
__device__ __forceinline__ void core_backward(const C& c, const PullSumsT<A>& S,
                                              const A tap[9], const CoreT<A>& o, const Mom<A>& lo,
                                              A lphi_o, A ldT2_in, PullSumsT<A>* lS, A* lkp,
                                              A* lkm, A* lpc, A g[3]) {
  A lrho = lo.rho, lu = lo.u, lv = lo.v, lphi = lphi_o;
  A ldT2 = ldT2_in;
  // Pg = (1 - omega_g) Qg, Pf = (1 - omega_f) Qf
  const A omg1 = A(1) - o.omega_g, omf1 = A(1) - (A)c.omega_f;
  const A lqgxx = omg1 * lo.pg[0], lqgyy = omg1 * lo.pg[1], lqgxy = omg1 * lo.pg[2];
  const A lomg = -(o.qgxx * lo.pg[0] + o.qgyy * lo.pg[1] + o.qgxy * lo.pg[2]);
  const A lqfxx = omf1 * lo.pf[0], lqfyy = omf1 * lo.pf[1], lqfxy = omf1 * lo.pf[2];
  // Qg = S_g - cs2 dT2 - T u u
  ldT2 += -A(kCs2) * (lqgxx + lqgyy);
  A lT = -(o.u * o.u * lqgxx + o.v * o.v * lqgyy + o.u * o.v * lqgxy);
  lu += -A(2) * o.T * o.u * lqgxx - o.T * o.v * lqgxy;
  lv += -A(2) * o.T * o.v * lqgyy - o.T * o.u * lqgxy;
  // Qf = S_f - cs2 rho - rho u u
  lrho += -(A(kCs2) + o.u * o.u) * lqfxx - (A(kCs2) + o.v * o.v) * lqfyy - o.u * o.v * lqfxy;
  lu += -A(2) * o.rho * o.u * lqfxx - o.rho * o.v * lqfxy;
  lv += -A(2) * o.rho * o.v * lqfyy - o.rho * o.u * lqfxy;
  // omega_g(phi, k_ratio) = 1 / (1/2 + alpha/cs2), alpha = alpha_l (fl +
  // k_ratio (1 - fl)): the k_ratio derivative is taken at k_ratio = 1 too,
  // where the forward uses the constant omega_g (the same value)
  {
    const A fl = A(0.5) * (A(1) + o.phi);
    const A domg_dalpha = -o.omega_g * o.omega_g * A(3);  // 1/cs2
    lphi += lomg * domg_dalpha * (A)c.alpha_l * (A(1) - (A)c.k_ratio) * A(0.5);
    g[1] += lomg * domg_dalpha * (A)c.alpha_l * (A(1) - fl);
  }
  // u = u2 + fx s, v = v2 + fy s, s = irho / omega_f
  const A s = o.irho * (A)c.inv_omega_f;
  const A fx = -(o.u2 * (A(1) - o.phi) * A(0.5));
  const A gb = (A)c.grav * (A)c.beta;
  const A fy0 = o.rho * gb * (o.T - (A)c.t_ref) * (A(1) + o.phi * A(0.5));
  const A fy = fy0 - o.v2 * (A(1) - o.phi) * A(0.5);
  A lu2 = lu, lv2 = lv;
  const A lfx = s * lu, lfy = s * lv;
  A lirho = (fx * lu + fy * lv) * (A)c.inv_omega_f;
  // fy = fy0 - v2 (1 - phi)/2, fx = -u2 (1 - phi)/2
  lv2 += -(A(1) - o.phi) * A(0.5) * lfy;
  lphi += o.v2 * A(0.5) * lfy;
  lu2 += -(A(1) - o.phi) * A(0.5) * lfx;
  lphi += o.u2 * A(0.5) * lfx;
  // fy0 = rho g beta (T - t_ref) (1 + phi/2)
  lrho += gb * (o.T - (A)c.t_ref) * (A(1) + o.phi * A(0.5)) * lfy;
  lT += o.rho * gb * (A(1) + o.phi * A(0.5)) * lfy;
  lphi += o.rho * gb * (o.T - (A)c.t_ref) * A(0.5) * lfy;
  // stored T - T_crit = dT; T = T_crit + dT; dT = dT2 - dt_st (phi - pc)/2
  const A ldT = lT + lo.th;
  ldT2 += ldT;
  lphi += -(A)c.dt_st * A(0.5) * ldT;
  A lpcv = (A)c.dt_st * A(0.5) * ldT;
  // dt_st = delta_T / St
  g[0] += (-(o.phi - o.pc) * A(0.5) * ldT) * (-(A)c.dt_st * (A)c.inv_stefan);
  // phi = pc + 0.5 kpr (1 - pc) - 0.5 kmr (1 + pc), kpr = kp/8 rp, kmr = km/8 rm
  const A rp = A(1) - o.rm;
  const A kpr = o.kp_sum * A(0.125) * rp, kmr = o.km_sum * A(0.125) * o.rm;
  const A lkpr = A(0.5) * (A(1) - o.pc) * lphi, lkmr = -A(0.5) * (A(1) + o.pc) * lphi;
  lpcv += lphi * (A(1) - A(0.5) * kpr - A(0.5) * kmr);
  *lkp = rp * A(0.125) * lkpr;
  *lkm = o.rm * A(0.125) * lkmr;
  const A lrm = o.km_sum * A(0.125) * lkmr - o.kp_sum * A(0.125) * lkpr;
  // rm = 1/(1 + exp((dT2 + crit_shift) * 2/T_width))
  ldT2 += -o.rm * rp * (A)c.two_over_tw * lrm;
  *lpc = lpcv;
  // u2 = jx irho, v2 = jy irho, irho = 1/rho
  lS->jx = o.irho * lu2;
  lS->jy = o.irho * lv2;
  lirho += S.jx * lu2 + S.jy * lv2;
  lrho += -o.irho * o.irho * lirho;
  lS->rho = lrho;
  lS->fxx = lqfxx;
  lS->fyy = lqfyy;
  lS->fxy = lqfxy;
  lS->dT = ldT2;
  lS->gxx = lqgxx;
  lS->gyy = lqgyy;
  lS->gxy = lqgxy;
}

// The constants in the arithmetic type: Consts for FP64, ConstsF for FP32.
template <typename A>
struct ConstsOf;
template <>
struct ConstsOf<double> {
  static __device__ __forceinline__ const Consts& of(const AdjArgs& a) { return a.c; }
};
template <>
struct ConstsOf<float> {
  static __device__ __forceinline__ const ConstsF& of(const AdjArgs& a) { return a.cf; }
};

// The adjoint of step n at node (i, j), window (ox, oy); the parameter
// adjoints are added to g.  EDGE = false for the tiles whose nodes and their
// neighbours are all interior nodes: no boundary terms and no tests.
// Arithmetic in the storage precision: FP64 for FP64 moments, FP32 for FP32
// moments (the Jacobians evaluated in FP32: the kernel is bound by its
// arithmetic).
template <typename R, bool EQF, bool EDGE>
__device__ __forceinline__ void adjoint_node(const AdjArgs& a, const Win<R, EQF>& W, int i, int j,
                                             int ox, int oy, R g[3]) {
  using A = R;
  const int nx = a.gm.nx, ny = a.gm.ny;
  const int64_t P = a.gm.pitch;
  const int64_t q = (int64_t)j * P + i;
  // this node's state after step n: recomputed from the window (the same
  // core_forward the forward kernels run), or the state itself for the
  // initial-state launch
  Mom<A> m;
  A ph;
  PullSumsT<A> S;
  A tap[9];
  CoreT<A> o;
  if (a.core) {
    pull_step<R, EQF, A, EDGE>(a, W, i, j, ox, oy, &S, tap);
    o = core_forward_t<A>(ConstsOf<A>::of(a), S, tap);
    m = stored_mom<R, EQF, A>(ConstsOf<A>::of(a), o);
    ph = o.phi;
  } else {
    m = win_mom<R, EQF, A, typename WinT<R>::A2>(W.a[oy][ox], W.b[oy][ox], W.c[oy][ox]);
    ph = (A)W.phi[oy][ox];
  }
  Mom<A> lm = {};
  A lphi = A(0);
  A lTrep = A(0);
  if (a.has_next) {
    using S = LSlots<EQF>;
    // ---- reverse of step n+1's pulls from this node: the adjoints of its
    // nine populations, then through the rebuild to its moments
    A lf[9], lg[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int xi = i + CX[k], xj = j + CY[k];
      if (!EDGE || (xi >= 1 && xi <= nx && xj >= 1 && xj <= ny)) {
        lf[k] = lam_f(W, k, ox + CX[k], oy + CY[k]);
        lg[k] = lam_g(W, k, ox + CX[k], oy + CY[k]);
      } else {
        lf[k] = lg[k] = A(0);
      }
    }
    if constexpr (EDGE) {
      // this node's own pull of direction k from a ghost: bounce-back of
      // its opposite population (not for the two frozen corner ghosts)
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const int si = i - CX[k], sj = j - CY[k];
        if (si < 1 || si > nx || sj < 1 || sj > ny) {
          const bool frozen = (k == 6 && si == nx + 1 && sj == 0) ||
                              (k == 7 && si == nx + 1 && sj == ny + 1);
          if (!frozen) lf[OPP[k]] += lam_f(W, k, ox, oy);
        }
      }
      // ghost rows of g: row 0 is a copy of row 1, row Ny+1 of row Ny
      if (j == 1) {
#pragma unroll
        for (int k = 0; k < 9; ++k)
          if (CY[k] > 0) {
            const int xi = i + CX[k];
            if (xi >= 1 && xi <= nx) lg[k] += lam_g(W, k, ox + CX[k], oy);
          }
      }
      if (j == ny) {
#pragma unroll
        for (int k = 0; k < 9; ++k)
          if (CY[k] < 0) {
            const int xi = i + CX[k];
            if (xi >= 1 && xi <= nx) lg[k] += lam_g(W, k, ox + CX[k], oy);
          }
      }
    }
    rebuild_adjoint<A>(m, lf, lg, &lm);
    // east ghost: equilibrium at this node's averaged T of step n (not
    // with the cold-wall extension: then a constant)
    if (EDGE && i == nx && !a.c.east_cold) {
      for (int xj = j - 1; xj <= j + 1; ++xj) {
        if (xj < 1 || xj > ny) continue;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          if (CX[k] >= 0) continue;  // directions 3, 6, 7 pull from column Nx+1
          const int sj = xj - CY[k];
          const int sjc = sj < 1 ? 1 : (sj > ny ? ny : sj);
          if (sjc != j) continue;
          const int64_t h = (int64_t)j * P + nx + 1;
          lTrep += ghost_dT<A>(k, (A)a.uh[h], (A)a.vh[h]) * lam_g(W, k, ox, oy + (xj - j));
        }
      }
    }
    // ---- phi of step n as a tap of step n+1 (C, E, N, W, S, NE, NW, SW, SE)
    constexpr int DX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    constexpr int DY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    A skp = A(0), skm = A(0);
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      // nodes x whose tap t reads this node: x = y - d, and in rows 1 / Ny
      // also those whose tap t reads the ghost row copied from this node
      const int xi = i - DX[t], xj = j - DY[t];
      if (!EDGE || (xi >= 1 && xi <= nx && xj >= 1 && xj <= ny)) {
        const int wx = ox - DX[t], wy = oy - DY[t];
        skp += W.lam[S::KP][wy][wx];
        skm += W.lam[S::KM][wy][wx];
      }
      if (EDGE && ((j == 1 && DY[t] < 0) || (j == ny && DY[t] > 0))) {
        if (xi >= 1 && xi <= nx) {
          const int wx = ox - DX[t];
          skp += W.lam[S::KP][oy][wx];
          skm += W.lam[S::KM][oy][wx];
        }
      }
    }
    lphi = A(0.5) * (A(1) + ph) * skp - A(0.5) * (A(1) - ph) * skm;
    lphi += (A)plane_at<R>(a.next, L_PC, q);
  }
  if (a.core) {
    // ---- loss seed at a probe node (averaged T after step n); column 0
    // of the index plane flags the rows that hold probes, so the other
    // rows skip the per-node index load
    if (a.seeds && a.probe_idx[(int64_t)j * P] >= 0) {
      const int p = a.probe_idx[q];
      if (p >= 0) lTrep += (A)a.seeds[p];
    }
    // averaged T = T_crit + (dT + dT2)/2: half on the stored dT, half on dT2
    lm.th += A(0.5) * lTrep;
    // ---- reverse of step n at this node
    PullSumsT<A> lS;
    A lkp, lkm, lpc;
    core_backward<A>(ConstsOf<A>::of(a), S, tap, o, lm, lphi, A(0.5) * lTrep, &lS, &lkp, &lkm,
                     &lpc, g);
    const AdjPlanes& v = a.cur;
    plane_set<R>(v, L_RHO, q, lS.rho);
    plane_set<R>(v, L_JX, q, lS.jx);
    plane_set<R>(v, L_JY, q, lS.jy);
    if constexpr (!EQF) {
      plane_set<R>(v, L_FXX, q, lS.fxx);
      plane_set<R>(v, L_FYY, q, lS.fyy);
      plane_set<R>(v, L_FXY, q, lS.fxy);
    }
    plane_set<R>(v, L_DT, q, lS.dT);
    plane_set<R>(v, L_GXX, q, lS.gxx);
    plane_set<R>(v, L_GYY, q, lS.gyy);
    plane_set<R>(v, L_GXY, q, lS.gxy);
    plane_set<R>(v, L_KP, q, lkp);
    plane_set<R>(v, L_KM, q, lkm);
    plane_set<R>(v, L_PC, q, lpc);
  }
  // this node's pulls from the west ghost (column 1): equilibria at the
  // hot-wall temperature (left for last: no register holds the sum while
  // the rest of the node is reversed)
  if (EDGE && a.has_next && i == 1) {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (CX[k] == 1) {
        const int sj = j - CY[k];
        const int sjc = sj < 1 ? 1 : (sj > ny ? ny : sj);
        const int64_t h = (int64_t)sjc * P;
        g[2] += ghost_dT<A>(k, (A)a.uh[h], (A)a.vh[h]) * lam_g(W, k, ox, oy);
      }
  }
  if (!a.core && i <= a.interf) {
    // initial state: T = t_hot in the liquid strip (initialization.f90:15-26)
    g[2] += lm.th;
  }
}

// Tile order.  gfx950 dispatches workgroup b to XCD b % 8 and each XCD has
// its own L2.  XCD x takes the x-th of eight horizontal bands of tile rows
// and walks it in groups of kGroup tile columns, row by row within a group:
// the cache lines of a window's halo (the 4- or 8-byte columns left and right
// of the tile and the rows above and below it) were then fetched into the
// same L2 by a tile a few places earlier on the same XCD, instead of by
// another XCD.  Returns false for the padding workgroups of the shorter bands.
constexpr int kGroup = 8;
__device__ __forceinline__ bool tile_of(int ntx, int nty, int b, int* tx, int* ty) {
  const int x = b % NXCD, s = b / NXCD;
  const int r0 = (int)((int64_t)x * nty / NXCD), r1 = (int)((int64_t)(x + 1) * nty / NXCD);
  const int rows = r1 - r0;
  if (s >= rows * ntx) return false;
  const int g = s / (kGroup * rows);
  const int r = s - g * kGroup * rows;
  const int gw = ntx - g * kGroup < kGroup ? ntx - g * kGroup : kGroup;
  *ty = r0 + r / gw;
  *tx = g * kGroup + r % gw;
  return true;
}
inline int tile_grid(int ntx, int nty) { return NXCD * ((nty + NXCD - 1) / NXCD) * ntx; }

// Waves per SIMD that the LDS of a CU (160 KiB) holds for k_adjoint: the
// register budget asked of the compiler (more registers would not raise the
// occupancy).
template <typename R, bool EQF>
constexpr int adj_waves() {
  constexpr int waves_per_wg = BX * tile_rows<R, EQF>() / 64;
  constexpr int lds = (int)sizeof(Win<R, EQF>) + 3 * 8 * waves_per_wg;
  constexpr int w = (160 * 1024 / lds) * waves_per_wg / 4;
  return w > 8 ? 8 : (w < 1 ? 1 : w);
}

template <typename R, bool EQF>
__global__ __launch_bounds__((BX * tile_rows<R, EQF>()), (adj_waves<R, EQF>())) void k_adjoint(
    AdjArgs a) {
  constexpr int BY = tile_rows<R, EQF>(), WY = BY + 2;
  const int nx = a.gm.nx, ny = a.gm.ny;
  const int64_t P = a.gm.pitch;
  int tx, ty;
  if (!tile_of(a.ntx, a.nty, (int)blockIdx.x, &tx, &ty)) return;  // whole workgroup
  const int i0 = tx * BX + 1, j0 = ty * BY + 1;
  const int ox = threadIdx.x + 1, oy = threadIdx.y + 1;  // this node in the window
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  // ---- the window: rows j0-1 .. j0+BY, columns i0-1 .. i0+BX, clamped to
  // the planes' extent (clamped entries are never read: every read below is
  // guarded as in the global form)
  __shared__ Win<R, EQF> W;
  {
    using S = LSlots<EQF>;
    for (int e = threadIdx.y * BX + threadIdx.x; e < WX * WY; e += BX * BY) {
      const int ex = e % WX, ey = e / WX;
      int xi = i0 - 1 + ex, xj = j0 - 1 + ey;
      xi = xi > nx + 1 ? nx + 1 : xi;
      xj = xj > ny + 1 ? ny + 1 : xj;
      const int64_t x = (int64_t)xj * P + xi;
      if (a.has_next)
#pragma unroll
        for (int sl = 0; sl < S::N; ++sl) W.lam[sl][ey][ex] = plane_at<R>(a.next, S::plane[sl], x);
      W.a[ey][ex] = WinT<R>::a2(a.sprev.a[x]);
      W.b[ey][ex] = b_at<R, EQF>(a.sprev.b, x);
      W.c[ey][ex] = c_at<R, EQF>(a.sprev.c, x);
      W.phi[ey][ex] = (typename WinT<R>::P)a.sprev.phi[x];
    }
  }
  __syncthreads();
  // parameter adjoints of this thread's node, in the arithmetic type (FP32
  // terms are summed in FP32 over the 64 nodes of a wave, in FP64 beyond)
  R g[3] = {R(0), R(0), R(0)};
  // tiles whose nodes all have four interior neighbours skip every
  // boundary test (a workgroup-uniform branch)
  const bool edge = !(i0 > 1 && i0 + BX <= nx && j0 > 1 && j0 + BY <= ny);
  if (i <= nx && j <= ny) {
    if (edge)
      adjoint_node<R, EQF, true>(a, W, i, j, ox, oy, g);
    else
      adjoint_node<R, EQF, false>(a, W, i, j, ox, oy, g);
  }
  // ---- parameter adjoints: workgroup sum into this workgroup's own slot
  // (butterfly within each wave, then the waves in order: deterministic)
  constexpr int NW = BX * BY / 64;
  __shared__ double red[3][NW];
  const int t = threadIdx.y * BX + threadIdx.x;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    for (int m = 32; m > 0; m >>= 1) g[c] += __shfl_xor(g[c], m, 64);
  if ((t & 63) == 0)
    for (int c = 0; c < 3; ++c) red[c][t >> 6] = (double)g[c];
  __syncthreads();
  if (t == 0) {
    const int64_t b = (int64_t)ty * a.ntx + tx;
    for (int c = 0; c < 3; ++c) {
      double sum = 0.0;
      for (int w = 0; w < NW; ++w) sum += red[c][w];
      a.partial[3 * b + c] += sum;
    }
  }
}

// Deterministic sum of the per-workgroup slots.
__global__ void k_adj_sum(const double* partial, int n, double* out) {
  __shared__ double red[3][256];
  const int t = threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  for (int b = t; b < n; b += 256)
    for (int c = 0; c < 3; ++c) s[c] += partial[3 * b + c];
  for (int c = 0; c < 3; ++c) red[c][t] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int c = 0; c < 3; ++c) red[c][t] += red[c][t + w];
    __syncthreads();
  }
  if (t == 0)
    for (int c = 0; c < 3; ++c) out[c] = red[c][0];
}

// probe index plane: -1 everywhere (k_probe_clear), then p at probe p's
// node and 0 in column 0 of its row (k_probe_set, a second launch on the
// same stream): the forward kernels read the row flag once per row
__global__ void k_probe_clear(int* idx, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) idx[q] = -1;
}

__global__ void k_probe_set(int* idx, const int* pij, int np, int64_t pitch) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) {
    idx[(int64_t)pij[2 * p + 1] * pitch + pij[2 * p]] = p;
    idx[(int64_t)pij[2 * p + 1] * pitch] = 0;
  }
}

}  // namespace

// ------------------------------------------------------------------ host

struct AdjCtx {
  int nsteps = 0, np = 0;
  int seg = 0;    // inner segment: every state of one is stored
  int seg1 = 0;   // outer segment (two levels), 0: one level
  int nckpt = 0;  // checkpoints of the forward run (every seg1, or seg, steps)
  int nckpt2 = 0; // checkpoints of one outer segment (two levels)
  int nsnap = 0;  // binomial schedule: snapshot slots (state 0's included); 0: levels
  void* pool = nullptr;
  size_t bytes = 0;
  std::vector<StatePtrs> slot;   // states 1..seg of an inner segment (slot[0] spare)
  std::vector<StatePtrs> ckpt;   // checkpoints of the forward run
  std::vector<StatePtrs> ckpt2;  // checkpoints of the outer segment being reversed
  std::vector<StatePtrs> snap;   // binomial: snapshot slots
  StatePtrs work[2] = {};        // binomial: states being advanced
  std::vector<int> chain;        // binomial: steps stored by the forward run (state
                                 // chain[k] in snap[k + 1], state 0 in snap[0])
  AdjPlanes lam[2];
  int* probe_idx = nullptr;
  int* probe_ij = nullptr;
  double* probe_T = nullptr;  // nsteps x np
  double* seeds = nullptr;    // nsteps x np
  double* macro[4] = {};      // scratch macroscopic planes for recomputed steps
  double* partial = nullptr;
  double* out3 = nullptr;
  int nblk = 0;
  int64_t recomputed = 0;     // forward steps the last backward pass ran
  bool forward_done = false;
};

namespace {

// ---- binomial checkpointing (Griewank's "revolve" schedule, written from
// the published method: A. Griewank, A. Walther, ACM TOMS 26 (2000) 19-45).
// Reversing l steps from a held state with s free snapshot slots costs, at
// best, T(l, s) = r l - C(S + r, S + 1) forward steps, S = s + 1 and r the
// least integer with C(S + r, S) >= l (T(l, 0) = l (l - 1)/2: every state
// advanced to from the held one).  An optimal schedule advances to a split
// m, stores that state, reverses [m, l) with s - 1 slots, frees the slot and
// reverses [0, m) with s slots; m minimises m + T(l - m, s - 1) + T(m, s).

int64_t rev_cost(int64_t l, int64_t s) {
  if (l <= 1) return 0;
  const int64_t S = s + 1;
  int64_t r = 0, beta = 1;  // beta = C(S + r, S)
  while (beta < l) {
    ++r;
    beta = beta * (S + r) / r;
  }
  return r * l - beta * r / (S + 1);  // C(S + r, S + 1) = C(S + r, S) r / (S + 1)
}

int64_t rev_split(int64_t l, int64_t s) {
  int64_t best = 1, cost = -1;
  for (int64_t m = 1; m < l; ++m) {
    const int64_t c = m + rev_cost(l - m, s - 1) + rev_cost(m, s);
    if (cost < 0 || c < cost) {
      cost = c;
      best = m;
    }
  }
  return best;
}

// Steps the forward run stores (the splits of the schedule's rightmost
// branch, which the recorded forward reaches anyway).
std::vector<int> binomial_chain(int nsteps, int nsnap) {
  std::vector<int> chain;
  int a = 0, freeslots = nsnap - 1;
  while (nsteps - a >= 2 && freeslots >= 1) {
    a += (int)rev_split(nsteps - a, freeslots);
    chain.push_back(a);
    --freeslots;
  }
  return chain;
}

// The operations of reversing steps a+1..b (step k maps state k-1 to k)
// from state a held in snapshot slot sa, with the snapshot slots `freev`
// free: `advance(src_state_slot_or_-1, from, to, dst)` and `adjoint(step,
// state_slot)` callbacks; work slots are -1 - w.  Returns false when a
// callback fails.
template <typename Advance, typename Adjoint>
bool reverse_steps(int a, int b, int sa, std::vector<int>& freev, Advance& advance,
                   Adjoint& adjoint) {
  const int l = b - a;
  if (l == 1) return adjoint(b, sa);
  if (freev.empty()) {
    for (int k = b; k > a; --k) {
      int st = sa;
      if (k - 1 > a) {
        st = -1 - ((k - 1 - a - 1) & 1);  // the work slot the last step lands in
        if (!advance(sa, a, k - 1, st)) return false;
      }
      if (!adjoint(k, st)) return false;
    }
    return true;
  }
  const int m = a + (int)rev_split(l, (int64_t)freev.size());
  const int sm = freev.back();
  freev.pop_back();
  if (!advance(sa, a, m, sm)) return false;
  if (!reverse_steps(m, b, sm, freev, advance, adjoint)) return false;
  freev.push_back(sm);
  return reverse_steps(a, m, sa, freev, advance, adjoint);
}

}  // namespace

// Forward steps the backward pass of a binomial schedule runs (the whole
// schedule's advances minus those of the stored chain, which the recorded
// forward run did).
int64_t binomial_recompute(int nsteps, int nsnap) {
  const std::vector<int> chain = binomial_chain(nsteps, nsnap);
  int64_t n = 0;
  auto advance = [&](int, int from, int to, int) {
    n += to - from;
    return true;
  };
  auto adjoint = [&](int, int) { return true; };
  int right = nsteps;
  for (int k = (int)chain.size(); k >= 0; --k) {
    const int a = k == 0 ? 0 : chain[(size_t)k - 1];
    std::vector<int> freev;
    for (int f = nsnap - 1; f > k; --f) freev.push_back(f);
    reverse_steps(a, right, k, freev, advance, adjoint);
    right = a;
  }
  return n;
}

namespace {

// bytes per node of the B and C moment planes of the handle's storage
size_t b_bytes(const bgk_handle* h) {
  return h->prm.precision == BGK_FP64_MOM ? sizeof(double4) : sizeof(float4);
}
size_t lam_bytes(const bgk_handle* h) {
  return h->prm.precision == BGK_FP64_MOM ? sizeof(double) : sizeof(float);
}
size_t c_bytes(const bgk_handle* h) {
  const bool f64 = h->prm.precision == BGK_FP64_MOM;
  if (h->mom_eqf) return f64 ? sizeof(double) : sizeof(float);
  return f64 ? sizeof(double4) : sizeof(float4);
}
// rows of the adjoint kernel's tile for the handle's instantiation
int tile_rows_of(const bgk_handle* h) {
  const bool f64 = h->prm.precision == BGK_FP64_MOM;
  if (h->mom_eqf) return f64 ? tile_rows<double, true>() : tile_rows<float, true>();
  return f64 ? tile_rows<double, false>() : tile_rows<float, false>();
}

struct Bump {
  char* base;
  size_t off;
  template <typename T>
  T* take(size_t n) {
    off = (off + 255) / 256 * 256;
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += n * sizeof(T);
    return p;
  }
  // a lattice plane of elements of esize bytes placed as the handle places
  // its own: element 1 (column 1 of row 0) on a 256-byte boundary, so that
  // with the 64-element pitch every row's column 1 is (the kernels' row
  // accesses then cover whole cache lines)
  template <typename T>
  T* plane(size_t n, size_t esize = sizeof(T)) {
    off = (off + 255) / 256 * 256 + 256 - esize;
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += n * esize;
    return p;
  }
};

void carve_state(Bump& b, const bgk_handle* h, StatePtrs* s) {
  const size_t n = (size_t)h->plane_elems;
  s->a = b.plane<double2>(n);
  s->b = b.plane<char>(n, b_bytes(h));
  s->c = b.plane<char>(n, c_bytes(h));
  s->phi = b.plane<double>(n);
  s->te = b.take<double>((size_t)h->ny + 2);
}

// device bytes of everything but the stored states
size_t carve_fixed(bgk_handle* h, AdjCtx* a, Bump& b) {
  const size_t n = (size_t)h->plane_elems;
  const size_t before = b.off;
  for (int l = 0; l < 2; ++l)
    for (int v = 0; v < 13; ++v) {
      const bool zero = h->mom_eqf && (v == L_FXX || v == L_FYY || v == L_FXY);
      a->lam[l].v[v] = zero ? nullptr : b.plane<char>(n, lam_bytes(h));
    }
  a->probe_idx = b.plane<int>(n);
  a->probe_ij = b.take<int>(2 * (size_t)std::max(a->np, 1));
  a->probe_T = b.take<double>((size_t)a->nsteps * std::max(a->np, 1));
  a->seeds = b.take<double>((size_t)a->nsteps * std::max(a->np, 1));
  for (int m = 0; m < 4; ++m) a->macro[m] = b.plane<double>(n);
  a->partial = b.take<double>(3 * (size_t)a->nblk);
  a->out3 = b.take<double>(3);
  return b.off - before;
}

size_t carve_ctx(bgk_handle* h, AdjCtx* a, char* base) {
  Bump b{base, 0};
  if (a->nsnap > 0) {
    a->slot.clear();
    a->ckpt.clear();
    a->ckpt2.clear();
    a->snap.assign((size_t)a->nsnap, StatePtrs{});
    for (auto& s : a->snap) carve_state(b, h, &s);
    for (auto& s : a->work) carve_state(b, h, &s);
  } else {
    a->slot.assign((size_t)a->seg + 1, StatePtrs{});
    a->ckpt.assign((size_t)a->nckpt, StatePtrs{});
    a->ckpt2.assign((size_t)a->nckpt2, StatePtrs{});
    for (auto& s : a->slot) carve_state(b, h, &s);
    for (auto& s : a->ckpt) carve_state(b, h, &s);
    for (auto& s : a->ckpt2) carve_state(b, h, &s);
  }
  carve_fixed(h, a, b);
  return b.off + 256;
}

// device bytes of one stored state and of the rest of a context
size_t state_bytes(bgk_handle* h) {
  Bump b{nullptr, 0};
  StatePtrs s;
  carve_state(b, h, &s);
  carve_state(b, h, &s);
  Bump b1{nullptr, 0};
  carve_state(b1, h, &s);
  return b.off - b1.off;
}

hipError_t copy_state(bgk_handle* h, const StatePtrs& dst, const double2* a, const void* bb,
                      const void* c, const double* phi, const double* te) {
  const size_t n = (size_t)h->plane_elems;
  hipError_t e = hipMemcpyAsync(dst.a, a, n * sizeof(double2), hipMemcpyDeviceToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dst.b, bb, n * b_bytes(h), hipMemcpyDeviceToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dst.c, c, n * c_bytes(h), hipMemcpyDeviceToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dst.phi, phi, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dst.te, te, ((size_t)h->ny + 2) * sizeof(double), hipMemcpyDeviceToDevice,
                       h->stream);
  return e;
}

// adjoint of the step from state `sprev` (core = 1), or the initial state's
// share (core = 0, sprev = the initial state)
hipError_t launch_adjoint(bgk_handle* h, AdjCtx* x, const StatePtrs& sprev, int lam_next,
                          int has_next, int core, const double* seeds) {
  AdjArgs a;
  a.sprev = sprev;
  a.uh = h->u;
  a.vh = h->v;
  a.next = x->lam[lam_next];
  a.cur = x->lam[1 - lam_next];
  a.has_next = has_next;
  a.core = core;
  a.probe_idx = x->probe_idx;
  a.seeds = seeds;
  a.partial = x->partial;
  a.c = h->c;
  a.cf = consts_f(h->c);
  a.gm = h->geom;
  a.interf = h->nx / 40;
  const int by = tile_rows_of(h);
  a.ntx = (h->nx + BX - 1) / BX;
  a.nty = (h->ny + by - 1) / by;
  const dim3 grd((unsigned)tile_grid(a.ntx, a.nty));
  const bool f64 = h->prm.precision == BGK_FP64_MOM;
  if (h->mom_eqf && f64)
    hipLaunchKernelGGL((k_adjoint<double, true>), grd, dim3(BX, by), 0, h->stream, a);
  else if (h->mom_eqf)
    hipLaunchKernelGGL((k_adjoint<float, true>), grd, dim3(BX, by), 0, h->stream, a);
  else if (f64)
    hipLaunchKernelGGL((k_adjoint<double, false>), grd, dim3(BX, by), 0, h->stream, a);
  else
    hipLaunchKernelGGL((k_adjoint<float, false>), grd, dim3(BX, by), 0, h->stream, a);
  return hipGetLastError();
}

// one forward step from state `in` to state `out` (explicit buffers)
hipError_t fwd_step(bgk_handle* h, const StatePtrs& in, const StatePtrs& out, double* probe_out,
                    const int* probe_idx, bool write_all, double* const* macro) {
  mom::MomStep st;
  st.in = mom::MomIn{in.a, in.b, in.c};
  st.out = mom::MomOut{out.a, out.b, out.c};
  st.pa = in.phi;
  st.pb = out.phi;
  st.te_in = in.te;
  st.te_out = out.te;
  st.write_all = write_all;
  st.probe_idx = probe_idx;
  st.probe_out = probe_out;
  for (int m = 0; m < 4; ++m) st.macro[m] = macro ? macro[m] : nullptr;
  return launch_mom_step(h, st);
}

}  // namespace

void adj_free(bgk_handle* h) {
  if (!h->adj) return;
  if (h->adj->pool) (void)hipFree(h->adj->pool);
  delete h->adj;
  h->adj = nullptr;
}

int64_t adj_state_bytes(bgk_handle* h) { return (int64_t)state_bytes(h); }

int64_t adj_bytes(bgk_handle* h, int nsteps, int np, int nsnap) {
  AdjCtx x;
  x.nsteps = nsteps;
  x.np = np;
  x.nsnap = nsnap;
  const int by = tile_rows_of(h);
  x.nblk = (int)(((h->nx + BX - 1) / BX) * ((h->ny + by - 1) / by));
  return (int64_t)carve_ctx(h, &x, nullptr);
}

// nsnap > 0: binomial schedule with nsnap snapshot slots; else the level
// schedule (seg_outer > 0: two levels)
hipError_t adj_setup(bgk_handle* h, int nsteps, int seg_outer, int seg, int nsnap, int np,
                     const int32_t* pi, const int32_t* pj) {
  adj_free(h);
  AdjCtx* x = new (std::nothrow) AdjCtx();
  if (!x) return hipErrorOutOfMemory;
  x->nsteps = nsteps;
  x->np = np;
  x->nsnap = nsnap;
  if (nsnap > 0) {
    x->chain = binomial_chain(nsteps, nsnap);
  } else {
    x->seg = seg;
    x->seg1 = seg_outer;
    const int lvl1 = seg_outer > 0 ? seg_outer : seg;
    x->nckpt = (nsteps + lvl1 - 1) / lvl1;
    // (the first checkpoint of an outer segment is the outer checkpoint)
    x->nckpt2 = seg_outer > 0 ? (std::min(seg_outer, nsteps) + seg - 1) / seg - 1 : 0;
  }
  const int by = tile_rows_of(h);
  x->nblk = (int)(((h->nx + BX - 1) / BX) * ((h->ny + by - 1) / by));
  x->bytes = carve_ctx(h, x, nullptr);
  hipError_t e = hipMalloc(&x->pool, x->bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // (the failed allocation's error is reported here)
    delete x;
    return e;
  }
  carve_ctx(h, x, static_cast<char*>(x->pool));
  h->adj = x;
  std::vector<int> ij(2 * (size_t)std::max(np, 1), 0);
  for (int p = 0; p < np; ++p) {
    ij[2 * p] = pi[p];
    ij[2 * p + 1] = pj[p];
  }
  e = hipMemsetAsync(x->pool, 0, x->bytes, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(x->probe_ij, ij.data(), ij.size() * sizeof(int), hipMemcpyHostToDevice,
                       h->stream);
  if (e == hipSuccess) {
    const int64_t n = h->plane_elems;
    hipLaunchKernelGGL(k_probe_clear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                       x->probe_idx, n);
    if (np > 0)
      hipLaunchKernelGGL(k_probe_set, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, h->stream,
                         x->probe_idx, x->probe_ij, np, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e;
}

bool adj_forward_done(const bgk_handle* h) { return h->adj && h->adj->forward_done; }

void adj_invalidate(bgk_handle* h) {
  if (h->adj) h->adj->forward_done = false;
}

hipError_t adj_forward(bgk_handle* h, double* probe_T_host) {
  AdjCtx* x = h->adj;
  x->forward_done = false;
  // halo velocities of the recorded run for the recomputed steps (the
  // ghost columns of u, v are constant over a run: they are the ones the
  // state holds now)
  hipError_t e = hipMemcpyAsync(x->macro[1], h->u, (size_t)h->plane_elems * sizeof(double),
                                hipMemcpyDeviceToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(x->macro[2], h->v, (size_t)h->plane_elems * sizeof(double),
                       hipMemcpyDeviceToDevice, h->stream);
  // east column of the initial state (T(Nx, :) as the first BC_g reads it)
  if (e == hipSuccess) e = launch_teast_prologue(h);
  if (e != hipSuccess) return e;
  int te = 0;  // teast[te] holds the current column
  size_t next_chain = 0;
  for (int n = 0; n < x->nsteps; ++n) {
    const int a = h->cur, b = 1 - h->cur;
    StatePtrs in{h->ma[a], h->mb[a], h->mc[a], h->phi[h->pc], h->teast[te]};
    StatePtrs out{h->ma[b], h->mb[b], h->mc[b], h->phi[1 - h->pc], h->teast[1 - te]};
    const StatePtrs* keep = nullptr;
    if (x->nsnap > 0) {
      if (n == 0) {
        keep = &x->snap[0];
      } else if (next_chain < x->chain.size() && x->chain[next_chain] == n) {
        keep = &x->snap[next_chain + 1];
        ++next_chain;
      }
    } else {
      const int lvl1 = x->seg1 > 0 ? x->seg1 : x->seg;
      if (n % lvl1 == 0) keep = &x->ckpt[(size_t)(n / lvl1)];
    }
    if (keep) {
      e = copy_state(h, *keep, in.a, in.b, in.c, in.phi, in.te);
      if (e != hipSuccess) return e;
    }
    e = fwd_step(h, in, out, x->probe_T + (size_t)n * x->np, x->probe_idx, n == x->nsteps - 1,
                 nullptr);
    if (e != hipSuccess) return e;
    h->cur = b;
    h->pc = 1 - h->pc;
    te = 1 - te;
  }
  // the handle's teast convention: bgk_step recomputes it from T at entry
  if (x->np > 0)
    e = hipMemcpyAsync(probe_T_host, x->probe_T, (size_t)x->nsteps * x->np * sizeof(double),
                       hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) x->forward_done = true;
  return e;
}

void adj_info(const bgk_handle* h, int64_t* bytes, int32_t* n_states) {
  const AdjCtx* x = h->adj;
  *bytes = (int64_t)x->bytes;
  *n_states = x->nsnap > 0 ? (int32_t)(x->snap.size() + 2)
                           : (int32_t)(x->slot.size() + x->ckpt.size() + x->ckpt2.size());
}

int64_t adj_recomputed(const bgk_handle* h) { return h->adj ? h->adj->recomputed : 0; }

hipError_t adj_backward(bgk_handle* h, const double* seeds_host, double grad[3]) {
  AdjCtx* x = h->adj;
  hipError_t e = hipSuccess;
  x->recomputed = 0;
  if (x->np > 0)
    e = hipMemcpyAsync(x->seeds, seeds_host, (size_t)x->nsteps * x->np * sizeof(double),
                       hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(x->partial, 0, 3 * (size_t)x->nblk * sizeof(double), h->stream);
  int ln = 0;  // lam[ln] holds the adjoint of the step after the current one
  bool has_next = false;
  // adjoint of step n (1-based) from the state before it
  auto adjoint_step = [&](int n, const StatePtrs& before) {
    const double* sd = x->np > 0 ? x->seeds + (size_t)(n - 1) * x->np : nullptr;
    e = launch_adjoint(h, x, before, ln, has_next ? 1 : 0, 1, sd);
    ln = 1 - ln;
    has_next = true;
    return e == hipSuccess;
  };
  auto step = [&](const StatePtrs& in, const StatePtrs& out) {
    ++x->recomputed;
    e = fwd_step(h, in, out, nullptr, nullptr, false, x->macro);
    return e == hipSuccess;
  };
  if (x->nsnap > 0) {
    // binomial schedule: slots >= 0 are snapshots, -1 and -2 the work states
    auto at = [&](int sl) -> const StatePtrs& { return sl >= 0 ? x->snap[(size_t)sl] : x->work[-1 - sl]; };
    auto advance = [&](int src, int from, int to, int dst) {
      for (int t = from; t < to; ++t) {
        const StatePtrs& in = t == from ? at(src) : x->work[(t - from - 1) & 1];
        const StatePtrs& out = t == to - 1 ? at(dst) : x->work[(t - from) & 1];
        if (!step(in, out)) return false;
      }
      return true;
    };
    auto adjoint = [&](int n, int sl) { return adjoint_step(n, at(sl)); };
    int right = x->nsteps;
    for (int k = (int)x->chain.size(); k >= 0 && e == hipSuccess; --k) {
      const int a = k == 0 ? 0 : x->chain[(size_t)k - 1];
      std::vector<int> freev;
      for (int f = x->nsnap - 1; f > k; --f) freev.push_back(f);
      reverse_steps(a, right, k, freev, advance, adjoint);
      right = a;
    }
  } else {
    // steps n0..n1-1 from `start` (the state before step n0): recompute every
    // state of the segment (state n0 is `start` itself), then reverse them
    auto inner = [&](const StatePtrs& start, int n0, int n1) {
      auto st = [&](int n) -> const StatePtrs& { return n == n0 ? start : x->slot[n - n0]; };
      for (int n = n0 + 1; n < n1 && e == hipSuccess; ++n) step(st(n - 1), st(n));
      for (int n = n1; n > n0 && e == hipSuccess; --n) adjoint_step(n, st(n - 1));
    };
    if (x->seg1 == 0) {
      for (int s = x->nckpt - 1; s >= 0 && e == hipSuccess; --s)
        inner(x->ckpt[s], s * x->seg, std::min((s + 1) * x->seg, x->nsteps));
    } else {
      for (int s1 = x->nckpt - 1; s1 >= 0 && e == hipSuccess; --s1) {
        const int m0 = s1 * x->seg1, m1 = std::min(m0 + x->seg1, x->nsteps);
        const int ns2 = (m1 - m0 + x->seg - 1) / x->seg;
        // checkpoints of this outer segment: c2(k) = state m0 + k seg, c2(0)
        // the outer checkpoint itself; the steps in between go through
        // slot[0] and slot[1]
        auto c2 = [&](int k) -> const StatePtrs& { return k == 0 ? x->ckpt[s1] : x->ckpt2[k - 1]; };
        for (int k = 1; k < ns2 && e == hipSuccess; ++k) {
          const StatePtrs* src = &c2(k - 1);
          for (int i = 0; i < x->seg && e == hipSuccess; ++i) {
            const StatePtrs* dst = (i == x->seg - 1) ? &x->ckpt2[k - 1] : &x->slot[i & 1];
            step(*src, *dst);
            src = dst;
          }
        }
        for (int s2 = ns2 - 1; s2 >= 0 && e == hipSuccess; --s2)
          inner(c2(s2), m0 + s2 * x->seg, std::min(m0 + (s2 + 1) * x->seg, m1));
      }
    }
  }
  // the initial state's share (hot-wall temperature of the liquid strip)
  const StatePtrs& s0 = x->nsnap > 0 ? x->snap[0] : x->ckpt[0];
  if (e == hipSuccess) e = launch_adjoint(h, x, s0, ln, has_next ? 1 : 0, 0, nullptr);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, h->stream, x->partial, x->nblk, x->out3);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(grad, x->out3, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e;
}

}  // namespace bgk
