// bgk_fused_mom.hip -- the fused time step with moment storage.
//
// The node state is the ten moments of bgk_mom.h (rho, T - T_crit in FP64;
// u, v and the scaled non-equilibrium second moments in R = float or
// double), 48 bytes per node with FP32 R (36 with omega_f = 1) instead of
// 144 for FP32 populations, and the pull rebuilds population k of the
// upstream node from its moments.  rho and T stay FP64: the rounding of rho
// (not of u) is what drives the velocity error of FP32 storage, since a
// neighbour's density error enters the momentum balance through
// sum_k c_k w_k rho(x - c_k).  u, v and the non-equilibrium moments carry
// their own relative precision.
//
// Tiling: a workgroup owns W columns and a strip of consecutive rows, and
// walks the strip row by row.  Row j+2 is fetched from HBM while row j is
// computed; rows j-1..j+2 live in an LDS ring, so every node's moments and
// phi are read from HBM once per step (plus the strip's two boundary rows
// and the two halo columns).  phi_prec's 9-point stencil comes from the
// same ring.  The boundary-adjacent nodes run in extra workgroups of the
// same launch with BC_f (no-slip) and BC_g folded in as in bgk_fused.hip:
// ghost rows of g and phi are the clamped rows, own opposite populations
// for bounce-back, the two frozen corner ghosts as constants.
//
// Arithmetic is FP64 throughout.  STEP_7's FP32 temporaries (ft, gt) cannot
// be represented by moments, so this path follows the clean rounding policy;
// beta follows bgk_params.rounding as everywhere.
#include <algorithm>

#include "bgk_mom.h"
#include "bgk_strip.h"

namespace bgk {
namespace {

using namespace mom;

constexpr int TX = 64, TY = 4;  // block of the initialization kernel

struct Macro {
  double *rho, *u, *v, *T;
};

// LDS ring of NS rows of W + 2 nodes (W columns and the two halo columns).
template <int W, int NS, typename R, bool EQF>
struct Ring {
  double2 a[NS][W + 2];
  typename Vec<R>::V4 b[NS][W + 2];
  typename CType<R, EQF>::T c[NS][W + 2];
  double phi[NS][W + 2];
};

template <typename R, bool EQF>
__device__ __forceinline__ Node<R, EQF> load_node(const MomIn& in, const double* __restrict__ pa,
                                                  const Geom& gm, int ic, int jr) {
  Node<R, EQF> n;
  n.a = make_double2(0.0, 0.0);
  n.b = make4<R>(R(0), R(0), R(0), R(0));
  n.c = typename CType<R, EQF>::T{};
  n.phi = 0.0;
  const int64_t q = (int64_t)jr * gm.pitch + ic;
  if (ic >= 1 && ic <= gm.nx) {
    n.a = stream_load(in.a + q);
    n.b = nt_load4<R>(static_cast<const typename Vec<R>::V4*>(in.b) + q);
    n.c = c_stream<R, EQF>(in.c, q);
  }
  if (ic >= 0 && ic <= gm.nx + 1) n.phi = pa[q];
  return n;
}

template <int W, int NS, typename R, bool EQF>
__device__ __forceinline__ void put(Ring<W, NS, R, EQF>& L, int r, int s, const Node<R, EQF>& n) {
  L.a[r][s] = n.a;
  L.b[r][s] = n.b;
  L.c[r][s] = n.c;
  L.phi[r][s] = n.phi;
}

// Node update from its pulled moments and phi_prec taps (C, E, N, W, S, NE,
// NW, SW, SE); stores the node's moments and phi, and T (averaged as
// STEP_7/moments.f90:11-25) at a probe node.  Returns the averaged moments.
template <typename R, bool EQF>
__device__ __forceinline__ NodeOut update_node(const Consts& c, const PullSums& S,
                                               const double tap[9], const MomOut& out,
                                               double* __restrict__ pb, int64_t q,
                                               const int* __restrict__ probe_idx,
                                               double* __restrict__ probe_out) {
  const Core o = core_forward(c, S, tap);
  const R us = (R)o.u, vs = (R)o.v;
  const R pgxx = (R)((1.0 - o.omega_g) * o.qgxx);
  const R pgyy = (R)((1.0 - o.omega_g) * o.qgyy);
  const R pgxy = (R)((1.0 - o.omega_g) * o.qgxy);
  stream_store(out.a + q, o.rho, o.dT);
  using V4 = typename Vec<R>::V4;
  if constexpr (EQF) {
    nt_store4<R>(static_cast<V4*>(out.b) + q, us, vs, pgxx, pgyy);
    __builtin_nontemporal_store(pgxy, static_cast<R*>(out.c) + q);
  } else {
    const double omf1 = 1.0 - c.omega_f;
    nt_store4<R>(static_cast<V4*>(out.b) + q, us, vs, (R)(omf1 * o.qfxx), (R)(omf1 * o.qfyy));
    nt_store4<R>(static_cast<V4*>(out.c) + q, (R)(omf1 * o.qfxy), pgxx, pgyy, pgxy);
  }
  stream_store(&pb[q], o.phi);
  NodeOut r;
  r.rho = o.rho;
  r.Tav = kTcrit + (o.dT + o.dT2) * 0.5;
  r.uav = ((double)us + o.u2) * 0.5;
  r.vav = ((double)vs + o.v2) * 0.5;
  if (probe_out) {
    const int p = probe_idx[q];
    if (p >= 0) probe_out[p] = r.Tav;
  }
  return r;
}

// Interior nodes (1 < i < Nx, 1 < j < Ny) by row strips through the LDS
// ring.  W columns per workgroup, NS rows in the ring (4: one barrier per
// row; 3: two, for less LDS).  Workgroups from sg.nmain on serve the
// boundary-adjacent nodes (mom_edge_nodes).
template <int W, int NS, typename R, bool EQF>
__device__ __forceinline__ void mom_interior(const MomIn& in, const MomOut& out,
                                             const double* __restrict__ pa,
                                             double* __restrict__ pb, const Macro& m,
                                             const Consts& c, const Geom& gm, const StripGrid& sg,
                                             int write_all, const int* __restrict__ probe_idx,
                                             double* __restrict__ probe_out) {
  __shared__ Ring<W, NS, R, EQF> L;
  const int nx = gm.nx, ny = gm.ny;
  const int64_t P = gm.pitch;
  int strip, seg;
  if (!strip_of(sg, (int)blockIdx.x, &strip, &seg)) return;  // whole workgroup
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
    put(L, slot(jj), t + 1, load_node<R, EQF>(in, pa, gm, i, clampj(jj)));
    if (t < 2) put(L, slot(jj), hslot, load_node<R, EQF>(in, pa, gm, hcol, clampj(jj)));
  }
  // rows j0 .. j0+63 of the strip that hold a probe (the flag in column 0
  // of the probe index plane): one load per lane here instead of a load and
  // its wait in every row
  uint64_t pmask = 0;
  if (probe_out) {
    const int jr = min(j0 + (t & 63), j1);
    pmask = __ballot(probe_idx[(int64_t)jr * P] >= 0);
  }
  __syncthreads();

  for (int j = j0; j <= j1; ++j) {
    const bool pre = (j + 2 <= j1 + 1);
    Node<R, EQF> nn, nh;
    if (pre) {
      nn = load_node<R, EQF>(in, pa, gm, i, clampj(j + 2));
      if (t < 2) nh = load_node<R, EQF>(in, pa, gm, hcol, clampj(j + 2));
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
        rebuild<R, EQF>(k, L.a[r][ss], L.b[r][ss], L.c[r][ss], &fk, &gk);
        pull_add(S, k, fk, gk);
      }
      const double tap[9] = {L.phi[rC][s],     L.phi[rC][s + 1], L.phi[rN][s],
                             L.phi[rC][s - 1], L.phi[rS][s],     L.phi[rN][s + 1],
                             L.phi[rN][s - 1], L.phi[rS][s - 1], L.phi[rS][s + 1]};
      const bool prow = probe_out && (j - j0 < 64 ? ((pmask >> (j - j0)) & 1) != 0
                                                   : probe_idx[(int64_t)j * P] >= 0);
      double* const po = prow ? probe_out : nullptr;
      const NodeOut o = update_node<R, EQF>(c, S, tap, out, pb, q, probe_idx, po);
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

// The boundary-adjacent nodes, one per thread, with BC_f type 4 and BC_g
// (STEP_7/BC_f.f90:62-68,101-107,125-136,181-192; BC_g.f90:9-64) folded into
// the pulls: bounce-back of the node's own opposite populations, the two
// frozen corner ghosts, W/E thermal ghosts at Tin and at the previous
// step's averaged T(Nx, j), N/S ghost rows of g and phi equal to rows Ny / 1;
// then the halo bookkeeping of the macroscopic fields.
template <typename R, bool EQF>
__device__ __forceinline__ void mom_edge_nodes(const MomIn& in, const MomOut& out,
                                               const double* __restrict__ pa,
                                               double* __restrict__ pb,
                                               const double* __restrict__ te_in,
                                               double* __restrict__ te_out, const Macro& m,
                                               const Consts& c, const Geom& gm, int e,
                                               const int* __restrict__ probe_idx,
                                               double* __restrict__ probe_out) {
  const int nx = gm.nx, ny = gm.ny;
  const int64_t P = gm.pitch;
  int i, j;
  if (!edge_node(e, nx, ny, &i, &j)) return;
  const int64_t q = (int64_t)j * P + i;
  const double2 oa = in.a[q];
  const typename Vec<R>::V4 ob = b_at<R, EQF>(in.b, q);
  const typename CType<R, EQF>::T oc = c_at<R, EQF>(in.c, q);
  double f[9], gd[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int si = i - CX[k], sj = j - CY[k];
    const int sjc = sj < 1 ? 1 : (sj > ny ? ny : sj);
    if (si >= 1 && si <= nx) {
      // interior source column: rebuild from the node (si, sjc); for a ghost
      // row this is the N/S copy of g, and f is replaced below
      const int64_t h = (int64_t)sjc * P + si;
      rebuild<R, EQF>(k, in.a[h], b_at<R, EQF>(in.b, h), c_at<R, EQF>(in.c, h), &f[k], &gd[k]);
    } else {
      const int64_t h = (int64_t)sjc * P + si;
      gd[k] = ghost_gdev(k, si == 0 ? c.tin : east_T(c, te_in, sjc), m.u[h], m.v[h]);
    }
    if (si < 1 || si > nx || sj < 1 || sj > ny) {
      if (k == 6 && si == nx + 1 && sj == 0)
        f[k] = c.frozen6;
      else if (k == 7 && si == nx + 1 && sj == ny + 1)
        f[k] = c.frozen7;
      else
        f[k] = rebuild_f<R, EQF>(OPP[k], oa, ob, oc);
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
  const NodeOut o = update_node<R, EQF>(c, S, tap, out, pb, q, probe_idx, probe_out);

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
    m.T[h] = east_T(c, te_in, j);
    pb[h] = pa[h];
    te_out[j] = o.Tav;
    if (s_edge) {
      m.T[nx + 1] = east_T(c, te_in, 1);
      pb[nx + 1] = pa[P + nx + 1];
    }
    if (n_edge) {
      const int64_t hn = (int64_t)(ny + 1) * P + nx + 1;
      m.T[hn] = east_T(c, te_in, ny);
      pb[hn] = pa[(int64_t)ny * P + nx + 1];
    }
  }
  m.rho[q] = o.rho;
  m.T[q] = o.Tav;
  m.u[q] = o.uav;
  m.v[q] = o.vav;
}

template <int W, int NS, typename R, bool EQF>
__global__ __launch_bounds__(W) void k_fused_mom(
    MomIn in, MomOut out, const double* __restrict__ pa, double* __restrict__ pb,
    const double* __restrict__ te_in, double* __restrict__ te_out, Macro m, Consts c, Geom gm,
    StripGrid sg, int write_all, const int* __restrict__ probe_idx,
    double* __restrict__ probe_out) {
  if ((int)blockIdx.x >= sg.nmain) {
    const int e = ((int)blockIdx.x - sg.nmain) * W + (int)threadIdx.x;
    mom_edge_nodes<R, EQF>(in, out, pa, pb, te_in, te_out, m, c, gm, e, probe_idx, probe_out);
    return;
  }
  mom_interior<W, NS, R, EQF>(in, out, pa, pb, m, c, gm, sg, write_all, probe_idx, probe_out);
}

// STEP_7/initialization.f90:15-152 for a moment handle: the initial
// populations are equilibria, i.e. moments with zero non-equilibrium part.
template <typename R, bool EQF>
__global__ void k_initialization_mom(MomOut st, Macro m, double* phi, double* phip, Consts c,
                                     Geom gm, int interf, int west_bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const NodeInit s0 = initial_node(c, gm, i, j, interf, west_bc);
  st.a[q] = make_double2(s0.rho, s0.T - kTcrit);
  static_cast<typename Vec<R>::V4*>(st.b)[q] = make4<R>((R)s0.u, (R)s0.v, R(0), R(0));
  static_cast<typename CType<R, EQF>::T*>(st.c)[q] = typename CType<R, EQF>::T{};
  m.rho[q] = s0.rho;
  m.u[q] = s0.u;
  m.v[q] = s0.v;
  m.T[q] = s0.T;
  phi[q] = s0.phi;
  phip[q] = s0.phi;
}

template <int W, int NS, typename R, bool EQF>
hipError_t launch_mom_t(bgk_handle* h, const MomStep& st) {
  // resident workgroups per CU of this instantiation (LDS and registers)
  static const int per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_mom<W, NS, R, EQF>, W, 0) !=
        hipSuccess)
      n = 0;
    return n;
  }();
  const StripGrid sg =
      make_strip_grid(h->nx, h->ny, W, EQF ? 64 : 32, per_cu * h->n_cu, h->strip_rows);
  const int nedge = edge_count(h->nx, h->ny);
  const int64_t nwg = (int64_t)sg.nmain + (nedge + W - 1) / W;
  Macro m{h->rho, h->u, h->v, h->T};
  if (st.macro[0]) m = Macro{st.macro[0], st.macro[1], st.macro[2], st.macro[3]};
  hipLaunchKernelGGL((k_fused_mom<W, NS, R, EQF>), dim3((unsigned)nwg), dim3(W), 0, h->stream,
                     st.in, st.out, st.pa, st.pb, st.te_in, st.te_out, m, h->c, h->geom, sg,
                     st.write_all ? 1 : 0, st.probe_idx, st.probe_out);
  return hipGetLastError();
}

// FP64 moments take at most 256 columns per workgroup (the 512-wide ring
// of 88-byte nodes exceeds the 160 KiB of LDS).
template <int W, int NS>
hipError_t launch_mom(bgk_handle* h, const MomStep& st) {
  constexpr int W64 = W > 256 ? 256 : W;
  const bool f64 = h->prm.precision == BGK_FP64_MOM;
  if (h->mom_eqf)
    return f64 ? launch_mom_t<W64, NS, double, true>(h, st)
               : launch_mom_t<W, NS, float, true>(h, st);
  return f64 ? launch_mom_t<W64, NS, double, false>(h, st)
             : launch_mom_t<W, NS, float, false>(h, st);
}

template <typename R, bool EQF>
void launch_init(bgk_handle* h, const MomOut& st, dim3 grd) {
  Macro m{h->rho, h->u, h->v, h->T};
  hipLaunchKernelGGL((k_initialization_mom<R, EQF>), grd, dim3(TX, TY), 0, h->stream, st, m,
                     h->phi[h->pc], h->phi[1 - h->pc], h->c, h->geom, h->geom.nx / 40,
                     h->prm.west_bc);
}

}  // namespace

hipError_t launch_mom_step(bgk_handle* h, const MomStep& st) {
  // Tiling (workgroup width W, ring rows NS).  Measured at 4096^2 on MI355X
  // with FP32 moments: omega_f = 1 (36-byte nodes, 64-row strips) W = 512 /
  // NS = 4 57.2 GLUPS, 512/3 53.1, 256/4 53.9 (32-row strips), 128/4 45.1,
  // 64/4 45.8; general layout (48 bytes, 32-row strips) 512/3 45.9, 512/4
  // 45.6, 256/3 42.0.  Wider segments stream longer contiguous runs of each
  // row.  FP64 moments take the same tiling with half the workgroups per CU.
  int v = h->mom_variant;
  if (v == 0)
    v = h->nx >= 2048 ? (h->mom_eqf ? 7 : 8) : h->nx >= 512 ? 5 : h->nx >= 128 ? 3 : 1;
  switch (v) {
    case 1: return launch_mom<64, 4>(h, st);
    case 2: return launch_mom<64, 3>(h, st);
    case 3: return launch_mom<128, 4>(h, st);
    case 4: return launch_mom<128, 3>(h, st);
    case 5: return launch_mom<256, 4>(h, st);
    case 6: return launch_mom<256, 3>(h, st);
    case 7: return launch_mom<512, 4>(h, st);
    default: return launch_mom<512, 3>(h, st);
  }
}

hipError_t launch_fused_step_mom(bgk_handle* h, int step_parity, bool write_all) {
  const int a = h->cur, b = 1 - h->cur;
  MomStep st;
  st.in = MomIn{h->ma[a], h->mb[a], h->mc[a]};
  st.out = MomOut{h->ma[b], h->mb[b], h->mc[b]};
  st.pa = h->phi[h->pc];
  st.pb = h->phi[1 - h->pc];
  st.te_in = h->teast[step_parity & 1];
  st.te_out = h->teast[(step_parity + 1) & 1];
  st.write_all = write_all;
  st.probe_idx = nullptr;
  st.probe_out = nullptr;
  for (int k = 0; k < 4; ++k) st.macro[k] = nullptr;
  hipError_t e = launch_mom_step(h, st);
  if (e == hipSuccess) {
    h->cur = b;
    h->pc = 1 - h->pc;
  }
  return e;
}

hipError_t launch_initialization_mom(bgk_handle* h) {
  const Geom& g = h->geom;
  MomOut st{h->ma[h->cur], h->mb[h->cur], h->mc[h->cur]};
  dim3 grd((unsigned)((g.nx + 2 + TX - 1) / TX), (unsigned)((g.ny + 2 + TY - 1) / TY));
  const bool f64 = h->prm.precision == BGK_FP64_MOM;
  if (h->mom_eqf)
    f64 ? launch_init<double, true>(h, st, grd) : launch_init<float, true>(h, st, grd);
  else
    f64 ? launch_init<double, false>(h, st, grd) : launch_init<float, false>(h, st, grd);
  return hipGetLastError();
}

}  // namespace bgk
