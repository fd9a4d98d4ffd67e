// bgk_fused32.hip -- the fused time step with FP32 population storage.
//
// Storage: f_k - w_k*rho_ref and g_k - w_k*T_crit in FP32, deviations from
// the rest equilibrium at a reference density of the node and at the melting
// temperature (T_crit = 273, STEP_7/phase_field.f90:16).  They are ~1e-4 of
// the populations, so FP32 keeps far more absolute precision than storing f
// and g themselves, and the temperature is most precise where the phase
// field is sensitive to it: at the melting front, where g_k - w_k*T_crit ~ 0.
// phi, the east-wall temperature column and all macroscopic fields stay FP64.
//
// Reference density: rho_ref(j, phi) = rho_s(j) + (1+phi)/2 (rho_l(j) -
// rho_s(j)) from a table of the solid and the liquid density of each row.
// The solid's drag (forces.f90:22-23) keeps the melt at its own density
// (~1.000 against the solid's ~1.024 at Ra 1e7, measured) and buoyancy
// builds a hydrostatic profile in each phase, so this reference follows the
// density to ~4e-5 where a constant rho0 or a row mean misses it by 2.6e-2;
// the density part of a deviation and its FP32 rounding, which drives the
// velocity error through sum_k c_k f_k, shrink accordingly.  The writer of a
// population evaluates the reference at its row and its new phi; the reader
// evaluates it at the source row and the source's phi_prec tap, the same
// inputs.  Every `rebase_every` steps the tables are reset to the rows' mean
// solid and liquid densities and the stored f re-expressed against them
// (k_phase_density, k_rebase).  Halo nodes (only the two corner ghosts of
// BC_f are ever read) keep rho0.
//
// Arithmetic: moments, temperature, phase field, latent heat and forces are
// FP64.  The collision runs in `RC` (double by default): on gfx950 unpacked
// FP32 and FP64 FMA issue at the same rate, so FP64 costs registers, not
// time, while FP32 collision arithmetic loses accuracy at the melting front.
// The algebra is the reference's written on deviations (m0 = rho_ref or
// T_crit, dm = m - m0):
//   feq_k - w_k m0 = w_k (dm + m P_k(u,v)),
//   P_k = 3 c_k.u + 4.5 (c_k.u)^2 - 1.5 u.u,
//   f_k' - w_k m0 = (1-omega) (ft_k - w_k m0) + omega (feq_k - w_k m0).
// The FP32 rounding of ft/gt done by STEP_7 (implicit REAL) is not modelled.
//
// Tiling: one node per thread in the XCD-aware row tiles of bgk_device.h;
// the 9-point phi_prec stencil comes through the L2.  Each phi row is then
// fetched into three XCDs' L2s (rows are interleaved over the XCDs), which
// the HBM counters show as 1.23x the algorithmic reads; staging phi in LDS
// removes that but costs more than it saves, measured at 4096^2 per step:
// these row tiles 0.472 ms; a row-strip kernel with phi in an LDS ring, as
// the moment kernel has, 0.61 ms (its barrier per row exposes the latency
// of the population loads); 64 x 4 and 64 x 8 tiles with a phi window in
// LDS, walked in XCD bands, 0.605 and 0.623 ms (a plane's rows are then
// streamed in 512-byte pieces instead of 2 KB).  Same folding of BC_f/BC_g and the same
// halo bookkeeping as the FP64 kernel (bgk_fused.hip).
#include "bgk_internal.h"

namespace bgk {
namespace {

constexpr int TX = 64, TY = 4;  // block of the whole-plane kernels

struct Macro {
  double *rho, *u, *v, *T;
};

inline Pops32 pops32(float* const* p) {
  Pops32 r;
  for (int k = 0; k < 9; ++k) r.p[k].base = p[k];
  return r;
}

// pair view of buffer b: plane k holds (f_k - w_k rho_ref, g_k - w_k T_crit)
inline Pairs32 pairs32(float* const* fplanes) {
  Pairs32 r;
  for (int k = 0; k < 9; ++k) r.p[k] = reinterpret_cast<float2*>(fplanes[k]);
  return r;
}

// P_k(u,v) of the equilibrium, in the operand order of the reference
// polynomial minus its leading 1.
template <typename R>
__device__ __forceinline__ R poly_k(int k, R u, R v) {
  const R uv = R(1.5) * (u * u + v * v);
  switch (k) {
    case 0: return -uv;
    case 1: return R(3) * u + R(4.5) * (u * u) - uv;
    case 2: return R(3) * v + R(4.5) * (v * v) - uv;
    case 3: return -R(3) * u + R(4.5) * (u * u) - uv;
    case 4: return -R(3) * v + R(4.5) * (v * v) - uv;
    case 5: { const R a = u + v; return R(3) * a + R(4.5) * (a * a) - uv; }
    case 6: { const R a = u - v; return -R(3) * a + R(4.5) * (a * a) - uv; }
    case 7: { const R a = u + v; return -R(3) * a + R(4.5) * (a * a) - uv; }
    default: { const R a = u - v; return R(3) * a + R(4.5) * (a * a) - uv; }
  }
}

// Regularised BGK on deviations: dp = population deviations, dm = moment
// deviation (drho or dT), m = moment (rho or T).  Writes deviations.
template <typename R>
__device__ __forceinline__ void regularised_dev(const Consts& c, R omega, R dm, R m, R u, R v,
                                                const R dp[9], R out[9]) {
  R deq[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) deq[k] = R(kW[k]) * (dm + m * poly_k<R>(k, u, v));
  R d[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) d[k] = dp[k] - deq[k];
  const R a = R(-kCs2), b = R(1.0 - kCs2);
  const R diag = d[5] + d[6] + d[7] + d[8];
  const R pxx = a * (d[0] + d[2] + d[4]) + b * (d[1] + d[3] + diag);
  const R pyy = a * (d[0] + d[1] + d[3]) + b * (d[2] + d[4] + diag);
  const R pxy = d[5] - d[6] + d[7] - d[8];
  const R h0 = a * pxx + a * pyy;
  const R h1 = b * pxx + a * pyy;
  const R h2 = a * pxx + b * pyy;
  const R hd = b * pxx + b * pyy;
  const R hx = R(2) * pxy;
  const R h[9] = {h0, h1, h2, h1, h2, hd + hx, hd - hx, hd + hx, hd - hx};
  const R om1 = R(1) - omega;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const R t = deq[k] + R(coef_of(k)) * h[k];
    out[k] = om1 * t + omega * deq[k];
  }
}

// Equilibrium deviation w_k*(m - m0) + w_k*m*P_k for a boundary ghost (FP64).
__device__ __forceinline__ double ghost_dev(const Consts& c, int k, double m, double m0, double u,
                                            double v) {
  return kW[k] * (m - m0) + kW[k] * m * poly_k<double>(k, u, v);
}

// What a node update leaves besides its stored populations: the new phi,
// the pre-force moments momentsPre gave (dT2 = T2 - T_crit, u2, v2) for the
// averages of STEP_7/moments.f90:11-25.
struct Node32 {
  double phi, dT2, u2, v2;
};

// momentsPre .. collision of one node (STEP_7/momentsPre.f90:8-21,
// phase_field.f90:24-52, forces.f90:18-26, collision.f90:14-150) from its
// pulled deviations df (against rho_ref(j) = r0) and dg (against T_crit) and
// its phi_prec taps (C, E, N, W, S, NE, NW, SW, SE); gives the
// post-collision f deviations against r0 (fd) and the stored g deviations.
template <typename RC>
__device__ __forceinline__ Node32 node_update(const Consts& c, double r0, const double df[9],
                                              const double dg[9], const double tap[9],
                                              double fd[9], float gs[9]) {
  const double drho = df[0] + df[1] + df[2] + df[3] + df[4] + df[5] + df[6] + df[7] + df[8];
  const double rho = r0 + drho;
  double dT = dg[0] + dg[1] + dg[2] + dg[3] + dg[4] + dg[5] + dg[6] + dg[7] + dg[8];
  double u = (df[1] - df[3] + df[5] - df[6] - df[7] + df[8]) / rho;
  double v = (df[2] - df[4] + df[5] + df[6] - df[7] - df[8]) / rho;
  Node32 o;
  o.dT2 = dT;
  o.u2 = u;
  o.v2 = v;
  // ---- phase field, FP64
  o.phi = phase_update(c, c.g_ref + dT, tap);
  // ---- latent heat and forces, FP64
  dT = dT - (c.dt_st * (o.phi - tap[0]) / 2.0);
  const double T = c.g_ref + dT;
  {
    double fy = buoyancy(c, rho, T, o.phi);
    const double fx = -(u * (1.0 - o.phi) * 0.5);
    fy = fy - v * (1.0 - o.phi) * 0.5;
    u = u + fx / (c.omega_f * rho);
    v = v + fy / (c.omega_f * rho);
  }
  // ---- collision
  RC dfr[9], dgr[9], fo[9], go[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    dfr[k] = (RC)df[k];
    dgr[k] = (RC)dg[k];
  }
  regularised_dev<RC>(c, (RC)c.omega_f, (RC)drho, (RC)rho, (RC)u, (RC)v, dfr, fo);
  regularised_dev<RC>(c, (RC)omega_g_at(c, o.phi), (RC)dT, (RC)T, (RC)u, (RC)v, dgr, go);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    fd[k] = (double)fo[k];
    gs[k] = (float)go[k];
  }
  return o;
}

// Averaged moments of the stored state (STEP_7/moments.f90:11-25): rho, T,
// u, v as the reference reports them.
struct Avg {
  double rho, T, u, v;
};

__device__ __forceinline__ Avg averaged(const Consts& c, double r0, const float fs[9],
                                        const float gs[9], const Node32& o) {
  double a[9], b[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    a[k] = (double)fs[k];
    b[k] = (double)gs[k];
  }
  const double dr = a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8];
  Avg r;
  r.rho = r0 + dr;
  const double dtt = b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + b[6] + b[7] + b[8];
  const double tu = (a[1] - a[3] + a[5] - a[6] - a[7] + a[8]) / r.rho;
  const double tv = (a[2] - a[4] + a[5] + a[6] - a[7] - a[8]) / r.rho;
  r.T = c.g_ref + (dtt + o.dT2) * 0.5;
  r.u = (tu + o.u2) * 0.5;
  r.v = (tv + o.v2) * 0.5;
  return r;
}

// Reference density of interior node (.., j) with phase field phi: the
// solid and liquid densities of row j (tab[j], tab[S + j], S = Ny + 2)
// blended by the liquid fraction (1 + phi)/2.  The writer of a population
// and its reader evaluate it on the same inputs (the writer's new phi is the
// reader's phi_prec tap), so the stored deviation decodes exactly.
__device__ __forceinline__ double rho_ref(const double* __restrict__ tab, int S, int j,
                                          double phi) {
  const double rs = tab[j];
  return rs + (0.5 * (1.0 + phi)) * (tab[S + j] - rs);
}

template <typename RC>
__global__ __launch_bounds__(TILE_W) void k_fused32(Pairs32 A, Pairs32 B,
                                                     const double* __restrict__ pa,
                                                     double* __restrict__ pb,
                                                     const double* __restrict__ te_in,
                                                     double* __restrict__ te_out, Macro m,
                                                     Consts c, Geom gm, TileGrid tg,
                                                     const double* __restrict__ tab,
                                                     int write_all) {
  const int nx = gm.nx, ny = gm.ny, S = ny + 2;
  int i, j;
  if (!tile_node(tg, gm, &i, &j)) return;
  const int64_t P = gm.pitch;
  const int64_t q = (int64_t)j * P + i;
  const bool w_edge = (i == 1), e_edge = (i == nx), s_edge = (j == 1), n_edge = (j == ny);
  const bool edge = w_edge || e_edge || s_edge || n_edge;

  // phi_prec taps C, E, N, W, S, NE, NW, SW, SE (rows clamped per BC_g); the
  // tap of direction OPP[k] is the phi of the node population k is pulled
  // from
  const int64_t rn = (int64_t)(j < ny ? j + 1 : ny) * P;
  const int64_t rs = (int64_t)(j > 1 ? j - 1 : 1) * P;
  const int64_t r0w = (int64_t)j * P;
  const double tap[9] = {pa[q],          pa[r0w + i + 1], pa[rn + i],     pa[r0w + i - 1],
                         pa[rs + i],     pa[rn + i + 1],  pa[rn + i - 1], pa[rs + i - 1],
                         pa[rs + i + 1]};
  const double r0 = rho_ref(tab, S, j, tap[0]);

  double df[9], dg[9];
  if (!edge) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const float2 fg = stream_load(&A.p[k][q - CX[k] - (int64_t)CY[k] * P]);
      const double sh = k == 0 ? 0.0 : rho_ref(tab, S, j - CY[k], tap[OPP[k]]) - r0;
      df[k] = (double)fg.x + kW[k] * sh;
      dg[k] = (double)fg.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int si = i - CX[k];
      const int sj = j - CY[k];
      if (si >= 1 && si <= nx && sj >= 1 && sj <= ny) {
        df[k] = (double)A.p[k][(int64_t)sj * P + si].x +
                kW[k] * (rho_ref(tab, S, sj, tap[OPP[k]]) - r0);
      } else if ((k == 6 && si == nx + 1 && sj == 0) || (k == 7 && si == nx + 1 && sj == ny + 1)) {
        // a corner ghost no BC_f branch writes: stored against rho0
        df[k] = (double)A.p[k][(int64_t)sj * P + si].x + kW[k] * (c.rho0 - r0);
      } else {
        df[k] = (double)A.p[OPP[k]][q].x;  // bounce-back from this node
      }
      const int sjc = sj < 1 ? 1 : (sj > ny ? ny : sj);
      if (si >= 1 && si <= nx) {
        dg[k] = (double)A.p[k][(int64_t)sjc * P + si].y;
      } else if (si == 0) {
        const int64_t h = (int64_t)sjc * P;
        dg[k] = ghost_dev(c, k, c.tin, c.g_ref, m.u[h], m.v[h]);
      } else {
        const int64_t h = (int64_t)sjc * P + nx + 1;
        dg[k] = ghost_dev(c, k, east_T(c, te_in, sjc), c.g_ref, m.u[h], m.v[h]);
      }
    }
  }

  double fd[9];
  float fs[9], gs[9];
  const Node32 o = node_update<RC>(c, r0, df, dg, tap, fd, gs);
  // store f against this node's reference at its new phi (what the readers
  // of the next step decode with)
  const double r1 = rho_ref(tab, S, j, o.phi);
  const double dr = r0 - r1;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    fs[k] = (float)(fd[k] + kW[k] * dr);
    stream_store(&B.p[k][q], fs[k], gs[k]);
  }
  stream_store(&pb[q], o.phi);

  if (!(edge || write_all)) return;
  // ---- moments of the stored state (STEP_7/moments.f90:11-25)
  const Avg a = averaged(c, r1, fs, gs, o);
  if (edge) {
    // halo as BC_f / BC_g left it at the start of this step
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
      m.T[r0w] = c.tin;
      pb[r0w] = pa[r0w];
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
      const int64_t h = r0w + nx + 1;
      m.T[h] = east_T(c, te_in, j);
      pb[h] = pa[h];
      te_out[j] = a.T;
      if (s_edge) {
        m.T[nx + 1] = east_T(c, te_in, 1);
        pb[nx + 1] = pa[P + nx + 1];
        B.p[6][nx + 1].x = A.p[6][nx + 1].x;
      }
      if (n_edge) {
        const int64_t hn = (int64_t)(ny + 1) * P + nx + 1;
        m.T[hn] = east_T(c, te_in, ny);
        pb[hn] = pa[(int64_t)ny * P + nx + 1];
        B.p[7][hn].x = A.p[7][hn].x;
      }
    }
  }
  m.rho[q] = a.rho;
  m.T[q] = a.T;
  m.u[q] = a.u;
  m.v[q] = a.v;
}

// New solid and liquid densities of each interior row j = blockIdx.x + 1:
// the mean density (decoded from the stored f) of its nodes with phi <= 0
// and phi > 0 (a phase absent from the row keeps its old value), summed in
// FP64 in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_phase_density(Pairs32 A, const double* __restrict__ pa,
                                                        const double* __restrict__ told,
                                                        double* __restrict__ tnew, Geom gm) {
  __shared__ double red[4][256];
  const int j = (int)blockIdx.x + 1, S = gm.ny + 2;
  const int64_t r = (int64_t)j * gm.pitch;
  double ss = 0.0, sl = 0.0, ns = 0.0, nl = 0.0;
  for (int i = 1 + (int)threadIdx.x; i <= gm.nx; i += 256) {
    const double ph = pa[r + i];
    double d = rho_ref(told, S, j, ph);
#pragma unroll
    for (int k = 0; k < 9; ++k) d += (double)A.p[k][r + i].x;
    if (ph > 0.0) {
      sl += d;
      nl += 1.0;
    } else {
      ss += d;
      ns += 1.0;
    }
  }
  const int t = (int)threadIdx.x;
  red[0][t] = ss;
  red[1][t] = sl;
  red[2][t] = ns;
  red[3][t] = nl;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int v = 0; v < 4; ++v) red[v][t] += red[v][t + w];
    __syncthreads();
  }
  if (t == 0) {
    tnew[j] = red[2][0] > 0.0 ? red[0][0] / red[2][0] : told[j];
    tnew[S + j] = red[3][0] > 0.0 ? red[1][0] / red[3][0] : told[S + j];
  }
}

// Re-express the stored f of every interior node against new reference
// densities: the tables told -> tnew at the same phi (pa), or the phi
// pa -> pnew with one table (then pnew is also copied into pa).
__global__ void k_rebase(Pairs32 A, double* __restrict__ pa, const double* __restrict__ pnew,
                         const double* __restrict__ told, const double* __restrict__ tnew,
                         Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int S = gm.ny + 2;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const double ph = pa[q];
  const double ph2 = pnew ? pnew[q] : ph;
  const double d = rho_ref(told, S, j, ph) - rho_ref(tnew, S, j, ph2);
#pragma unroll
  for (int k = 0; k < 9; ++k) A.p[k][q].x = (float)((double)A.p[k][q].x + kW[k] * d);
  if (pnew) pa[q] = ph2;
}

// Reference density tables at initialization: rho0 (the solid) and rho1 (the
// melt) in every row, the densities initialization.f90:15-26 gives them.
__global__ void k_init_tab(double* __restrict__ tab, int S, double rho0, double rho1) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 2 * S) tab[k] = k < S ? rho0 : rho1;
}

// Whole-plane conversions between an FP64 plane and one member of a pair
// plane (stride 2): FP32 deviations or FP64 values.  The base of a node is
// `base`, or with a reference table (the f members of an FP32 handle) base *
// rho_ref(j, phi) on interior nodes and base * rho0 on the halo.
struct DevBase {
  double base;
  const double* tab;  // null: the constant base
  const double* phi;  // phi the table is evaluated at
  double rho0;
};

__device__ __forceinline__ double base_at(const DevBase& b, const Geom& gm, int i, int j,
                                          int64_t q) {
  if (!b.tab) return b.base;
  if (i < 1 || i > gm.nx || j < 1 || j > gm.ny) return b.base * b.rho0;
  return b.base * rho_ref(b.tab, gm.ny + 2, j, b.phi[q]);
}

__global__ void k_dev_to_f64(const float* __restrict__ src, int sstride, double* __restrict__ dst,
                             DevBase b, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  dst[q] = base_at(b, gm, i, j, q) + (double)src[q * sstride];
}

__global__ void k_f64_to_dev(const double* __restrict__ src, float* __restrict__ dst, int dstride,
                             DevBase b, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  dst[q * dstride] = (float)(src[q] - base_at(b, gm, i, j, q));
}

__global__ void k_copy64(const double* __restrict__ src, int sstride, double* __restrict__ dst,
                         int dstride, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  dst[q * dstride] = src[q * sstride];
}

// STEP_7/initialization.f90:15-152 for an FP32 handle: FP64 equilibria,
// stored as deviations (f of interior nodes against rho_ref(j, phi) of the
// initial tables, of halo nodes against rho0; g against T_crit).
__global__ void k_initialization32(Pops32 f, Pops32 g, Macro m, double* phi, double* phip,
                                   const double* __restrict__ tab, Consts c, Geom gm, int interf,
                                   int west_bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const NodeInit s0 = initial_node(c, gm, i, j, interf, west_bc);
  const bool inner = i >= 1 && i <= gm.nx && j >= 1 && j <= gm.ny;
  const double rr = inner ? rho_ref(tab, gm.ny + 2, j, s0.phi) : c.rho0;
  double eq[9];
  equilibria(c, s0.rho, s0.u, s0.v, eq);
#pragma unroll
  for (int k = 0; k < 9; ++k) f.p[k][q] = (float)(eq[k] - kW[k] * rr);
  equilibria(c, s0.T, s0.u, s0.v, eq);
#pragma unroll
  for (int k = 0; k < 9; ++k) g.p[k][q] = (float)(eq[k] - kW[k] * c.g_ref);
  m.rho[q] = s0.rho;
  m.u[q] = s0.u;
  m.v[q] = s0.v;
  m.T[q] = s0.T;
  phi[q] = s0.phi;
  phip[q] = s0.phi;
}

inline dim3 full_grid(const Geom& g) {
  return dim3((unsigned)((g.nx + 2 + TX - 1) / TX), (unsigned)((g.ny + 2 + TY - 1) / TY));
}

inline dim3 interior_grid(const Geom& g) {
  return dim3((unsigned)((g.nx + TX - 1) / TX), (unsigned)((g.ny + TY - 1) / TY));
}

}  // namespace

hipError_t launch_rebase32(bgk_handle* h) {
  const Geom& g = h->geom;
  const Pairs32 A = pairs32(h->f32[h->cur]);
  double* pa = h->phi[h->pc];
  const double* told = h->rref[h->rc];
  double* tnew = h->rref[1 - h->rc];
  // rows 0 and Ny+1 of the tables are never read; they keep the new table's
  hipLaunchKernelGGL(k_phase_density, dim3((unsigned)g.ny), dim3(256), 0, h->stream, A, pa, told,
                     tnew, g);
  hipLaunchKernelGGL(k_rebase, interior_grid(g), dim3(TX, TY), 0, h->stream, A, pa,
                     (const double*)nullptr, told, (const double*)tnew, g);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    h->rc = 1 - h->rc;
    h->since_rebase = 0;
  }
  return e;
}

hipError_t launch_rephase32(bgk_handle* h, const double* phi_new) {
  const Geom& g = h->geom;
  const double* tab = h->rref[h->rc];
  hipLaunchKernelGGL(k_rebase, interior_grid(g), dim3(TX, TY), 0, h->stream,
                     pairs32(h->f32[h->cur]), h->phi[h->pc], phi_new, tab, tab, g);
  return hipGetLastError();
}

hipError_t launch_fused_step32(bgk_handle* h, int step_parity, bool write_all) {
  if (h->rebase_every > 0 && h->since_rebase >= h->rebase_every) {
    hipError_t e = launch_rebase32(h);
    if (e != hipSuccess) return e;
  }
  const Geom& g = h->geom;
  const TileGrid tg = make_tile_grid(g);
  dim3 blk(TILE_W);
  dim3 grd((unsigned)(tg.per * NXCD));
  Macro m{h->rho, h->u, h->v, h->T};
  const double* te_in = h->teast[step_parity & 1];
  double* te_out = h->teast[(step_parity + 1) & 1];
  const int a = h->cur, b = 1 - h->cur;
  const double* tab = h->rref[h->rc];
  if (h->prm.fp32_collision)
    hipLaunchKernelGGL(k_fused32<float>, grd, blk, 0, h->stream, pairs32(h->f32[a]),
                       pairs32(h->f32[b]), h->phi[h->pc], h->phi[1 - h->pc], te_in, te_out, m,
                       h->c, g, tg, tab, write_all ? 1 : 0);
  else
    hipLaunchKernelGGL(k_fused32<double>, grd, blk, 0, h->stream, pairs32(h->f32[a]),
                       pairs32(h->f32[b]), h->phi[h->pc], h->phi[1 - h->pc], te_in, te_out, m,
                       h->c, g, tg, tab, write_all ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    h->cur = b;
    h->pc = 1 - h->pc;
    ++h->since_rebase;
  }
  return e;
}

hipError_t launch_dev_to_f64(bgk_handle* h, const float* src, int sstride, double* dst,
                             double base, bool rho_table) {
  const DevBase b{base, rho_table ? h->rref[h->rc] : nullptr, h->phi[h->pc], h->c.rho0};
  hipLaunchKernelGGL(k_dev_to_f64, full_grid(h->geom), dim3(TX, TY), 0, h->stream, src, sstride,
                     dst, b, h->geom);
  return hipGetLastError();
}

hipError_t launch_f64_to_dev(bgk_handle* h, const double* src, float* dst, int dstride,
                             double base, bool rho_table) {
  const DevBase b{base, rho_table ? h->rref[h->rc] : nullptr, h->phi[h->pc], h->c.rho0};
  hipLaunchKernelGGL(k_f64_to_dev, full_grid(h->geom), dim3(TX, TY), 0, h->stream, src, dst,
                     dstride, b, h->geom);
  return hipGetLastError();
}

hipError_t launch_copy64(bgk_handle* h, const double* src, int sstride, double* dst, int dstride) {
  hipLaunchKernelGGL(k_copy64, full_grid(h->geom), dim3(TX, TY), 0, h->stream, src, sstride, dst,
                     dstride, h->geom);
  return hipGetLastError();
}

hipError_t launch_initialization32(bgk_handle* h) {
  Macro m{h->rho, h->u, h->v, h->T};
  const int S = h->ny + 2;
  hipLaunchKernelGGL(k_init_tab, dim3((unsigned)((2 * S + 255) / 256)), dim3(256), 0, h->stream,
                     h->rref[h->rc], S, h->c.rho0, h->c.rho1);
  hipLaunchKernelGGL(k_initialization32, full_grid(h->geom), dim3(TX, TY), 0, h->stream,
                     pops32(h->f32[h->cur]), pops32(h->g32[h->cur]), m, h->phi[h->pc],
                     h->phi[1 - h->pc], (const double*)h->rref[h->rc], h->c, h->geom,
                     h->geom.nx / 40, h->prm.west_bc);
  h->since_rebase = 0;
  return hipGetLastError();
}

}  // namespace bgk
