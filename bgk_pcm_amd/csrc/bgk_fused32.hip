// bgk_fused32.hip -- the fused time step with FP32 population storage.
//
// Storage: f_k - w_k*rho0 and g_k - w_k*T_crit in FP32, deviations from the
// rest equilibrium at the solid density and at the melting temperature
// (T_crit = 273, STEP_7/phase_field.f90:16).  They are ~1e-2 of the
// populations, so FP32 keeps far more absolute precision than storing f and
// g themselves, and the temperature is most precise where the phase field
// is sensitive to it: at the melting front, where g_k - w_k*T_crit ~ 0.  phi,
// the east-wall temperature column and all macroscopic fields stay FP64.
//
// Arithmetic: moments, temperature, phase field, latent heat and forces are
// FP64.  The collision runs in `RC` (double by default): on gfx950 unpacked
// FP32 and FP64 FMA issue at the same rate, so FP64 costs registers, not
// time, while FP32 collision arithmetic loses accuracy at the melting front.
// Same folding of BC_f/BC_g and the same halo bookkeeping as the FP64 kernel
// (bgk_fused.hip); the algebra is the reference's written on deviations
// (m0 = rho0 or T_crit, dm = m - m0):
//   feq_k - w_k m0 = w_k (dm + m P_k(u,v)),
//   P_k = 3 c_k.u + 4.5 (c_k.u)^2 - 1.5 u.u,
//   f_k' - w_k m0 = (1-omega) (ft_k - w_k m0) + omega (feq_k - w_k m0).
// The FP32 rounding of ft/gt done by STEP_7 (implicit REAL) is not modelled.
#include "bgk_internal.h"

namespace bgk {
namespace {

constexpr int TX = 64, TY = 4;  // block of the whole-plane conversion kernels


struct Macro {
  double *rho, *u, *v, *T;
};

inline Pops32 pops32(float* const* p) {
  Pops32 r;
  for (int k = 0; k < 9; ++k) r.p[k].base = p[k];
  return r;
}

// pair view of buffer b: plane k holds (f_k - w_k rho0, g_k - w_k T0) per node
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
  for (int k = 0; k < 9; ++k) deq[k] = R(c.w[k]) * (dm + m * poly_k<R>(k, u, v));
  R d[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) d[k] = dp[k] - deq[k];
  const R a = R(-c.cs2), b = R(1.0 - c.cs2);
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
    const R t = deq[k] + R(c.coef[k]) * h[k];
    out[k] = om1 * t + omega * deq[k];
  }
}

// Equilibrium deviation w_k*(m - m0) + w_k*m*P_k for a boundary ghost (FP64).
__device__ __forceinline__ double ghost_dev(const Consts& c, int k, double m, double m0, double u,
                                            double v) {
  return c.w[k] * (m - m0) + c.w[k] * m * poly_k<double>(k, u, v);
}

template <typename RC>
__global__ __launch_bounds__(TILE_W) void k_fused32(Pairs32 A, Pairs32 B,
                                                     const double* __restrict__ pa,
                                                     double* __restrict__ pb,
                                                     const double* __restrict__ te_in,
                                                     double* __restrict__ te_out, Macro m,
                                                     Consts c, Geom gm, TileGrid tg,
                                                     int write_all) {
  const int nx = gm.nx, ny = gm.ny;
  int i, j;
  if (!tile_node(tg, gm, &i, &j)) return;
  const int64_t P = gm.pitch;
  const int64_t q = (int64_t)j * P + i;
  const bool w_edge = (i == 1), e_edge = (i == nx), s_edge = (j == 1), n_edge = (j == ny);
  const bool edge = w_edge || e_edge || s_edge || n_edge;

  double df[9], dg[9];
  if (!edge) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const float2 fg = stream_load(&A.p[k][q - CX[k] - (int64_t)CY[k] * P]);
      df[k] = (double)fg.x;
      dg[k] = (double)fg.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int si = i - CX[k];
      const int sj = j - CY[k];
      if (si >= 1 && si <= nx && sj >= 1 && sj <= ny) {
        df[k] = (double)A.p[k][(int64_t)sj * P + si].x;
      } else if ((k == 6 && si == nx + 1 && sj == 0) || (k == 7 && si == nx + 1 && sj == ny + 1)) {
        df[k] = (double)A.p[k][(int64_t)sj * P + si].x;
      } else {
        df[k] = (double)A.p[OPP[k]][q].x;
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

  // ---- momentsPre, FP64
  const double drho = df[0] + df[1] + df[2] + df[3] + df[4] + df[5] + df[6] + df[7] + df[8];
  const double rho = c.rho0 + drho;
  double dT = dg[0] + dg[1] + dg[2] + dg[3] + dg[4] + dg[5] + dg[6] + dg[7] + dg[8];
  double u = (df[1] - df[3] + df[5] - df[6] - df[7] + df[8]) / rho;
  double v = (df[2] - df[4] + df[5] + df[6] - df[7] - df[8]) / rho;
  const double dT2 = dT, u2 = u, v2 = v;

  // ---- phase field, FP64
  const int64_t rn = (int64_t)(j < ny ? j + 1 : ny) * P;
  const int64_t rs = (int64_t)(j > 1 ? j - 1 : 1) * P;
  const int64_t r0 = (int64_t)j * P;
  const double tap[9] = {pa[q],          pa[r0 + i + 1], pa[rn + i],     pa[r0 + i - 1], pa[rs + i],
                         pa[rn + i + 1], pa[rn + i - 1], pa[rs + i - 1], pa[rs + i + 1]};
  const double phi = phase_update(c, c.g_ref + dT, tap);
  const double pc = tap[0];

  // ---- latent heat and forces, FP64
  dT = dT - (c.dt_st * (phi - pc) / 2.0);
  const double T = c.g_ref + dT;
  {
    double fy = buoyancy(c, rho, T, phi);
    const double fx = -(u * (1.0 - phi) * 0.5);
    fy = fy - v * (1.0 - phi) * 0.5;
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
  regularised_dev<RC>(c, (RC)omega_g_at(c, phi), (RC)dT, (RC)T, (RC)u, (RC)v, dgr, go);
  float fs[9], gs[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    fs[k] = (float)fo[k];
    gs[k] = (float)go[k];
    stream_store(&B.p[k][q], fs[k], gs[k]);
  }
  stream_store(&pb[q], phi);

  // ---- moments of the stored state, FP64
  if (edge || write_all) {
    double a[9], b[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      a[k] = (double)fs[k];
      b[k] = (double)gs[k];
    }
    const double dr = a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8];
    const double r = c.rho0 + dr;
    const double dtt = b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + b[6] + b[7] + b[8];
    const double tu = (a[1] - a[3] + a[5] - a[6] - a[7] + a[8]) / r;
    const double tv = (a[2] - a[4] + a[5] + a[6] - a[7] - a[8]) / r;
    const double Tav = c.g_ref + (dtt + dT2) * 0.5;
    const double uav = (tu + u2) * 0.5;
    const double vav = (tv + v2) * 0.5;
    if (edge) {
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
        te_out[j] = Tav;
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
    m.rho[q] = r;
    m.T[q] = Tav;
    m.u[q] = uav;
    m.v[q] = vav;
  }
}

// Whole-plane conversions between an FP64 plane and one member of a pair
// plane (stride 2): FP32 deviations or FP64 values.
__global__ void k_dev_to_f64(const float* __restrict__ src, int sstride, double* __restrict__ dst,
                             double base, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  dst[q] = base + (double)src[q * sstride];
}

__global__ void k_f64_to_dev(const double* __restrict__ src, float* __restrict__ dst, int dstride,
                             double base, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  dst[q * dstride] = (float)(src[q] - base);
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
// stored as deviations.
__global__ void k_initialization32(Pops32 f, Pops32 g, Macro m, double* phi, double* phip,
                                   Consts c, Geom gm, int interf, int west_bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const NodeInit s0 = initial_node(c, gm, i, j, interf, west_bc);
  double eq[9];
  equilibria(c, s0.rho, s0.u, s0.v, eq);
#pragma unroll
  for (int k = 0; k < 9; ++k) f.p[k][q] = (float)(eq[k] - c.w[k] * c.rho0);
  equilibria(c, s0.T, s0.u, s0.v, eq);
#pragma unroll
  for (int k = 0; k < 9; ++k) g.p[k][q] = (float)(eq[k] - c.w[k] * c.g_ref);
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

}  // namespace

hipError_t launch_fused_step32(bgk_handle* h, int step_parity, bool write_all) {
  const Geom& g = h->geom;
  const TileGrid tg = make_tile_grid(g);
  dim3 blk(TILE_W);
  dim3 grd((unsigned)(tg.per * NXCD));
  Macro m{h->rho, h->u, h->v, h->T};
  const double* te_in = h->teast[step_parity & 1];
  double* te_out = h->teast[(step_parity + 1) & 1];
  const int a = h->cur, b = 1 - h->cur;
  if (h->prm.fp32_collision)
    hipLaunchKernelGGL(k_fused32<float>, grd, blk, 0, h->stream, pairs32(h->f32[a]),
                       pairs32(h->f32[b]), h->phi[h->pc],
                       h->phi[1 - h->pc], te_in, te_out, m, h->c, g, tg, write_all ? 1 : 0);
  else
    hipLaunchKernelGGL(k_fused32<double>, grd, blk, 0, h->stream, pairs32(h->f32[a]),
                       pairs32(h->f32[b]), h->phi[h->pc],
                       h->phi[1 - h->pc], te_in, te_out, m, h->c, g, tg, write_all ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    h->cur = b;
    h->pc = 1 - h->pc;
  }
  return e;
}

hipError_t launch_dev_to_f64(bgk_handle* h, const float* src, int sstride, double* dst,
                             double base) {
  hipLaunchKernelGGL(k_dev_to_f64, full_grid(h->geom), dim3(TX, TY), 0, h->stream, src, sstride,
                     dst, base, h->geom);
  return hipGetLastError();
}

hipError_t launch_f64_to_dev(bgk_handle* h, const double* src, float* dst, int dstride,
                             double base) {
  hipLaunchKernelGGL(k_f64_to_dev, full_grid(h->geom), dim3(TX, TY), 0, h->stream, src, dst,
                     dstride, base, h->geom);
  return hipGetLastError();
}

hipError_t launch_copy64(bgk_handle* h, const double* src, int sstride, double* dst, int dstride) {
  hipLaunchKernelGGL(k_copy64, full_grid(h->geom), dim3(TX, TY), 0, h->stream, src, sstride, dst,
                     dstride, h->geom);
  return hipGetLastError();
}

hipError_t launch_initialization32(bgk_handle* h) {
  Macro m{h->rho, h->u, h->v, h->T};
  hipLaunchKernelGGL(k_initialization32, full_grid(h->geom), dim3(TX, TY), 0, h->stream,
                     pops32(h->f32[h->cur]), pops32(h->g32[h->cur]), m, h->phi[h->pc],
                     h->phi[1 - h->pc], h->c, h->geom, h->geom.nx / 40, h->prm.west_bc);
  return hipGetLastError();
}

}  // namespace bgk
