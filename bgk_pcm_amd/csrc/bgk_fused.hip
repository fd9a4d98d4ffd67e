// bgk_fused.hip -- the fused FP64 time step for the no-slip cavity.
//
// One kernel does the whole block STEP_7/main.f90:60-69 for one node per
// thread: pull streaming with BC_f/BC_g folded in, momentsPre, phase_field,
// forces, collision and moments, reading buffer A (cur) and writing buffer
// B (1-cur).  Compiled with -ffp-contract=off; the arithmetic is operand for
// operand the reference's, so results equal STEP_7's bit for bit.
//
// How the boundary routines fold into the pull when all four walls are
// no-slip (BC type 4, STEP_7/BC_f.f90:62-68,101-107,125-136,181-192):
//  * every ghost population the pull reads was written by BC_f from the
//    reading node's own opposite population (half-way bounce-back), except
//    f6(Nx+1,0) and f7(Nx+1,Ny+1), which no BC writes: they keep their
//    initial value.  Those two are read from memory and copied into B.
//  * BC_g (STEP_7/BC_g.f90:9-64) makes the west ghosts the equilibrium at
//    Tin with u(0,j), v(0,j), the east ghosts the equilibrium at the
//    previous step's averaged T(Nx,j) with u(Nx+1,j), v(Nx+1,j), and then
//    copies rows 1 and Ny to rows 0 and Ny+1 for all columns: a pull from
//    row 0 / Ny+1 reads row 1 / Ny (clamped row), and a pull from column 0 /
//    Nx+1 evaluates that equilibrium.  T(Nx,j) of the previous step is kept
//    in the small ping-pong array teast.
//  * phi_prec rows 0 / Ny+1 are BC_g copies of rows 1 / Ny (clamped rows);
//    columns 0 / Nx+1 are never written after initialization and are copied
//    from A to B.
//  * the halo of rho, T, u, v and phi after a step holds what BC_f/BC_g
//    wrote at the start of that step, i.e. previous-step values; the edge
//    nodes write them so a download sees the reference's full arrays.
#include "bgk_internal.h"

namespace bgk {
namespace {

struct Macro {
  double *rho, *u, *v, *T;
};

// pair view of buffer b: plane k holds (f_k, g_k) per node
inline Pairs64 pairs(double* const* fplanes) {
  Pairs64 r;
  for (int k = 0; k < 9; ++k) r.p[k] = reinterpret_cast<double2*>(fplanes[k]);
  return r;
}

__global__ void k_teast_prologue(const double* __restrict__ T, double* __restrict__ te, Geom gm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > gm.ny + 1) return;
  te[j] = T[(int64_t)j * gm.pitch + gm.nx];
}

template <bool STEP7>
__global__ __launch_bounds__(TILE_W) void k_fused64(Pairs64 A, Pairs64 B,
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

  // ---- streaming (STEP_7/streaming.f90:7-25) with the boundaries folded in
  double fp[9], gp[9];
  if (!edge) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double2 fg = stream_load(&A.p[k][q - CX[k] - (int64_t)CY[k] * P]);
      fp[k] = fg.x;
      gp[k] = fg.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int si = i - CX[k];
      const int sj = j - CY[k];
      if (si >= 1 && si <= nx && sj >= 1 && sj <= ny) {
        fp[k] = A.p[k][(int64_t)sj * P + si].x;
      } else if ((k == 6 && si == nx + 1 && sj == 0) || (k == 7 && si == nx + 1 && sj == ny + 1)) {
        fp[k] = A.p[k][(int64_t)sj * P + si].x;  // never written by BC_f
      } else {
        fp[k] = A.p[OPP[k]][q].x;  // bounce-back
      }
      const int sjc = sj < 1 ? 1 : (sj > ny ? ny : sj);
      if (si >= 1 && si <= nx) {
        gp[k] = A.p[k][(int64_t)sjc * P + si].y;
      } else if (si == 0) {
        const int64_t h = (int64_t)sjc * P;
        gp[k] = equilibrium_k(c, k, c.tin, m.u[h], m.v[h]);
      } else {
        const int64_t h = (int64_t)sjc * P + nx + 1;
        gp[k] = equilibrium_k(c, k, east_T(c, te_in, sjc), m.u[h], m.v[h]);
      }
    }
  }

  // ---- momentsPre (STEP_7/momentsPre.f90:8-21)
  const double rho = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
  double T = gp[0] + gp[1] + gp[2] + gp[3] + gp[4] + gp[5] + gp[6] + gp[7] + gp[8];
  double u, v;
  {
    const double irho = 1.0 / rho;
    u = irho * (fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]);
    v = irho * (fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]);
  }
  const double T2 = T, u2 = u, v2 = v;

  // ---- phase_field (STEP_7/phase_field.f90:19-52), rows clamped per BC_g
  const int64_t rn = (int64_t)(j < ny ? j + 1 : ny) * P;
  const int64_t rs = (int64_t)(j > 1 ? j - 1 : 1) * P;
  const int64_t r0 = (int64_t)j * P;
  const double tap[9] = {pa[q],          pa[r0 + i + 1], pa[rn + i],     pa[r0 + i - 1], pa[rs + i],
                         pa[rn + i + 1], pa[rn + i - 1], pa[rs + i - 1], pa[rs + i + 1]};
  const double phi = phase_update(c, T, tap);
  const double pc = tap[0];

  // ---- forces (STEP_7/forces.f90:18-26)
  T = T - (c.dt_st * (phi - pc) / 2.0);
  {
    double fy = buoyancy(c, rho, T, phi);
    const double fx = -(u * (1.0 - phi) * 0.5);
    fy = fy - v * (1.0 - phi) * 0.5;
    u = u + fx / (c.omega_f * rho);
    v = v + fy / (c.omega_f * rho);
  }

  // ---- collision (STEP_7/collision.f90:14-150)
  double fo[9], go[9];
  regularised_bgk<STEP7>(c, c.omega_f, rho, u, v, fp, fo);
  regularised_bgk<STEP7>(c, omega_g_at(c, phi), T, u, v, gp, go);
#pragma unroll
  for (int k = 0; k < 9; ++k) stream_store(&B.p[k][q], fo[k], go[k]);
  stream_store(&pb[q], phi);

  // ---- moments (STEP_7/moments.f90:11-25), where somebody reads them
  if (edge || write_all) {
    const double r = fo[0] + fo[1] + fo[2] + fo[3] + fo[4] + fo[5] + fo[6] + fo[7] + fo[8];
    const double irho = 1.0 / r;
    const double tt = go[0] + go[1] + go[2] + go[3] + go[4] + go[5] + go[6] + go[7] + go[8];
    const double tu = irho * (fo[1] - fo[3] + fo[5] - fo[6] - fo[7] + fo[8]);
    const double tv = irho * (fo[2] - fo[4] + fo[5] + fo[6] - fo[7] - fo[8]);
    const double Tav = (tt + T2) * 0.5;
    const double uav = (tu + u2) * 0.5;
    const double vav = (tv + v2) * 0.5;
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

}  // namespace

hipError_t launch_teast_prologue(bgk_handle* h) {
  const Geom& g = h->geom;
  hipLaunchKernelGGL(k_teast_prologue, dim3((unsigned)((g.ny + 2 + 255) / 256)), dim3(256), 0,
                     h->stream, h->T, h->teast[0], g);
  return hipGetLastError();
}

hipError_t launch_fused_step64(bgk_handle* h, int step_parity, bool write_all) {
  const Geom& g = h->geom;
  const TileGrid tg = make_tile_grid(g);
  dim3 blk(TILE_W);
  dim3 grd((unsigned)(tg.per * NXCD));
  Macro m{h->rho, h->u, h->v, h->T};
  const double* te_in = h->teast[step_parity & 1];
  double* te_out = h->teast[(step_parity + 1) & 1];
  const int a = h->cur, b = 1 - h->cur;
  if (h->prm.rounding == BGK_ROUND_STEP7)
    hipLaunchKernelGGL(k_fused64<true>, grd, blk, 0, h->stream, pairs(h->f[a]), pairs(h->f[b]), h->phi[h->pc], h->phi[1 - h->pc], te_in,
                       te_out, m, h->c, g, tg, write_all ? 1 : 0);
  else
    hipLaunchKernelGGL(k_fused64<false>, grd, blk, 0, h->stream, pairs(h->f[a]), pairs(h->f[b]), h->phi[h->pc], h->phi[1 - h->pc], te_in,
                       te_out, m, h->c, g, tg, write_all ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    h->cur = b;
    h->pc = 1 - h->pc;
  }
  return e;
}

}  // namespace bgk
