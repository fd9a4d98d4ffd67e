// bgk_stage.hip -- one kernel (or kernel pair) per reference subroutine.
//
// This is the per-subroutine path of the drop-in boundary: bgk_bc_f,
// bgk_bc_g, bgk_streaming, ... each launch the kernels below on the handle's
// stream, on the same fields the reference's `module shared` holds.  It
// serves every boundary type of the reference and is the path bgk_step takes
// when the walls are not all no-slip.  Compiled with -ffp-contract=off: the
// FP64 results equal the reference's bit for bit.
#include "bgk_internal.h"

namespace bgk {
namespace {

struct Macro {
  double *rho, *u, *v, *T, *rho2, *u2, *v2, *T2;
};

inline Macro macro_of(bgk_handle* h) {
  return Macro{h->rho, h->u, h->v, h->T, h->rho2, h->u2, h->v2, h->T2};
}

inline Pops64 pops(double* const* p) {
  Pops64 r;
  for (int k = 0; k < 9; ++k) r.p[k].base = p[k];
  return r;
}

constexpr int TX = 64, TY = 4;

inline dim3 interior_grid(const Geom& g) {
  return dim3((unsigned)((g.nx + TX - 1) / TX), (unsigned)((g.ny + TY - 1) / TY));
}

// ------------------------------------------------------------ initialization
// STEP_7/initialization.f90:15-152, one thread per node including the halo.
__global__ void k_initialization(Pops64 f, Pops64 g, Macro m, double* phi, double* phip,
                                 Consts c, Geom gm, int interf, int west_bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > gm.nx + 1 || j > gm.ny + 1) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const NodeInit s0 = initial_node(c, gm, i, j, interf, west_bc);
  const double T = s0.T, rho = s0.rho, u = s0.u, v = s0.v;
  double eq[9];
  equilibria(c, rho, u, v, eq);
#pragma unroll
  for (int k = 0; k < 9; ++k) f.p[k][q] = eq[k];
  equilibria(c, T, u, v, eq);
#pragma unroll
  for (int k = 0; k < 9; ++k) g.p[k][q] = eq[k];
  m.rho[q] = rho;
  m.u[q] = u;
  m.v[q] = v;
  m.T[q] = T;
  m.rho2[q] = 0.0;
  m.u2[q] = 0.0;
  m.v2[q] = 0.0;
  m.T2[q] = 0.0;
  phi[q] = s0.phi;
  phip[q] = s0.phi;
}

// ----------------------------------------------------------------------- BC_f
// West and east sides, one thread per row j = 0..Ny+1 (STEP_7/BC_f.f90:9-104).
__global__ void k_bc_f_we(Pops64 F, Macro m, Consts c, Geom gm, int wbc, int ebc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = gm.nx, ny = gm.ny;
  if (j > ny + 1) return;
  const int64_t P = gm.pitch;
  auto at = [&](int i, int jj) { return (int64_t)jj * P + i; };
  const Strided2<double>* f = F.p;
  const bool inner = (j >= 1 && j <= ny);
  double uu, vv, uv, upv, umv;
  // WEST
  if (wbc == 0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k][at(0, j)] = f[k][at(nx, j)];
  } else if (wbc == 1) {
    if (inner) {
      const double u0 = c.u0, v0 = c.v0;
      uu = u0 * u0; vv = v0 * v0; uv = 1.5 * (uu + vv); upv = u0 + v0; umv = u0 - v0;
      const double r1 = m.rho[at(1, j)];
      f[1][at(0, j)] = c.w[1] * r1 * (1.0 + 3.0 * u0 + 4.5 * uu - uv);
      f[5][at(0, j)] = c.w[5] * r1 * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
      f[8][at(0, j)] = c.w[8] * r1 * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
      m.u[at(0, j)] = u0;
      m.v[at(0, j)] = v0;
      m.rho[at(0, j)] = r1;
    }
  } else if (wbc == 2) {
    const double uw = m.u[at(0, j)], v0 = c.v0;
    uu = uw * uw; vv = v0 * v0; uv = 1.5 * (uu + vv); upv = uw + v0; umv = uw - v0;
    const double r1 = m.rho[at(1, j)];
    f[1][at(0, j)] = c.w[1] * r1 * (1.0 + 3.0 * uw + 4.5 * uu - uv);
    f[5][at(0, j)] = c.w[5] * r1 * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
    f[8][at(0, j)] = c.w[8] * r1 * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
  } else if (wbc == 3) {
    const double u1 = m.u[at(1, j)], v1 = m.v[at(1, j)];
    uu = u1 * u1; vv = v1 * v1; uv = 1.5 * (uu + vv); upv = u1 + v1; umv = u1 - v1;
    f[1][at(0, j)] = c.w[1] * c.rho1 * (1.0 + 3.0 * u1 + 4.5 * uu - uv);
    f[5][at(0, j)] = c.w[5] * c.rho1 * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
    f[8][at(0, j)] = c.w[8] * c.rho1 * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
    m.u[at(0, j)] = u1;
    m.v[at(0, j)] = v1;
    m.rho[at(0, j)] = c.rho1;
  } else if (wbc == 4) {
    if (inner) {
      f[1][at(0, j)] = f[3][at(1, j)];
      f[8][at(0, j)] = f[6][at(1, j - 1)];
      f[5][at(0, j)] = f[7][at(1, j + 1)];
    }
  }
  // EAST
  if (ebc == 0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k][at(nx + 1, j)] = f[k][at(1, j)];
  } else if (ebc == 3) {
    if (inner) {
      const double ue = m.u[at(nx, j)], ve = m.v[at(nx, j)];
      uu = ue * ue; vv = ve * ve; uv = 1.5 * (uu + vv); upv = ue + ve; umv = ue - ve;
      f[3][at(nx + 1, j)] = c.w[3] * c.rho0 * (1.0 - 3.0 * ue + 4.5 * uu - uv);
      f[6][at(nx + 1, j)] = c.w[6] * c.rho0 * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
      f[7][at(nx + 1, j)] = c.w[7] * c.rho0 * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
      m.u[at(nx + 1, j)] = ue;
      m.v[at(nx + 1, j)] = ve;
      m.rho[at(nx + 1, j)] = c.rho0;
    }
  } else if (ebc == 4) {
    if (inner) {
      f[3][at(nx + 1, j)] = f[1][at(nx, j)];
      f[6][at(nx + 1, j)] = f[8][at(nx, j + 1)];
      f[7][at(nx + 1, j)] = f[5][at(nx, j - 1)];
    }
  }
}

// North and south sides, one thread per column i = 0..Nx+1
// (STEP_7/BC_f.f90:108-196); runs after k_bc_f_we.
__global__ void k_bc_f_ns(Pops64 F, Macro m, Consts c, Geom gm, int nbc, int sbc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = gm.nx, ny = gm.ny;
  if (i > nx + 1) return;
  const int64_t P = gm.pitch;
  auto at = [&](int ii, int jj) { return (int64_t)jj * P + ii; };
  const Strided2<double>* f = F.p;
  const bool inner = (i >= 1 && i <= nx);
  // NORTH
  if (nbc == 0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k][at(i, ny + 1)] = f[k][at(i, 1)];
  } else if (nbc == 4) {
    if (inner) {
      f[4][at(i, ny + 1)] = f[2][at(i, ny)];
      f[7][at(i, ny + 1)] = f[5][at(i - 1, ny)];
      f[8][at(i, ny + 1)] = f[6][at(i + 1, ny)];
      m.rho[at(i, ny + 1)] = m.rho[at(i, ny)];
      m.u[at(i, ny + 1)] = 0.0;
      m.v[at(i, ny + 1)] = 0.0;
    }
    if (i == 0) f[8][at(0, ny + 1)] = f[6][at(1, ny)];
  } else if (nbc == 5) {
    f[4][at(i, ny + 1)] = f[2][at(i, ny - 1)];
    f[7][at(i, ny + 1)] = f[6][at(i, ny - 1)];
    f[8][at(i, ny + 1)] = f[5][at(i, ny - 1)];
  } else if (nbc == 6) {
    f[4][at(i, ny + 1)] = f[4][at(i, ny)];
    f[7][at(i, ny + 1)] = f[7][at(i, ny)];
    f[8][at(i, ny + 1)] = f[8][at(i, ny)];
  } else if (nbc == 1) {
    const double u0 = c.u0, v0 = c.v0;
    const double uu = u0 * u0, vv = v0 * v0, uv = 1.5 * (uu + vv), upv = u0 + v0, umv = u0 - v0;
    const double rn = m.rho[at(i, ny)];
    f[4][at(i, ny + 1)] = c.w[4] * rn * (1.0 - 3.0 * v0 + 4.5 * vv - uv);
    f[7][at(i, ny + 1)] = c.w[7] * rn * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
    f[8][at(i, ny + 1)] = c.w[8] * rn * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
  }
  // SOUTH
  if (sbc == 0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k][at(i, 0)] = f[k][at(i, ny)];
  } else if (sbc == 4) {
    if (inner) {
      f[2][at(i, 0)] = f[4][at(i, 1)];
      f[5][at(i, 0)] = f[7][at(i + 1, 1)];
      f[6][at(i, 0)] = f[8][at(i - 1, 1)];
      m.rho[at(i, 0)] = m.rho[at(i, 1)];
      m.u[at(i, 0)] = 0.0;
      m.v[at(i, 0)] = 0.0;
    }
    if (i == 0) f[5][at(0, 0)] = f[7][at(1, 1)];
  } else if (sbc == 6) {
    f[2][at(i, 0)] = f[2][at(i, 1)];
    f[5][at(i, 0)] = f[5][at(i, 1)];
    f[6][at(i, 0)] = f[6][at(i, 1)];
  }
}

// ----------------------------------------------------------------------- BC_g
// West and east, rows j = 1..Ny (STEP_7/BC_g.f90:9-31).
__global__ void k_bc_g_we(Pops64 G, Macro m, Consts c, Geom gm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int nx = gm.nx, ny = gm.ny;
  if (j > ny) return;
  const int64_t P = gm.pitch;
  auto at = [&](int ii, int jj) { return (int64_t)jj * P + ii; };
  const Strided2<double>* g = G.p;
  {
    const double uw = m.u[at(0, j)], vw = m.v[at(0, j)];
    const double uu = uw * uw, vv = vw * vw, uv = 1.5 * (uu + vv), upv = uw + vw, umv = uw - vw;
    g[1][at(0, j)] = c.w[1] * c.tin * (1.0 + 3.0 * uw + 4.5 * uu - uv);
    g[5][at(0, j)] = c.w[5] * c.tin * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
    g[8][at(0, j)] = c.w[8] * c.tin * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
    m.T[at(0, j)] = c.tin;
  }
  {
    // the averaged T(Nx, j), or the cold-wall extension's fixed temperature
    const double ue = m.u[at(nx + 1, j)], ve = m.v[at(nx + 1, j)];
    const double te = c.east_cold ? c.t_cold : m.T[at(nx, j)];
    const double uu = ue * ue, vv = ve * ve, uv = 1.5 * (uu + vv), upv = ue + ve, umv = ue - ve;
    g[3][at(nx + 1, j)] = c.w[3] * te * (1.0 - 3.0 * ue + 4.5 * uu - uv);
    g[6][at(nx + 1, j)] = c.w[6] * te * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
    g[7][at(nx + 1, j)] = c.w[7] * te * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
    m.T[at(nx + 1, j)] = te;
  }
}

// North and south zero-gradient copies, i = 0..Nx+1 (STEP_7/BC_g.f90:36-64).
__global__ void k_bc_g_ns(Pops64 G, Macro m, double* phi, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ny = gm.ny;
  if (i > gm.nx + 1) return;
  const int64_t P = gm.pitch;
  const int64_t qn = (int64_t)ny * P + i, qn1 = (int64_t)(ny + 1) * P + i;
  const int64_t q1 = P + i, q0 = i;
#pragma unroll
  for (int k = 0; k < 9; ++k) G.p[k][qn1] = G.p[k][qn];
  m.T[qn1] = m.T[qn];
  phi[qn1] = phi[qn];
#pragma unroll
  for (int k = 0; k < 9; ++k) G.p[k][q0] = G.p[k][q1];
  m.T[q0] = m.T[q1];
  phi[q0] = phi[q1];
}

// ------------------------------------------------------------------ streaming
// Pull of f and g into fp and gp, STEP_7/streaming.f90:5-27.
__global__ void k_streaming(Pops64 f, Pops64 fp, Pops64 g, Pops64 gp, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int64_t P = gm.pitch;
  const int64_t q = (int64_t)j * P + i;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int64_t s = q - CX[k] - (int64_t)CY[k] * P;
    fp.p[k][q] = f.p[k][s];
    gp.p[k][q] = g.p[k][s];
  }
}

// ----------------------------------------------------------------- momentsPre
// STEP_7/momentsPre.f90:8-21.
__global__ void k_moments_pre(Pops64 fp, Pops64 gp, Macro m, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  double a[9], b[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    a[k] = fp.p[k][q];
    b[k] = gp.p[k][q];
  }
  const double r = a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8];
  const double t = b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + b[6] + b[7] + b[8];
  const double irho = 1.0 / r;
  const double u = irho * (a[1] - a[3] + a[5] - a[6] - a[7] + a[8]);
  const double v = irho * (a[2] - a[4] + a[5] + a[6] - a[7] - a[8]);
  m.rho[q] = r;
  m.T[q] = t;
  m.u[q] = u;
  m.v[q] = v;
  m.rho2[q] = r;
  m.T2[q] = t;
  m.u2[q] = u;
  m.v2[q] = v;
}

// ---------------------------------------------------------------- phase_field
// Interior update from phi_prec (STEP_7/phase_field.f90:21-55); the copy
// phi_prec(:,:) = phi(:,:) of :19 is a device copy before the launch.
__global__ void k_phase_field(double* phi, const double* pp, const double* T, Consts c,
                              Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int64_t P = gm.pitch;
  const int64_t q = (int64_t)j * P + i;
  const double tap[9] = {pp[q],         pp[q + 1],     pp[q + P],     pp[q - 1], pp[q - P],
                         pp[q + 1 + P], pp[q - 1 + P], pp[q - 1 - P], pp[q + 1 - P]};
  phi[q] = phase_update(c, T[q], tap);
}

// --------------------------------------------------------------------- forces
// STEP_7/forces.f90:15-26.
__global__ void k_forces(Macro m, const double* phi, const double* pp, Consts c, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const double ph = phi[q];
  const double rho = m.rho[q];
  const double t = m.T[q] - (c.dt_st * (ph - pp[q]) / 2.0);
  double fy = buoyancy(c, rho, t, ph);
  const double u = m.u[q], v = m.v[q];
  const double fx = -(u * (1.0 - ph) * 0.5);
  fy = fy - v * (1.0 - ph) * 0.5;
  m.T[q] = t;
  m.u[q] = u + fx / (c.omega_f * rho);
  m.v[q] = v + fy / (c.omega_f * rho);
}

// ------------------------------------------------------------------ collision
// STEP_7/collision.f90:14-152.
template <bool STEP7>
__global__ void k_collision(Pops64 fp, Pops64 gp, Pops64 f, Pops64 g, Macro m,
                            const double* __restrict__ phi, Consts c, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  const double u = m.u[q], v = m.v[q];
  double pop[9], out[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) pop[k] = fp.p[k][q];
  regularised_bgk<STEP7>(c, c.omega_f, m.rho[q], u, v, pop, out);
#pragma unroll
  for (int k = 0; k < 9; ++k) f.p[k][q] = out[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) pop[k] = gp.p[k][q];
  regularised_bgk<STEP7>(c, omega_g_at(c, phi[q]), m.T[q], u, v, pop, out);
#pragma unroll
  for (int k = 0; k < 9; ++k) g.p[k][q] = out[k];
}

// -------------------------------------------------------------------- moments
// STEP_7/moments.f90:11-25.
__global__ void k_moments(Pops64 f, Pops64 g, Macro m, Geom gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > gm.nx || j > gm.ny) return;
  const int64_t q = (int64_t)j * gm.pitch + i;
  double a[9], b[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    a[k] = f.p[k][q];
    b[k] = g.p[k][q];
  }
  const double r = a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8];
  const double irho = 1.0 / r;
  const double tt = b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + b[6] + b[7] + b[8];
  const double tu = irho * (a[1] - a[3] + a[5] - a[6] - a[7] + a[8]);
  const double tv = irho * (a[2] - a[4] + a[5] + a[6] - a[7] - a[8]);
  m.rho[q] = r;
  m.T[q] = (tt + m.T2[q]) * 0.5;
  m.u[q] = (tu + m.u2[q]) * 0.5;
  m.v[q] = (tv + m.v2[q]) * 0.5;
}

inline unsigned blocks1d(int n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_initialization(bgk_handle* h) {
  const Geom& g = h->geom;
  dim3 blk(TX, TY);
  dim3 grd((unsigned)((g.nx + 2 + TX - 1) / TX), (unsigned)((g.ny + 2 + TY - 1) / TY));
  hipLaunchKernelGGL(k_initialization, grd, blk, 0, h->stream, pops(h->f[h->cur]),
                     pops(h->g[h->cur]), macro_of(h), h->phi[h->pc], h->phi[1 - h->pc], h->c,
                     g, g.nx / 40, h->prm.west_bc);
  return hipGetLastError();
}

hipError_t launch_bc_f(bgk_handle* h) {
  const Geom& g = h->geom;
  hipLaunchKernelGGL(k_bc_f_we, dim3(blocks1d(g.ny + 2, 256)), dim3(256), 0, h->stream,
                     pops(h->f[h->cur]), macro_of(h), h->c, g, h->prm.west_bc, h->prm.east_bc);
  hipLaunchKernelGGL(k_bc_f_ns, dim3(blocks1d(g.nx + 2, 256)), dim3(256), 0, h->stream,
                     pops(h->f[h->cur]), macro_of(h), h->c, g, h->prm.north_bc,
                     h->prm.south_bc);
  return hipGetLastError();
}

hipError_t launch_bc_g(bgk_handle* h) {
  const Geom& g = h->geom;
  hipLaunchKernelGGL(k_bc_g_we, dim3(blocks1d(g.ny, 256)), dim3(256), 0, h->stream,
                     pops(h->g[h->cur]), macro_of(h), h->c, g);
  hipLaunchKernelGGL(k_bc_g_ns, dim3(blocks1d(g.nx + 2, 256)), dim3(256), 0, h->stream,
                     pops(h->g[h->cur]), macro_of(h), h->phi[h->pc], g);
  return hipGetLastError();
}

hipError_t launch_streaming(bgk_handle* h) {
  hipLaunchKernelGGL(k_streaming, interior_grid(h->geom), dim3(TX, TY), 0, h->stream,
                     pops(h->f[h->cur]), pops(h->f[1 - h->cur]), pops(h->g[h->cur]),
                     pops(h->g[1 - h->cur]), h->geom);
  return hipGetLastError();
}

hipError_t launch_moments_pre(bgk_handle* h) {
  hipLaunchKernelGGL(k_moments_pre, interior_grid(h->geom), dim3(TX, TY), 0, h->stream,
                     pops(h->f[1 - h->cur]), pops(h->g[1 - h->cur]), macro_of(h), h->geom);
  return hipGetLastError();
}

hipError_t launch_phase_field(bgk_handle* h) {
  const size_t bytes = (size_t)h->plane_elems * sizeof(double);
  hipError_t e = hipMemcpyAsync(h->phi[1 - h->pc], h->phi[h->pc], bytes,
                                hipMemcpyDeviceToDevice, h->stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_phase_field, interior_grid(h->geom), dim3(TX, TY), 0, h->stream,
                     h->phi[h->pc], h->phi[1 - h->pc], h->T, h->c, h->geom);
  return hipGetLastError();
}

hipError_t launch_forces(bgk_handle* h) {
  hipLaunchKernelGGL(k_forces, interior_grid(h->geom), dim3(TX, TY), 0, h->stream, macro_of(h),
                     h->phi[h->pc], h->phi[1 - h->pc], h->c, h->geom);
  return hipGetLastError();
}

hipError_t launch_collision(bgk_handle* h) {
  if (h->prm.rounding == BGK_ROUND_STEP7)
    hipLaunchKernelGGL(k_collision<true>, interior_grid(h->geom), dim3(TX, TY), 0, h->stream,
                       pops(h->f[1 - h->cur]), pops(h->g[1 - h->cur]), pops(h->f[h->cur]),
                       pops(h->g[h->cur]), macro_of(h), h->phi[h->pc], h->c, h->geom);
  else
    hipLaunchKernelGGL(k_collision<false>, interior_grid(h->geom), dim3(TX, TY), 0, h->stream,
                       pops(h->f[1 - h->cur]), pops(h->g[1 - h->cur]), pops(h->f[h->cur]),
                       pops(h->g[h->cur]), macro_of(h), h->phi[h->pc], h->c, h->geom);
  return hipGetLastError();
}

hipError_t launch_moments(bgk_handle* h) {
  hipLaunchKernelGGL(k_moments, interior_grid(h->geom), dim3(TX, TY), 0, h->stream,
                     pops(h->f[h->cur]), pops(h->g[h->cur]), macro_of(h), h->geom);
  return hipGetLastError();
}

}  // namespace bgk
