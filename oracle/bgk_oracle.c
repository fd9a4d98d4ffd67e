/*
 * bgk_oracle.c -- CPU restatement of the reference STEP_7 time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see
 * bgk_oracle.h for who may load it and what pins it).
 *
 * Evaluation-order notes that matter for bit parity with the amdflang build
 * of STEP_7:
 *  - Fortran evaluates a*b*c as (a*b)*c and a+b-c as (a+b)-c; C does the
 *    same, so the expressions below keep the source's operand order.
 *  - x**2.0d0 is lowered by amdflang to x*x (checked in its LLVM IR).
 *  - `1.`, `2.`, `273.0` are default-REAL literals with exact FP64 values.
 *  - ft0..ft8, gt0..gt8 (STEP_7/collision.f90:94-130) and beta
 *    (STEP_7/forces.f90:13) are not declared and the files have no
 *    `implicit none`, so they are default REAL: the value is rounded to
 *    FP32 on assignment.  r4() reproduces that under BGK_ROUND_STEP7.
 *  - This file must be compiled without FMA contraction (-ffp-contract=off)
 *    and without -ffast-math.
 */
#include "bgk_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

struct oracle {
  bgk_params p;
  int nx, ny, ld;
  size_t n;
  int step7;
  double w[9];
  double cs2, omega_f, omega_g, kvisc, alpha;
  double u0, v0, tin;
  int interf;
  /* extensions (off: t_hot = T0, k_ratio = 1, no cold wall, no noise) */
  double t_hot, k_ratio;
  int east_cold;
  double t_cold, noise_amp;
  uint64_t noise_seed;
  double t_crit, t_ref; /* phase-field T_crit (273), buoyancy reference (0) */
  double* fld[BGK_NFIELDS];
};

#define IX(i, j) ((size_t)(i) + (size_t)ld * (size_t)(j))

static inline double r4(int step7, double x) { return step7 ? (double)(float)x : x; }

double oracle_tanh(double x) { return tanh(x); }

/* ---------------------------------------------------------------- setup */

oracle_t* oracle_create(const bgk_params* p) {
  if (!p || p->nx < 1 || p->ny < 1) return NULL;
  oracle_t* s = (oracle_t*)calloc(1, sizeof(oracle_t));
  if (!s) return NULL;
  s->p = *p;
  s->nx = p->nx;
  s->ny = p->ny;
  s->ld = p->nx + 2;
  s->n = (size_t)(p->nx + 2) * (size_t)(p->ny + 2);
  s->step7 = (p->rounding == BGK_ROUND_STEP7);
  for (int f = 0; f < BGK_NFIELDS; ++f) {
    s->fld[f] = (double*)calloc(s->n, sizeof(double));
    if (!s->fld[f]) {
      oracle_destroy(s);
      return NULL;
    }
  }
  /* STEP_7/setup.f90:79-105 */
  s->w[0] = 4.0 / 9.0;
  for (int k = 1; k <= 4; ++k) s->w[k] = 1.0 / 9.0;
  for (int k = 5; k <= 8; ++k) s->w[k] = 1.0 / 36.0;
  s->cs2 = 1.0 / 3.0;
  s->omega_f = 1.0 / p->tau_f;
  s->omega_g = 1.0 / p->tau_g;
  s->kvisc = s->cs2 * (p->tau_f - 0.5);
  s->alpha = s->cs2 * (p->tau_g - 0.5);
  /* Scalars the reference sets in initialization.f90:5-12; they depend on
   * the parameters only, so a state loaded without initialization (a
   * snapshot) gets them too. */
  s->interf = p->nx / 40;
  s->t_hot = p->t0;
  s->k_ratio = 1.0;
  s->t_crit = 273.0;
  s->t_ref = 0.0;
  s->tin = s->t_hot;
  s->u0 = p->re * s->kvisc / p->length;
  s->v0 = 0.0;
  return s;
}

/* Physics extensions (not in the reference; off = STEP_7): the hot-wall
 * and initial liquid temperature t_hot in place of T0 (the temperature
 * scale T0 - T_in_sol of forces.f90:12-13 is kept), and a solid/liquid
 * conductivity ratio through a phase-dependent thermal relaxation time,
 * alpha(phi) = alpha_l (fl + k_ratio (1 - fl)), fl = (1 + phi)/2,
 * tau_g = 1/2 + alpha/cs2.  Call before oracle_initialization. */
void oracle_set_ext(oracle_t* s, double t_hot, double k_ratio) {
  s->t_hot = t_hot;
  s->k_ratio = k_ratio;
  s->tin = t_hot;
}

/* Extensions: east wall at a fixed temperature t_cold (BC_g's east ghost
 * equilibrium at t_cold instead of the averaged T(Nx, j)), and uniform
 * noise on the initial interior temperature, noise_amp (T0 - T_in_sol)
 * (2U - 1) with U the SplitMix64 output of state seed + (q+1) *
 * 0x9E3779B97F4A7C15 for node q = i + (Nx+2) j.  Call before
 * oracle_initialization. */
void oracle_set_ext2(oracle_t* s, int east_cold, double t_cold, double noise_amp,
                     uint64_t noise_seed) {
  s->east_cold = east_cold;
  s->t_cold = t_cold;
  s->noise_amp = noise_amp;
  s->noise_seed = noise_seed;
}

/* Extensions: the phase field's critical temperature t_crit in place of
 * phase_field.f90:12's 273, and a reference temperature t_ref of the
 * buoyancy, forces.f90:20 with T - t_ref in place of T (Boussinesq form;
 * 0 keeps STEP_7's absolute T).  Call before the first step. */
void oracle_set_ext3(oracle_t* s, double t_crit, double t_ref) {
  s->t_crit = t_crit;
  s->t_ref = t_ref;
}

static double splitmix_uniform(uint64_t seed, uint64_t q) {
  uint64_t z = seed + (q + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void oracle_destroy(oracle_t* s) {
  if (!s) return;
  for (int f = 0; f < BGK_NFIELDS; ++f) free(s->fld[f]);
  free(s);
}

double* oracle_field(oracle_t* s, int id) {
  if (!s || id < 0 || id >= BGK_NFIELDS) return NULL;
  return s->fld[id];
}

/* Second-order equilibria w_k*m*(...) of all nine directions, operand order
 * of STEP_7/collision.f90:17-41 (and initialization.f90:104-128). */
static inline void equilibria(const double* w, double m, double u, double v, double* eq) {
  double uu = u * u;
  double vv = v * v;
  double uv = 1.5 * (uu + vv);
  double upv = u + v;
  double umv = u - v;
  eq[0] = w[0] * m * (1.0 - uv);
  eq[1] = w[1] * m * (1.0 + 3.0 * u + 4.5 * uu - uv);
  eq[2] = w[2] * m * (1.0 + 3.0 * v + 4.5 * vv - uv);
  eq[3] = w[3] * m * (1.0 - 3.0 * u + 4.5 * uu - uv);
  eq[4] = w[4] * m * (1.0 - 3.0 * v + 4.5 * vv - uv);
  eq[5] = w[5] * m * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
  eq[6] = w[6] * m * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
  eq[7] = w[7] * m * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
  eq[8] = w[8] * m * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
}

/* ------------------------------------------------------- initialization */

void oracle_initialization(oracle_t* s) {
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const bgk_params* p = &s->p;
  double** F = s->fld;
  double *rho = F[BGK_RHO], *u = F[BGK_U], *v = F[BGK_V], *T = F[BGK_T];
  double *phi = F[BGK_PHI], *phip = F[BGK_PHI_PREC];

  s->interf = nx / 40;                    /* initialization.f90:5   */
  s->tin = s->t_hot;                      /* :9 (T0 unless extended) */
  s->u0 = p->re * s->kvisc / p->length;   /* :11                    */
  s->v0 = 0.0;                            /* :12                    */
  for (int j = 0; j <= ny + 1; ++j) {     /* :15-26                 */
    for (int i = 0; i <= s->interf && i <= nx + 1; ++i) {
      T[IX(i, j)] = s->t_hot;
      rho[IX(i, j)] = p->rho1;
    }
    for (int i = s->interf + 1; i <= nx + 1; ++i) {
      T[IX(i, j)] = p->t_in_sol;
      rho[IX(i, j)] = p->rho0;
    }
  }
  for (size_t q = 0; q < s->n; ++q) {     /* :49-50                 */
    u[q] = s->u0;
    v[q] = s->v0;
  }
  for (int j = 0; j <= ny + 1; ++j) T[IX(0, j)] = s->tin;   /* :72-74 */
  if (s->noise_amp > 0.0) { /* extension: seeded initial noise */
    const double a = s->noise_amp * (p->t0 - p->t_in_sol);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        T[IX(i, j)] += a * (2.0 * splitmix_uniform(s->noise_seed, IX(i, j)) - 1.0);
  }
  for (int i = 0; i <= nx + 1; ++i) {                        /* :76-81 */
    u[IX(i, ny + 1)] = 0.0;
    v[IX(i, ny + 1)] = 0.0;
    u[IX(i, 0)] = 0.0;
    v[IX(i, 0)] = 0.0;
  }
  if (p->west_bc == 3) {                                     /* :83-89 */
    for (size_t q = 0; q < s->n; ++q) {
      u[q] = 0.0;
      v[q] = 0.0;
    }
    for (int j = 0; j <= ny + 1; ++j) rho[IX(0, j)] = p->rho1;
  }
  for (int j = 1; j <= ny; ++j) {                            /* :91-99 */
    if (p->west_bc == 1) {
      u[IX(0, j)] = s->u0;
      v[IX(0, j)] = s->v0;
    } else if (p->west_bc == 2) {
      double yj = (double)j;                  /* y(j), setup.f90:93-98 */
      double np1 = (double)(ny + 1);
      u[IX(0, j)] = -(6.0 * s->u0 / (np1 * np1)) * (yj * yj - np1 * yj);
      v[IX(0, j)] = s->v0;
    }
  }
  for (int j = 0; j <= ny + 1; ++j) {                        /* :101-131 */
    for (int i = 0; i <= nx + 1; ++i) {
      double eq[9];
      size_t q = IX(i, j);
      equilibria(s->w, rho[q], u[q], v[q], eq);
      for (int k = 0; k < 9; ++k) F[BGK_F0 + k][q] = eq[k];
      equilibria(s->w, T[q], u[q], v[q], eq);
      for (int k = 0; k < 9; ++k) F[BGK_G0 + k][q] = eq[k];
    }
  }
  for (size_t q = 0; q < s->n; ++q) {                        /* :133-141 */
    phi[q] = -1.0;
    phip[q] = -1.0;
  }
  for (int j = 0; j <= ny + 1; ++j)                          /* :143-152 */
    for (int i = 0; i <= s->interf && i <= nx + 1; ++i) {
      phi[IX(i, j)] = 1.0;
      phip[IX(i, j)] = 1.0;
    }
}

/* ------------------------------------------------------------------ BC_f */

void oracle_bc_f(oracle_t* s) {
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const bgk_params* p = &s->p;
  const double* w = s->w;
  double** F = s->fld;
  double *f0 = F[0], *f1 = F[1], *f2 = F[2], *f3 = F[3], *f4 = F[4];
  double *f5 = F[5], *f6 = F[6], *f7 = F[7], *f8 = F[8];
  double *rho = F[BGK_RHO], *u = F[BGK_U], *v = F[BGK_V];
  double uu, vv, uv, upv, umv;
  const double u0 = s->u0, v0 = s->v0;
  (void)f0;

  /* WEST, BC_f.f90:9-61 */
  if (p->west_bc == 0) {
    for (int j = 0; j <= ny + 1; ++j)
      for (int k = 0; k < 9; ++k) F[k][IX(0, j)] = F[k][IX(nx, j)];
  } else if (p->west_bc == 1) {
    uu = u0 * u0; vv = v0 * v0; uv = 1.5 * (uu + vv); upv = u0 + v0; umv = u0 - v0;
    for (int j = 1; j <= ny; ++j) {
      f1[IX(0, j)] = w[1] * rho[IX(1, j)] * (1.0 + 3.0 * u0 + 4.5 * uu - uv);
      f5[IX(0, j)] = w[5] * rho[IX(1, j)] * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
      f8[IX(0, j)] = w[8] * rho[IX(1, j)] * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
      u[IX(0, j)] = u0;
      v[IX(0, j)] = v0;
      rho[IX(0, j)] = rho[IX(1, j)];
    }
  } else if (p->west_bc == 2) {
    for (int j = 0; j <= ny + 1; ++j) {
      double uw = u[IX(0, j)];
      uu = uw * uw; vv = v0 * v0; uv = 1.5 * (uu + vv); upv = uw + v0; umv = uw - v0;
      f1[IX(0, j)] = w[1] * rho[IX(1, j)] * (1.0 + 3.0 * uw + 4.5 * uu - uv);
      f5[IX(0, j)] = w[5] * rho[IX(1, j)] * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
      f8[IX(0, j)] = w[8] * rho[IX(1, j)] * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
    }
  } else if (p->west_bc == 3) {
    for (int j = 0; j <= ny + 1; ++j) {
      double u1 = u[IX(1, j)], v1 = v[IX(1, j)];
      uu = u1 * u1; vv = v1 * v1; uv = 1.5 * (uu + vv); upv = u1 + v1; umv = u1 - v1;
      f1[IX(0, j)] = w[1] * p->rho1 * (1.0 + 3.0 * u1 + 4.5 * uu - uv);
      f5[IX(0, j)] = w[5] * p->rho1 * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
      f8[IX(0, j)] = w[8] * p->rho1 * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
      u[IX(0, j)] = u1;
      v[IX(0, j)] = v1;
      rho[IX(0, j)] = p->rho1;
    }
  } else if (p->west_bc == 4) {
    for (int j = 1; j <= ny; ++j) {
      f1[IX(0, j)] = f3[IX(1, j)];
      f8[IX(0, j)] = f6[IX(1, j - 1)];
      f5[IX(0, j)] = f7[IX(1, j + 1)];
    }
  }

  /* EAST, BC_f.f90:65-104 */
  if (p->east_bc == 0) {
    for (int j = 0; j <= ny + 1; ++j)
      for (int k = 0; k < 9; ++k) F[k][IX(nx + 1, j)] = F[k][IX(1, j)];
  } else if (p->east_bc == 3) {
    for (int j = 1; j <= ny; ++j) {
      double ue = u[IX(nx, j)], ve = v[IX(nx, j)];
      uu = ue * ue; vv = ve * ve; uv = 1.5 * (uu + vv); upv = ue + ve; umv = ue - ve;
      f3[IX(nx + 1, j)] = w[3] * p->rho0 * (1.0 - 3.0 * ue + 4.5 * uu - uv);
      f6[IX(nx + 1, j)] = w[6] * p->rho0 * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
      f7[IX(nx + 1, j)] = w[7] * p->rho0 * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
      u[IX(nx + 1, j)] = ue;
      v[IX(nx + 1, j)] = ve;
      rho[IX(nx + 1, j)] = p->rho0;
    }
  } else if (p->east_bc == 4) {
    for (int j = 1; j <= ny; ++j) {
      f3[IX(nx + 1, j)] = f1[IX(nx, j)];
      f6[IX(nx + 1, j)] = f8[IX(nx, j + 1)];
      f7[IX(nx + 1, j)] = f5[IX(nx, j - 1)];
    }
  }

  /* NORTH, BC_f.f90:108-160 */
  if (p->north_bc == 0) {
    for (int i = 0; i <= nx + 1; ++i)
      for (int k = 0; k < 9; ++k) F[k][IX(i, ny + 1)] = F[k][IX(i, 1)];
  } else if (p->north_bc == 4) {
    for (int i = 1; i <= nx; ++i) {
      f4[IX(i, ny + 1)] = f2[IX(i, ny)];
      f7[IX(i, ny + 1)] = f5[IX(i - 1, ny)];
      f8[IX(i, ny + 1)] = f6[IX(i + 1, ny)];
      rho[IX(i, ny + 1)] = rho[IX(i, ny)];
      u[IX(i, ny + 1)] = 0.0;
      v[IX(i, ny + 1)] = 0.0;
    }
    f8[IX(0, ny + 1)] = f6[IX(1, ny)];
  } else if (p->north_bc == 5) {
    for (int i = 0; i <= nx + 1; ++i) {
      f4[IX(i, ny + 1)] = f2[IX(i, ny - 1)];
      f7[IX(i, ny + 1)] = f6[IX(i, ny - 1)];
      f8[IX(i, ny + 1)] = f5[IX(i, ny - 1)];
    }
  } else if (p->north_bc == 6) {
    for (int i = 0; i <= nx + 1; ++i) {
      f4[IX(i, ny + 1)] = f4[IX(i, ny)];
      f7[IX(i, ny + 1)] = f7[IX(i, ny)];
      f8[IX(i, ny + 1)] = f8[IX(i, ny)];
    }
  } else if (p->north_bc == 1) {
    uu = u0 * u0; vv = v0 * v0; uv = 1.5 * (uu + vv); upv = u0 + v0; umv = u0 - v0;
    for (int i = 0; i <= nx + 1; ++i) {
      f4[IX(i, ny + 1)] = w[4] * rho[IX(i, ny)] * (1.0 - 3.0 * v0 + 4.5 * vv - uv);
      f7[IX(i, ny + 1)] = w[7] * rho[IX(i, ny)] * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
      f8[IX(i, ny + 1)] = w[8] * rho[IX(i, ny)] * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
    }
  }

  /* SOUTH, BC_f.f90:164-196 */
  if (p->south_bc == 0) {
    for (int i = 0; i <= nx + 1; ++i)
      for (int k = 0; k < 9; ++k) F[k][IX(i, 0)] = F[k][IX(i, ny)];
  } else if (p->south_bc == 4) {
    for (int i = 1; i <= nx; ++i) {
      f2[IX(i, 0)] = f4[IX(i, 1)];
      f5[IX(i, 0)] = f7[IX(i + 1, 1)];
      f6[IX(i, 0)] = f8[IX(i - 1, 1)];
      rho[IX(i, 0)] = rho[IX(i, 1)];
      u[IX(i, 0)] = 0.0;
      v[IX(i, 0)] = 0.0;
    }
    f5[IX(0, 0)] = f7[IX(1, 1)];
  } else if (p->south_bc == 6) {
    for (int i = 0; i <= nx + 1; ++i) {
      f2[IX(i, 0)] = f2[IX(i, 1)];
      f5[IX(i, 0)] = f5[IX(i, 1)];
      f6[IX(i, 0)] = f6[IX(i, 1)];
    }
  }
}

/* ------------------------------------------------------------------ BC_g */

void oracle_bc_g(oracle_t* s) {
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const double* w = s->w;
  double** F = s->fld;
  double *u = F[BGK_U], *v = F[BGK_V], *T = F[BGK_T], *phi = F[BGK_PHI];
  double *g1 = F[BGK_G0 + 1], *g3 = F[BGK_G0 + 3], *g5 = F[BGK_G0 + 5];
  double *g6 = F[BGK_G0 + 6], *g7 = F[BGK_G0 + 7], *g8 = F[BGK_G0 + 8];
  const double tin = s->tin;
  double uu, vv, uv, upv, umv;

  for (int j = 1; j <= ny; ++j) { /* WEST, BC_g.f90:9-17 */
    double uw = u[IX(0, j)], vw = v[IX(0, j)];
    uu = uw * uw; vv = vw * vw; uv = 1.5 * (uu + vv); upv = uw + vw; umv = uw - vw;
    g1[IX(0, j)] = w[1] * tin * (1.0 + 3.0 * uw + 4.5 * uu - uv);
    g5[IX(0, j)] = w[5] * tin * (1.0 + 3.0 * upv + 4.5 * upv * upv - uv);
    g8[IX(0, j)] = w[8] * tin * (1.0 + 3.0 * umv + 4.5 * umv * umv - uv);
    T[IX(0, j)] = tin;
  }
  for (int j = 1; j <= ny; ++j) { /* EAST, BC_g.f90:21-31 */
    double ue = u[IX(nx + 1, j)], ve = v[IX(nx + 1, j)];
    double te = s->east_cold ? s->t_cold : T[IX(nx, j)]; /* cold wall: extension */
    uu = ue * ue; vv = ve * ve; uv = 1.5 * (uu + vv); upv = ue + ve; umv = ue - ve;
    g3[IX(nx + 1, j)] = w[3] * te * (1.0 - 3.0 * ue + 4.5 * uu - uv);
    g6[IX(nx + 1, j)] = w[6] * te * (1.0 - 3.0 * umv + 4.5 * umv * umv - uv);
    g7[IX(nx + 1, j)] = w[7] * te * (1.0 - 3.0 * upv + 4.5 * upv * upv - uv);
    T[IX(nx + 1, j)] = te;
  }
  for (int i = 0; i <= nx + 1; ++i) { /* NORTH, BC_g.f90:36-48 */
    for (int k = 0; k < 9; ++k) F[BGK_G0 + k][IX(i, ny + 1)] = F[BGK_G0 + k][IX(i, ny)];
    T[IX(i, ny + 1)] = T[IX(i, ny)];
    phi[IX(i, ny + 1)] = phi[IX(i, ny)];
  }
  for (int i = 0; i <= nx + 1; ++i) { /* SOUTH, BC_g.f90:52-64 */
    for (int k = 0; k < 9; ++k) F[BGK_G0 + k][IX(i, 0)] = F[BGK_G0 + k][IX(i, 1)];
    T[IX(i, 0)] = T[IX(i, 1)];
    phi[IX(i, 0)] = phi[IX(i, 1)];
  }
}

/* ------------------------------------------------------------- streaming */

static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

void oracle_streaming(oracle_t* s) { /* streaming.f90:5-27 (pull) */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  double** F = s->fld;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < 9; ++k) {
        F[BGK_FP0 + k][IX(i, j)] = F[BGK_F0 + k][IX(i - CX[k], j - CY[k])];
        F[BGK_GP0 + k][IX(i, j)] = F[BGK_G0 + k][IX(i - CX[k], j - CY[k])];
      }
}

/* ------------------------------------------------------------ momentsPre */

void oracle_moments_pre(oracle_t* s) { /* momentsPre.f90:8-21 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  double** F = s->fld;
  double** fp = F + BGK_FP0;
  double** gp = F + BGK_GP0;
  double *rho = F[BGK_RHO], *u = F[BGK_U], *v = F[BGK_V], *T = F[BGK_T];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      size_t q = IX(i, j);
      double r = fp[0][q] + fp[1][q] + fp[2][q] + fp[3][q] + fp[4][q] + fp[5][q] + fp[6][q] +
                 fp[7][q] + fp[8][q];
      double t = gp[0][q] + gp[1][q] + gp[2][q] + gp[3][q] + gp[4][q] + gp[5][q] + gp[6][q] +
                 gp[7][q] + gp[8][q];
      double irho = 1.0 / r;
      double uu = irho * (fp[1][q] - fp[3][q] + fp[5][q] - fp[6][q] - fp[7][q] + fp[8][q]);
      double vv = irho * (fp[2][q] - fp[4][q] + fp[5][q] + fp[6][q] - fp[7][q] - fp[8][q]);
      rho[q] = r;
      T[q] = t;
      u[q] = uu;
      v[q] = vv;
      F[BGK_RHO2][q] = r;
      F[BGK_T2][q] = t;
      F[BGK_U2][q] = uu;
      F[BGK_V2][q] = vv;
    }
}

/* ----------------------------------------------------------- phase_field */

void oracle_phase_field(oracle_t* s) { /* phase_field.f90:15-52 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  double *phi = s->fld[BGK_PHI], *pp = s->fld[BGK_PHI_PREC], *T = s->fld[BGK_T];
  const double t_act = 0.0, t_crit = s->t_crit, f_plus = 0.5, f_minus = 0.5;
  const double tw = s->p.t_width;
  memcpy(pp, phi, s->n * sizeof(double)); /* :19 phi_prec(:,:) = phi(:,:) */
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      /* taps in source order: C, E, N, W, S, NE, NW, SW, SE (:27-45) */
      const double tap[9] = {pp[IX(i, j)],         pp[IX(i + 1, j)],     pp[IX(i, j + 1)],
                             pp[IX(i - 1, j)],     pp[IX(i, j - 1)],     pp[IX(i + 1, j + 1)],
                             pp[IX(i - 1, j + 1)], pp[IX(i - 1, j - 1)], pp[IX(i + 1, j - 1)]};
      double tq = T[IX(i, j)];
      double react_plus = 0.5 * (1.0 + tanh(((tq - t_crit) - t_act) / tw));
      double react_minus = 0.5 * (1.0 - tanh(((tq - t_crit) + t_act) / tw));
      double a = 0.5 * (1.0 + tap[0]);
      double kp = a * a;
      for (int m = 1; m < 9; ++m) {
        a = 0.5 * (1.0 + tap[m]);
        kp = kp + a * a;
      }
      double b = 0.5 * (1.0 - tap[0]);
      double km = b * b;
      for (int m = 1; m < 9; ++m) {
        b = 0.5 * (1.0 - tap[m]);
        km = km + b * b;
      }
      kp = (kp / 8.0) * react_plus;
      km = (km / 8.0) * react_minus;
      double pc = tap[0];
      phi[IX(i, j)] = pc + (f_plus * kp * (1.0 - pc) - f_minus * km * (1.0 + pc));
    }
}

/* ---------------------------------------------------------------- forces */

void oracle_forces(oracle_t* s) { /* forces.f90:12-26 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const bgk_params* p = &s->p;
  double *rho = s->fld[BGK_RHO], *u = s->fld[BGK_U], *v = s->fld[BGK_V], *T = s->fld[BGK_T];
  double *phi = s->fld[BGK_PHI], *pp = s->fld[BGK_PHI_PREC];
  const double delta_t = p->t0 - p->t_in_sol;
  const double beta = r4(s->step7, (1.0 - p->rho1 / p->rho0) / delta_t);
  const double st = p->stefan_number, grav = p->grav, omf = s->omega_f;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      size_t q = IX(i, j);
      double t = T[q] - (delta_t / st * (phi[q] - pp[q]) / 2.0);
      double fy = rho[q] * grav * beta * (t - s->t_ref) * (1.0 + phi[q] / 2.0);
      double fx = -(1.0 * u[q] * (1.0 - phi[q]) * 0.5);
      fy = fy - 1.0 * v[q] * (1.0 - phi[q]) * 0.5;
      T[q] = t;
      u[q] = u[q] + fx / (omf * rho[q]);
      v[q] = v[q] + fy / (omf * rho[q]);
    }
}

/* ------------------------------------------------------------- collision */

/* Regularised BGK of one population set (collision.f90:44-150).  m is rho
 * (flow) or T (temperature); pop holds fp or gp; out receives f or g. */
static inline void regularised_bgk(const double* w, double cs2, double omega, int step7,
                                   double m, double u, double v, const double* pop,
                                   double* out) {
  double eq[9];
  equilibria(w, m, u, v, eq);
  const double a = -cs2, b = 1.0 - cs2;
  double d[9];
  for (int k = 0; k < 9; ++k) d[k] = pop[k] - eq[k];
  double pxx = a * d[0] + b * d[1] + a * d[2] + b * d[3] + a * d[4] + b * d[5] + b * d[6] +
               b * d[7] + b * d[8];
  double pyy = a * d[0] + a * d[1] + b * d[2] + a * d[3] + b * d[4] + b * d[5] + b * d[6] +
               b * d[7] + b * d[8];
  double pxy = d[5] - d[6] + d[7] - d[8];
  double c[9];
  for (int k = 0; k < 9; ++k) c[k] = (0.5 * w[k]) / (cs2 * cs2);
  double t[9];
  t[0] = r4(step7, eq[0] + c[0] * (a * pxx + a * pyy));
  t[1] = r4(step7, eq[1] + c[1] * (b * pxx + a * pyy));
  t[2] = r4(step7, eq[2] + c[2] * (a * pxx + b * pyy));
  t[3] = r4(step7, eq[3] + c[3] * (b * pxx + a * pyy));
  t[4] = r4(step7, eq[4] + c[4] * (a * pxx + b * pyy));
  t[5] = r4(step7, eq[5] + c[5] * (b * pxx + b * pyy + 2.0 * pxy));
  t[6] = r4(step7, eq[6] + c[6] * (b * pxx + b * pyy - 2.0 * pxy));
  t[7] = r4(step7, eq[7] + c[7] * (b * pxx + b * pyy + 2.0 * pxy));
  t[8] = r4(step7, eq[8] + c[8] * (b * pxx + b * pyy - 2.0 * pxy));
  for (int k = 0; k < 9; ++k) out[k] = t[k] - omega * (t[k] - eq[k]);
}

void oracle_collision(oracle_t* s) { /* collision.f90:14-152 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  double** F = s->fld;
  double *rho = F[BGK_RHO], *u = F[BGK_U], *v = F[BGK_V], *T = F[BGK_T];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      size_t q = IX(i, j);
      double pop[9], out[9];
      for (int k = 0; k < 9; ++k) pop[k] = F[BGK_FP0 + k][q];
      regularised_bgk(s->w, s->cs2, s->omega_f, s->step7, rho[q], u[q], v[q], pop, out);
      for (int k = 0; k < 9; ++k) F[BGK_F0 + k][q] = out[k];
      for (int k = 0; k < 9; ++k) pop[k] = F[BGK_GP0 + k][q];
      double omg = s->omega_g;
      if (s->k_ratio != 1.0) { /* extension: phase-dependent conductivity */
        double fl = 0.5 * (1.0 + F[BGK_PHI][q]);
        double alpha = s->alpha * (fl + s->k_ratio * (1.0 - fl));
        omg = 1.0 / (0.5 + alpha / s->cs2);
      }
      regularised_bgk(s->w, s->cs2, omg, s->step7, T[q], u[q], v[q], pop, out);
      for (int k = 0; k < 9; ++k) F[BGK_G0 + k][q] = out[k];
    }
}

/* --------------------------------------------------------------- moments */

void oracle_moments(oracle_t* s) { /* moments.f90:11-25 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  double** F = s->fld;
  double** f = F + BGK_F0;
  double** g = F + BGK_G0;
  double *rho = F[BGK_RHO], *u = F[BGK_U], *v = F[BGK_V], *T = F[BGK_T];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      size_t q = IX(i, j);
      double r = f[0][q] + f[1][q] + f[2][q] + f[3][q] + f[4][q] + f[5][q] + f[6][q] + f[7][q] +
                 f[8][q];
      double irho = 1.0 / r;
      double tt = g[0][q] + g[1][q] + g[2][q] + g[3][q] + g[4][q] + g[5][q] + g[6][q] + g[7][q] +
                  g[8][q];
      double tu = irho * (f[1][q] - f[3][q] + f[5][q] - f[6][q] - f[7][q] + f[8][q]);
      double tv = irho * (f[2][q] - f[4][q] + f[5][q] + f[6][q] - f[7][q] - f[8][q]);
      rho[q] = r;
      T[q] = (tt + F[BGK_T2][q]) * 0.5;
      u[q] = (tu + F[BGK_U2][q]) * 0.5;
      v[q] = (tv + F[BGK_V2][q]) * 0.5;
    }
}

void oracle_step(oracle_t* s, int nsteps) { /* main.f90:60-69 */
  for (int n = 0; n < nsteps; ++n) {
    oracle_bc_f(s);
    oracle_bc_g(s);
    oracle_streaming(s);
    oracle_moments_pre(s);
    oracle_phase_field(s);
    oracle_forces(s);
    oracle_collision(s);
    oracle_moments(s);
  }
}

/* ----------------------------------------------------------- reductions */

void oracle_field_check(const oracle_t* s, double out[3]) { /* field_check.f90:11-21 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const double *rho = s->fld[BGK_RHO], *T = s->fld[BGK_T], *phi = s->fld[BGK_PHI];
  double m = 0.0, t = 0.0, ph = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      m = m + rho[IX(i, j)];
      t = t + T[IX(i, j)];
      ph = ph + phi[IX(i, j)];
    }
  double nn = (double)(int32_t)((uint32_t)nx * (uint32_t)ny);
  out[0] = m / nn;
  out[1] = t / nn;
  out[2] = ph / nn;
}

void oracle_mass_check(const oracle_t* s, double out[3]) { /* mass_check.f90:12-29 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const double *rho = s->fld[BGK_RHO], *T = s->fld[BGK_T];
  double mass = 0.0, maver = 0.0, taver = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      mass = mass + rho[IX(i, j)];
      maver = maver + rho[IX(i, j)];
    }
  double nn = (double)(int32_t)((uint32_t)nx * (uint32_t)ny);
  mass = (nn * s->p.rho0 - mass) * 100.0 / (nn * s->p.rho0);
  maver = maver / nn;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) taver = taver + T[IX(i, j)];
  taver = taver / nn;
  out[0] = mass;
  out[1] = maver;
  out[2] = taver;
}

void oracle_diag_sums(const oracle_t* s, double out[4]) { /* diag.f90:63-104 */
  const int nx = s->nx, ny = s->ny, ld = s->ld;
  const double *T = s->fld[BGK_T], *phi = s->fld[BGK_PHI];
  const double dT = s->p.t0 - s->p.t_in_sol;
  const double dny = (double)ny;
  double nc = 0.0, nc2 = 0.0, melt = 0.0, front = 0.0;
  for (int j = 1; j <= ny; ++j) nc = nc + dny * ((s->tin - T[IX(2, j)]) / 2.0) / dT;
  for (int j = 1; j <= ny; ++j) {
    double gt = fabs((-(3.0 * T[IX(1, j)]) + 4.0 * T[IX(2, j)] - T[IX(3, j)]) / 2.0);
    nc2 = nc2 + dny * gt / dT;
  }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      if (phi[IX(i, j)] > 0.0) melt = melt + 1.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      if (phi[IX(i, j)] == 0.0) front = front + (double)i;
  out[0] = nc;
  out[1] = nc2;
  out[2] = melt;
  out[3] = front;
}

void oracle_diag(const oracle_t* s, int it, double out[12], int* valid) { /* diag.f90:15-114 */
  const bgk_params* p = &s->p;
  const int nx = s->nx, ny = s->ny;
  double sums[4];
  oracle_diag_sums(s, sums);
  const double dT = p->t0 - p->t_in_sol;
  const double beta_gras = (p->rho0 / p->rho1 - 1.0) / dT;
  /* Ny**3 is default-INTEGER arithmetic and wraps past 2**31 */
  const int32_t ny3 = (int32_t)((uint32_t)ny * (uint32_t)ny * (uint32_t)ny);
  const int32_t ny2 = (int32_t)((uint32_t)ny * (uint32_t)ny);
  const double gras = p->grav * beta_gras * dT * (double)ny3 / (s->kvisc * s->kvisc);
  const double pr = (2.0 * p->tau_f - 1.0) / (2.0 * p->tau_g - 1.0);
  const double ray = pr * gras;
  const double fou = s->alpha * (double)it / (double)ny2;
  const double theta = p->stefan_number * fou;
  const double theta1 = pow(ray, -0.5);
  const double theta2 = (double)(nx / ny) * pow(ray, -0.25);
  const double nus = pow(2.0 * theta, -0.5) +
                     (0.35 * pow(ray, 0.25) - pow(2.0 * theta, -0.5)) *
                         pow(1.0 + pow(0.0175 * pow(ray, 0.75) * pow(theta, 1.5), -2.0), -0.5);
  const double nus_comp = sums[0] / (double)ny;
  const double nus_comp_2 = sums[1] / (double)ny;
  double melt = 0.0, front = 0.0, front_exp = 0.0;
  int v = 0;
  if (it % 100 == 0) {
    melt = sums[2] / (double)(int32_t)((uint32_t)nx * (uint32_t)ny);
    if (theta <= theta1) {
      front = (double)ny / pow(2.0 * theta, -0.5);
      front_exp = (double)ny / nus;
      v = 1;
    } else if (theta > theta1 && theta < theta2) {
      front = sums[3] / (double)ny;
      front_exp = (double)ny * theta * pow(ray, 0.25);
      v = 1;
    }
  }
  out[0] = (double)it;
  out[1] = theta;
  out[2] = nus;
  out[3] = nus_comp;
  out[4] = nus_comp_2;
  out[5] = melt;
  out[6] = front;
  out[7] = front_exp;
  out[8] = gras;
  out[9] = pr;
  out[10] = ray;
  out[11] = fou;
  if (valid) *valid = v;
}
