/*
 * bgk_oracle.h -- CPU restatement of the reference STEP_7 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (bgk_pcm_amd) never links or calls it.
 *
 * Each function restates one subroutine of giacomofalcucci/BGK_PCM STEP_7
 * (file:line beside it) with the evaluation order Fortran gives the source
 * expressions (left to right within a precedence level, parentheses kept),
 * no contraction into FMA, and the implicit default-REAL (FP32) temporaries
 * of STEP_7 under BGK_ROUND_STEP7.  Built with gcc -O2 -ffp-contract=off on
 * x86-64 it reproduces the amdflang -O3 build of STEP_7 bit for bit; the
 * pinning evidence lives in tests/test_oracle_*.py:
 *   - against oracle/_ref (STEP_7 compiled from its own sources), stage by
 *     stage and over many steps, both rounding policies;
 *   - against the committed golden fixtures tests/golden/*.npz written by
 *     that build;
 *   - against closed forms (equilibrium at rest, bounce-back, permutation
 *     of streaming, conservation in the clean policy, phase-field fixed
 *     points).
 * Reductions (field_check, mass_check, diag sums) are serial in the
 * reference's j-outer/i-inner order.
 */
#ifndef BGK_ORACLE_H
#define BGK_ORACLE_H

#include <stdint.h>

#include "../include/bgk_pcm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

/* STEP_7/setup.f90:34-115 (constants and allocation). */
oracle_t* oracle_create(const bgk_params* p);
void oracle_destroy(oracle_t* s);
/* Physics extensions of this library (not in the reference; t_hot = T0 and
 * k_ratio = 1 give STEP_7): hot-wall / initial liquid temperature and the
 * solid/liquid conductivity ratio.  Call before oracle_initialization. */
void oracle_set_ext(oracle_t* s, double t_hot, double k_ratio);
/* Cold east wall at t_cold and seeded initial temperature noise (see
 * bgk_ext in include/bgk_pcm.h). */
void oracle_set_ext2(oracle_t* s, int east_cold, double t_cold, double noise_amp,
                     uint64_t noise_seed);
/* Phase-field critical temperature (273 in STEP_7) and buoyancy reference
 * temperature (0: STEP_7's absolute T). */
void oracle_set_ext3(oracle_t* s, double t_crit, double t_ref);

/* STEP_7/initialization.f90:5-152 */
void oracle_initialization(oracle_t* s);

void oracle_bc_f(oracle_t* s);        /* STEP_7/BC_f.f90:7-200        */
void oracle_bc_g(oracle_t* s);        /* STEP_7/BC_g.f90:7-64         */
void oracle_streaming(oracle_t* s);   /* STEP_7/streaming.f90:5-27    */
void oracle_moments_pre(oracle_t* s); /* STEP_7/momentsPre.f90:5-23   */
void oracle_phase_field(oracle_t* s); /* STEP_7/phase_field.f90:11-55 */
void oracle_forces(oracle_t* s);      /* STEP_7/forces.f90:12-28      */
void oracle_collision(oracle_t* s);   /* STEP_7/collision.f90:14-152  */
void oracle_moments(oracle_t* s);     /* STEP_7/moments.f90:8-27      */

/* nsteps of STEP_7/main.f90:60-69 */
void oracle_step(oracle_t* s, int nsteps);

/* Pointer to field `id` (BGK_* field ids), (Nx+2)*(Ny+2) doubles, ld=Nx+2. */
double* oracle_field(oracle_t* s, int id);

/* STEP_7/field_check.f90:11-21 -> {<rho>,<T>,<phi>} */
void oracle_field_check(const oracle_t* s, double out[3]);
/* STEP_7/mass_check.f90:12-29 -> {mass, <rho>, <T>} */
void oracle_mass_check(const oracle_t* s, double out[3]);
/* STEP_7/diag.f90:63-104 field sums, same layout as bgk_diag_sums. */
void oracle_diag_sums(const oracle_t* s, double out[4]);
/* STEP_7/diag.f90:15-114: the 12 values of format 640 for iteration `it`
 * (it, theta, Nus, Nus_comp, Nus_comp_2, melt, melt_front,
 * melt_front_expected, Gras, Pr, Ray, Fou).  melt and the melt fronts are
 * only defined when mod(it,100)==0; `valid` flags which fronts were set. */
void oracle_diag(const oracle_t* s, int it, double out[12], int* valid);

/* glibc tanh, as called by the reference (STEP_7/phase_field.f90:24-25) */
double oracle_tanh(double x);

#ifdef __cplusplus
}
#endif
#endif
