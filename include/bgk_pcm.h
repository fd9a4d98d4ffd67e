/*
 * bgk_pcm.h -- C ABI of the MI355X (gfx950) BGK_PCM time step.
 *
 * The reference program (giacomofalcucci/BGK_PCM, STEP_7) keeps its whole
 * state in Fortran `module shared` (STEP_7/shared.f90:1-104) and advances it
 * with argument-less subroutines called from the time loop of
 * STEP_7/main.f90:60-69.  This header is the drop-in boundary for that loop:
 * a handle owns the same state resident in HBM of one GPU, and each entry
 * point below replaces one reference subroutine (cited beside it).  A Fortran
 * caller binds these with ISO_C_BINDING (`bind(C, name="...")`).
 *
 * Conventions
 *  - Every field is a (0:Nx+1, 0:Ny+1) array in Fortran column-major order:
 *    node (i,j) lives at host[i + ld*j] with ld >= Nx+2 (interior i=1..Nx,
 *    j=1..Ny, one halo layer all around), exactly as allocated in
 *    STEP_7/setup.f90:34-77.
 *  - Every entry point returns BGK_OK (0) or a negative BGK_E_* code; the
 *    message of the last failure on the calling thread is bgk_last_error().
 *  - Calls on one handle are issued in order on the handle's HIP stream.
 *    bgk_download, the reductions and bgk_synchronize wait for completion;
 *    all other calls return once the work is queued.  A handle must be used
 *    by one host thread at a time; different handles are independent.
 *  - No entry point falls back to host computation: without a usable GPU,
 *    bgk_create fails with BGK_E_NODEVICE.
 */
#ifndef BGK_PCM_H
#define BGK_PCM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BGK_ABI_VERSION 4

/* return codes */
enum {
  BGK_OK = 0,
  BGK_E_INVALID = -1,     /* bad argument (null pointer, size, field id, ...) */
  BGK_E_HIP = -2,         /* a HIP runtime call failed                         */
  BGK_E_NOMEM = -3,       /* device allocation failed                          */
  BGK_E_UNSUPPORTED = -4, /* valid request this handle cannot serve            */
  BGK_E_NODEVICE = -5,    /* no GPU visible / wrong device ordinal             */
  BGK_E_NONFINITE = -6    /* a reduction met a NaN or an infinity in the fields
                             it sums; its outputs are written (non-finite)     */
};

/* storage precision of the populations (bgk_params.precision) */
enum {
  BGK_FP64 = 0,     /* f, g stored and computed in FP64                        */
  BGK_FP32_DEV = 1, /* f - w*rho0 and g - w*T_crit (T_crit = 273) stored in
                       FP32; phase field, enthalpy update and moments in FP64  */
  BGK_FP32_MOM = 2, /* moment storage: per node rho and T - T_crit in FP64,
                       u, v and the non-equilibrium second moments of f and g
                       (times 1 - omega) in FP32, 48 bytes instead of 144;
                       populations are rebuilt from them at each pull.  Clean
                       rounding of the collision (STEP_7's FP32 temporaries
                       ft, gt are not representable); no-slip cavity only;
                       populations are not addressable (download/upload of
                       f, g, fp, gp, rho2..T2 and all uploads: unsupported) */
  BGK_FP64_MOM = 3  /* the same moment storage with every moment in FP64: the
                       exact forward model of the adjoint (bgk_adj_*)        */
};

/* rounding policy of the FP64 arithmetic (bgk_params.rounding) */
enum {
  BGK_ROUND_STEP7 = 0, /* as compiled STEP_7: the regularised populations
                          ft0..ft8, gt0..gt8 (STEP_7/collision.f90:94-130) and
                          beta (STEP_7/forces.f90:13) are implicitly typed
                          default REAL, i.e. rounded to FP32                   */
  BGK_ROUND_CLEAN = 1  /* the same expressions with those temporaries in FP64
                          (STEP_7 built with default REAL = REAL(8))           */
};

/* Field identifiers for bgk_upload / bgk_download / bgk_gather.  Names and
 * meaning follow STEP_7/shared.f90:36-79. */
enum {
  BGK_F0 = 0,   /* f0..f8   : 0..8   post-collision flow populations        */
  BGK_FP0 = 9,  /* fp0..fp8 : 9..17  post-streaming flow populations        */
  BGK_G0 = 18,  /* g0..g8   : 18..26 post-collision temperature populations */
  BGK_GP0 = 27, /* gp0..gp8 : 27..35 post-streaming temperature populations */
  BGK_RHO = 36,
  BGK_U = 37,
  BGK_V = 38,
  BGK_T = 39,
  BGK_RHO2 = 40,
  BGK_U2 = 41,
  BGK_V2 = 42,
  BGK_T2 = 43,
  BGK_PHI = 44,
  BGK_PHI_PREC = 45,
  BGK_NFIELDS = 46
};

/* Simulation parameters.  The first 21 members are the 21 records of
 * input.dat, in file order, with the meaning STEP_7/setup.f90:7-31 gives
 * them; the rest select options of this library. */
typedef struct bgk_params {
  int32_t iter;          /* id(1)  number of time steps                     */
  int32_t noutput;       /* id(2)  number of VTK outputs                    */
  int32_t dump;          /* id(3)  restart flag                             */
  int32_t nx;            /* id(4)  interior nodes along x                   */
  int32_t ny;            /* id(5)  interior nodes along y                   */
  int32_t reserved0;     /* keeps the doubles 8-byte aligned; set to 0      */
  double re;             /* id(6)  Reynolds number, sets u0                 */
  double rho0;           /* id(7)  density of the solid phase               */
  double rho1;           /* id(8)  density of the melted phase              */
  double tau_f;          /* id(9)  flow relaxation time                     */
  double tau_g;          /* id(10) thermal relaxation time                  */
  double t0;             /* id(11) hot (west) temperature                   */
  double t_in_sol;       /* id(12) initial solid temperature                */
  double grav;           /* id(13) gravity                                  */
  double eps;            /* id(14) read by the reference, never used        */
  double stefan_number;  /* id(15)                                          */
  double t_width;        /* id(16) width of the melting tanh                */
  double length;         /* id(17) characteristic length (sets u0)          */
  int32_t west_bc;       /* id(18) 0 periodic,1 velocity,2 parabolic,3 pressure,4 no-slip */
  int32_t east_bc;       /* id(19) 0 periodic,3 pressure,4 no-slip          */
  int32_t north_bc;      /* id(20) 0 periodic,1 velocity,4 no-slip,5 symmetric,6 open */
  int32_t south_bc;      /* id(21) 0 periodic,4 no-slip,6 open              */
  int32_t precision;     /* BGK_FP64, BGK_FP32_DEV, BGK_FP32_MOM, BGK_FP64_MOM */
  int32_t rounding;      /* BGK_ROUND_STEP7 or BGK_ROUND_CLEAN              */
  int32_t device;        /* HIP device ordinal                              */
  int32_t fp32_collision;/* FP32 handles: 1 = collision arithmetic in FP32,
                            0 (default) = FP64 arithmetic on FP32 storage   */
} bgk_params;

/* Memory report (replaces STEP_7/get_mem.c:23-38, which reports ru_maxrss). */
typedef struct bgk_info {
  int64_t device_bytes;     /* HBM held by this handle                    */
  int64_t device_free;      /* hipMemGetInfo free bytes after allocation  */
  int64_t device_total;     /* hipMemGetInfo total bytes                  */
  double host_maxrss_mb;    /* getrusage ru_maxrss of this process, MB     */
  int32_t pitch;            /* device row pitch in elements               */
  int32_t fused_path;       /* 1 if bgk_step runs the fused kernel        */
} bgk_info;

/* Physics extensions (not in the reference; the defaults t_hot = 0 ->
 * T0, k_ratio = 1, east_cold = 0, noise_amp = 0, t_crit = 0 -> 273 and
 * t_ref = 0 give STEP_7's step).  Every handle honours every extension (the
 * FP64 population kernels, fused and per-subroutine, bit for bit as the
 * oracle's restatement with the same extensions). */
typedef struct bgk_ext {
  double t_hot;       /* hot (west) wall and initial liquid temperature in
                         place of T0 (0: T0); the scale T0 - T_in_sol of
                         STEP_7/forces.f90:12-13 is unchanged               */
  double k_ratio;     /* solid/liquid thermal conductivity ratio: alpha(phi)
                         = alpha_l (fl + k_ratio (1 - fl)), fl = (1+phi)/2,
                         tau_g = 1/2 + alpha/cs2 (alpha_l from tau_g)       */
  int32_t east_cold;  /* 1: east wall at the fixed temperature t_cold
                         (Dirichlet) in place of STEP_7's averaged T(Nx, j)
                         of the previous step (BC_g.f90:21-31)              */
  int32_t reserved;   /* set to 0                                           */
  double t_cold;      /* east wall temperature when east_cold = 1           */
  double noise_amp;   /* initial interior T += noise_amp (T0 - T_in_sol)
                         (2U - 1), U in [0,1) from SplitMix64 with state
                         noise_seed + (q+1) 0x9E3779B97F4A7C15, q = i +
                         (Nx+2) j; 0: none                                  */
  uint64_t noise_seed;
  double t_crit;      /* critical temperature of the phase field in place of
                         STEP_7/phase_field.f90:12's 273 (0: 273)           */
  double t_ref;       /* buoyancy reference temperature: forces.f90:20 with
                         T - t_ref in place of T (the Boussinesq form: no
                         hydrostatic part at T = t_ref); 0: STEP_7's
                         absolute T                                          */
} bgk_ext;

typedef struct bgk_handle bgk_handle;

int bgk_abi_version(void);
const char* bgk_last_error(void);

/* Allocation and constants: STEP_7/setup.f90:34-115 (the input.dat read of
 * setup.f90:5-32 stays with the caller, who fills bgk_params).  Validates the
 * parameters and allocates every field on the device. */
int bgk_create(const bgk_params* params, bgk_handle** out);
int bgk_destroy(bgk_handle* h);

/* Use an existing HIP stream (hipStream_t passed as void*) for all later
 * work of this handle; NULL returns to the handle's own stream. */
int bgk_set_stream(bgk_handle* h, void* hip_stream);

/* STEP_7/initialization.f90:5-152: initial rho, u, v, T, phi, phi_prec and
 * equilibrium populations on every node, halo included. */
int bgk_initialization(bgk_handle* h);

/* Set the physics extensions (bgk_ext); takes effect from the next
 * bgk_initialization.  No reference counterpart: the extensions express
 * the configurations STEP_7 cannot (hot-wall temperature as a parameter of
 * its own, conductivity ratio). */
int bgk_set_extensions(bgk_handle* h, const bgk_ext* e);

/* Copy one field between host and device.  src/dst hold (Nx+2) x (Ny+2)
 * doubles with leading dimension ld >= Nx+2.  Replaces the host<->device
 * copies the GPU variants make around the time loop (STEP_CUF/main.f90:52-101,
 * 129-158) and the restart read of STEP_7/dump_in.f90:5-20. */
int bgk_upload(bgk_handle* h, int32_t field, const double* src, int64_t ld);
int bgk_download(bgk_handle* h, int32_t field, double* dst, int64_t ld);

/* One subroutine of the time step each; same order of calls as
 * STEP_7/main.f90:60-69.  FP64 handles only. */
int bgk_bc_f(bgk_handle* h);        /* STEP_7/BC_f.f90:7-200       */
int bgk_bc_g(bgk_handle* h);        /* STEP_7/BC_g.f90:7-64        */
int bgk_streaming(bgk_handle* h);   /* STEP_7/streaming.f90:5-27   */
int bgk_moments_pre(bgk_handle* h); /* STEP_7/momentsPre.f90:5-23  */
int bgk_phase_field(bgk_handle* h); /* STEP_7/phase_field.f90:11-55 */
int bgk_forces(bgk_handle* h);      /* STEP_7/forces.f90:12-28     */
int bgk_collision(bgk_handle* h);   /* STEP_7/collision.f90:14-152 */
int bgk_moments(bgk_handle* h);     /* STEP_7/moments.f90:8-27     */

/* nsteps iterations of the block STEP_7/main.f90:60-69 (BC_f, BC_g,
 * streaming, momentsPre, phase_field, forces, collision, moments).  With
 * no-slip walls on all four sides (BC type 4, the only walls used by the
 * reference inputs UTIL/input.dat.0..3) this is one fused kernel per step;
 * other boundary types run the per-subroutine kernels (FP64 handles only).
 * Afterwards f, g, rho, u, v, T and phi (halo included) hold what the
 * reference holds after the same steps; fp, gp, rho2, u2, v2, T2 and
 * phi_prec are scratch of the per-subroutine path and are left unspecified. */
int bgk_step(bgk_handle* h, int32_t nsteps);

/* STEP_7/field_check.f90:11-24: out = { <rho>, <T>, <phi> } over the
 * interior.  Returns BGK_E_NONFINITE (out still written) when a mean is not
 * finite: the reference prints NaN and runs on; a caller that wants that
 * behaviour ignores this code. */
int bgk_field_check(bgk_handle* h, double out[3]);

/* Allocation-free variant of bgk_field_check.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc),
 * which cannot hold a handle either.  The same means { <rho>, <T>, <phi> }
 * over the interior, of three device planes the caller owns (node (i,j) at
 * plane[i + pitch*j], i = 0..Nx+1, j = 0..Ny+1, pitch >= Nx+2), computed on
 * the current device on `hip_stream` (NULL: the null stream), which this call
 * then waits for.  The partial sums live in the library's static device
 * scratch (64 KiB per GPU, any lattice size: the first pass uses at most 1024
 * workgroups); it calls no hipMalloc/hipFree.  Every call on a GPU shares
 * that scratch: calls that overlap on one GPU (other threads, other streams)
 * race and must be serialised by the caller.  BGK_E_NONFINITE as
 * bgk_field_check. */
int bgk_field_check_planes(const double* rho, const double* T, const double* phi, int32_t nx,
                           int32_t ny, int64_t pitch, void* hip_stream, double out[3]);

/* STEP_7/mass_check.f90:8-32: out = { mass (percent defect),
 * <rho>, <T> } over the interior; BGK_E_NONFINITE as bgk_field_check. */
int bgk_mass_check(bgk_handle* h, double out[3]);

/* Field sums used by STEP_7/diag.f90:63-104:
 * out[0] = sum_j Ny*((Tin - T(2,j))/2)/Delta_T            (before /Ny)
 * out[1] = sum_j Ny*|(-3T(1,j)+4T(2,j)-T(3,j))/2|/Delta_T (before /Ny)
 * out[2] = number of interior nodes with phi > 0
 * out[3] = sum of i over interior nodes with phi == 0
 * BGK_E_NONFINITE (out still written) when out[0] or out[1] is not finite. */
int bgk_diag_sums(bgk_handle* h, double out[4]);

/* Values of one field at n nodes (i[k], j[k]) (0-based halo-inclusive
 * indices as in Fortran: 0..Nx+1, 0..Ny+1).  Serves STEP_7/probe.f90:7-15,
 * prof_x.f90:7-17 and prof_y.f90:7-17. */
int bgk_gather(bgk_handle* h, int32_t field, int32_t n, const int32_t* i,
               const int32_t* j, double* out);

/* Wait for all queued work of the handle. */
int bgk_synchronize(bgk_handle* h);

/* ---- Adjoint of the time step (no reference counterpart: the reference
 * has no derivatives).  Differentiates the moment step (precision
 * BGK_FP64_MOM, or BGK_FP32_MOM: then the adjoint of the step evaluated at
 * the FP32-stored states, the storage rounding taken as the identity;
 * no-slip cavity) for a loss that depends on the averaged temperature T
 * (STEP_7/moments.f90:11-25) at n_probes interior nodes (pi[p], pj[p]) after
 * each of nsteps steps.  The state is checkpointed every seg_len steps and
 * recomputed segment by segment.
 *
 * bgk_adj_setup     allocates the checkpoints and adjoint state.
 * bgk_adj_setup_levels  the same with two levels of checkpoints: every
 *                   seg_outer steps during the forward run, every seg_inner
 *                   steps of one outer segment while it is recomputed, and
 *                   every step of one inner segment: nsteps/seg_outer +
 *                   seg_outer/seg_inner + seg_inner + 1 states instead of
 *                   nsteps/seg_inner + seg_inner + 1, for one more forward
 *                   recomputation.  seg_outer = 0: one level (bgk_adj_setup).
 * bgk_adj_info      device bytes of the adjoint's checkpoints and state, and
 *                   the number of stored lattice states among them.
 * bgk_adj_forward   runs nsteps steps from the current state (the
 *                   initialization, for the t_hot gradient to include the
 *                   initial liquid temperature), writing the probe record
 *                   probe_T[n*n_probes + p] = T after step n+1.
 * bgk_adj_backward  takes seeds[n*n_probes + p] = dL/dprobe_T[...] and
 *                   returns grad = dL/d(stefan_number, k_ratio, t_hot)
 *                   (k_ratio and t_hot: bgk_ext). */
int bgk_adj_setup(bgk_handle* h, int32_t nsteps, int32_t seg_len, int32_t n_probes,
                  const int32_t* pi, const int32_t* pj);
int bgk_adj_setup_levels(bgk_handle* h, int32_t nsteps, int32_t seg_outer, int32_t seg_inner,
                         int32_t n_probes, const int32_t* pi, const int32_t* pj);
int bgk_adj_info(bgk_handle* h, int64_t* device_bytes, int32_t* n_states);

/* Binomial checkpointing (the "revolve" schedule of Griewank and Walther,
 * ACM TOMS 26 (2000)) in place of checkpoint levels: n_states snapshot slots
 * (the initial state's included; 0: as many as the free device memory holds,
 * 2 GiB kept back) plus two working states.  The forward run stores the
 * states the schedule's first branch needs; the backward pass recomputes the
 * minimal number of forward steps for that memory (bgk_adj_plan): 1.05 per
 * step for 5000 steps with 95 snapshots (8192^2 FP32 moments in an MI355X's
 * free memory), where two checkpoint levels recompute 2.  bgk_adj_forward
 * and bgk_adj_backward as above; the gradient
 * equals the level schedules' bit for bit (the same states are recomputed). */
int bgk_adj_setup_binomial(bgk_handle* h, int32_t nsteps, int32_t n_states, int32_t n_probes,
                           const int32_t* pi, const int32_t* pj);

/* Forward steps the backward pass of a binomial schedule of nsteps steps
 * with n_states snapshot slots recomputes (host arithmetic only, no handle
 * or device needed): to choose the memory a run spends. */
int bgk_adj_plan(int32_t nsteps, int32_t n_states, int64_t* recomputed_steps);
int bgk_adj_forward(bgk_handle* h, double* probe_T);
int bgk_adj_backward(bgk_handle* h, const double* seeds, double grad[3]);

/* Device memory held, device totals and host RSS. */
int bgk_get_info(bgk_handle* h, bgk_info* info);

#ifdef __cplusplus
}
#endif

#endif /* BGK_PCM_H */
