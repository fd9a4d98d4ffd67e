// bgk_internal.h -- handle layout and kernel launchers of libbgk_pcm.
#ifndef BGK_INTERNAL_H
#define BGK_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bgk_pcm.h"
#include "bgk_device.h"

namespace bgk {
struct AdjCtx;  // adjoint context (bgk_adjoint.hip)
}

// One lattice resident in HBM.
//
// Layout: every field is a plane of (Ny+2) rows of `pitch` nodes; node
// (i,j) of a plane is at plane[j*pitch + i] (pair planes: elements
// 2*(j*pitch+i) and +1).  The pitch is a multiple of 64 nodes and not of
// 256, and each plane starts one node before a 256-byte boundary, so node
// (1,j) of every row starts a cache line.  All planes come from one
// allocation (`pool`).
//
// Populations are double-buffered: f[cur] is the reference's f (post
// collision) and f[1-cur] its fp (post streaming); the fused step reads
// f[cur], writes f[1-cur] and flips cur.  f_k and g_k of a buffer share one
// pair plane (f at element 2q, g at 2q+1): f[b][k] points at the plane and
// g[b][k] = f[b][k] + 1.  phi[pc] is phi and phi[1-pc] phi_prec.  FP32
// handles hold f32/g32 = f - w*rho_ref, g - w*273 (T_crit) the same way,
// rho_ref the node's reference density from the tables rref[rc] (the solid
// and the liquid density of each row, blended by the liquid fraction).
// Moment handles hold per node ma = (rho, T - 273) FP64, mb = (u, v, Pf_xx,
// Pf_yy) and mc = (Pf_xy, Pg_xx, Pg_yy, Pg_xy) FP32, double-buffered by cur
// (with omega_f = 1: mb = (u, v, Pg_xx, Pg_yy), mc = Pg_xy).
struct bgk_handle {
  bgk_params prm;
  bgk_ext ext;  // physics extensions (defaults: STEP_7)
  int nx, ny;
  int64_t pitch;        // elements per row
  int64_t plane_elems;  // nodes per plane: (ny+2)*pitch
  int device;
  int n_cu;  // compute units of the device
  hipStream_t own_stream;
  hipStream_t stream;

  void* pool;
  size_t pool_bytes;

  // FP64 state (FP64 handles)
  double* f[2][9];
  double* g[2][9];
  // FP32 deviation state (FP32 handles)
  float* f32[2][9];
  float* g32[2][9];
  // FP32 moment state (moment handles)
  double2* ma[2];
  void* mb[2];    // float4 (FP32 moments) or double4 (FP64 moments) planes
  void* mc[2];    // the same, or float / double planes when mom_eqf
  bool mom_eqf;   // omega_f == 1: Pf vanishes, mb = (u, v, Pg_xx, Pg_yy), mc = Pg_xy
  int strip_rows;   // rows per workgroup strip of the moment kernel (0: default)
  int mom_variant;  // tiling variant of the moment kernel (0: default)
  int cur;

  double* phi[2];
  int pc;
  double *rho, *u, *v, *T, *rho2, *u2, *v2, *T2;
  double* teast[2];  // averaged T(Nx, 0..Ny+1) of the previous step, ping-pong
  // FP32 population handles: solid and liquid reference densities of rows
  // 0..Ny+1 ([0, Ny+2) and [Ny+2, 2(Ny+2)) of rref[rc]; the other is the
  // next table) the f members are stored against, and the re-basing cadence
  double* rref[2];
  int rc;
  int since_rebase;
  int rebase_every;
  double* tmp;       // one FP64 plane: staging for population transfers
  double* partial;   // reduction partials
  double* small;     // small device outputs (reductions, gathers)
  int64_t small_elems;

  bgk::Consts c;
  bgk::Geom geom;
  bool fused_ok;  // all four walls no-slip (BC type 4)
  bgk::AdjCtx* adj;  // adjoint context, null until bgk_adj_setup
};

namespace bgk {

// ---- per-subroutine kernels (bgk_stage.hip), FP64 handles
hipError_t launch_initialization(bgk_handle* h);
hipError_t launch_bc_f(bgk_handle* h);
hipError_t launch_bc_g(bgk_handle* h);
hipError_t launch_streaming(bgk_handle* h);
hipError_t launch_moments_pre(bgk_handle* h);
hipError_t launch_phase_field(bgk_handle* h);
hipError_t launch_forces(bgk_handle* h);
hipError_t launch_collision(bgk_handle* h);
hipError_t launch_moments(bgk_handle* h);

// ---- fused step (bgk_fused.hip / bgk_fused32.hip)
// Copy averaged T(Nx, :) into teast[cur] before the first fused step of a call.
hipError_t launch_teast_prologue(bgk_handle* h);
// One fused FP64 step from buffer cur to 1-cur; write_all: write the
// moments of every node (else only the boundary-adjacent ones).
hipError_t launch_fused_step64(bgk_handle* h, int step_parity, bool write_all);
hipError_t launch_fused_step32(bgk_handle* h, int step_parity, bool write_all);
hipError_t launch_fused_step_mom(bgk_handle* h, int step_parity, bool write_all);
hipError_t launch_initialization_mom(bgk_handle* h);

// ---- plane conversions (bgk_fused32.hip); strides in elements.  The base
// of a node is `base`, or with rho_table (the f members of an FP32 handle)
// base * the node's reference density (rho0 on the halo).
hipError_t launch_dev_to_f64(bgk_handle* h, const float* src, int sstride, double* dst,
                             double base, bool rho_table);
hipError_t launch_f64_to_dev(bgk_handle* h, const double* src, float* dst, int dstride,
                             double base, bool rho_table);
// FP32 handles: reset the reference-density tables to the rows' mean solid
// and liquid densities, re-expressing the stored f; and re-express the
// stored f for a new phi (copied into phi[pc]) before an upload of phi
hipError_t launch_rebase32(bgk_handle* h);
hipError_t launch_rephase32(bgk_handle* h, const double* phi_new);
hipError_t launch_copy64(bgk_handle* h, const double* src, int sstride, double* dst, int dstride);
// initialization of an FP32 handle (FP64 equilibria rounded to deviations)
hipError_t launch_initialization32(bgk_handle* h);

// ---- reductions and gathers (bgk_reduce.hip); results land in h->small
// sums over the interior: out[0..2] = sum rho, sum T, sum phi
hipError_t launch_field_sums(bgk_handle* h, double* out_dev);
// diag sums (bgk_diag_sums layout)
hipError_t launch_diag_sums(bgk_handle* h, double* out_dev);
hipError_t launch_gather(bgk_handle* h, const double* plane, int n, const int32_t* ij_dev,
                         double* out_dev);

int reduce_partial_count(const bgk_handle* h);
// field_check's sums over the interior of three caller-owned planes, with the
// partials in the library's static device scratch (allocates nothing); waits
// for the result on `stream` and writes the three sums to out_host
hipError_t field_sums_planes(const double* a, const double* b, const double* c3, const Geom& g,
                             hipStream_t stream, double out_host[3]);
int max_scratch_devices();

// ---- adjoint (bgk_adjoint.hip), FP64 moment handles
// nsnap > 0: binomial checkpointing with nsnap snapshot slots; else the
// level schedule (seg_outer > 0: two levels)
hipError_t adj_setup(bgk_handle* h, int nsteps, int seg_outer, int seg_inner, int nsnap, int np,
                     const int32_t* pi, const int32_t* pj);
// device bytes of a binomial context with nsnap snapshot slots
int64_t adj_bytes(bgk_handle* h, int nsteps, int np, int nsnap);
// forward steps the backward pass of a binomial schedule recomputes (host)
int64_t binomial_recompute(int nsteps, int nsnap);
// forward steps the last backward pass ran
int64_t adj_recomputed(const bgk_handle* h);
void adj_info(const bgk_handle* h, int64_t* bytes, int32_t* n_states);
hipError_t adj_forward(bgk_handle* h, double* probe_T_host);
hipError_t adj_backward(bgk_handle* h, const double* seeds_host, double grad[3]);
void adj_free(bgk_handle* h);
// true once adj_forward has recorded the checkpoints the backward pass
// replays; calls that change the state or the constants clear it
bool adj_forward_done(const bgk_handle* h);
void adj_invalidate(bgk_handle* h);

}  // namespace bgk

#endif  // BGK_INTERNAL_H
