// bgk_api.hip -- the C ABI of include/bgk_pcm.h.
//
// Host side of the library: validates arguments, owns the device memory of
// a handle, issues the kernels of bgk_stage.hip / bgk_fused*.hip /
// bgk_reduce.hip on the handle's stream, and reports errors as return codes
// plus a thread-local message.  There is no host fallback: every computation
// runs on the GPU.
#include <sys/resource.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bgk_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? BGK_E_NOMEM : BGK_E_HIP, "%s: %s (%d)", what,
              hipGetErrorString(e), (int)e);
}

#define HIP_TRY(expr, what)                       \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

bool valid_bc(int v, std::initializer_list<int> allowed) {
  return std::find(allowed.begin(), allowed.end(), v) != allowed.end();
}

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Plane layout: each plane pointer sits one node before a 256-byte boundary
// so that node (1, j) of every row starts a cache line.  The pitch (nodes
// per row) is a multiple of 64 and not of 256: measured on MI355X, row
// strides that are odd multiples of 32 nodes, or multiples of 256, lose
// 5..10% of the fused step's bandwidth to HBM channel conflicts between the
// eight rows the eight XCDs stream at once.
constexpr int64_t kPitchAlign = 64;
constexpr int64_t kPlaneAlign = 4096;

int64_t row_pitch(int nx) {
  int64_t p = (nx + 2 + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
  if (p % 256 == 0) p += kPitchAlign;
  return p;
}

struct Carver {
  char* base;
  size_t off;
  size_t bytes_for(int64_t elems, size_t esize) const {
    return (size_t)round_up((int64_t)(elems * esize) + 256, 256);
  }
  template <typename T>
  T* plane(int64_t elems) {
    const size_t start = (size_t)round_up((int64_t)off, kPlaneAlign) + 256 - sizeof(T);
    off = (size_t)round_up((int64_t)(start + elems * sizeof(T)), 256);
    return base ? reinterpret_cast<T*>(base + start) : nullptr;
  }
  template <typename T>
  T* block(int64_t elems) {
    const size_t start = (size_t)round_up((int64_t)off, 256);
    off = start + (size_t)(elems * sizeof(T));
    return base ? reinterpret_cast<T*>(base + start) : nullptr;
  }
};

// Assign every device array of h from `base` (null: only measure).
size_t carve(bgk_handle* h, char* base) {
  Carver cv{base, 0};
  const int64_t n = h->plane_elems;
  const int prec = h->prm.precision;
  const bool fp32 = prec == BGK_FP32_DEV;
  const bool mom = prec == BGK_FP32_MOM || prec == BGK_FP64_MOM;
  const bool m64 = prec == BGK_FP64_MOM;
  for (int b = 0; b < 2; ++b) {
    h->ma[b] = nullptr;
    h->mb[b] = nullptr;
    h->mc[b] = nullptr;
    for (int k = 0; k < 9; ++k) {
      h->f[b][k] = h->g[b][k] = nullptr;
      h->f32[b][k] = h->g32[b][k] = nullptr;
    }
    if (mom) {
      h->ma[b] = cv.plane<double2>(n);
      if (m64) {
        h->mb[b] = cv.plane<double4>(n);
        h->mc[b] = h->mom_eqf ? static_cast<void*>(cv.plane<double>(n))
                              : static_cast<void*>(cv.plane<double4>(n));
      } else {
        h->mb[b] = cv.plane<float4>(n);
        h->mc[b] = h->mom_eqf ? static_cast<void*>(cv.plane<float>(n))
                              : static_cast<void*>(cv.plane<float4>(n));
      }
      continue;
    }
    for (int k = 0; k < 9; ++k) {
      if (fp32) {
        float* pr = reinterpret_cast<float*>(cv.plane<float2>(n));
        h->f32[b][k] = pr;
        h->g32[b][k] = pr ? pr + 1 : nullptr;
      } else {
        double* pr = reinterpret_cast<double*>(cv.plane<double2>(n));
        h->f[b][k] = pr;
        h->g[b][k] = pr ? pr + 1 : nullptr;
      }
    }
  }
  h->phi[0] = cv.plane<double>(n);
  h->phi[1] = cv.plane<double>(n);
  h->rho = cv.plane<double>(n);
  h->u = cv.plane<double>(n);
  h->v = cv.plane<double>(n);
  h->T = cv.plane<double>(n);
  if (fp32 || mom) {
    h->rho2 = h->u2 = h->v2 = h->T2 = nullptr;
  } else {
    h->rho2 = cv.plane<double>(n);
    h->u2 = cv.plane<double>(n);
    h->v2 = cv.plane<double>(n);
    h->T2 = cv.plane<double>(n);
  }
  h->tmp = cv.plane<double>(n);
  h->teast[0] = cv.block<double>(h->ny + 2);
  h->teast[1] = cv.block<double>(h->ny + 2);
  h->rref[0] = fp32 ? cv.block<double>(2 * ((int64_t)h->ny + 2)) : nullptr;
  h->rref[1] = fp32 ? cv.block<double>(2 * ((int64_t)h->ny + 2)) : nullptr;
  h->partial = cv.block<double>(bgk::reduce_partial_count(h));
  h->small = cv.block<double>(h->small_elems);
  return cv.off + 256;
}

void set_constants(bgk_handle* h) {
  const bgk_params& p = h->prm;
  bgk::Consts& c = h->c;
  // STEP_7/setup.f90:79-105
  c.w[0] = 4.0 / 9.0;
  for (int k = 1; k <= 4; ++k) c.w[k] = 1.0 / 9.0;
  for (int k = 5; k <= 8; ++k) c.w[k] = 1.0 / 36.0;
  c.cs2 = 1.0 / 3.0;
  for (int k = 0; k < 9; ++k) c.coef[k] = (0.5 * c.w[k]) / (c.cs2 * c.cs2);
  c.omega_f = 1.0 / p.tau_f;
  c.omega_g = 1.0 / p.tau_g;
  c.inv_omega_f = 1.0 / c.omega_f;
  c.rho0 = p.rho0;
  c.rho1 = p.rho1;
  c.t0 = p.t0;
  c.t_in_sol = p.t_in_sol;
  // STEP_7/initialization.f90:9-12 (the hot wall and the liquid strip at
  // T0 unless the t_hot extension moves them)
  c.tin = h->ext.t_hot > 0.0 ? h->ext.t_hot : p.t0;
  c.t_liquid0 = c.tin;
  c.k_ratio = h->ext.k_ratio;
  c.alpha_l = c.cs2 * (p.tau_g - 0.5);
  c.east_cold = h->ext.east_cold ? 1 : 0;
  c.t_cold = h->ext.t_cold;
  c.noise_amp = h->ext.noise_amp * (p.t0 - p.t_in_sol);
  c.noise_seed = h->ext.noise_seed;
  c.crit_shift = h->ext.t_crit > 0.0 ? 273.0 - h->ext.t_crit : 0.0;
  c.t_crit = h->ext.t_crit > 0.0 ? h->ext.t_crit : 273.0;
  c.t_ref = h->ext.t_ref;
  const double kvisc = c.cs2 * (p.tau_f - 0.5);
  c.u0 = p.re * kvisc / p.length;
  c.v0 = 0.0;
  // STEP_7/forces.f90:12-13 (beta is an implicit default-REAL variable)
  c.delta_t = p.t0 - p.t_in_sol;
  c.dt_st = c.delta_t / p.stefan_number;
  const double beta = (1.0 - p.rho1 / p.rho0) / c.delta_t;
  c.beta = (p.rounding == BGK_ROUND_STEP7) ? (double)(float)beta : beta;
  c.grav = p.grav;
  c.t_width = p.t_width;
  c.two_over_tw = 2.0 / p.t_width;
  c.inv_stefan = 1.0 / p.stefan_number;
  c.g_ref = 273.0;  // T_crit of STEP_7/phase_field.f90:16
  // The initial equilibria of the two ghosts BC_f never writes
  // (initialization.f90:22-26,76-81: rho0 right of the interface, u = v = 0
  // on rows 0 and Ny+1), in the operand order of collision.f90:23-41.
  c.frozen6 = c.w[6] * p.rho0 * (1.0 - 0.0);
  c.frozen7 = c.w[7] * p.rho0 * (1.0 - 0.0);
}

int check_handle(bgk_handle* h) {
  if (!h) return fail(BGK_E_INVALID, "null handle");
  hipError_t e = hipSetDevice(h->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return BGK_OK;
}

int need_fp64(bgk_handle* h, const char* what) {
  if (h->prm.precision != BGK_FP64)
    return fail(BGK_E_UNSUPPORTED, "%s needs an FP64 handle (this one stores %s)", what,
                h->prm.precision == BGK_FP32_DEV ? "FP32 deviations" : "moments");
  return BGK_OK;
}

int launched(hipError_t e, const char* what) {
  if (e != hipSuccess) return hip_fail(e, what);
  return BGK_OK;
}

// Device plane of an FP64 field, or null with *fp32 set for the FP32
// populations (k in 0..8, base w_k*rho0 or w_k*T0), or null on error.
struct FieldRef {
  double* d64 = nullptr;  // FP64 field; element stride `stride`
  float* d32 = nullptr;   // FP32 deviation member of a pair plane (stride 2)
  double base = 0.0;      // deviation base (times the reference density with
  bool rho_table = false; // rho_table: the f members of an FP32 handle)
  int stride = 1;
};

int resolve(bgk_handle* h, int field, FieldRef* r) {
  const bool fp32 = h->prm.precision == BGK_FP32_DEV;
  if (field < 0 || field >= BGK_NFIELDS) return fail(BGK_E_INVALID, "field id %d", field);
  const bool mom = h->prm.precision == BGK_FP32_MOM || h->prm.precision == BGK_FP64_MOM;
  if (mom && field < BGK_RHO)
    return fail(BGK_E_UNSUPPORTED, "a moment handle stores no populations (field %d)", field);
  if (field < BGK_FP0 || (field >= BGK_G0 && field < BGK_GP0)) {
    const bool isf = field < BGK_FP0;
    const int k = isf ? field - BGK_F0 : field - BGK_G0;
    if (fp32) {
      r->d32 = isf ? h->f32[h->cur][k] : h->g32[h->cur][k];
      r->base = isf ? h->c.w[k] : h->c.w[k] * h->c.g_ref;
      r->rho_table = isf;
    } else {
      r->d64 = isf ? h->f[h->cur][k] : h->g[h->cur][k];
      r->stride = 2;
    }
    return BGK_OK;
  }
  if (field < BGK_G0 || (field >= BGK_GP0 && field < BGK_RHO)) {
    if (fp32) return fail(BGK_E_UNSUPPORTED, "fp/gp fields exist only on FP64 handles");
    const bool isf = field < BGK_G0;
    const int k = isf ? field - BGK_FP0 : field - BGK_GP0;
    r->d64 = isf ? h->f[1 - h->cur][k] : h->g[1 - h->cur][k];
    r->stride = 2;
    return BGK_OK;
  }
  switch (field) {
    case BGK_RHO: r->d64 = h->rho; break;
    case BGK_U: r->d64 = h->u; break;
    case BGK_V: r->d64 = h->v; break;
    case BGK_T: r->d64 = h->T; break;
    case BGK_PHI: r->d64 = h->phi[h->pc]; break;
    case BGK_PHI_PREC: r->d64 = h->phi[1 - h->pc]; break;
    default:
      if (h->prm.precision != BGK_FP64)
        return fail(BGK_E_UNSUPPORTED, "rho2/u2/v2/T2 exist only on FP64 handles");
      r->d64 = field == BGK_RHO2 ? h->rho2 : field == BGK_U2 ? h->u2 : field == BGK_V2 ? h->v2 : h->T2;
  }
  return BGK_OK;
}

int run_stage(bgk_handle* h, hipError_t (*fn)(bgk_handle*), const char* name) {
  int rc = check_handle(h);
  if (rc) return rc;
  if ((rc = need_fp64(h, name))) return rc;
  return launched(fn(h), name);
}

int copy_small_back(bgk_handle* h, double* out, int n) {
  HIP_TRY(hipMemcpyAsync(out, h->small, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream),
          "hipMemcpyAsync(D2H)");
  HIP_TRY(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  return BGK_OK;
}

// Probes of the adjoint: interior nodes, each at most once (each node
// carries one probe slot on the device: a repeated node would leave one of
// the two records unwritten).
int check_probes(bgk_handle* h, int32_t n_probes, const int32_t* pi, const int32_t* pj) {
  for (int32_t p = 0; p < n_probes; ++p)
    if (pi[p] < 1 || pi[p] > h->nx || pj[p] < 1 || pj[p] > h->ny)
      return fail(BGK_E_INVALID, "probe %d at (%d,%d) is not an interior node", p, pi[p], pj[p]);
  std::vector<int64_t> q((size_t)n_probes);
  for (int32_t p = 0; p < n_probes; ++p) q[(size_t)p] = (int64_t)pj[p] * (h->nx + 2) + pi[p];
  std::sort(q.begin(), q.end());
  if (std::adjacent_find(q.begin(), q.end()) != q.end())
    return fail(BGK_E_INVALID, "two probes share one node; give each node at most once");
  return BGK_OK;
}

// BGK_E_NONFINITE when one of out[0..n) is NaN or infinite: a sum over the
// interior is non-finite exactly when a summed field holds a NaN or an
// infinity (or the sum overflows), so the host test of the n results is
// the detection.
int nonfinite_check(const double* out, int n, const char* what) {
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(out[k]))
      return fail(BGK_E_NONFINITE, "%s: result %d is %g: the fields hold a NaN or an infinity",
                  what, k, out[k]);
  return BGK_OK;
}

}  // namespace

extern "C" {

int bgk_abi_version(void) { return BGK_ABI_VERSION; }

const char* bgk_last_error(void) { return g_last_error.c_str(); }

int bgk_create(const bgk_params* p, bgk_handle** out) {
  if (!p || !out) return fail(BGK_E_INVALID, "null argument");
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1) return fail(BGK_E_INVALID, "Nx=%d Ny=%d: both must be >= 1", p->nx, p->ny);
  if ((int64_t)(p->nx + 2) * (p->ny + 2) > ((int64_t)1 << 40))
    return fail(BGK_E_INVALID, "grid %d x %d too large", p->nx, p->ny);
  if (!valid_bc(p->west_bc, {0, 1, 2, 3, 4}) || !valid_bc(p->east_bc, {0, 3, 4}) ||
      !valid_bc(p->north_bc, {0, 1, 4, 5, 6}) || !valid_bc(p->south_bc, {0, 4, 6}))
    return fail(BGK_E_INVALID, "boundary types W=%d E=%d N=%d S=%d not in the reference's sets",
                p->west_bc, p->east_bc, p->north_bc, p->south_bc);
  if (p->tau_f == 0.0 || p->tau_g == 0.0 || p->stefan_number == 0.0 || p->t_width == 0.0 ||
      p->rho0 == 0.0 || p->t0 == p->t_in_sol)
    return fail(BGK_E_INVALID, "tau_f, tau_g, stefan_number, T_width, rho0 and T0-T_in_sol "
                               "must be nonzero");
  if (p->precision < BGK_FP64 || p->precision > BGK_FP64_MOM)
    return fail(BGK_E_INVALID, "precision %d", p->precision);
  if (p->rounding != BGK_ROUND_STEP7 && p->rounding != BGK_ROUND_CLEAN)
    return fail(BGK_E_INVALID, "rounding %d", p->rounding);
  const bool fused_ok = p->west_bc == 4 && p->east_bc == 4 && p->north_bc == 4 && p->south_bc == 4;
  if (p->precision != BGK_FP64 && !fused_ok)
    return fail(BGK_E_UNSUPPORTED, "FP32 storage is implemented for the no-slip cavity "
                                   "(all four boundary types 4) only");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(BGK_E_NODEVICE, "no HIP device visible (%s)", hipGetErrorString(e));
  if (p->device < 0 || p->device >= ndev)
    return fail(BGK_E_NODEVICE, "device %d out of range (%d visible)", p->device, ndev);
  HIP_TRY(hipSetDevice(p->device), "hipSetDevice");
  int n_cu = 0;
  HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device),
          "hipDeviceGetAttribute");

  bgk_handle* h = new (std::nothrow) bgk_handle();
  if (!h) return fail(BGK_E_NOMEM, "host allocation");
  h->prm = *p;
  h->nx = p->nx;
  h->ny = p->ny;
  h->device = p->device;
  h->n_cu = n_cu;
  h->pitch = row_pitch(p->nx);
  h->plane_elems = (int64_t)(p->ny + 2) * h->pitch;
  h->small_elems = std::max<int64_t>(4096, 3 * (int64_t)(p->nx + p->ny + 4));
  h->geom = bgk::Geom{p->nx, p->ny, h->pitch};
  h->fused_ok = fused_ok;
  h->adj = nullptr;
  h->cur = 0;
  h->pc = 0;
  h->rc = 0;
  h->since_rebase = 0;
  // steps between re-basings of the FP32 populations' row densities
  // (BGK_REBASE_EVERY, 0: never); a tuning knob, every value is correct
  h->rebase_every = 500;
  if (const char* env = std::getenv("BGK_REBASE_EVERY"))
    h->rebase_every = std::max(0, std::atoi(env));
  std::memset(&h->ext, 0, sizeof h->ext);
  h->ext.k_ratio = 1.0;
  // tiling of the moment kernel: BGK_MOM_ROWS (rows per strip; default
  // about 64, or 32 for tau_f != 1, fitted to the device) and
  // BGK_MOM_VARIANT (1..8: workgroup width 64/128/256/512 with a 4- or
  // 3-row ring; 0: chosen from Nx) override the defaults; tuning knobs,
  // every value is correct
  h->strip_rows = 0;
  h->mom_variant = 0;
  h->mom_eqf = (1.0 / p->tau_f) == 1.0;
  if (const char* env = std::getenv("BGK_MOM_ROWS")) h->strip_rows = std::max(0, std::atoi(env));
  if (const char* env = std::getenv("BGK_MOM_VARIANT")) h->mom_variant = std::max(0, std::atoi(env));
  set_constants(h);

  h->pool_bytes = carve(h, nullptr);
  e = hipMalloc(&h->pool, h->pool_bytes);
  if (e != hipSuccess) {
    const size_t want = h->pool_bytes;
    delete h;
    return fail(e == hipErrorOutOfMemory ? BGK_E_NOMEM : BGK_E_HIP,
                "hipMalloc(%zu bytes) for a %d x %d lattice: %s", want, p->nx, p->ny,
                hipGetErrorString(e));
  }
  carve(h, static_cast<char*>(h->pool));
  e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipFree(h->pool);
    delete h;
    return hip_fail(e, "hipStreamCreate");
  }
  h->stream = h->own_stream;
  e = hipMemsetAsync(h->pool, 0, h->pool_bytes, h->stream);
  if (e == hipSuccess && h->rref[0]) {
    // the reference densities start at rho0 (what an upload encodes against
    // before any initialization)
    const std::vector<double> r0(2 * ((size_t)h->ny + 2), p->rho0);
    for (int b = 0; b < 2 && e == hipSuccess; ++b)
      e = hipMemcpyAsync(h->rref[b], r0.data(), r0.size() * sizeof(double),
                         hipMemcpyHostToDevice, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(h->own_stream);
    (void)hipFree(h->pool);
    delete h;
    return hip_fail(e, "hipMemset");
  }
  *out = h;
  return BGK_OK;
}

int bgk_destroy(bgk_handle* h) {
  if (!h) return BGK_OK;
  (void)hipSetDevice(h->device);
  hipError_t e1 = hipStreamSynchronize(h->stream);
  bgk::adj_free(h);
  hipError_t e2 = hipStreamDestroy(h->own_stream);
  hipError_t e3 = hipFree(h->pool);
  delete h;
  if (e1 != hipSuccess) return hip_fail(e1, "hipStreamSynchronize");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamDestroy");
  if (e3 != hipSuccess) return hip_fail(e3, "hipFree");
  return BGK_OK;
}

int bgk_set_stream(bgk_handle* h, void* s) {
  int rc = check_handle(h);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  h->stream = s ? static_cast<hipStream_t>(s) : h->own_stream;
  return BGK_OK;
}

int bgk_initialization(bgk_handle* h) {
  int rc = check_handle(h);
  if (rc) return rc;
  bgk::adj_invalidate(h);
  if (h->prm.precision == BGK_FP32_DEV) return launched(bgk::launch_initialization32(h), "initialization");
  if (h->prm.precision == BGK_FP32_MOM || h->prm.precision == BGK_FP64_MOM)
    return launched(bgk::launch_initialization_mom(h), "initialization");
  return launched(bgk::launch_initialization(h), "initialization");
}

int bgk_set_extensions(bgk_handle* h, const bgk_ext* e) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!e) return fail(BGK_E_INVALID, "null extensions");
  if (!(e->t_hot >= 0.0) || !(e->k_ratio > 0.0) || !(e->noise_amp >= 0.0) ||
      (e->east_cold != 0 && e->east_cold != 1) || (e->east_cold && !(e->t_cold > 0.0)) ||
      !(e->t_crit >= 0.0) || !(e->t_ref >= 0.0))
    return fail(BGK_E_INVALID,
                "t_hot=%g k_ratio=%g east_cold=%d t_cold=%g noise_amp=%g t_crit=%g t_ref=%g: "
                "need t_hot >= 0, k_ratio > 0, east_cold 0/1 with t_cold > 0, noise_amp >= 0, "
                "t_crit >= 0, t_ref >= 0",
                e->t_hot, e->k_ratio, e->east_cold, e->t_cold, e->noise_amp, e->t_crit, e->t_ref);
  h->ext = *e;
  set_constants(h);
  bgk::adj_invalidate(h);
  return BGK_OK;
}

int bgk_upload(bgk_handle* h, int32_t field, const double* src, int64_t ld) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!src) return fail(BGK_E_INVALID, "null source");
  if (ld < h->nx + 2) return fail(BGK_E_INVALID, "ld=%lld < Nx+2=%d", (long long)ld, h->nx + 2);
  if (h->prm.precision == BGK_FP32_MOM || h->prm.precision == BGK_FP64_MOM)
    return fail(BGK_E_UNSUPPORTED, "a moment handle's state is set by bgk_initialization only");
  FieldRef r;
  if ((rc = resolve(h, field, &r))) return rc;
  bgk::adj_invalidate(h);
  // the stored f of an FP32 handle depend on phi through their reference
  // density: a new phi is staged and the f re-expressed before it lands
  const bool phi_via_tmp = h->prm.precision == BGK_FP32_DEV && field == BGK_PHI;
  const bool direct = r.d64 && r.stride == 1 && !phi_via_tmp;
  double* dst = direct ? r.d64 : h->tmp;
  HIP_TRY(hipMemcpy2DAsync(dst, (size_t)h->pitch * sizeof(double), src, (size_t)ld * sizeof(double),
                           (size_t)(h->nx + 2) * sizeof(double), (size_t)(h->ny + 2),
                           hipMemcpyHostToDevice, h->stream),
          "hipMemcpy2DAsync(H2D)");
  if (r.d32) HIP_TRY(bgk::launch_f64_to_dev(h, h->tmp, r.d32, 2, r.base, r.rho_table), "f64_to_dev");
  if (!direct && r.d64 && !phi_via_tmp)
    HIP_TRY(bgk::launch_copy64(h, h->tmp, 1, r.d64, r.stride), "copy64");
  if (phi_via_tmp)
    HIP_TRY(bgk::launch_rephase32(h, h->tmp), "re-express f for the new phi");
  HIP_TRY(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  return BGK_OK;
}

int bgk_download(bgk_handle* h, int32_t field, double* dst, int64_t ld) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!dst) return fail(BGK_E_INVALID, "null destination");
  if (ld < h->nx + 2) return fail(BGK_E_INVALID, "ld=%lld < Nx+2=%d", (long long)ld, h->nx + 2);
  FieldRef r;
  if ((rc = resolve(h, field, &r))) return rc;
  const double* src = r.d64;
  if (r.d32) {
    HIP_TRY(bgk::launch_dev_to_f64(h, r.d32, 2, h->tmp, r.base, r.rho_table), "dev_to_f64");
    src = h->tmp;
  } else if (r.stride != 1) {
    HIP_TRY(bgk::launch_copy64(h, r.d64, r.stride, h->tmp, 1), "copy64");
    src = h->tmp;
  }
  HIP_TRY(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), src, (size_t)h->pitch * sizeof(double),
                           (size_t)(h->nx + 2) * sizeof(double), (size_t)(h->ny + 2),
                           hipMemcpyDeviceToHost, h->stream),
          "hipMemcpy2DAsync(D2H)");
  HIP_TRY(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  return BGK_OK;
}

int bgk_bc_f(bgk_handle* h) { return run_stage(h, bgk::launch_bc_f, "BC_f"); }
int bgk_bc_g(bgk_handle* h) { return run_stage(h, bgk::launch_bc_g, "BC_g"); }
int bgk_streaming(bgk_handle* h) { return run_stage(h, bgk::launch_streaming, "streaming"); }
int bgk_moments_pre(bgk_handle* h) { return run_stage(h, bgk::launch_moments_pre, "momentsPre"); }
int bgk_phase_field(bgk_handle* h) { return run_stage(h, bgk::launch_phase_field, "phase_field"); }
int bgk_forces(bgk_handle* h) { return run_stage(h, bgk::launch_forces, "forces"); }
int bgk_collision(bgk_handle* h) { return run_stage(h, bgk::launch_collision, "collision"); }
int bgk_moments(bgk_handle* h) { return run_stage(h, bgk::launch_moments, "moments"); }

int bgk_step(bgk_handle* h, int32_t nsteps) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (nsteps < 0) return fail(BGK_E_INVALID, "nsteps=%d", nsteps);
  if (nsteps == 0) return BGK_OK;
  bgk::adj_invalidate(h);
  if (h->fused_ok) {
    if ((rc = launched(bgk::launch_teast_prologue(h), "teast prologue"))) return rc;
    const int prec = h->prm.precision;
    for (int s = 0; s < nsteps; ++s) {
      const bool last = (s == nsteps - 1);
      hipError_t e = prec == BGK_FP32_DEV   ? bgk::launch_fused_step32(h, s, last)
                     : prec >= BGK_FP32_MOM ? bgk::launch_fused_step_mom(h, s, last)
                                            : bgk::launch_fused_step64(h, s, last);
      if (e != hipSuccess) return hip_fail(e, "fused step");
    }
    return BGK_OK;
  }
  if ((rc = need_fp64(h, "bgk_step with these boundary types"))) return rc;
  for (int s = 0; s < nsteps; ++s) {
    hipError_t e;
    if ((e = bgk::launch_bc_f(h)) != hipSuccess) return hip_fail(e, "BC_f");
    if ((e = bgk::launch_bc_g(h)) != hipSuccess) return hip_fail(e, "BC_g");
    if ((e = bgk::launch_streaming(h)) != hipSuccess) return hip_fail(e, "streaming");
    if ((e = bgk::launch_moments_pre(h)) != hipSuccess) return hip_fail(e, "momentsPre");
    if ((e = bgk::launch_phase_field(h)) != hipSuccess) return hip_fail(e, "phase_field");
    if ((e = bgk::launch_forces(h)) != hipSuccess) return hip_fail(e, "forces");
    if ((e = bgk::launch_collision(h)) != hipSuccess) return hip_fail(e, "collision");
    if ((e = bgk::launch_moments(h)) != hipSuccess) return hip_fail(e, "moments");
  }
  return BGK_OK;
}

// Nx*Ny as the reference computes it: default-INTEGER (32-bit) product.
static double nxny_as_reference(const bgk_handle* h) {
  return (double)(int32_t)((uint32_t)h->nx * (uint32_t)h->ny);
}

int bgk_field_check(bgk_handle* h, double out[3]) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!out) return fail(BGK_E_INVALID, "null output");
  if ((rc = launched(bgk::launch_field_sums(h, h->small), "field sums"))) return rc;
  double s[3];
  if ((rc = copy_small_back(h, s, 3))) return rc;
  const double nn = nxny_as_reference(h);
  out[0] = s[0] / nn;
  out[1] = s[1] / nn;
  out[2] = s[2] / nn;
  return nonfinite_check(out, 3, "field_check");
}

int bgk_field_check_planes(const double* rho, const double* T, const double* phi, int32_t nx,
                           int32_t ny, int64_t pitch, void* hip_stream, double out[3]) {
  if (!rho || !T || !phi || !out) return fail(BGK_E_INVALID, "null argument");
  if (nx < 1 || ny < 1 || pitch < (int64_t)nx + 2 ||
      (int64_t)(ny + 2) * pitch > ((int64_t)1 << 40))
    return fail(BGK_E_INVALID, "Nx=%d Ny=%d pitch=%lld: need Nx, Ny >= 1 and pitch >= Nx+2", nx,
                ny, (long long)pitch);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev >= bgk::max_scratch_devices())
    return fail(BGK_E_UNSUPPORTED, "device %d: the scratch table holds devices 0..%d", dev,
                bgk::max_scratch_devices() - 1);
  const bgk::Geom g{nx, ny, pitch};
  double s[3];
  HIP_TRY(bgk::field_sums_planes(rho, T, phi, g, static_cast<hipStream_t>(hip_stream), s),
          "field sums");
  const double nn = (double)(int32_t)((uint32_t)nx * (uint32_t)ny);
  out[0] = s[0] / nn;
  out[1] = s[1] / nn;
  out[2] = s[2] / nn;
  return nonfinite_check(out, 3, "field_check");
}

int bgk_mass_check(bgk_handle* h, double out[3]) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!out) return fail(BGK_E_INVALID, "null output");
  if ((rc = launched(bgk::launch_field_sums(h, h->small), "field sums"))) return rc;
  double s[3];
  if ((rc = copy_small_back(h, s, 3))) return rc;
  const double nn = nxny_as_reference(h);
  out[0] = (nn * h->prm.rho0 - s[0]) * 100.0 / (nn * h->prm.rho0);
  out[1] = s[0] / nn;
  out[2] = s[1] / nn;
  return nonfinite_check(out, 3, "mass_check");
}

int bgk_diag_sums(bgk_handle* h, double out[4]) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!out) return fail(BGK_E_INVALID, "null output");
  if (h->nx < 2)
    return fail(BGK_E_UNSUPPORTED, "diag reads T(1..3,j): needs Nx >= 2 (Nx=%d)", h->nx);
  if ((rc = launched(bgk::launch_diag_sums(h, h->small), "diag sums"))) return rc;
  if ((rc = copy_small_back(h, out, 4))) return rc;
  return nonfinite_check(out, 2, "diag sums");
}

int bgk_gather(bgk_handle* h, int32_t field, int32_t n, const int32_t* i, const int32_t* j,
               double* out) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!i || !j || !out))) return fail(BGK_E_INVALID, "bad gather arguments");
  FieldRef r;
  if ((rc = resolve(h, field, &r))) return rc;
  for (int32_t k = 0; k < n; ++k)
    if (i[k] < 0 || i[k] > h->nx + 1 || j[k] < 0 || j[k] > h->ny + 1)
      return fail(BGK_E_INVALID, "node (%d,%d) outside 0..%d x 0..%d", i[k], j[k], h->nx + 1,
                  h->ny + 1);
  const double* plane = r.d64;
  if (r.d32) {
    HIP_TRY(bgk::launch_dev_to_f64(h, r.d32, 2, h->tmp, r.base, r.rho_table), "dev_to_f64");
    plane = h->tmp;
  } else if (r.stride != 1) {
    HIP_TRY(bgk::launch_copy64(h, r.d64, r.stride, h->tmp, 1), "copy64");
    plane = h->tmp;
  }
  // small buffer: first half outputs, second half (i,j) pairs
  const int64_t chunk = h->small_elems / 2;
  int32_t* ij_dev = reinterpret_cast<int32_t*>(h->small + chunk);
  int32_t* pairs = new (std::nothrow) int32_t[2 * (size_t)std::min<int64_t>(chunk, n > 0 ? n : 1)];
  if (!pairs) return fail(BGK_E_NOMEM, "host allocation");
  for (int64_t base = 0; base < n; base += chunk) {
    const int m = (int)std::min<int64_t>(chunk, n - base);
    for (int k = 0; k < m; ++k) {
      pairs[2 * k] = i[base + k];
      pairs[2 * k + 1] = j[base + k];
    }
    hipError_t e = hipMemcpyAsync(ij_dev, pairs, (size_t)m * 2 * sizeof(int32_t),
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = bgk::launch_gather(h, plane, m, ij_dev, h->small);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + base, h->small, (size_t)m * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      delete[] pairs;
      return hip_fail(e, "gather");
    }
  }
  delete[] pairs;
  return BGK_OK;
}

int bgk_adj_setup_levels(bgk_handle* h, int32_t nsteps, int32_t seg_outer, int32_t seg_inner,
                         int32_t n_probes, const int32_t* pi, const int32_t* pj) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (h->prm.precision != BGK_FP64_MOM && h->prm.precision != BGK_FP32_MOM)
    return fail(BGK_E_UNSUPPORTED, "the adjoint differentiates the moment step: create the "
                                   "handle with precision BGK_FP64_MOM or BGK_FP32_MOM");
  if (nsteps < 1 || seg_inner < 1 || seg_outer < 0 || (seg_outer > 0 && seg_outer < seg_inner) ||
      n_probes < 0 || (n_probes > 0 && (!pi || !pj)))
    return fail(BGK_E_INVALID, "nsteps=%d seg_outer=%d seg_inner=%d n_probes=%d", nsteps,
                seg_outer, seg_inner, n_probes);
  if ((rc = check_probes(h, n_probes, pi, pj))) return rc;
  hipError_t e = bgk::adj_setup(h, nsteps, seg_outer, seg_inner, 0, n_probes, pi, pj);
  if (e == hipErrorOutOfMemory)
    return fail(BGK_E_NOMEM, "adjoint checkpoints of %d steps (segments %d, %d) do not fit",
                nsteps, seg_outer, seg_inner);
  if (e != hipSuccess) return hip_fail(e, "adjoint setup");
  return BGK_OK;
}

int bgk_adj_setup_binomial(bgk_handle* h, int32_t nsteps, int32_t n_states, int32_t n_probes,
                           const int32_t* pi, const int32_t* pj) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (h->prm.precision != BGK_FP64_MOM && h->prm.precision != BGK_FP32_MOM)
    return fail(BGK_E_UNSUPPORTED, "the adjoint differentiates the moment step: create the "
                                   "handle with precision BGK_FP64_MOM or BGK_FP32_MOM");
  if (nsteps < 1 || n_states < 0 || n_probes < 0 || (n_probes > 0 && (!pi || !pj)))
    return fail(BGK_E_INVALID, "nsteps=%d n_states=%d n_probes=%d", nsteps, n_states, n_probes);
  if ((rc = check_probes(h, n_probes, pi, pj))) return rc;
  int nsnap = n_states;
  if (nsnap == 0) {
    // as many snapshots as the free device memory holds, 2 GiB kept back
    bgk::adj_free(h);
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    const int64_t margin = (int64_t)2 << 30;
    const int64_t one = bgk::adj_bytes(h, nsteps, n_probes, 1);
    const int64_t per = bgk::adj_bytes(h, nsteps, n_probes, 2) - one;
    const int64_t avail = (int64_t)fr - margin - one;
    if (avail < 0)
      return fail(BGK_E_NOMEM, "the adjoint of a %d x %d lattice needs %lld bytes; %zu free",
                  h->nx, h->ny, (long long)(one + margin), fr);
    nsnap = (int)std::min<int64_t>(1 + avail / per, nsteps);
  }
  nsnap = std::min(nsnap, nsteps);
  hipError_t e = bgk::adj_setup(h, nsteps, 0, 0, nsnap, n_probes, pi, pj);
  if (e == hipErrorOutOfMemory)
    return fail(BGK_E_NOMEM, "binomial checkpoints of %d steps with %d states do not fit",
                nsteps, nsnap);
  if (e != hipSuccess) return hip_fail(e, "adjoint setup");
  return BGK_OK;
}

int bgk_adj_plan(int32_t nsteps, int32_t n_states, int64_t* recomputed_steps) {
  if (nsteps < 1 || n_states < 1 || !recomputed_steps)
    return fail(BGK_E_INVALID, "nsteps=%d n_states=%d (both >= 1), output non-null", nsteps,
                n_states);
  *recomputed_steps = bgk::binomial_recompute(nsteps, std::min(n_states, nsteps));
  return BGK_OK;
}

int bgk_adj_setup(bgk_handle* h, int32_t nsteps, int32_t seg_len, int32_t n_probes,
                  const int32_t* pi, const int32_t* pj) {
  return bgk_adj_setup_levels(h, nsteps, 0, seg_len, n_probes, pi, pj);
}

int bgk_adj_info(bgk_handle* h, int64_t* device_bytes, int32_t* n_states) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!h->adj) return fail(BGK_E_INVALID, "bgk_adj_setup first");
  if (!device_bytes || !n_states) return fail(BGK_E_INVALID, "null output");
  bgk::adj_info(h, device_bytes, n_states);
  return BGK_OK;
}

int bgk_adj_forward(bgk_handle* h, double* probe_T) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!h->adj) return fail(BGK_E_INVALID, "bgk_adj_setup first");
  if (!probe_T) return fail(BGK_E_INVALID, "null probe record");
  hipError_t e = bgk::adj_forward(h, probe_T);
  if (e != hipSuccess) return hip_fail(e, "adjoint forward");
  return BGK_OK;
}

int bgk_adj_backward(bgk_handle* h, const double* seeds, double grad[3]) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!h->adj) return fail(BGK_E_INVALID, "bgk_adj_setup first");
  if (!seeds || !grad) return fail(BGK_E_INVALID, "null seeds or gradient");
  if (!bgk::adj_forward_done(h))
    return fail(BGK_E_INVALID, "bgk_adj_forward first: no recorded forward run matches the "
                               "current state and constants");
  hipError_t e = bgk::adj_backward(h, seeds, grad);
  if (e != hipSuccess) return hip_fail(e, "adjoint backward");
  return BGK_OK;
}

int bgk_synchronize(bgk_handle* h) {
  int rc = check_handle(h);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  return BGK_OK;
}

int bgk_get_info(bgk_handle* h, bgk_info* info) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!info) return fail(BGK_E_INVALID, "null info");
  size_t fr = 0, tot = 0;
  HIP_TRY(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  struct rusage ru;
  getrusage(RUSAGE_SELF, &ru);
  info->device_bytes = (int64_t)h->pool_bytes;
  info->device_free = (int64_t)fr;
  info->device_total = (int64_t)tot;
  info->host_maxrss_mb = (double)ru.ru_maxrss / 1024.0;
  info->pitch = (int32_t)h->pitch;
  info->fused_path = h->fused_ok ? 1 : 0;
  return BGK_OK;
}

}  // extern "C"
