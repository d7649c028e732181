// Host side of libwavecore_hip: the C ABI of include/wavecore_hip.h over the gfx950 kernels.
//
// A wc_solver is the reference's Solver (crates/wavecore/src/solver.rs:260-280) plus the one
// State it steps, resident on one GPU. Everything a step needs lives in device memory; the host
// only evaluates what depends on nothing but the time and dt (the wave-maker's ramp and phase,
// the sponges' exp(-rate dt)), with the same libm calls the reference makes, and ships those few
// numbers each step.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/wavecore_hip.h"
#include "../../include/wavecore_unet.h"
#include "wc_kernels.h"
#include "wc_run.h"

namespace wc {
int unet_body(wcu_net* n, hipStream_t st, const int* skip);
void unet_views(wcu_net* n, void** x0, const void** y0, int* w0, int* h0, const float** wo);
}  // namespace wc

using namespace wc;

namespace {

constexpr double TAU = 6.283185307179586476925286766559;
constexpr double PI = 3.14159265358979323846264338327950;
constexpr int PRM_SLOTS = 3;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_OK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? WC_ENOMEM : WC_EDEVICE,                     \
                  std::string(#expr) + " failed: " + hipGetErrorString(e_));              \
  } while (0)

enum Family {
  F_FILL,
  F_SWITCH,
  F_FLUX,
  F_EULER,
  F_SGN_SETUP,
  F_CG_INIT,
  F_CG_MATVEC,
  F_CG_UPDATE,
  F_CG_PUPDATE,
  F_CG_STATS,
  F_SGN_FINISH,
  F_REDUCE,
  F_CG_REDUCE,
  F_RELAX,
  F_CG_ITER,
  F_NET,
  F_FRAME,
  F_CG_FLUSH,
  F_CG_PAIR,
  F_COUNT
};
const char* kFamilyNames[F_COUNT] = {"fill_state", "switch",    "flux",      "finish_euler",
                                     "sgn_setup",  "cg_init",   "cg_matvec", "cg_update",
                                     "cg_pupdate", "cg_stats",  "sgn_finish", "reduce",
                                     "cg_reduce",  "relax",     "cg_iter",   "net",
                                     "frame",      "cg_flush",  "cg_pair"};

struct HostSponge {
  uint64_t width;
  double strength, level;
  bool west, east, south, north;
  long long offset, len;
};

// linear_wave (forcing.rs:53-75).
bool linear_wave(double omega, double h, bool dispersive, wc_linear_wave_t* out) {
  if (!dispersive) {
    const double c = sqrt(WC_G * h);
    out->k = omega / c;
    out->phase_speed = c;
    out->group_speed = c;
    return true;
  }
  const double nu = omega * omega * h / WC_G;
  if (!(nu < 3.0)) return false;
  const double kh = sqrt(nu / (1.0 - nu / 3.0));
  const double k = kh / h;
  const double phase = omega / k;
  out->k = k;
  out->phase_speed = phase;
  out->group_speed = phase / (1.0 + kh * kh / 3.0);
  return true;
}

}  // namespace

struct wc_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  Geo geo{};
  size_t cells = 0;
  std::vector<double> bed;

  double cfl = 0.4;
  int order = WC_ORDER_SECOND;
  double manning = 0.0;
  bool periodic = false;
  int precision = WC_PRECISION_FP64;

  // wave-maker and its Drive (forcing.rs:77-139)
  bool has_maker = false;
  wc_wavemaker_t maker{};
  bool has_drive = false;
  double omega = 0.0, ky = 0.0, strength = 0.0;
  int src_i0 = 0, src_i1 = 0;

  std::vector<HostSponge> sponges;
  std::vector<double> sponge_rates;  // concatenated tables, static

  bool dispersive = false;
  wc_dispersion_t disp{};
  // breaking fronts and their eddy viscosity (Breaking::viscosity, eddy.rs)
  bool eddy_on = false;
  EddyP eddy{};
  double *e_rise = nullptr, *e_mem = nullptr, *e_nu = nullptr, *e_hu = nullptr, *e_hv = nullptr;
  bool has_held = false;
  double time = 0.0;
  unsigned long long prev_it[2] = {0, 0};

  // device arrays (padded, FP64)
  double *h = nullptr, *hu = nullptr, *hv = nullptr, *br = nullptr;
  double *h1 = nullptr, *hu1 = nullptr, *hv1 = nullptr;
  double *b = nullptr, *phi = nullptr, *lh = nullptr, *lhu = nullptr, *lhv = nullptr;
  double *c = nullptr, *dgx = nullptr, *dgy = nullptr, *bx = nullptr, *by = nullptr;
  double *wx = nullptr, *wy = nullptr, *rx = nullptr, *ry = nullptr, *px = nullptr,
         *py = nullptr, *ax = nullptr, *ay = nullptr;
  double *zx = nullptr, *zy = nullptr, *qx = nullptr, *qy = nullptr;  // fused PCG: z, 2nd p
  // the default FP64 PCG (wc_pcg2.hip): the diagonal, r and the two p buffers as interleaved
  // (x, y) pairs, one 16-byte load per halo point; allocated at its first use
  double2 *v_dg = nullptr, *v_r = nullptr, *v_pa = nullptr, *v_pb = nullptr;

  // the FP32 chosen form
  float *f_sh = nullptr, *f_shu = nullptr, *f_shv = nullptr, *f_slh = nullptr, *f_slhu = nullptr,
        *f_slhv = nullptr;
  float *f_s1h = nullptr, *f_s1hl = nullptr, *f_s1hu = nullptr, *f_s1hv = nullptr,
        *f_s1hul = nullptr, *f_s1hvl = nullptr;
  double *f_l1h = nullptr, *f_lh = nullptr;  // mass rates stay FP64
  float *f_l1hu = nullptr, *f_l1hv = nullptr;
  double *f_lhu = nullptr, *f_lhv = nullptr, *f_bx = nullptr, *f_by = nullptr;  // FP64
  float *f_phi = nullptr;
  float *f_c = nullptr, *f_wx = nullptr, *f_wy = nullptr,
        *f_rx = nullptr, *f_ry = nullptr, *f_px = nullptr, *f_py = nullptr, *f_qx = nullptr,
        *f_qy = nullptr, *f_zx = nullptr, *f_zy = nullptr, *f_ax = nullptr, *f_ay = nullptr;
  // the one-kernel PCG's own vectors (wc_pcg1.hip), interleaved (x, y)
  float2 *g_r[2] = {nullptr, nullptr}, *g_q[2] = {nullptr, nullptr},
         *g_s[2] = {nullptr, nullptr}, *g_p = nullptr, *g_di = nullptr;
  // the extrapolated warm start of the FP32 form's solves (WarmStart in wc_kernels.h)
  float2* ws_ring = nullptr;
  double2* ws_ring64 = nullptr;  // the FP64 path's, allocated when "fp64_warm_start" is set
  int ws_order64 = 0;            // the FP64 path's warm start: 0, the reference's
  int ws_order = 3;   // exact for steps of degree < 3; 0: the reference's warm start
  int ws_window = 6;  // past steps weighed (at least ws_order)
  int ws_guard = 0;   // per mille: cells whose switch reads below start from x_{n-1}

  // the learned preconditioner (wc_fpcg.hip) and the capture of solves for its training
  int precond = WC_PRECOND_JACOBI;
  wcu_net* net = nullptr;
  float *bgx = nullptr, *bgy = nullptr;
  int capture = 0;
  static constexpr int CAP_SLOTS = 4;  // the solve's state before iterations 0, 2, 5, 10
  float *cap_wx[CAP_SLOTS] = {}, *cap_wy[CAP_SLOTS] = {}, *cap_h = nullptr;
  float2* cap_r[CAP_SLOTS] = {};
  unsigned long long cap_last_it = 0;
  int precond_stall = 2;
  int precond_ratio = 70;  // percent
  double* history = nullptr;  // r.r per iteration of the last network-preconditioned solve

  // implementation options (wc_solver_set_option)
  int pcg_kernels = 0;  // 0: the default of each precision (FP64 3, FP32 4)
  double* lanes = nullptr;  // 8 x 512 lane sums of the lane PCGs
  int pcg_lanes = 256;      // blocks of the one-kernel PCG's launches: 256 or 512
  int pcg_march = 0;        // one-kernel PCG: blocks march down column strips (1) or take
                            // 64x16 tiles (0)
  int recon64 = 1;          // FP32 form: reconstruct the FP32 state in FP64 (1) or FP32 (0)
  int pcg_walk = 1;         // one-kernel PCG on 64x16 tiles: tiles per block to aim for
  int flux_kernel = 3;
  int march_rows = 0;  // band height of the marching flux launches (0: not yet chosen)
  int flux_rows = 0;   // option: a fixed band height, or 0 for the tuner's choice
  std::vector<std::pair<int, double>> march_tried;  // (band height, us per launch) as timed
  unsigned long long pcg_batch = 4;
  uint8_t* held = nullptr;
  double* part = nullptr;
  CgScalars* sc = nullptr;
  double* src_col = nullptr;
  double* sponge_rate_d = nullptr;
  double* prm = nullptr;
  size_t prm_len = 0;
  unsigned long long* red64 = nullptr;  // [0] max rate bits, [1] max momentum, [2] min-h key
  unsigned int* red32 = nullptr;        // [1] first NaN, [2] first max
  double* redf = nullptr;               // [0] volume sum
  std::vector<void*> owned;
  // statistics and relaxation zones built on this solver: freeing the solver releases their
  // device memory and leaves them detached, whichever order a caller frees things in
  std::vector<wc_stats*> stats;
  std::vector<wc_relaxation*> relaxations;

  // pinned host staging
  double* prm_host[PRM_SLOTS] = {nullptr, nullptr, nullptr};
  hipEvent_t prm_ev[PRM_SLOTS] = {nullptr, nullptr, nullptr};
  bool prm_used[PRM_SLOTS] = {false, false, false};
  int prm_slot = 0;
  size_t prm_cap = 0;
  CgScalars* sc_host = nullptr;
  unsigned long long* red64_host = nullptr;
  unsigned int* red32_host = nullptr;
  double* redf_host = nullptr;

  // profiling: prof times every launch with its own pair of events; prof_batch times only the
  // PCG iterations (the FP32 form's one kernel, the FP64 path's pair), one pair of events around
  // each batch of back-to-back launches
  bool prof = false, prof_batch = false;
  double prof_ms[F_COUNT] = {};
  unsigned long long prof_n[F_COUNT] = {};
  struct PendingTime {
    int fam;
    hipEvent_t e0, e1;
    unsigned long long launches;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<PendingTime> ev_pending;
};

namespace {

int set_device(wc_solver* s) {
  HIP_OK(hipSetDevice(s->device));
  return WC_OK;
}

template <class T>
int dalloc(wc_solver* s, T** p, size_t n) {
  void* v = nullptr;
  HIP_OK(hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T)));
  HIP_OK(hipMemsetAsync(v, 0, std::max<size_t>(n, 1) * sizeof(T), s->stream));
  s->owned.push_back(v);
  *p = static_cast<T*>(v);
  return WC_OK;
}

hipEvent_t take_event(wc_solver* s) {
  if (!s->ev_pool.empty()) {
    hipEvent_t e = s->ev_pool.back();
    s->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

template <typename... KArgs, typename... Args>
int launch(wc_solver* s, int fam, void (*k)(KArgs...), dim3 grid, dim3 block, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->prof) {
    e0 = take_event(s);
    e1 = take_event(s);
    if (e0) HIP_OK(hipEventRecord(e0, s->stream));
  }
  hipLaunchKernelGGL(k, grid, block, 0, s->stream, args...);
  HIP_OK(hipGetLastError());
  if (s->prof && e0 && e1) {
    HIP_OK(hipEventRecord(e1, s->stream));
    s->ev_pending.push_back({fam, e0, e1, 1});
  }
  return WC_OK;
}

#define TRY(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != WC_OK) return rc_; \
  } while (0)

int prof_collect(wc_solver* s) {
  if (s->ev_pending.empty()) return WC_OK;
  HIP_OK(hipStreamSynchronize(s->stream));
  for (auto& p : s->ev_pending) {
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, p.e0, p.e1));
    s->prof_ms[p.fam] += ms;
    s->prof_n[p.fam] += p.launches;
    s->ev_pool.push_back(p.e0);
    s->ev_pool.push_back(p.e1);
  }
  s->ev_pending.clear();
  return WC_OK;
}

dim3 tiles(const wc_solver* s) { return dim3((unsigned)(s->geo.ntx * s->geo.nty)); }

// The marching flux launch (wc_flux.hip): one wave per 58-column strip and band of rows, a
// tile of 58 x rows cells per wave. Every band height computes the same bits; which is fastest
// depends on the grid and the GPU (taller bands recompute fewer halo rows, shorter ones give
// more waves to hide latency with). The tuner picks it at a solver's first flux launch: it
// times the launch itself (which only writes rates from the state, so extra launches change
// nothing) at the height where all waves of a launch fit on the GPU at once and at a few others
// around it, and keeps the fastest. Only a launch that may run again unchanged is timed (one
// that writes rates or stage 1 from other arrays; stage 2 averages into the state in place):
// otherwise, and under stream capture, the fitting height serves untimed, and a launch that
// may not be repeated leaves the choice to the next one that may. Option flux_rows fixes the
// height instead.
constexpr int MARCH_MW = TX - 6;
template <typename... KArgs, class Run>
int march_launch(wc_solver* s, void (*kernel)(KArgs...), bool repeatable, dim3* grid, int* rows,
                 Run run) {
  const int strips = (s->geo.nx + MARCH_MW - 1) / MARCH_MW;
  auto grid_of = [&](int r) { return dim3((unsigned)(strips * ((s->geo.ny + r - 1) / r))); };
  if (s->flux_rows > 0) s->march_rows = std::min(s->flux_rows, s->geo.ny);
  if (!s->march_rows) {
    int per_cu = 0, cus = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TX, 0));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    const int slots = std::max(per_cu, 1) * std::max(cus, 1);
    const int bands = std::max(1, slots / strips);
    const int fit = std::min(s->geo.ny, std::max(8, (s->geo.ny + bands - 1) / bands));
    if (!repeatable) {
      *grid = grid_of(fit);
      *rows = fit;
      return WC_OK;
    }
    s->march_rows = fit;
    s->march_tried.clear();
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_OK(hipStreamIsCapturing(s->stream, &cap));
    if (cap == hipStreamCaptureStatusNone) {
      std::vector<int> cand;
      for (int r : {fit, fit / 2, (fit * 3) / 4, (fit * 3) / 2, 2 * fit, 16, 32, 64, 128}) {
        r = std::min(s->geo.ny, std::max(8, r));
        if (std::find(cand.begin(), cand.end(), r) == cand.end()) cand.push_back(r);
      }
      hipEvent_t e0, e1;
      HIP_OK(hipEventCreate(&e0));
      HIP_OK(hipEventCreate(&e1));
      double best = 1e30;
      for (int r : cand) {
        run(grid_of(r), r);  // warm
        HIP_OK(hipEventRecord(e0, s->stream));
        for (int q = 0; q < 3; ++q) run(grid_of(r), r);
        HIP_OK(hipEventRecord(e1, s->stream));
        HIP_OK(hipEventSynchronize(e1));
        HIP_OK(hipGetLastError());
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / 3.0;
        s->march_tried.push_back({r, us});
        if (us < best) best = us, s->march_rows = r;
      }
      HIP_OK(hipEventDestroy(e0));
      HIP_OK(hipEventDestroy(e1));
    }
  }
  *grid = grid_of(s->march_rows);
  *rows = s->march_rows;
  return WC_OK;
}
dim3 over(long long n) { return dim3((unsigned)((n + NT - 1) / NT)); }
dim3 rate_blocks(long long n) { return dim3((unsigned)((n + RATE_CELLS - 1) / RATE_CELLS)); }

Breaking breaking_of(const wc_solver* s) {
  // Solver::breaking_config (solver.rs:619-623): the dispersive model's (with the breaking
  // memory when the fronts are followed), or the defaults.
  if (s->dispersive)
    return Breaking{s->disp.breaking.eta_ratio_lo, s->disp.breaking.eta_ratio_hi,
                    s->disp.breaking.slope_lo, s->disp.breaking.slope_hi,
                    s->eddy_on ? s->br : nullptr};
  return Breaking{0.30, 0.55, 0.40, 0.65, nullptr};
}

// Drive::new (forcing.rs:88-117) and the static per-column part of Drive::source.
int build_drive(wc_solver* s, double still_depth) {
  const wc_wavemaker_t& m = s->maker;
  if (!(still_depth > 0.0)) return fail(WC_EINVAL, "wave-maker must sit in water");
  if (!(m.period > 0.0 && m.sigma > 0.0 && m.amplitude >= 0.0))
    return fail(WC_EINVAL, "wave-maker needs period > 0, sigma > 0 and amplitude >= 0");
  const double omega = TAU / m.period;
  wc_linear_wave_t lw;
  if (!linear_wave(omega, still_depth, s->dispersive, &lw))
    return fail(WC_EINVAL, "a wave of this angular frequency is too short for the dispersive "
                           "model at the wave-maker's depth");
  const double k = lw.k, c = lw.group_speed;
  // angle.sin() and angle.cos() of one argument: LLVM lowers the pair to glibc's sincos on
  // x86_64-linux-gnu, whose last bit can differ from sin's and cos's, so call it by name.
  double sa, ca;
  sincos(m.angle, &sa, &ca);
  const double ky = k * sa;
  const double kx = k * ca;
  const double rolloff = exp(0.5 * kx * kx * m.sigma * m.sigma);
  const double flux = 2.0 * c * m.amplitude * ca * rolloff;
  const size_t nx = (size_t)s->geo.nx;
  double norm = -0.0;
  for (size_t i = 0; i < nx; ++i) {
    const double x = ((double)i + 0.5) * s->geo.dx;
    const double d = x - m.x;
    norm = norm + exp(-(d * d) / (2.0 * m.sigma * m.sigma)) * s->geo.dx;
  }
  const double strength = flux / norm;
  // strength * profile per column where the profile is at least 1e-12 (a Gaussian: one run).
  std::vector<double> col(nx, 0.0);
  int i0 = -1, i1 = -1;
  for (size_t i = 0; i < nx; ++i) {
    const double x = ((double)i + 0.5) * s->geo.dx;  // Grid::centre (grid.rs:52-57)
    const double d = x - m.x;
    const double profile = exp(-(d * d) / (2.0 * m.sigma * m.sigma));
    if (!(profile < 1e-12)) {
      if (i0 < 0) i0 = (int)i;
      i1 = (int)i + 1;
      col[i] = strength * profile;
    }
  }
  for (int i = std::max(i0, 0); i < i1; ++i) {
    const double x = ((double)i + 0.5) * s->geo.dx;
    const double d = x - m.x;
    if (exp(-(d * d) / (2.0 * m.sigma * m.sigma)) < 1e-12)
      return fail(WC_EINVAL, "internal: the wave-maker's columns are not one run");
  }
  if (!s->src_col) TRY(dalloc(s, &s->src_col, nx));
  HIP_OK(hipMemcpyAsync(s->src_col, col.data(), nx * sizeof(double), hipMemcpyHostToDevice,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  s->omega = omega;
  s->ky = ky;
  s->strength = strength;
  s->src_i0 = i0 < 0 ? 0 : i0;
  s->src_i1 = i0 < 0 ? 0 : i1;
  s->has_drive = true;
  return WC_OK;
}

// solver.rs:421-429
int check_periodic(wc_solver* s) {
  if (s->periodic && s->has_drive) {
    const double turns = s->ky * (double)s->geo.ny * s->geo.dy / TAU;
    if (!(fabs(turns - round(turns)) < 1e-6)) {
      char buf[200];
      snprintf(buf, sizeof buf,
               "wave-maker does not fit the periodic domain: k_y * Ly / 2pi = %.17g, must be a "
               "whole number",
               turns);
      return fail(WC_EINVAL, buf);
    }
  }
  return WC_OK;
}

// solver.rs:400-409
int rebuild_drive(wc_solver* s) {
  if (!s->has_maker) return WC_OK;
  const Geo& g = s->geo;
  double fi = round(s->maker.x / g.dx - 0.5);
  fi = wc_max(fi, 0.0);
  size_t i = fi >= 1.8e19 ? (size_t)-1 : (size_t)fi;
  if (i > (size_t)g.nx - 1) i = (size_t)g.nx - 1;
  i += GH;
  const double depth = -s->bed[(size_t)(GH + g.ny / 2) * g.w + i];
  TRY(build_drive(s, depth));
  return check_periodic(s);
}

// Per-step values the host evaluates (glibc exp, sin, cos, as the reference calls them): the
// ramp and phase of the wave-maker at t and t + dt, and each sponge's exp(-rate dt).
void fill_step_params(const wc_solver* s, double t, double dt, double* p) {
  const int ny = s->geo.ny;
  if (s->has_drive) {
    const wc_wavemaker_t& m = s->maker;
    for (int st = 0; st < 2; ++st) {
      const double tt = st == 0 ? t : t + dt;
      const double ramp_time = m.ramp_periods * m.period;
      p[st] = tt >= ramp_time ? 1.0 : 0.5 * (1.0 - cos(PI * tt / ramp_time));
      double* row = p + 2 + (size_t)st * ny;
      for (int j = 0; j < ny; ++j) {
        const double y = ((double)j + 0.5) * s->geo.dy;
        row[j] = sin(s->omega * tt - s->ky * y);
      }
    }
  }
  double* f = p + 2 + 2 * (size_t)ny;
  for (size_t q = 0; q < s->sponge_rates.size(); ++q) {
    const double rate = s->sponge_rates[q];
    f[q] = rate > 0.0 ? exp(-rate * dt) : 1.0;
  }
}

int upload_step_params(wc_solver* s, double t, double dt) {
  const size_t need = s->prm_len;
  if (need == 0) return WC_OK;
  const int slot = s->prm_slot;
  s->prm_slot = (s->prm_slot + 1) % PRM_SLOTS;
  if (s->prm_used[slot]) HIP_OK(hipEventSynchronize(s->prm_ev[slot]));
  double* p = s->prm_host[slot];
  fill_step_params(s, t, dt, p);
  HIP_OK(hipMemcpyAsync(s->prm, p, need * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipEventRecord(s->prm_ev[slot], s->stream));
  s->prm_used[slot] = true;
  return WC_OK;
}

int resize_params(wc_solver* s) {
  const size_t need = 2 + 2 * (size_t)s->geo.ny + s->sponge_rates.size();
  s->prm_len = need;
  if (need <= s->prm_cap) return WC_OK;
  HIP_OK(hipStreamSynchronize(s->stream));
  for (int q = 0; q < PRM_SLOTS; ++q) {
    if (s->prm_host[q]) HIP_OK(hipHostFree(s->prm_host[q]));
    HIP_OK(hipHostMalloc((void**)&s->prm_host[q], need * sizeof(double), hipHostMallocDefault));
    memset(s->prm_host[q], 0, need * sizeof(double));
    s->prm_used[q] = false;
  }
  void* d = nullptr;
  HIP_OK(hipMalloc(&d, need * sizeof(double)));
  HIP_OK(hipMemsetAsync(d, 0, need * sizeof(double), s->stream));
  if (s->prm) {
    auto it = std::find(s->owned.begin(), s->owned.end(), (void*)s->prm);
    if (it != s->owned.end()) s->owned.erase(it);
    HIP_OK(hipFree(s->prm));
  }
  s->prm = static_cast<double*>(d);
  s->owned.push_back(d);
  s->prm_cap = need;
  return WC_OK;
}

StepArgs step_args(wc_solver* s, double dt, int stage) {
  StepArgs a{};
  a.g = s->geo;
  a.order = s->order;
  a.dt = dt;
  a.b = s->b;
  a.phi = s->phi;
  a.s_h = s->h;
  a.s_hu = s->hu;
  a.s_hv = s->hv;
  a.s1_h = s->h1;
  a.s1_hu = s->hu1;
  a.s1_hv = s->hv1;
  a.in_h = stage == 0 ? s->h : s->h1;
  a.in_hu = stage == 0 ? s->hu : s->hu1;
  a.in_hv = stage == 0 ? s->hv : s->hv1;
  a.lh = s->lh;
  a.lhu = s->lhu;
  a.lhv = s->lhv;
  a.has_drive = s->has_drive ? 1 : 0;
  a.src_i0 = s->src_i0;
  a.src_i1 = s->src_i1;
  a.src_col = s->src_col;
  a.prm = s->prm;
  a.ramp_prm = stage;
  a.sin_prm = 2 + stage * s->geo.ny;
  a.sponge_prm = 2 + 2 * s->geo.ny;
  a.n_sponges = (int)s->sponges.size();
  for (size_t m = 0; m < s->sponges.size(); ++m) {
    const HostSponge& hs = s->sponges[m];
    a.sponges[m] = SpongeDev{hs.offset, hs.len, hs.level, hs.west ? 1 : 0, hs.east ? 1 : 0,
                             hs.south ? 1 : 0, hs.north ? 1 : 0};
  }
  a.sponge_rate = s->sponge_rate_d;
  a.manning = s->manning;
  return a;
}

WarmStart ws_args(wc_solver* s);  // below, with the FP32 form

// Whether the FP64 path's solve runs the 64x16-tile kernels of wc_pcg2.hip (pcg dispatches the
// same way).
bool uses_pcg2(const wc_solver* s) { return s->pcg_kernels == 0 || s->pcg_kernels >= 4; }

int alloc_pcg2(wc_solver* s) {
  if (s->v_dg) return WC_OK;
  for (double2** p : {&s->v_dg, &s->v_r, &s->v_pa, &s->v_pb}) TRY(dalloc(s, p, s->cells));
  return WC_OK;
}

SgnArgs sgn_args(wc_solver* s, int stage) {
  SgnArgs a{};
  a.g = s->geo;
  a.h = stage == 0 ? s->h : s->h1;
  a.hu = stage == 0 ? s->hu : s->hu1;
  a.hv = stage == 0 ? s->hv : s->hv1;
  a.phi = s->phi;
  a.lh = s->lh;
  a.lhu = s->lhu;
  a.lhv = s->lhv;
  a.c = s->c;
  a.dgx = s->dgx;
  a.dgy = s->dgy;
  a.bx = s->bx;
  a.by = s->by;
  a.wx = s->wx;
  a.wy = s->wy;
  a.rx = s->rx;
  a.ry = s->ry;
  a.px = s->px;
  a.py = s->py;
  a.ax = s->ax;
  a.ay = s->ay;
  a.zx = s->zx;
  a.zy = s->zy;
  a.dg2 = uses_pcg2(s) ? s->v_dg : nullptr;
  a.part = s->part;
  a.sc = s->sc;
  a.nonlinear = s->disp.nonlinear ? 1 : 0;
  a.tolerance = s->disp.tolerance;
  a.max_iterations = s->disp.max_iterations;
  a.ws = ws_args(s);
  return a;
}

// The 256-lane PCG loop (wc_pcg.hip): two kernels per iteration, batches of iterations
// between looks at the device-side flag, a one-block check closing every batch.
template <typename V>
int pcg3(wc_solver* s, int stage, CgV<V> a, V* pA_x, V* pA_y, V* pB_x, V* pB_y) {
  a.acc_init = s->lanes;
  a.acc_pap = s->lanes + 3 * 256;
  a.acc_update = s->lanes + 4 * 256;
  a.sc = s->sc;
  a.tolerance = s->disp.tolerance;
  a.max_iterations = s->disp.max_iterations;
  const dim3 grid(256), block(1024);
  TRY(launch(s, F_CG_INIT, k_cg3_init<V>, grid, block, a));
  const unsigned long long maxit = s->disp.max_iterations;
  unsigned long long n = 1;
  unsigned long long batch = std::max<unsigned long long>(s->prev_it[stage], 1);
  while (true) {
    const unsigned long long stop = std::min(n + batch, maxit + 1);
    for (; n < stop; ++n) {
      // p of iteration 1 is what init stored in A; then p_n goes to B (n even) or A (n odd)
      V* cur_x = (n % 2 == 1) ? pA_x : pB_x;
      V* cur_y = (n % 2 == 1) ? pA_y : pB_y;
      const V* old_x = (n == 1) ? pA_x : ((n % 2 == 1) ? pB_x : pA_x);
      const V* old_y = (n == 1) ? pA_y : ((n % 2 == 1) ? pB_y : pA_y);
      TRY(launch(s, F_CG_MATVEC, k_cg3_matvec<V>, grid, block, a, n, old_x, old_y, cur_x,
                 cur_y));
      TRY(launch(s, F_CG_UPDATE, k_cg3_update<V>, grid, block, a, n, (const V*)cur_x,
                 (const V*)cur_y));
    }
    TRY(launch(s, F_CG_REDUCE, k_cg3_check<V>, dim3(1), dim3(64), a, n));
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    if (s->sc_host->done || n > maxit) break;
    batch = s->pcg_batch;
  }
  s->prev_it[stage] = s->sc_host->it;
  TRY(launch(s, F_CG_STATS, k_cg_stats, dim3(1), dim3(64), s->sc, 0ull));
  return WC_OK;
}

// The FP64 path's default PCG loop (wc_pcg2.hip): 64x16 tiles, two kernels per iteration,
// batches of iterations between looks at the device-side flag, a one-wave check closing every
// batch. p alternates between the two buffers (A, B) by iteration parity: M_n reads p_{n-2}
// where neighbouring tiles still read it while it writes p_{n-1}.
int pcg2(wc_solver* s, int stage) {
  if (!s->v_dg) return fail(WC_EINVAL, "the FP64 PCG's vectors were not allocated before the set-up");
  CgV<double> a{};
  a.g = s->geo;
  a.h = stage == 0 ? s->h : s->h1;
  a.c = s->c;
  a.dg2 = s->v_dg;
  a.r2 = s->v_r;
  a.bx = s->bx;
  a.by = s->by;
  a.wx = s->wx;
  a.wy = s->wy;
  a.acc_init = s->lanes;
  a.acc_pap = s->lanes + 3 * 256;
  a.acc_update = s->lanes + 4 * 256;
  a.sc = s->sc;
  a.tolerance = s->disp.tolerance;
  a.max_iterations = s->disp.max_iterations;
  double2 *pA = s->v_pa, *pB = s->v_pb;
  const dim3 grid(256), block(1024);
  TRY(launch(s, F_CG_INIT, k_cg2_init, grid, block, a));
  const unsigned long long maxit = s->disp.max_iterations;
  unsigned long long n = 1;
  unsigned long long batch = std::max<unsigned long long>(s->prev_it[stage], 1);
  while (true) {
    const unsigned long long stop = std::min(n + batch, maxit + 1);
    hipEvent_t b0 = nullptr, b1 = nullptr;
    if (s->prof_batch) {
      b0 = take_event(s);
      b1 = take_event(s);
      if (b0) HIP_OK(hipEventRecord(b0, s->stream));
    }
    const unsigned long long first = n;
    for (; n < stop; ++n) {
      double2* cur = (n % 2 == 1) ? pA : pB;
      const double2* old = (n % 2 == 1) ? pB : pA;
      TRY(launch(s, F_CG_MATVEC, k_cg2_matvec, grid, block, a, n, old, cur));
      TRY(launch(s, F_CG_UPDATE, n % 2 == 0 ? k_cg2_update<1> : k_cg2_update<0>, grid, block, a,
                 n, (const double2*)cur, (const double2*)old));
    }
    if (s->prof_batch && b0 && b1) {
      HIP_OK(hipEventRecord(b1, s->stream));
      s->ev_pending.push_back({F_CG_PAIR, b0, b1, n - first});
    }
    TRY(launch(s, F_CG_REDUCE, k_cg2_check, dim3(1), dim3(64), a, n));
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    if (s->sc_host->done || n > maxit) break;
    batch = s->pcg_batch;
  }
  const unsigned long long it = s->sc_host->it;
  // U_it of an odd it left its step for the next iteration, which did not run: w += alpha p_{it-1}
  // (p_{it-1} is in the buffer that iteration it's matvec wrote, pA for odd it)
  if (it % 2 == 1)
    TRY(launch(s, F_CG_FLUSH, k_cg2_flush, over((long long)s->geo.nx * s->geo.ny), dim3(NT), a,
               it, (const double2*)pA));
  s->prev_it[stage] = it;
  TRY(launch(s, F_CG_STATS, k_cg_stats, dim3(1), dim3(64), s->sc, 0ull));
  return WC_OK;
}

// The PCG loop (dispersion.rs:493-545). Iterations run in batches; every kernel of an
// iteration returns at once when the device-side flag says the solve is over, so the host
// only looks at the flag between batches. The first batch is as long as the last solve.
int pcg(wc_solver* s, int stage) {
  // 0 (default), 4 (the FP32 form's one-kernel value) and 5: the 64x16-tile kernels
  if (s->pcg_kernels == 0 || s->pcg_kernels >= 4) return pcg2(s, stage);
  const SgnArgs a = sgn_args(s, stage);
  if (s->pcg_kernels == 3) {
    CgV<double> v{};
    v.g = s->geo;
    v.h = a.h;
    v.c = s->c;
    v.bx = s->bx;
    v.by = s->by;
    v.wx = s->wx;
    v.wy = s->wy;
    v.rx = s->rx;
    v.ry = s->ry;
    v.px = s->px;
    v.py = s->py;
    v.zx = s->zx;
    v.zy = s->zy;
    v.ax = s->ax;
    v.ay = s->ay;
    return pcg3<double>(s, stage, v, s->px, s->py, s->qx, s->qy);
  }
  TRY(launch(s, F_CG_INIT, k_cg_init, tiles(s), dim3(NT), a));
  TRY(launch(s, F_CG_REDUCE, k_cg_reduce, dim3(1), dim3(NT), a, (int)REDUCE_INIT));
  const unsigned long long maxit = s->disp.max_iterations;
  unsigned long long launched = 0;
  unsigned long long batch = std::max<unsigned long long>(s->prev_it[stage], 1);
  bool seen = false;
  while (launched < maxit) {
    const unsigned long long n = std::min(batch, maxit - launched);
    for (unsigned long long q = 0; q < n; ++q) {
      if (s->pcg_kernels == 1) {
        TRY(launch(s, F_CG_MATVEC, k_cg_matvec, tiles(s), dim3(NT), a));
        TRY(launch(s, F_CG_REDUCE, k_cg_reduce, dim3(1), dim3(NT), a, (int)REDUCE_ALPHA));
        TRY(launch(s, F_CG_UPDATE, k_cg_update, tiles(s), dim3(NT), a));
        TRY(launch(s, F_CG_REDUCE, k_cg_reduce, dim3(1), dim3(NT), a, (int)REDUCE_BETA));
        TRY(launch(s, F_CG_PUPDATE, k_cg_pupdate, tiles(s), dim3(NT), a));
        continue;
      }
      // p of iteration 1 is what cg_init stored in (px, py); afterwards the new p alternates
      // between (qx, qy) and (px, py).
      const unsigned long long it = launched + q + 1;
      const int first = it == 1 ? 1 : 0;
      double* cur_x = (it % 2 == 1) ? s->px : s->qx;
      double* cur_y = (it % 2 == 1) ? s->py : s->qy;
      const double* old_x = (it % 2 == 1) ? s->qx : s->px;
      const double* old_y = (it % 2 == 1) ? s->qy : s->py;
      if (first) {
        old_x = s->px;
        old_y = s->py;
      }
      TRY(launch(s, F_CG_MATVEC, k_cg_matvec2, tiles(s), dim3(NT), a, first, old_x, old_y,
                 cur_x, cur_y));
      TRY(launch(s, F_CG_REDUCE, k_cg_reduce, dim3(1), dim3(NT), a, (int)REDUCE_ALPHA));
      TRY(launch(s, F_CG_UPDATE, k_cg_update2, tiles(s), dim3(NT), a, (const double*)cur_x,
                 (const double*)cur_y));
      TRY(launch(s, F_CG_REDUCE, k_cg_reduce, dim3(1), dim3(NT), a, (int)REDUCE_BETA));
    }
    launched += n;
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    seen = true;
    if (s->sc_host->done) break;
    batch = s->pcg_batch;
  }
  if (!seen) {
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
  }
  s->prev_it[stage] = s->sc_host->it;
  TRY(launch(s, F_CG_STATS, k_cg_stats, dim3(1), dim3(64), s->sc, 0ull));
  return WC_OK;
}

// One evaluation of the rates at stage `stage` and the update they drive (Solver::rhs,
// solver.rs:625-733, then advance or the RK2 average).
int stage(wc_solver* s, double dt, int stage, int mode) {
  const Geo& g = s->geo;
  const double* in_h = stage == 0 ? s->h : s->h1;
  if (s->order == WC_ORDER_FIFTH || s->dispersive)
    TRY(launch(s, F_SWITCH, k_switch, over((long long)g.w * g.rows), dim3(NT), g, in_h,
               (const double*)s->b, breaking_of(s), (const uint8_t*)(s->has_held ? s->held : nullptr),
               s->phi));
  const StepArgs a = step_args(s, dt, stage);
  const int flux_mode = s->dispersive ? (int)MODE_RATES : mode;
  if (s->flux_kernel == 2) {
    FluxTileArgs<double> f{};
    f.g = a.g;
    f.order = a.order;
    f.h = a.in_h;
    f.h_lo = nullptr;
    f.hu = a.in_hu;
    f.hv = a.in_hv;
    f.b = a.b;
    f.phi = a.phi;
    f.has_drive = a.has_drive;
    f.src_i0 = a.src_i0;
    f.src_i1 = a.src_i1;
    f.src_col = a.src_col;
    f.prm = a.prm;
    f.ramp_prm = a.ramp_prm;
    f.sin_prm = a.sin_prm;
    TRY(launch(s, F_FLUX, k_flux_tile64, tiles(s), dim3(NT), f, a, flux_mode));
  } else if (s->flux_kernel == 3) {
    FluxTileArgs<double> f{};
    f.g = a.g;
    f.order = a.order;
    f.h = a.in_h;
    f.h_lo = nullptr;
    f.hu = a.in_hu;
    f.hv = a.in_hv;
    f.b = a.b;
    f.phi = a.phi;
    f.has_drive = a.has_drive;
    f.src_i0 = a.src_i0;
    f.src_i1 = a.src_i1;
    f.src_col = a.src_col;
    f.prm = a.prm;
    f.ramp_prm = a.ramp_prm;
    f.sin_prm = a.sin_prm;
    dim3 grid;
    int rows;
    TRY(march_launch(s, k_flux_march64, flux_mode != MODE_FINAL, &grid, &rows,
                     [&](dim3 gr, int r) {
                       hipLaunchKernelGGL(k_flux_march64, gr, dim3(TX), 0, s->stream, f, a,
                                          flux_mode, r);
                     }));
    TRY(launch(s, F_FLUX, k_flux_march64, grid, dim3(TX), f, a, flux_mode, rows));
  } else {
    TRY(launch(s, F_FLUX, k_flux, tiles(s), dim3(NT), a, flux_mode));
  }
  if (!s->dispersive) return WC_OK;
  if (uses_pcg2(s)) TRY(alloc_pcg2(s));
  TRY(launch(s, F_SGN_SETUP, k_sgn_setup, tiles(s), dim3(NT), sgn_args(s, stage)));
  TRY(pcg(s, stage));
  TRY(launch(s, F_SGN_FINISH, k_sgn_finish, tiles(s), dim3(NT), a, (const double*)s->wx,
             (const double*)s->wy, mode, ws_args(s)));
  return WC_OK;
}

int fill(wc_solver* s, double* h, double* hu, double* hv, double* br, double* ah, double* ahu,
         double* ahv) {
  const Geo& g = s->geo;
  return launch(s, F_FILL, k_fill_state, over(6LL * g.w + 6LL * g.ny), dim3(NT), g, h, hu, hv,
                br, ah, ahu, ahv);
}

// Solver::follow_breaking (solver.rs:512-532) on the FP64 state: the switch read afresh
// (without the fronts), the fronts followed (eddy.rs:163-221), their viscosity applied in
// explicit sub-steps (eddy.rs:225-302), dry cells cleaned. The memory and the momentum swap
// buffers with the scratch arrays, as the reference swaps its vectors.
int follow_breaking(wc_solver* s, double dt) {
  if (!(s->dispersive && s->eddy_on)) return WC_OK;
  const Geo& g = s->geo;
  TRY(fill(s, s->h, s->hu, s->hv, s->br, nullptr, nullptr, nullptr));
  Breaking fresh = breaking_of(s);
  fresh.memory = nullptr;
  TRY(launch(s, F_SWITCH, k_switch, over((long long)g.w * g.rows), dim3(NT), g,
             (const double*)s->h, (const double*)s->b, fresh, (const uint8_t*)nullptr, s->phi));
  TRY(launch(s, F_SWITCH, k_eddy_rise, over((long long)g.w * g.rows), dim3(NT), g,
             (const double*)s->hu, (const double*)s->hv, s->e_rise));
  TRY(launch(s, F_SWITCH, k_eddy_follow, over((long long)g.w * g.rows), dim3(NT), g, s->eddy,
             (const double*)s->h, (const double*)s->br, (const double*)s->phi,
             (const double*)s->e_rise, dt, s->e_mem, s->e_nu));
  std::swap(s->br, s->e_mem);
  if (s->eddy.mixing > 0.0) {
    HIP_OK(hipMemsetAsync(s->red64, 0, sizeof(unsigned long long), s->stream));
    TRY(launch(s, F_REDUCE, k_eddy_top, over((long long)g.nx * g.ny), dim3(NT), g,
               (const double*)s->e_nu, s->red64));
    HIP_OK(hipMemcpyAsync(s->red64_host, s->red64, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    double top;
    memcpy(&top, &s->red64_host[0], sizeof top);
    if (top > 0.0) {
      const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
      const double limit = 0.5 / (top * wc_max(2.0 * idx2 + idy2, idx2 + 2.0 * idy2));
      const unsigned long long steps = (unsigned long long)wc_max(ceil(dt / limit), 1.0);
      const double sub = dt / (double)steps;
      for (unsigned long long q = 0; q < steps; ++q) {
        if (q > 0) TRY(fill(s, s->h, s->hu, s->hv, s->br, nullptr, nullptr, nullptr));
        TRY(launch(s, F_SGN_FINISH, k_eddy_diffuse, over((long long)g.nx * g.ny), dim3(NT), g,
                   (const double*)s->h, (const double*)s->hu, (const double*)s->hv,
                   (const double*)s->e_nu, sub, s->e_hu, s->e_hv));
        std::swap(s->hu, s->e_hu);
        std::swap(s->hv, s->e_hv);
      }
    }
  }
  TRY(launch(s, F_SGN_FINISH, k_clean_dry, over((long long)g.nx * g.ny), dim3(NT), g, s->h, s->hu,
             s->hv));
  return WC_OK;
}

// Solver::step (solver.rs:481-507).
int do_step(wc_solver* s, double dt) {
  const double t = s->time;
  TRY(upload_step_params(s, t, dt));
  TRY(fill(s, s->h, s->hu, s->hv, s->br, nullptr, nullptr, nullptr));
  if (s->order == WC_ORDER_FIRST) {
    TRY(stage(s, dt, 0, MODE_S1));
    TRY(launch(s, F_EULER, k_finish_euler, tiles(s), dim3(NT), step_args(s, dt, 0)));
  } else {
    TRY(stage(s, dt, 0, MODE_S1));
    TRY(fill(s, s->h1, s->hu1, s->hv1, nullptr, s->h, s->hu, s->hv));
    TRY(stage(s, dt, 1, MODE_FINAL));
  }
  TRY(follow_breaking(s, dt));
  s->time = t + dt;
  return WC_OK;
}

int stable_dt(wc_solver* s, double* dt) {
  const Geo& g = s->geo;
  HIP_OK(hipMemsetAsync(s->red64, 0, sizeof(unsigned long long), s->stream));
  TRY(launch(s, F_REDUCE, k_max_rate, rate_blocks((long long)g.nx * g.ny), dim3(NT), g,
             (const double*)s->h, (const double*)s->hu, (const double*)s->hv, s->red64));
  HIP_OK(hipMemcpyAsync(s->red64_host, s->red64, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  double rate;
  const unsigned long long bits = s->red64_host[0];
  memcpy(&rate, &bits, sizeof rate);
  if (isnan(rate))
    *dt = NAN;
  else if (rate > 0.0)
    *dt = s->cfl / rate;
  else
    *dt = INFINITY;
  return WC_OK;
}

int alloc_dispersion(wc_solver* s) {
  if (s->c) return WC_OK;
  const size_t n = s->cells;
  for (double** p : {&s->c, &s->dgx, &s->dgy, &s->bx, &s->by, &s->wx, &s->wy, &s->rx, &s->ry,
                     &s->px, &s->py, &s->ax, &s->ay, &s->zx, &s->zy, &s->qx, &s->qy})
    TRY(dalloc(s, p, n));
  return WC_OK;
}

int alloc_fp32(wc_solver* s) {
  if (s->f_sh) return WC_OK;
  const size_t n = s->cells;
  for (double** p : {&s->f_l1h, &s->f_lh, &s->f_lhu, &s->f_lhv, &s->f_bx, &s->f_by})
    TRY(dalloc(s, p, n));
  for (float** p : {&s->f_sh, &s->f_shu, &s->f_shv, &s->f_slh, &s->f_slhu, &s->f_slhv, &s->f_s1h,
                    &s->f_s1hl, &s->f_s1hu, &s->f_s1hv, &s->f_s1hul, &s->f_s1hvl,
                    &s->f_l1hu, &s->f_l1hv,
                    &s->f_phi, &s->f_c, &s->f_wx,
                    &s->f_wy, &s->f_rx, &s->f_ry, &s->f_px, &s->f_py, &s->f_qx, &s->f_qy,
                    &s->f_zx, &s->f_zy, &s->f_ax, &s->f_ay})
    TRY(dalloc(s, p, n));
  for (float2** p : {&s->g_r[0], &s->g_r[1], &s->g_q[0], &s->g_q[1], &s->g_s[0], &s->g_s[1],
                     &s->g_p, &s->g_di})
    TRY(dalloc(s, p, n));
  TRY(dalloc(s, &s->ws_ring, WS_SLOTS * n));
  return WC_OK;
}

// The warm start's record of past solution steps, cleared: the next solves start as the
// reference's (from the previous solution) until the record refills.
int ws_reset(wc_solver* s) {
  if (s->ws_ring)
    HIP_OK(hipMemsetAsync(s->ws_ring, 0, WS_SLOTS * s->cells * sizeof(float2), s->stream));
  if (s->ws_ring64)
    HIP_OK(hipMemsetAsync(s->ws_ring64, 0, WS_SLOTS * s->cells * sizeof(double2), s->stream));
  return WC_OK;
}

// The warm start's weights c_1..c_K (WarmStart): exact for steps of degree < m, of least norm.
// K = m: the binomial weights (-1)^(j+1) C(m, j). K > m: c = V^T (V V^T)^-1 e_0 with
// V[r][j] = (-j)^r, r < m, j = 1..K (the least-norm solution of V c = e_0), by Gaussian
// elimination with partial pivoting on the m x m system.
void ws_coefficients(int m, int K, double* c) {
  if (K == m) {
    double cf = 1.0;
    for (int j = 1; j <= m; ++j) {
      cf = cf * (double)(m - j + 1) / (double)j;
      c[j - 1] = (j & 1) ? cf : -cf;
    }
    return;
  }
  double V[WS_MAX_ORDER][WS_MAX_WINDOW], A[WS_MAX_ORDER][WS_MAX_ORDER + 1];
  for (int r = 0; r < m; ++r)
    for (int j = 1; j <= K; ++j) V[r][j - 1] = std::pow(-(double)j, (double)r);
  for (int r = 0; r < m; ++r) {
    for (int q = 0; q < m; ++q) {
      double t = 0.0;
      for (int j = 0; j < K; ++j) t += V[r][j] * V[q][j];
      A[r][q] = t;
    }
    A[r][m] = r == 0 ? 1.0 : 0.0;
  }
  for (int col = 0; col < m; ++col) {
    int piv = col;
    for (int r = col + 1; r < m; ++r)
      if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
    for (int q = 0; q <= m; ++q) std::swap(A[col][q], A[piv][q]);
    for (int r = 0; r < m; ++r) {
      if (r == col) continue;
      const double f = A[r][col] / A[col][col];
      for (int q = col; q <= m; ++q) A[r][q] -= f * A[col][q];
    }
  }
  double y[WS_MAX_ORDER];
  for (int r = 0; r < m; ++r) y[r] = A[r][m] / A[r][r];
  for (int j = 0; j < K; ++j) {
    double t = 0.0;
    for (int r = 0; r < m; ++r) t += V[r][j] * y[r];
    c[j] = t;
  }
}

WarmStart ws_args(wc_solver* s) {
  WarmStart w{};
  if (s->ws_order64 > 0 && s->dispersive && s->precision == WC_PRECISION_FP64 && s->ws_ring64) {
    w.ring64 = s->ws_ring64;
    w.stride = s->cells;
    w.sc = s->sc;
    w.terms = std::max(s->ws_window, s->ws_order64);
    w.period = s->order == WC_ORDER_FIRST ? 1 : 2;
    ws_coefficients(s->ws_order64, w.terms, w.coef);
  }
  if (s->ws_order > 0 && s->dispersive && s->precision == WC_PRECISION_FP32 && s->ws_ring) {
    w.ring = s->ws_ring;
    w.stride = s->cells;
    w.sc = s->sc;
    w.terms = std::max(s->ws_window, s->ws_order);
    w.period = s->order == WC_ORDER_FIRST ? 1 : 2;
    ws_coefficients(s->ws_order, w.terms, w.coef);
    w.guard = (float)s->ws_guard * 1e-3f;
  }
  return w;
}

// FP64 state arrays -> FP32 pairs (after an upload, or when switching to FP32).
int split_state(wc_solver* s) {
  const long long n = (long long)s->cells;
  TRY(launch(s, F_REDUCE, k_split32, over(n), dim3(NT), n, (const double*)s->h, s->f_sh,
             s->f_slh));
  TRY(launch(s, F_REDUCE, k_split32, over(n), dim3(NT), n, (const double*)s->hu, s->f_shu,
             s->f_slhu));
  TRY(launch(s, F_REDUCE, k_split32, over(n), dim3(NT), n, (const double*)s->hv, s->f_shv,
             s->f_slhv));
  return WC_OK;
}

// FP32 pairs -> the FP64 state arrays, ghosts refilled from the interior (before a download
// or a diagnostic, or when switching back to FP64).
int join_state(wc_solver* s) {
  if (s->precision != WC_PRECISION_FP32) return WC_OK;
  const long long n = (long long)s->cells;
  TRY(launch(s, F_REDUCE, k_join32, over(n), dim3(NT), n, (const float*)s->f_sh,
             (const float*)s->f_slh, s->h));
  TRY(launch(s, F_REDUCE, k_join32, over(n), dim3(NT), n, (const float*)s->f_shu,
             (const float*)s->f_slhu, s->hu));
  TRY(launch(s, F_REDUCE, k_join32, over(n), dim3(NT), n, (const float*)s->f_shv,
             (const float*)s->f_slhv, s->hv));
  TRY(fill(s, s->h, s->hu, s->hv, nullptr, nullptr, nullptr, nullptr));
  return WC_OK;
}

Sgn32 sgn32_args(wc_solver* s, int stage) {
  Sgn32 a{};
  a.g = s->geo;
  a.h = stage == 0 ? s->f_sh : s->f_s1h;
  a.hu = stage == 0 ? s->f_shu : s->f_s1hu;
  a.hv = stage == 0 ? s->f_shv : s->f_s1hv;
  a.hl = stage == 0 ? s->f_slh : s->f_s1hl;
  a.hul = stage == 0 ? s->f_slhu : s->f_s1hul;
  a.hvl = stage == 0 ? s->f_slhv : s->f_s1hvl;
  a.phi = s->f_phi;
  a.lh = s->f_lh;
  a.lhu = s->f_lhu;
  a.lhv = s->f_lhv;
  a.c = s->f_c;
  a.bx = s->f_bx;
  a.by = s->f_by;
  a.wx = s->f_wx;
  a.wy = s->f_wy;
  a.solx = s->wx;
  a.soly = s->wy;
  a.rx = s->f_rx;
  a.ry = s->f_ry;
  a.px = s->f_px;
  a.py = s->f_py;
  a.ax = s->f_ax;
  a.ay = s->f_ay;
  a.zx = s->f_zx;
  a.zy = s->f_zy;
  a.part = s->part;
  a.sc = s->sc;
  a.nonlinear = s->disp.nonlinear ? 1 : 0;
  a.tolerance = s->disp.tolerance;
  a.max_iterations = s->disp.max_iterations;
  a.ws = ws_args(s);
  return a;
}

// The one-kernel PCG on 64x16 tiles: the blocks that take tiles, at most its lanes and few
// enough that each takes pcg_walk tiles (block b takes b, b + stride, ...). The lanes past them
// read zero.
int tile_stride(const wc_solver* s) {
  const int nt = ((s->geo.nx + 63) / 64) * ((s->geo.ny + 15) / 16);
  return std::max(1, std::min(s->pcg_lanes, (nt + s->pcg_walk - 1) / s->pcg_walk));
}

// The one-kernel PCG (wc_pcg1.hip), FP32 form: start kernels, then one kernel per iteration in
// batches, a one-wave check closing every batch.
CgG cg1_args(wc_solver* s, int stage) {
  const Sgn32 sa = sgn32_args(s, stage);
  CgG a{};
  a.g = s->geo;
  a.h = sa.h;
  a.c = s->f_c;
  a.bx = s->f_bx;
  a.by = s->f_by;
  a.wx = s->f_wx;
  a.wy = s->f_wy;
  a.solx = s->wx;
  a.soly = s->wy;
  a.p = s->g_p;
  a.di = s->g_di;
  for (int m = 0; m < 2; ++m) {
    a.r[m] = s->g_r[m];
    a.q[m] = s->g_q[m];
    a.s[m] = s->g_s[m];
  }
  a.lanes_b = s->lanes;
  a.lanes = s->lanes + s->pcg_lanes;
  a.sc = s->sc;
  a.tolerance = s->disp.tolerance;
  a.max_iterations = s->disp.max_iterations;
  a.defer = s->capture ? 0 : 1;  // captured iterates must hold every step
  a.stride = tile_stride(s);
  return a;
}

template <int L>
int pcg1_32l(wc_solver* s, int stage, unsigned long long earlier) {
  const Sgn32 sa = sgn32_args(s, stage);
  const CgG a = cg1_args(s, stage);
  const dim3 grid(L), block(1024);
  const size_t fbytes = s->cells * sizeof(float);
  static constexpr unsigned long long cap_it[wc_solver::CAP_SLOTS] = {0, 2, 5, 10};
  // training pairs: the depth, and (r_n, w_n) before iteration n of the solve
  auto capture_at = [&](int slot, const float2* r) -> int {
    HIP_OK(hipMemcpyAsync(s->cap_wx[slot], s->f_wx, fbytes, hipMemcpyDeviceToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->cap_wy[slot], s->f_wy, fbytes, hipMemcpyDeviceToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->cap_r[slot], r, 2 * fbytes, hipMemcpyDeviceToDevice, s->stream));
    return WC_OK;
  };
  if (s->capture) HIP_OK(hipMemcpyAsync(s->cap_h, sa.h, fbytes, hipMemcpyDeviceToDevice, s->stream));
  const bool mar = s->pcg_march != 0;
  TRY(mar ? launch(s, F_CG_INIT, k_cg1m_init<L>, grid, block, a)
          : launch(s, F_CG_INIT, k_cg1_init<L>, grid, block, a));
  if (s->capture) TRY(capture_at(0, s->g_r[0]));
  TRY(mar ? launch(s, F_CG_INIT, k_cg1m_start<L>, grid, block, a)
          : launch(s, F_CG_INIT, k_cg1_start<L>, grid, block, a));
  const unsigned long long maxit = s->disp.max_iterations;
  unsigned long long n = 1;
  unsigned long long batch = std::max<unsigned long long>(s->prev_it[stage], 1);
  while (true) {
    const unsigned long long stop = std::min(n + batch, maxit + 1);
    hipEvent_t b0 = nullptr, b1 = nullptr;
    if (s->prof_batch) {
      b0 = take_event(s);
      b1 = take_event(s);
      if (b0) HIP_OK(hipEventRecord(b0, s->stream));
    }
    const unsigned long long first = n;
    for (; n < stop; ++n) {
      TRY(mar ? launch(s, F_CG_ITER, k_cg1m_iter<L>, grid, block, a, n)
              : launch(s, F_CG_ITER, k_cg1_iter<L>, grid, block, a, n));
      if (s->capture)  // K_n leaves r_n in r[n & 1] and w_n in w
        for (int slot = 1; slot < wc_solver::CAP_SLOTS; ++slot)
          if (n == cap_it[slot]) TRY(capture_at(slot, s->g_r[n & 1]));
    }
    if (s->prof_batch && b0 && b1) {
      HIP_OK(hipEventRecord(b1, s->stream));
      s->ev_pending.push_back({F_CG_ITER, b0, b1, n - first});
    }
    TRY(launch(s, F_CG_REDUCE, k_cg1_check<L>, dim3(1), dim3(64), a, n));
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    if (s->sc_host->done || n > maxit) break;
    batch = s->pcg_batch;
  }
  s->prev_it[stage] = s->sc_host->it;
  s->cap_last_it = s->sc_host->it;
  TRY(launch(s, F_CG_FLUSH, k_cg1_flush, tiles(s), dim3(NT), a));
  TRY(launch(s, F_CG_STATS, k_cg_stats, dim3(1), dim3(64), s->sc, earlier));
  return WC_OK;
}

int pcg1_32(wc_solver* s, int stage, unsigned long long earlier = 0) {
  return s->pcg_lanes == 512 ? pcg1_32l<512>(s, stage, earlier) : pcg1_32l<256>(s, stage, earlier);
}

// Flexible PCG with the network (wc_fpcg.hip), FP32 form: batches of iterations, each the input
// kernel, the network, the output kernel, the matvec and the update; a one-wave check ends a
// batch.
int pcg_net32(wc_solver* s, int stage) {
  const Sgn32 sa = sgn32_args(s, stage);
  FcgArgs a{};
  a.g = s->geo;
  a.h = sa.h;
  a.c = s->f_c;
  a.bx = s->f_bx;
  a.by = s->f_by;
  a.wx = s->f_wx;
  a.wy = s->f_wy;
  a.solx = s->wx;
  a.soly = s->wy;
  a.r = s->g_r[0];
  a.q = s->g_q[0];
  a.z = s->g_s[0];
  a.di = s->g_di;
  a.p[0] = s->g_p;
  a.p[1] = s->g_r[1];
  a.bgx = s->bgx;
  a.bgy = s->bgy;
  void* x0;
  const void* y0;
  int w0, h0;
  const float* wo;
  unet_views(s->net, &x0, &y0, &w0, &h0, &wo);
  a.x0 = (__bf16*)x0;
  a.y0 = (const __bf16*)y0;
  a.nw = w0;
  a.wo = wo;
  a.hs = 1.0 / (4.0 * s->geo.dx);
  a.lanes = s->lanes;
  a.sc = s->sc;
  a.tolerance = s->disp.tolerance;
  a.max_iterations = s->disp.max_iterations;
  a.stall_limit = s->precond_stall;
  a.min_ratio = 0.01 * s->precond_ratio;
  a.history = s->history;
  a.history_len = 1024;
  const dim3 grid(256), block(1024);
  TRY(launch(s, F_CG_INIT, k_fcg_init, grid, block, a));
  const unsigned long long maxit = s->disp.max_iterations;
  unsigned long long n = 1;
  unsigned long long batch = 4;
  while (true) {
    const unsigned long long stop = std::min(n + batch, maxit + 1);
    for (; n < stop; ++n) {
      TRY(launch(s, F_NET, k_fcg_in, grid, block, a, n));
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (s->prof) {
        e0 = take_event(s);
        e1 = take_event(s);
        if (e0) HIP_OK(hipEventRecord(e0, s->stream));
      }
      TRY(unet_body(s->net, s->stream, &s->sc->net_off));
      if (s->prof && e0 && e1) {
        HIP_OK(hipEventRecord(e1, s->stream));
        s->ev_pending.push_back({F_NET, e0, e1, 1});
      }
      TRY(launch(s, F_NET, k_fcg_out, grid, block, a, n));
      TRY(launch(s, F_CG_MATVEC, k_fcg_mv, grid, block, a, n));
      TRY(launch(s, F_CG_UPDATE, k_fcg_up, grid, block, a, n));
    }
    TRY(launch(s, F_CG_REDUCE, k_fcg_check, dim3(1), dim3(64), a, n));
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    if (s->sc_host->done || n > maxit) break;
    // Fallen back: the rest of the solve is the reference's Jacobi PCG from the current w (the
    // fallback restarts the directions anyway), in the one-kernel form, with the network's
    // iterations counted in.
    if (s->sc_host->fallback) return pcg1_32(s, stage, s->sc_host->it);
  }
  TRY(launch(s, F_CG_STATS, k_cg_stats, dim3(1), dim3(64), s->sc, 0ull));
  return WC_OK;
}

int pcg32(wc_solver* s, int stage) {
  if (s->precond == WC_PRECOND_NETWORK) return pcg_net32(s, stage);
  if (s->pcg_kernels == 0 || s->pcg_kernels == 4) return pcg1_32(s, stage);
  const Sgn32 a = sgn32_args(s, stage);
  if (s->pcg_kernels == 3) {
    CgV<float> v{};
    v.g = s->geo;
    v.h = a.h;
    v.c = s->f_c;
    v.bx = s->f_bx;
    v.by = s->f_by;
    v.solx = s->wx;
    v.soly = s->wy;
    v.wx = s->f_wx;
    v.wy = s->f_wy;
    v.rx = s->f_rx;
    v.ry = s->f_ry;
    v.px = s->f_px;
    v.py = s->f_py;
    v.zx = s->f_zx;
    v.zy = s->f_zy;
    v.ax = s->f_ax;
    v.ay = s->f_ay;
    return pcg3<float>(s, stage, v, s->f_px, s->f_py, s->f_qx, s->f_qy);
  }
  TRY(launch(s, F_CG_INIT, k_cg_init32, tiles(s), dim3(NT), a));
  TRY(launch(s, F_CG_REDUCE, k_cg_reduce32, dim3(1), dim3(NT), a, (int)REDUCE_INIT));
  const unsigned long long maxit = s->disp.max_iterations;
  unsigned long long launched = 0;
  unsigned long long batch = std::max<unsigned long long>(s->prev_it[stage], 1);
  bool seen = false;
  while (launched < maxit) {
    const unsigned long long n = std::min(batch, maxit - launched);
    for (unsigned long long q = 0; q < n; ++q) {
      const unsigned long long it = launched + q + 1;
      const int first = it == 1 ? 1 : 0;
      float* cur_x = (it % 2 == 1) ? s->f_px : s->f_qx;
      float* cur_y = (it % 2 == 1) ? s->f_py : s->f_qy;
      const float* old_x = first ? s->f_px : ((it % 2 == 1) ? s->f_qx : s->f_px);
      const float* old_y = first ? s->f_py : ((it % 2 == 1) ? s->f_qy : s->f_py);
      TRY(launch(s, F_CG_MATVEC, k_cg_matvec32, tiles(s), dim3(NT), a, first, old_x, old_y,
                 cur_x, cur_y));
      TRY(launch(s, F_CG_REDUCE, k_cg_reduce32, dim3(1), dim3(NT), a, (int)REDUCE_ALPHA));
      TRY(launch(s, F_CG_UPDATE, k_cg_update32, tiles(s), dim3(NT), a, (const float*)cur_x,
                 (const float*)cur_y));
      TRY(launch(s, F_CG_REDUCE, k_cg_reduce32, dim3(1), dim3(NT), a, (int)REDUCE_BETA));
    }
    launched += n;
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    seen = true;
    if (s->sc_host->done) break;
    batch = s->pcg_batch;
  }
  if (!seen) {
    HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
  }
  s->prev_it[stage] = s->sc_host->it;
  TRY(launch(s, F_CG_STATS, k_cg_stats, dim3(1), dim3(64), s->sc, 0ull));
  return WC_OK;
}

// The argument blocks of one FP32 stage, shared by the single launches below and the batched
// ones (wc_batch_*).
BSwitch switch_args32(wc_solver* s, int stage) {
  BSwitch a{};
  a.g = s->geo;
  a.h = stage == 0 ? s->f_sh : s->f_s1h;
  a.hl = stage == 0 ? s->f_slh : s->f_s1hl;
  a.b = s->b;
  a.br = breaking_of(s);
  a.held = s->has_held ? s->held : nullptr;
  a.phi = s->f_phi;
  return a;
}

BFlux flux_args32(wc_solver* s, int stage) {
  const Geo& g = s->geo;
  BFlux a{};
  FluxTileArgs<float>& f = a.f;
  f.g = g;
  f.order = s->order;
  f.h = stage == 0 ? s->f_sh : s->f_s1h;
  f.h_lo = stage == 0 ? s->f_slh : s->f_s1hl;
  f.hu = stage == 0 ? s->f_shu : s->f_s1hu;
  f.hv = stage == 0 ? s->f_shv : s->f_s1hv;
  f.hu_lo = stage == 0 ? s->f_slhu : s->f_s1hul;
  f.hv_lo = stage == 0 ? s->f_slhv : s->f_s1hvl;
  f.b = s->b;
  f.phi = s->f_phi;
  f.has_drive = s->has_drive ? 1 : 0;
  f.src_i0 = s->src_i0;
  f.src_i1 = s->src_i1;
  f.src_col = s->src_col;
  f.prm = s->prm;
  f.ramp_prm = stage;
  f.sin_prm = 2 + stage * g.ny;
  a.lh = s->f_lh;
  a.lhu = s->f_lhu;
  a.lhv = s->f_lhv;
  a.recon64 = s->recon64;
  return a;
}

Stage32 stage_args32(wc_solver* s, double dt, int stage) {
  Stage32 a{};
  a.g = s->geo;
  a.dt = dt;
  a.b = s->b;
  a.sh = s->f_sh;
  a.shu = s->f_shu;
  a.shv = s->f_shv;
  a.sl_h = s->f_slh;
  a.sl_hu = s->f_slhu;
  a.sl_hv = s->f_slhv;
  a.s1h = s->f_s1h;
  a.s1hl = s->f_s1hl;
  a.s1hu = s->f_s1hu;
  a.s1hv = s->f_s1hv;
  a.s1hul = s->f_s1hul;
  a.s1hvl = s->f_s1hvl;
  a.l1h = s->f_l1h;
  a.l1hu = s->f_l1hu;
  a.l1hv = s->f_l1hv;
  a.lh = s->f_lh;
  a.lhu = s->f_lhu;
  a.lhv = s->f_lhv;
  a.wx = s->f_wx;
  a.wy = s->f_wy;
  a.solx = s->wx;
  a.soly = s->wy;
  a.dispersive = s->dispersive ? 1 : 0;
  a.damp = step_args(s, dt, stage);
  a.ws = ws_args(s);
  return a;
}

bool switch_on(const wc_solver* s) { return s->order == WC_ORDER_FIFTH || s->dispersive; }

int stage32(wc_solver* s, double dt, int stage, int mode) {
  const Geo& g = s->geo;
  if (switch_on(s)) {
    const BSwitch w = switch_args32(s, stage);
    TRY(launch(s, F_SWITCH, k_switch32, over((long long)g.w * g.rows), dim3(NT), g, w.h, w.hl,
               w.b, w.br, w.held, w.phi));
  }
  const BFlux f = flux_args32(s, stage);
  if (s->flux_kernel == 3) {
    dim3 grid;
    int rows;
    TRY(march_launch(s, s->recon64 ? k_flux_march32r : k_flux_march32, true, &grid, &rows,
                     [&](dim3 gr, int r) {
                       if (s->recon64)
                         hipLaunchKernelGGL(k_flux_march32r, gr, dim3(TX), 0, s->stream, f.f,
                                            f.lh, f.lhu, f.lhv, r);
                       else
                         hipLaunchKernelGGL(k_flux_march32, gr, dim3(TX), 0, s->stream, f.f,
                                            f.lh, f.lhu, f.lhv, r);
                     }));
    TRY(s->recon64
            ? launch(s, F_FLUX, k_flux_march32r, grid, dim3(TX), f.f, f.lh, f.lhu, f.lhv, rows)
            : launch(s, F_FLUX, k_flux_march32, grid, dim3(TX), f.f, f.lh, f.lhu, f.lhv, rows));
  } else {
    TRY(s->recon64
            ? launch(s, F_FLUX, k_flux_tile32r, tiles(s), dim3(NT), f.f, f.lh, f.lhu, f.lhv)
            : launch(s, F_FLUX, k_flux_tile32, tiles(s), dim3(NT), f.f, f.lh, f.lhu, f.lhv));
  }
  if (s->dispersive) {
    TRY(launch(s, F_SGN_SETUP, k_sgn_setup32, tiles(s), dim3(NT), sgn32_args(s, stage)));
    TRY(pcg32(s, stage));
  }
  TRY(launch(s, F_SGN_FINISH, k_stage32, tiles(s), dim3(NT), stage_args32(s, dt, stage), mode));
  return WC_OK;
}

int do_step32(wc_solver* s, double dt) {
  const double t = s->time;
  const Geo& g = s->geo;
  TRY(upload_step_params(s, t, dt));
  TRY(launch(s, F_FILL, k_fill32, over(6LL * g.w + 6LL * g.ny), dim3(NT), g, s->f_sh, s->f_slh,
             s->f_shu, s->f_shv, s->f_slhu, s->f_slhv));
  if (s->order == WC_ORDER_FIRST) {
    TRY(stage32(s, dt, 0, MODE_EULER));
  } else {
    TRY(stage32(s, dt, 0, MODE_S1));
    TRY(launch(s, F_FILL, k_fill32, over(6LL * g.w + 6LL * g.ny), dim3(NT), g, s->f_s1h,
               s->f_s1hl, s->f_s1hu, s->f_s1hv, s->f_s1hul, s->f_s1hvl));
    TRY(stage32(s, dt, 1, MODE_FINAL));
  }
  if (s->dispersive && s->eddy_on) {  // the fronts and their viscosity run in FP64
    TRY(join_state(s));
    TRY(follow_breaking(s, dt));
    TRY(split_state(s));
  }
  s->time = t + dt;
  return WC_OK;
}

int stable_dt32(wc_solver* s, double* dt) {
  const Geo& g = s->geo;
  HIP_OK(hipMemsetAsync(s->red64, 0, sizeof(unsigned long long), s->stream));
  TRY(launch(s, F_REDUCE, k_max_rate32, rate_blocks((long long)g.nx * g.ny), dim3(NT), g,
             (const float*)s->f_sh, (const float*)s->f_slh, (const float*)s->f_shu,
             (const float*)s->f_slhu, (const float*)s->f_shv, (const float*)s->f_slhv,
             s->red64));
  HIP_OK(hipMemcpyAsync(s->red64_host, s->red64, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  double rate;
  const unsigned long long bits = s->red64_host[0];
  memcpy(&rate, &bits, sizeof rate);
  if (isnan(rate))
    *dt = NAN;
  else if (rate > 0.0)
    *dt = s->cfl / rate;
  else
    *dt = INFINITY;
  return WC_OK;
}

}  // namespace

// ================================================================== the C ABI
namespace wc {
// The error setter for the other sources of the library (wc_unet.hip).
int api_fail(int code, const char* msg) { return fail(code, msg); }
}  // namespace wc

namespace wc {
void detach(wc_stats* st);       // releases the device memory; the object stays, inert
void detach(wc_relaxation* r);
}  // namespace wc

extern "C" {

const char* wc_last_error(void) { return g_err.c_str(); }

int wc_solver_new(uint64_t nx, uint64_t ny, double dx, double dy, const double* bed,
                  wc_solver** out) {
  if (!out) return fail(WC_EINVAL, "null output pointer");
  *out = nullptr;
  if (!(nx > 0 && ny > 0)) return fail(WC_EINVAL, "grid must have at least one interior cell");
  if (!(dx > 0.0 && dy > 0.0)) return fail(WC_EINVAL, "cell size must be positive");
  if (nx < 3 || ny < 3)
    return fail(WC_EINVAL, "the library needs at least 3 interior cells along x and along y");
  if ((nx + 6) * (ny + 6) > (uint64_t)0x7fffffff / 2)
    return fail(WC_EINVAL, "grid too large for 32-bit cell indices");
  if (!bed) return fail(WC_EINVAL, "null bed");
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(WC_EDEVICE, std::string("libwavecore_hip is built for gfx950, device is ") +
                                prop.gcnArchName);
  auto* s = new wc_solver();
  s->device = dev;
  Geo& g = s->geo;
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.w = (int)nx + 2 * GH;
  g.rows = (int)ny + 2 * GH;
  g.dx = dx;
  g.dy = dy;
  g.periodic = 0;
  g.ntx = (g.nx + TX - 1) / TX;
  g.nty = (g.ny + TY - 1) / TY;
  s->cells = (size_t)g.w * g.rows;
  s->bed.assign(bed, bed + s->cells);
  auto cleanup = [&](int rc) {
    wc_solver_free(s);
    return rc;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(WC_EDEVICE, "hipStreamCreate failed"));
  const size_t n = s->cells;
  int rc = WC_OK;
  for (double** p : {&s->h, &s->hu, &s->hv, &s->br, &s->h1, &s->hu1, &s->hv1, &s->b, &s->phi,
                     &s->lh, &s->lhu, &s->lhv})
    if ((rc = dalloc(s, p, n)) != WC_OK) return cleanup(rc);
  const size_t ntiles = (size_t)g.ntx * g.nty;
  if ((rc = dalloc(s, &s->part, 4 * ntiles)) != WC_OK) return cleanup(rc);
  if ((rc = dalloc(s, &s->sc, 1)) != WC_OK) return cleanup(rc);
  if ((rc = dalloc(s, &s->lanes, 8 * 512)) != WC_OK) return cleanup(rc);
  if ((rc = dalloc(s, &s->red64, 4)) != WC_OK) return cleanup(rc);
  if ((rc = dalloc(s, &s->red32, 4)) != WC_OK) return cleanup(rc);
  if ((rc = dalloc(s, &s->redf, 4)) != WC_OK) return cleanup(rc);
  if (hipMemcpyAsync(s->b, s->bed.data(), n * sizeof(double), hipMemcpyHostToDevice,
                     s->stream) != hipSuccess)
    return cleanup(fail(WC_EDEVICE, "bed upload failed"));
  if (hipHostMalloc((void**)&s->sc_host, sizeof(CgScalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->red64_host, 4 * sizeof(unsigned long long),
                    hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->red32_host, 4 * sizeof(unsigned int), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&s->redf_host, 4 * sizeof(double), hipHostMallocDefault) !=
          hipSuccess)
    return cleanup(fail(WC_EDEVICE, "pinned allocation failed"));
  for (int q = 0; q < PRM_SLOTS; ++q)
    if (hipEventCreateWithFlags(&s->prm_ev[q], hipEventDisableTiming) != hipSuccess)
      return cleanup(fail(WC_EDEVICE, "hipEventCreate failed"));
  if ((rc = resize_params(s)) != WC_OK) return cleanup(rc);
  if (hipStreamSynchronize(s->stream) != hipSuccess)
    return cleanup(fail(WC_EDEVICE, "initial synchronisation failed"));
  *out = s;
  return WC_OK;
}

void wc_solver_free(wc_solver* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (wc_stats* st : std::vector<wc_stats*>(s->stats)) wc::detach(st);
  for (wc_relaxation* r : std::vector<wc_relaxation*>(s->relaxations)) wc::detach(r);
  if (s->net) wcu_net_free(s->net);
  for (void* p : s->owned) (void)hipFree(p);
  for (int q = 0; q < PRM_SLOTS; ++q) {
    if (s->prm_host[q]) (void)hipHostFree(s->prm_host[q]);
    if (s->prm_ev[q]) (void)hipEventDestroy(s->prm_ev[q]);
  }
  if (s->sc_host) (void)hipHostFree(s->sc_host);
  if (s->red64_host) (void)hipHostFree(s->red64_host);
  if (s->red32_host) (void)hipHostFree(s->red32_host);
  if (s->redf_host) (void)hipHostFree(s->redf_host);
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  for (auto& p : s->ev_pending) {
    (void)hipEventDestroy(p.e0);
    (void)hipEventDestroy(p.e1);
  }
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int wc_solver_set_cfl(wc_solver* s, double cfl) {
  s->cfl = cfl;
  return WC_OK;
}

int wc_solver_cfl(const wc_solver* s, double* cfl) {
  *cfl = s->cfl;
  return WC_OK;
}

int wc_solver_with_order(wc_solver* s, int32_t order) {
  if (order < WC_ORDER_FIRST || order > WC_ORDER_FIFTH)
    return fail(WC_EINVAL, "order must be 0 (first), 1 (second) or 2 (fifth)");
  s->order = order;
  TRY(set_device(s));
  TRY(ws_reset(s));  // the rhythm of the solves changes with the order
  return WC_OK;
}

int wc_solver_with_manning(wc_solver* s, double n) {
  s->manning = n;
  return WC_OK;
}

int wc_solver_with_sponge(wc_solver* s, const wc_sponge_t* sp) {
  TRY(set_device(s));
  if ((int)s->sponges.size() >= MAX_SPONGES)
    return fail(WC_EINVAL, "at most 16 sponges per solver");
  HostSponge hs{};
  hs.width = sp->width;
  hs.strength = sp->strength;
  hs.level = sp->level;
  hs.west = sp->west != 0;
  hs.east = sp->east != 0;
  hs.south = sp->south != 0;
  hs.north = sp->north != 0;
  // Sponge::rate (forcing.rs:181-206) only ever sees distances below max(nx, ny).
  const uint64_t span = (uint64_t)std::max(s->geo.nx, s->geo.ny);
  hs.len = (long long)std::min<uint64_t>(hs.width, span);
  hs.offset = (long long)s->sponge_rates.size();
  const double w = (double)hs.width;
  for (long long d = 0; d < hs.len; ++d) {
    const double f = (w - (double)d) / w;
    s->sponge_rates.push_back(hs.strength * (f * f));
  }
  s->sponges.push_back(hs);
  HIP_OK(hipStreamSynchronize(s->stream));
  if (s->sponge_rate_d) {
    auto it = std::find(s->owned.begin(), s->owned.end(), (void*)s->sponge_rate_d);
    if (it != s->owned.end()) s->owned.erase(it);
    HIP_OK(hipFree(s->sponge_rate_d));
    s->sponge_rate_d = nullptr;
  }
  TRY(dalloc(s, &s->sponge_rate_d, s->sponge_rates.size()));
  HIP_OK(hipMemcpyAsync(s->sponge_rate_d, s->sponge_rates.data(),
                        s->sponge_rates.size() * sizeof(double), hipMemcpyHostToDevice,
                        s->stream));
  TRY(resize_params(s));
  HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

int wc_solver_with_wavemaker(wc_solver* s, const wc_wavemaker_t* m) {
  TRY(set_device(s));
  s->maker = *m;
  s->has_maker = true;
  return rebuild_drive(s);
}

int wc_solver_with_dispersion(wc_solver* s, const wc_dispersion_t* d) {
  TRY(set_device(s));
  if (d->reserved != 0) return fail(WC_EINVAL, "wc_dispersion_t.reserved must be 0");
  s->disp = *d;
  s->dispersive = true;
  s->eddy_on = false;  // Breaking::viscosity comes with wc_solver_with_eddy_viscosity
  TRY(alloc_dispersion(s));
  // Work::new (dispersion.rs:237-266): fresh scratch and statistics.
  for (double* p : {s->c, s->dgx, s->dgy, s->bx, s->by, s->wx, s->wy, s->rx, s->ry, s->px, s->py,
                    s->ax, s->ay})
    HIP_OK(hipMemsetAsync(p, 0, s->cells * sizeof(double), s->stream));
  HIP_OK(hipMemsetAsync(s->sc, 0, sizeof(CgScalars), s->stream));
  TRY(ws_reset(s));
  HIP_OK(hipStreamSynchronize(s->stream));
  s->prev_it[0] = s->prev_it[1] = 0;
  return rebuild_drive(s);
}

int wc_solver_with_eddy_viscosity(wc_solver* s, const wc_eddy_t* e) {
  if (!s) return fail(WC_EINVAL, "null solver");
  if (!s->dispersive)
    return fail(WC_EINVAL, "the eddy viscosity is part of the dispersion: set it after "
                           "wc_solver_with_dispersion");
  TRY(set_device(s));
  if (!e) {
    s->eddy_on = false;
    return WC_OK;
  }
  if (!s->e_rise) {  // EddyWork::new (eddy.rs:102-113): zeros
    for (double** p : {&s->e_rise, &s->e_mem, &s->e_nu, &s->e_hu, &s->e_hv})
      TRY(dalloc(s, p, s->cells));
  }
  s->eddy = EddyP{e->initial, e->settled, e->duration, e->mixing, e->linger};
  s->eddy_on = true;
  HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

int wc_solver_with_periodic_y(wc_solver* s) {
  TRY(set_device(s));
  s->periodic = true;
  s->geo.periodic = 1;
  // wrap_ghosts_y on the bed (grid.rs:79-87)
  const int w = s->geo.w, ny = s->geo.ny;
  for (int i = 0; i < w; ++i)
    for (int k = 0; k < GH; ++k) {
      s->bed[(size_t)(GH - 1 - k) * w + i] = s->bed[(size_t)(GH + ny - 1 - k) * w + i];
      s->bed[(size_t)(GH + ny + k) * w + i] = s->bed[(size_t)(GH + k) * w + i];
    }
  HIP_OK(hipMemcpyAsync(s->b, s->bed.data(), s->cells * sizeof(double), hipMemcpyHostToDevice,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  return check_periodic(s);
}

int wc_solver_with_shallow_cells(wc_solver* s, const uint8_t* cells, uint64_t n) {
  TRY(set_device(s));
  if (n != (uint64_t)s->geo.nx * (uint64_t)s->geo.ny)
    return fail(WC_EINVAL, "one entry per interior cell");
  if (!s->held) TRY(dalloc(s, &s->held, n));
  std::vector<uint8_t> v(n);
  for (uint64_t q = 0; q < n; ++q) v[q] = cells[q] ? 1 : 0;
  HIP_OK(hipMemcpyAsync(s->held, v.data(), n, hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  s->has_held = true;
  return WC_OK;
}

int wc_solver_bed(const wc_solver* s, double* out) {
  memcpy(out, s->bed.data(), s->cells * sizeof(double));
  return WC_OK;
}

int wc_warm_start_weights(int32_t order, int32_t window, double* out) {
  if (order < 1 || order > WS_MAX_ORDER || window < 0 || window > WS_MAX_WINDOW || !out)
    return fail(WC_EINVAL, "order must be in [1, 5], window in [0, 12], out not null");
  ws_coefficients(order, std::max<int>(window, order), out);
  return WC_OK;
}

int wc_solver_set_option(wc_solver* s, const char* key, int64_t value) {
  if (!key) return fail(WC_EINVAL, "null option name");
  if (strcmp(key, "pcg_kernels") == 0) {
    if (value < 1 || value > 5) return fail(WC_EINVAL, "pcg_kernels must be 1, 2, 3, 4 or 5");
    s->pcg_kernels = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "pcg_tiles_per_block") == 0) {
    if (value < 1 || value > 64) return fail(WC_EINVAL, "pcg_tiles_per_block must be in [1, 64]");
    TRY(set_device(s));
    HIP_OK(hipStreamSynchronize(s->stream));
    s->pcg_walk = (int)value;
    HIP_OK(hipMemsetAsync(s->lanes, 0, 8 * 512 * sizeof(double), s->stream));
    return WC_OK;
  }
  if (strcmp(key, "pcg_lanes") == 0) {
    if (value != 256 && value != 512) return fail(WC_EINVAL, "pcg_lanes must be 256 or 512");
    TRY(set_device(s));
    HIP_OK(hipStreamSynchronize(s->stream));
    s->pcg_lanes = (int)value;
    // lanes past a launch's blocks must read zero
    HIP_OK(hipMemsetAsync(s->lanes, 0, 8 * 512 * sizeof(double), s->stream));
    return WC_OK;
  }
  if (strcmp(key, "capture") == 0) {
    if (value != 0 && value != 1) return fail(WC_EINVAL, "capture must be 0 or 1");
    if (value && !s->cap_h) {
      TRY(set_device(s));
      TRY(dalloc(s, &s->cap_h, s->cells));
      for (int slot = 0; slot < wc_solver::CAP_SLOTS; ++slot) {
        TRY(dalloc(s, &s->cap_wx[slot], s->cells));
        TRY(dalloc(s, &s->cap_wy[slot], s->cells));
        TRY(dalloc(s, &s->cap_r[slot], s->cells));
      }
    }
    s->capture = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "precond_ratio") == 0) {
    if (value < 1 || value > 100) return fail(WC_EINVAL, "precond_ratio must be in [1, 100]");
    s->precond_ratio = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "precond_stall") == 0) {
    if (value < 0 || value > 1000) return fail(WC_EINVAL, "precond_stall must be in [0, 1000]");
    s->precond_stall = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "flux_kernel") == 0) {
    if (value < 1 || value > 3) return fail(WC_EINVAL, "flux_kernel must be 1, 2 or 3");
    s->flux_kernel = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "flux_rows") == 0) {
    if (value < 0 || value > (1 << 30))
      return fail(WC_EINVAL, "flux_rows must be 0 (tuned) or a band height of rows");
    s->flux_rows = (int)value;
    s->march_rows = 0;
    return WC_OK;
  }
  if (strcmp(key, "flux_reconstruction") == 0) {
    if (value != 32 && value != 64) return fail(WC_EINVAL, "flux_reconstruction must be 32 or 64");
    s->recon64 = value == 64 ? 1 : 0;
    s->march_rows = 0;
    return WC_OK;
  }
  if (strcmp(key, "pcg_march") == 0) {
    if (value != 0 && value != 1) return fail(WC_EINVAL, "pcg_march must be 0 or 1");
    s->pcg_march = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "pcg_batch") == 0) {
    if (value < 1 || value > 1000) return fail(WC_EINVAL, "pcg_batch must be in [1, 1000]");
    s->pcg_batch = (unsigned long long)value;
    return WC_OK;
  }
  if (strcmp(key, "warm_start") == 0) {
    if (value < 0 || value > WS_MAX_ORDER)
      return fail(WC_EINVAL, "warm_start must be in [0, 5] (0: the previous solution)");
    TRY(set_device(s));
    s->ws_order = (int)value;
    TRY(ws_reset(s));
    HIP_OK(hipStreamSynchronize(s->stream));
    return WC_OK;
  }
  if (strcmp(key, "fp64_warm_start") == 0) {
    if (value < 0 || value > WS_MAX_ORDER)
      return fail(WC_EINVAL, "fp64_warm_start must be in [0, 5] (0: the previous solution)");
    TRY(set_device(s));
    if (value > 0 && !s->ws_ring64) TRY(dalloc(s, &s->ws_ring64, WS_SLOTS * s->cells));
    s->ws_order64 = (int)value;
    TRY(ws_reset(s));
    HIP_OK(hipStreamSynchronize(s->stream));
    return WC_OK;
  }
  if (strcmp(key, "warm_start_guard") == 0) {
    if (value < 0 || value > 1001)
      return fail(WC_EINVAL, "warm_start_guard must be in [0, 1001] (per mille of the switch)");
    s->ws_guard = (int)value;
    return WC_OK;
  }
  if (strcmp(key, "warm_start_window") == 0) {
    if (value < 0 || value > WS_MAX_WINDOW)
      return fail(WC_EINVAL, "warm_start_window must be in [0, 12] (0: the order)");
    TRY(set_device(s));
    s->ws_window = (int)value;
    TRY(ws_reset(s));
    HIP_OK(hipStreamSynchronize(s->stream));
    return WC_OK;
  }
  return fail(WC_EINVAL, std::string("unknown option ") + key);
}

int wc_solver_flux_rows(const wc_solver* s, int32_t* rows, int32_t* tried_rows, double* tried_us,
                        int32_t capacity, int32_t* tried) {
  if (!s) return fail(WC_EINVAL, "null solver");
  if (capacity < 0) return fail(WC_EINVAL, "capacity must be >= 0");
  if (rows) *rows = s->march_rows;
  const int n = (int)s->march_tried.size();
  for (int q = 0; q < std::min(n, (int)capacity); ++q) {
    if (tried_rows) tried_rows[q] = s->march_tried[q].first;
    if (tried_us) tried_us[q] = s->march_tried[q].second;
  }
  if (tried) *tried = n;
  return WC_OK;
}

int wc_solver_set_precision(wc_solver* s, int32_t precision) {
  TRY(set_device(s));
  if (precision != WC_PRECISION_FP64 && precision != WC_PRECISION_FP32)
    return fail(WC_EINVAL, "precision must be WC_PRECISION_FP64 or WC_PRECISION_FP32");
  if (precision == s->precision) return WC_OK;
  s->march_rows = 0;  // the other precision's flux kernel has its own occupancy
  if (precision == WC_PRECISION_FP32) {
    TRY(alloc_fp32(s));
    s->precision = precision;
    TRY(split_state(s));
    TRY(ws_reset(s));
  } else {
    TRY(join_state(s));
    s->precision = precision;
  }
  HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

int wc_state_set(wc_solver* s, const double* h, const double* hu, const double* hv,
                 const double* breaking, double time) {
  TRY(set_device(s));
  const size_t bytes = s->cells * sizeof(double);
  HIP_OK(hipMemcpyAsync(s->h, h, bytes, hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemcpyAsync(s->hu, hu, bytes, hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipMemcpyAsync(s->hv, hv, bytes, hipMemcpyHostToDevice, s->stream));
  if (breaking)
    HIP_OK(hipMemcpyAsync(s->br, breaking, bytes, hipMemcpyHostToDevice, s->stream));
  else
    HIP_OK(hipMemsetAsync(s->br, 0, bytes, s->stream));
  if (s->precision == WC_PRECISION_FP32) TRY(split_state(s));
  TRY(ws_reset(s));
  HIP_OK(hipStreamSynchronize(s->stream));
  s->time = time;
  return WC_OK;
}

int wc_state_get(wc_solver* s, double* h, double* hu, double* hv, double* breaking,
                 double* time) {
  TRY(set_device(s));
  const size_t bytes = s->cells * sizeof(double);
  TRY(join_state(s));
  if (h) HIP_OK(hipMemcpyAsync(h, s->h, bytes, hipMemcpyDeviceToHost, s->stream));
  if (hu) HIP_OK(hipMemcpyAsync(hu, s->hu, bytes, hipMemcpyDeviceToHost, s->stream));
  if (hv) HIP_OK(hipMemcpyAsync(hv, s->hv, bytes, hipMemcpyDeviceToHost, s->stream));
  if (breaking) HIP_OK(hipMemcpyAsync(breaking, s->br, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  if (time) *time = s->time;
  return WC_OK;
}

int wc_state_set_time(wc_solver* s, double time) {
  s->time = time;
  return WC_OK;
}

int wc_stable_dt(wc_solver* s, double* dt) {
  TRY(set_device(s));
  return s->precision == WC_PRECISION_FP32 ? stable_dt32(s, dt) : stable_dt(s, dt);
}

int wc_step(wc_solver* s, double dt) {
  TRY(set_device(s));
  TRY(s->precision == WC_PRECISION_FP32 ? do_step32(s, dt) : do_step(s, dt));
  if (!s->prof) HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

// ---- the learned preconditioner ----------------------------------------------------------------

int wc_solver_set_preconditioner(wc_solver* s, int32_t kind, const float* params, uint64_t n) {
  if (!s) return fail(WC_EINVAL, "null solver");
  if (kind == WC_PRECOND_JACOBI) {
    s->precond = kind;
    return WC_OK;
  }
  if (kind != WC_PRECOND_NETWORK) return fail(WC_EINVAL, "unknown preconditioner");
  if (!params || n != wcu_param_count())
    return fail(WC_EINVAL, "the network needs wcu_param_count() parameters");
  TRY(set_device(s));
  if (!s->net) {
    int rc = wcu_net_new((uint64_t)s->geo.nx, (uint64_t)s->geo.ny, &s->net);
    if (rc != WC_OK) return rc;
    TRY(dalloc(s, &s->bgx, s->cells));
    TRY(dalloc(s, &s->bgy, s->cells));
    TRY(dalloc(s, &s->history, 1024));
  }
  float* d = nullptr;
  HIP_OK(hipMalloc(&d, n * sizeof(float)));
  HIP_OK(hipMemcpyAsync(d, params, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
  int rc = wcu_net_set_params(s->net, d, s->stream);
  HIP_OK(hipStreamSynchronize(s->stream));
  HIP_OK(hipFree(d));
  if (rc != WC_OK) return rc;
  TRY(launch(s, F_FILL, k_bed_features, over((long long)s->cells), dim3(NT), s->geo,
             (const double*)s->b, WC_BED_FEATURE_SCALE, s->bgx, s->bgy));
  HIP_OK(hipStreamSynchronize(s->stream));
  s->precond = kind;
  return WC_OK;
}

int wc_solver_precond_fallbacks(wc_solver* s, uint64_t* out) {
  if (!s || !out) return fail(WC_EINVAL, "null argument");
  TRY(set_device(s));
  HIP_OK(hipMemcpy(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost));
  *out = s->sc_host->fallbacks;
  return WC_OK;
}

int wc_solver_capture(wc_solver* s, int32_t slot, float* h, float* rx, float* ry, float* ex,
                      float* ey, int32_t* valid) {
  if (!s || !s->cap_h) return fail(WC_EINVAL, "capture is not on");
  if (slot < 0 || slot >= wc_solver::CAP_SLOTS) return fail(WC_EINVAL, "capture slot out of range");
  static constexpr unsigned long long cap_it[wc_solver::CAP_SLOTS] = {0, 2, 5, 10};
  TRY(set_device(s));
  const size_t n = s->cells;
  std::vector<float> w0x(n), w0y(n), wx(n), wy(n), r(2 * n);
  HIP_OK(hipStreamSynchronize(s->stream));
  if (h) HIP_OK(hipMemcpy(h, s->cap_h, n * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(r.data(), s->cap_r[slot], 2 * n * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(w0x.data(), s->cap_wx[slot], n * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(w0y.data(), s->cap_wy[slot], n * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(wx.data(), s->f_wx, n * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(wy.data(), s->f_wy, n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < n; ++k) {
    if (rx) rx[k] = r[2 * k];
    if (ry) ry[k] = r[2 * k + 1];
    if (ex) ex[k] = wx[k] - w0x[k];
    if (ey) ey[k] = wy[k] - w0y[k];
  }
  if (valid) *valid = s->cap_last_it > cap_it[slot] ? 1 : 0;
  return WC_OK;
}

int wc_solver_residual_history(wc_solver* s, double* out, uint64_t max, uint64_t* n) {
  if (!s || !n) return fail(WC_EINVAL, "null argument");
  *n = 0;
  if (!s->history) return WC_OK;
  TRY(set_device(s));
  HIP_OK(hipStreamSynchronize(s->stream));
  HIP_OK(hipMemcpy(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost));
  const uint64_t count = std::min<uint64_t>(std::min<uint64_t>(s->sc_host->it + 1, 1024), max);
  if (out && count) HIP_OK(hipMemcpy(out, s->history, count * sizeof(double), hipMemcpyDeviceToHost));
  *n = count;
  return WC_OK;
}

// ---- the run's statistics and the driver step (crates/wavesim/src/lib.rs:849-937,1099-1434) --

struct wc_stats {
  wc_solver* s = nullptr;
  double from = 0.0, period = 0.0, next_wave = 0.0, seconds = 0.0, omega = 0.0;
  double moments[6] = {0, 0, 0, 0, 0, 0};
  uint64_t waves = 0;
  double *breaking_seconds = nullptr, *eta_max = nullptr, *eta_min = nullptr,
         *wave_max = nullptr, *wave_min = nullptr, *harm = nullptr, *wave_sum = nullptr,
         *wave_first = nullptr, *wave_last = nullptr;
  float* elevation = nullptr;
  std::vector<void*> owned;
};

namespace {

__global__ void k_fill_value(double* a, long long n, double v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

}  // namespace

int wc_stats_new(wc_solver* s, double from, double period, const float* elevation,
                 wc_stats** out) {
  if (!s || !out) return fail(WC_EINVAL, "null argument");
  if (!(period > 0.0)) return fail(WC_EINVAL, "the wave period must be positive");
  *out = nullptr;
  TRY(set_device(s));
  auto st = std::make_unique<wc_stats>();
  st->s = s;
  s->stats.push_back(st.get());
  struct Undo {  // a failed construction leaves nothing behind
    std::unique_ptr<wc_stats>& st;
    ~Undo() {
      if (st) wc::detach(st.get());
    }
  } undo{st};
  st->from = from;
  st->period = period;
  st->next_wave = from + period;
  st->omega = 2.0 * M_PI / period;
  const long long n = (long long)s->geo.nx * s->geo.ny;
  auto alloc = [&](double** p, long long count, double v) -> int {
    HIP_OK(hipMalloc(p, std::max<long long>(count, 1) * sizeof(double)));
    st->owned.push_back(*p);
    hipLaunchKernelGGL(k_fill_value, over(count), dim3(NT), 0, s->stream, *p, count, v);
    HIP_OK(hipGetLastError());
    return WC_OK;
  };
  TRY(alloc(&st->breaking_seconds, n, 0.0));
  TRY(alloc(&st->eta_max, n, -INFINITY));
  TRY(alloc(&st->eta_min, n, INFINITY));
  TRY(alloc(&st->wave_max, n, -INFINITY));
  TRY(alloc(&st->wave_min, n, INFINITY));
  TRY(alloc(&st->harm, 3 * n, 0.0));
  TRY(alloc(&st->wave_sum, n, 0.0));
  TRY(alloc(&st->wave_first, n, 0.0));
  TRY(alloc(&st->wave_last, n, 0.0));
  if (elevation) {
    HIP_OK(hipMalloc(&st->elevation, n * sizeof(float)));
    st->owned.push_back(st->elevation);
    HIP_OK(hipMemcpyAsync(st->elevation, elevation, n * sizeof(float), hipMemcpyHostToDevice,
                          s->stream));
  }
  HIP_OK(hipStreamSynchronize(s->stream));
  *out = st.release();
  return WC_OK;
}

void wc_stats_free(wc_stats* st) {
  if (!st) return;
  wc::detach(st);
  delete st;
}

// Stage::record (lib.rs:1276-1300) after a step of dt that has just ended.
int wc_stats_record(wc_stats* st, double dt) {
  if (!st) return fail(WC_EINVAL, "null statistics");
  if (!st->s) return fail(WC_EINVAL, "the statistics' solver has been freed");
  wc_solver* s = st->s;
  const double inside = std::min(dt, s->time - st->from);
  if (inside <= 0.0) return WC_OK;
  TRY(set_device(s));
  TRY(join_state(s));
  RecordArgs a{};
  a.g = s->geo;
  a.h = s->h;
  a.b = s->b;
  a.br = breaking_of(s);
  a.elevation = st->elevation;
  if (s->time > st->next_wave + 1e-9) {  // Stats::close_wave (lib.rs:918-932)
    a.close_wave = 1;
    st->waves += 1;
    a.first_wave = st->waves == 1 ? 1 : 0;
    st->next_wave += st->period;
  }
  // FirstHarmonic::weights (lib.rs:1118-1128)
  // (omega t).sin_cos(): glibc's sincos, as the wave-maker's angle
  double sn, cs;
  sincos(st->omega * s->time, &sn, &cs);
  const double v[6] = {1.0, cs, sn, cs * cs, cs * sn, sn * sn};
  for (int m = 0; m < 6; ++m) st->moments[m] += inside * v[m];
  a.inside = inside;
  a.w0 = inside;
  a.w1 = inside * cs;
  a.w2 = inside * sn;
  a.breaking_seconds = st->breaking_seconds;
  a.eta_max = st->eta_max;
  a.eta_min = st->eta_min;
  a.wave_max = st->wave_max;
  a.wave_min = st->wave_min;
  a.harm = st->harm;
  a.wave_sum = st->wave_sum;
  a.wave_first = st->wave_first;
  a.wave_last = st->wave_last;
  TRY(launch(s, F_REDUCE, k_record, over((long long)s->geo.nx * s->geo.ny), dim3(NT), a));
  st->seconds += inside;
  if (!s->prof) HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

int wc_stats_read(wc_stats* st, double* seconds, uint64_t* waves, double* breaking_seconds,
                  double* eta_max, double* eta_min, double* wave_mean, double* wave_first,
                  double* wave_last) {
  if (!st) return fail(WC_EINVAL, "null statistics");
  if (!st->s) return fail(WC_EINVAL, "the statistics' solver has been freed");
  wc_solver* s = st->s;
  TRY(set_device(s));
  const size_t n = (size_t)s->geo.nx * s->geo.ny;
  HIP_OK(hipStreamSynchronize(s->stream));
  auto get = [&](double* dst, const double* src) -> int {
    if (dst) HIP_OK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
    return WC_OK;
  };
  if (seconds) *seconds = st->seconds;
  if (waves) *waves = st->waves;
  TRY(get(breaking_seconds, st->breaking_seconds));
  TRY(get(eta_max, st->eta_max));
  TRY(get(eta_min, st->eta_min));
  TRY(get(wave_first, st->wave_first));
  TRY(get(wave_last, st->wave_last));
  if (wave_mean) {
    TRY(get(wave_mean, st->wave_sum));
    for (size_t k = 0; k < n; ++k)
      wave_mean[k] = st->waves ? wave_mean[k] / (double)st->waves : 0.0;
  }
  return WC_OK;
}

// FirstHarmonic::amplitudes (lib.rs:1136-1160): (a, -b) per cell, *ok = 0 where the reference
// returns None.
int wc_stats_harmonic(wc_stats* st, int32_t* ok, double* a_out, double* b_out) {
  if (!st || !ok) return fail(WC_EINVAL, "null argument");
  if (!st->s) return fail(WC_EINVAL, "the statistics' solver has been freed");
  wc_solver* s = st->s;
  TRY(set_device(s));
  const size_t n = (size_t)s->geo.nx * s->geo.ny;
  const double s0 = st->moments[0], c = st->moments[1], sn = st->moments[2];
  const double cc = st->moments[3], cs = st->moments[4], ss = st->moments[5];
  const double m[3][3] = {{s0, c, sn}, {c, cc, cs}, {sn, cs, ss}};
  auto cof = [&](int r, int k) {
    const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    return m[r1][k1] * m[r2][k2] - m[r1][k2] * m[r2][k1];
  };
  double det = 0.0;
  for (int k = 0; k < 3; ++k) det += m[0][k] * cof(0, k);
  if (std::isnan(det) || det <= 1e-3 * s0 * s0 * s0 / 4.0) {
    *ok = 0;
    return WC_OK;
  }
  *ok = 1;
  std::vector<double> e(3 * n);
  HIP_OK(hipStreamSynchronize(s->stream));
  HIP_OK(hipMemcpy(e.data(), st->harm, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  auto inv = [&](int r, int k) { return cof(k, r) / det; };
  for (size_t q = 0; q < n; ++q) {
    auto row = [&](int r) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += inv(r, k) * e[3 * q + k];
      return t;
    };
    if (a_out) a_out[q] = row(1);
    if (b_out) b_out[q] = -row(2);
  }
  return WC_OK;
}

// The end of a run (lib.rs:1451-1453): the wave in progress is closed if it is due.
int wc_stats_finish(wc_stats* st) {
  if (!st) return fail(WC_EINVAL, "null statistics");
  if (!st->s) return fail(WC_EINVAL, "the statistics' solver has been freed");
  wc_solver* s = st->s;
  if (!(s->time >= st->next_wave - 1e-9)) return WC_OK;
  TRY(set_device(s));
  RecordArgs a{};
  a.g = s->geo;
  st->waves += 1;
  a.first_wave = st->waves == 1 ? 1 : 0;
  st->next_wave += st->period;
  a.wave_max = st->wave_max;
  a.wave_min = st->wave_min;
  a.wave_sum = st->wave_sum;
  a.wave_first = st->wave_first;
  a.wave_last = st->wave_last;
  TRY(launch(s, F_REDUCE, k_close_wave, over((long long)s->geo.nx * s->geo.ny), dim3(NT), a));
  if (!s->prof) HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

// Stage::frame_if_due's frame (lib.rs:1348-1380) from the solver's state, through a pinned slot
// of the run writer; the disk write is the writer thread's.
int wc_run_frame_from(wc_run_writer* w, wc_solver* s, double tide, double* rise,
                      int32_t* status) {
  if (!w || !s || !rise || !status) return fail(WC_EINVAL, "null argument");
  TRY(set_device(s));
  TRY(join_state(s));
  wc::run::DeviceFrame f;
  TRY(wc::run::begin_device_frame(w, s->device, s->geo.nx, s->geo.ny, &f));
  const size_t n = f.cells;
  const int rc = [&]() -> int {
    HIP_OK(hipMemsetAsync(f.scalars, 0, 2 * sizeof(unsigned long long), s->stream));
    FrameArgs a{};
    a.g = s->geo;
    a.h = s->h;
    a.b = s->b;
    a.br = breaking_of(s);
    a.elevation = f.elevation;
    a.tide = tide;
    a.eta = f.device + (size_t)f.eta * n;
    a.breaking = f.device + (size_t)f.breaking * n;
    a.rise_bits = f.scalars;
    a.nonfinite = (unsigned int*)(f.scalars + 1);
    TRY(launch(s, F_FRAME, k_frame, over((long long)n), dim3(NT), a));
    HIP_OK(hipMemcpyAsync(f.pinned, f.device, 2 * n * sizeof(float), hipMemcpyDeviceToHost,
                          s->stream));
    HIP_OK(hipMemcpyAsync(f.scalars_host, f.scalars, 2 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return WC_OK;
  }();
  if (rc) {
    wc::run::abandon_frame(w, f);
    return rc;
  }
  double r;
  memcpy(&r, &f.scalars_host[0], sizeof r);
  const bool nonfinite = (unsigned int)f.scalars_host[1] != 0;
  *rise = r;
  if (!std::isfinite(r) || r > 100.0 || nonfinite) {
    wc::run::abandon_frame(w, f);
    *status = WC_STAGE_UNSTABLE;
    return WC_OK;
  }
  *status = WC_STAGE_OK;
  return wc::run::commit_frame(w, f, s->time);
}

// Stage::finish's statistics (lib.rs:1436-1509) without the break line.
int wc_run_stats_from(wc_run_writer* w, wc_stats* st) {
  if (!w || !st) return fail(WC_EINVAL, "null argument");
  if (!st->s) return fail(WC_EINVAL, "the statistics' solver has been freed");
  wc_solver* s = st->s;
  TRY(wc_stats_finish(st));
  if (!(st->seconds > 0.0)) return WC_OK;
  const long long nx = s->geo.nx, ny = s->geo.ny;
  const size_t n = (size_t)nx * ny;
  std::vector<double> bs(n), hi(n), lo(n), mean(n), first(n), last(n), ha(n), hb(n);
  double seconds = 0.0;
  uint64_t waves = 0;
  TRY(wc_stats_read(st, &seconds, &waves, bs.data(), hi.data(), lo.data(), mean.data(),
                    first.data(), last.data()));
  int32_t ok = 0;
  TRY(wc_stats_harmonic(st, &ok, ha.data(), hb.data()));
  auto f32s = [&](const std::vector<double>& v) {
    std::vector<float> o(v.size());
    for (size_t q = 0; q < v.size(); ++q) o[q] = (float)v[q];
    return o;
  };
  // Stage::breaking_fraction (lib.rs:1337-1341)
  const double t = std::max(seconds, std::numeric_limits<double>::min());
  std::vector<float> fraction(n);
  for (size_t q = 0; q < n; ++q) fraction[q] = (float)(bs[q] / t);
  std::vector<float> high = f32s(hi), low = f32s(lo), height = f32s(mean), re = f32s(ha),
                     im = f32s(hb);
  std::vector<const char*> names = {"break_fraction", "eta_max", "eta_min"};
  std::vector<const float*> blocks = {fraction.data(), high.data(), low.data()};
  if (waves > 0) {
    wc::run::note(w, "waves_recorded", wc::json::Value::uint(waves));
    // Stage::unsettled (lib.rs:1315-1335): the median over the middle half of the width of
    // |last / first - 1| per period, over the cells whose first wave exceeded 0.3 m
    if (waves >= 2) {
      std::vector<double> change;
      for (long long j = ny / 4; j < ny - ny / 4; ++j)
        for (long long i = 0; i < nx; ++i) {
          const size_t m = (size_t)(j * nx + i);
          if (first[m] > 0.3)
            change.push_back(std::fabs(last[m] / first[m] - 1.0) / (double)(waves - 1));
        }
      if (!change.empty()) {
        std::sort(change.begin(), change.end(), [](double a, double b) {
          return wc::json::total_order(a) < wc::json::total_order(b);
        });
        wc::run::note(w, "unsettled_per_period", wc::json::Value::number(change[change.size() / 2]));
      }
    }
    names.push_back("wave_height");
    blocks.push_back(height.data());
    if (ok) {
      names.push_back("harmonic_re");
      names.push_back("harmonic_im");
      blocks.push_back(re.data());
      blocks.push_back(im.data());
    }
  }
  return wc_run_write_stats(w, names.data(), blocks.data(), names.size(), nullptr, st->from,
                            s->time);
}

// StepWatch (lib.rs:861-896)
int wc_watch_init(wc_watch_t* w, double reference) {
  if (!w) return fail(WC_EINVAL, "null watch");
  w->reference = std::isfinite(reference) ? reference : 0.0;
  w->steps = 0;
  w->seconds = 0.0;
  return WC_OK;
}

int wc_watch_collapsed(wc_watch_t* w, double stable, double* ratio) {
  if (!w || !ratio) return fail(WC_EINVAL, "null argument");
  *ratio = 0.0;
  w->steps += 1;
  w->seconds += stable;
  if (w->steps < WC_WATCH_STEPS) return WC_OK;
  const double mean = w->seconds / (double)w->steps;
  w->steps = 0;
  w->seconds = 0.0;
  if (mean * WC_COLLAPSE < w->reference) *ratio = w->reference / mean;
  return WC_OK;
}

// Stage::next_dt + advance + record (lib.rs:1385-1434), without frames: the next step as long
// as is stable and ending no later than `until`.
int wc_stage_step(wc_solver* s, wc_stats* st, wc_watch_t* w, double until, double* dt,
                  int32_t* status, double* ratio) {
  if (!s || !dt || !status) return fail(WC_EINVAL, "null argument");
  if (st && st->s != s) return fail(WC_EINVAL, "the statistics belong to another solver");
  *status = WC_STAGE_OK;
  if (ratio) *ratio = 0.0;
  double stable = 0.0;
  TRY(wc_stable_dt(s, &stable));
  *dt = stable;
  if (std::isnan(stable) || stable <= 0.0) {
    *status = WC_STAGE_UNSTABLE;
    return WC_OK;
  }
  if (std::isinf(stable)) {
    *status = WC_STAGE_NO_WATER;
    return WC_OK;
  }
  if (w) {
    double r = 0.0;
    TRY(wc_watch_collapsed(w, stable, &r));
    if (r > 0.0) {
      *status = WC_STAGE_COLLAPSED;
      if (ratio) *ratio = r;
      return WC_OK;
    }
  }
  *dt = std::min(stable, until - s->time);
  TRY(wc_step(s, *dt));
  if (st) TRY(wc_stats_record(st, *dt));
  return WC_OK;
}

// ---- relaxation zones (forcing.rs:227-284) ---------------------------------------------------

struct wc_relaxation {
  wc_solver* s = nullptr;
  std::vector<uint64_t> ci, cj;       // padded (i, j), the reference's order
  std::vector<double> rates;          // the distinct rates, in order of first appearance
  int* cell = nullptr;                // device: padded index of each cell
  int* rate = nullptr;                // device: index into rates
  double* factor = nullptr;           // device: exp(-rate dt) per distinct rate
  double* target = nullptr;           // device: staging for host targets
  double* factor_host = nullptr;      // pinned
  hipEvent_t factor_ev = nullptr;     // the last copy out of factor_host
  bool factor_pending = false;
};

int wc_relaxation_new(wc_solver* s, const wc_sponge_t* zones, uint64_t n_zones,
                      wc_relaxation** out) {
  if (!s || !out || (n_zones > 0 && !zones)) return fail(WC_EINVAL, "null argument");
  *out = nullptr;
  TRY(set_device(s));
  const Geo& g = s->geo;
  auto r = std::make_unique<wc_relaxation>();
  r->s = s;
  std::vector<int> cell, rate;
  for (uint64_t j = GH; j < (uint64_t)g.ny + GH; ++j)
    for (uint64_t i = GH; i < (uint64_t)g.nx + GH; ++i) {
      // Sponge::rate (forcing.rs:181-206) of each zone, folded with max from 0.0
      double best = 0.0;
      for (uint64_t z = 0; z < n_zones; ++z) {
        const wc_sponge_t& sp = zones[z];
        const uint64_t a = i - GH, b = j - GH;
        const double w = (double)sp.width;
        auto ramp = [&](uint64_t d) {
          if (d < sp.width) {
            const double f = (w - (double)d) / w;
            return f * f;
          }
          return 0.0;
        };
        double m = 0.0;
        if (sp.west) m = std::max(m, ramp(a));
        if (sp.east) m = std::max(m, ramp((uint64_t)g.nx - 1 - a));
        if (sp.south) m = std::max(m, ramp(b));
        if (sp.north) m = std::max(m, ramp((uint64_t)g.ny - 1 - b));
        best = std::max(best, sp.strength * m);
      }
      if (!(best > 0.0)) continue;
      size_t q = 0;
      while (q < r->rates.size() && r->rates[q] != best) ++q;
      if (q == r->rates.size()) r->rates.push_back(best);
      r->ci.push_back(i);
      r->cj.push_back(j);
      cell.push_back((int)(j * g.w + i));
      rate.push_back((int)q);
    }
  const size_t n = cell.size();
  s->relaxations.push_back(r.get());
  struct Undo {  // a failed construction leaves nothing behind
    std::unique_ptr<wc_relaxation>& r;
    ~Undo() {
      if (r) wc::detach(r.get());
    }
  } undo{r};
  HIP_OK(hipMalloc(&r->cell, std::max<size_t>(n, 1) * sizeof(int)));
  HIP_OK(hipMalloc(&r->rate, std::max<size_t>(n, 1) * sizeof(int)));
  HIP_OK(hipMalloc(&r->target, std::max<size_t>(3 * n, 1) * sizeof(double)));
  HIP_OK(hipMalloc(&r->factor, std::max<size_t>(r->rates.size(), 1) * sizeof(double)));
  HIP_OK(hipHostMalloc(&r->factor_host, std::max<size_t>(r->rates.size(), 1) * sizeof(double),
                       hipHostMallocDefault));
  HIP_OK(hipEventCreateWithFlags(&r->factor_ev, hipEventDisableTiming));
  if (n) {
    HIP_OK(hipMemcpy(r->cell, cell.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(r->rate, rate.data(), n * sizeof(int), hipMemcpyHostToDevice));
  }
  *out = r.release();
  return WC_OK;
}

void wc_relaxation_free(wc_relaxation* r) {
  if (!r) return;
  wc::detach(r);
  delete r;
}

}  // extern "C"

namespace wc {

void detach(wc_stats* st) {
  wc_solver* s = st->s;
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->stream);
  for (void* p : st->owned) (void)hipFree(p);
  st->owned.clear();
  s->stats.erase(std::remove(s->stats.begin(), s->stats.end(), st), s->stats.end());
  st->s = nullptr;
}

void detach(wc_relaxation* r) {
  wc_solver* s = r->s;
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (r->factor_ev) {
    (void)hipEventSynchronize(r->factor_ev);
    (void)hipEventDestroy(r->factor_ev);
    r->factor_ev = nullptr;
  }
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(r->cell);
  (void)hipFree(r->rate);
  (void)hipFree(r->target);
  (void)hipFree(r->factor);
  (void)hipHostFree(r->factor_host);
  r->cell = r->rate = nullptr;
  r->target = r->factor = r->factor_host = nullptr;
  s->relaxations.erase(std::remove(s->relaxations.begin(), s->relaxations.end(), r),
                       s->relaxations.end());
  r->s = nullptr;
}

}  // namespace wc

extern "C" {

uint64_t wc_relaxation_len(const wc_relaxation* r) { return r ? (uint64_t)r->ci.size() : 0; }

int wc_relaxation_cells(const wc_relaxation* r, uint64_t* i, uint64_t* j) {
  if (!r) return fail(WC_EINVAL, "null relaxation");
  for (size_t m = 0; m < r->ci.size(); ++m) {
    if (i) i[m] = r->ci[m];
    if (j) j[m] = r->cj[m];
  }
  return WC_OK;
}

namespace {

int relax_apply(wc_relaxation* r, double dt, const double* targets_device) {
  wc_solver* s = r->s;
  const size_t n = r->ci.size();
  if (n == 0) return WC_OK;
  // exp(-rate dt) with the reference's expression, once per distinct rate; the pinned buffer
  // is rewritten only after its previous copy has left it.
  if (r->factor_pending) HIP_OK(hipEventSynchronize(r->factor_ev));
  for (size_t q = 0; q < r->rates.size(); ++q) r->factor_host[q] = exp(-r->rates[q] * dt);
  HIP_OK(hipMemcpyAsync(r->factor, r->factor_host, r->rates.size() * sizeof(double),
                        hipMemcpyHostToDevice, s->stream));
  HIP_OK(hipEventRecord(r->factor_ev, s->stream));
  r->factor_pending = true;
  const RelaxArgs a{(long long)n, r->cell, r->rate, r->factor, targets_device, s->b};
  const dim3 grid((unsigned)((n + NT - 1) / NT));
  if (s->precision == WC_PRECISION_FP32)
    TRY(launch(s, F_RELAX, k_relax32, grid, dim3(NT), a, s->f_sh, s->f_slh, s->f_shu, s->f_slhu,
               s->f_shv, s->f_slhv));
  else
    TRY(launch(s, F_RELAX, k_relax64, grid, dim3(NT), a, s->h, s->hu, s->hv));
  return WC_OK;
}

}  // namespace

int wc_relaxation_apply(wc_relaxation* r, double dt, const double* targets) {
  if (!r || (!targets && !r->ci.empty())) return fail(WC_EINVAL, "null argument");
  if (!r->s) return fail(WC_EINVAL, "the relaxation's solver has been freed");
  TRY(set_device(r->s));
  const size_t n = r->ci.size();
  if (n)
    HIP_OK(hipMemcpyAsync(r->target, targets, 3 * n * sizeof(double), hipMemcpyHostToDevice,
                          r->s->stream));
  TRY(relax_apply(r, dt, r->target));
  if (!r->s->prof) HIP_OK(hipStreamSynchronize(r->s->stream));
  return WC_OK;
}

int wc_relaxation_apply_device(wc_relaxation* r, double dt, const double* targets) {
  if (!r || (!targets && !r->ci.empty())) return fail(WC_EINVAL, "null argument");
  if (!r->s) return fail(WC_EINVAL, "the relaxation's solver has been freed");
  TRY(set_device(r->s));
  TRY(relax_apply(r, dt, targets));
  if (!r->s->prof) HIP_OK(hipStreamSynchronize(r->s->stream));
  return WC_OK;
}

}  // extern "C"

// ---- the coarse-to-fine sampler of nested runs (crates/wavesim/src/lib.rs:1548-1616) --------

struct wc_sampler {
  int device = 0;
  long long n = 0;       // zone cells
  int cnx = 0, cny = 0;  // the coarse grid the corners index
  int* idx = nullptr;    // device: four padded coarse indices per zone cell
  double* w = nullptr;   // device: their bilinear weights
  double* smp[2] = {nullptr, nullptr};  // device: (eta, u, v) per zone cell, then and now
  double t[2] = {NAN, NAN};
  double* targets = nullptr;  // device: the interpolated targets of the last apply
  std::vector<void*> owned;
};

extern "C" {

// Sampler::new (lib.rs:1559-1590): the fine zone cell (i, j) sits at x = (x0 + i - GHOST + 0.5)
// fine cells of the bed file; its coarse neighbours are clamped to one ghost layer.
int wc_sampler_new(const wc_relaxation* zone, uint64_t fine_x0, double fine_dx, double fine_dy,
                   uint64_t coarse_nx, uint64_t coarse_ny, double coarse_dx, double coarse_dy,
                   uint64_t crop, wc_sampler** out) {
  if (!zone || !out) return fail(WC_EINVAL, "null argument");
  if (!zone->s) return fail(WC_EINVAL, "the relaxation's solver has been freed");
  if (coarse_nx < 3 || coarse_ny < 3 || !(coarse_dx > 0.0) || !(coarse_dy > 0.0))
    return fail(WC_EINVAL, "the coarse grid needs at least 3 x 3 cells of positive size");
  *out = nullptr;
  const size_t n = zone->ci.size();
  const long long cw = (long long)coarse_nx + 2 * GH;
  std::vector<int> idx(4 * std::max<size_t>(n, 1), 0);
  std::vector<double> w(4 * std::max<size_t>(n, 1), 0.0);
  for (size_t m = 0; m < n; ++m) {
    const double x = (double)(fine_x0 + zone->ci[m] - GH) * fine_dx + 0.5 * fine_dx;
    const double y = (double)(zone->cj[m] - GH) * fine_dy + 0.5 * fine_dy;
    const double cx = x / coarse_dx - (double)crop - 0.5;
    const double cy = y / coarse_dy - 0.5;
    const double ix = std::floor(cx), iy = std::floor(cy);
    const double wx = cx - ix, wy = cy - iy;
    auto at = [&](double di, double dj) {
      const long long pi = std::clamp((long long)(ix + di) + GH, (long long)GH - 1,
                                      (long long)coarse_nx + GH);
      const long long pj = std::clamp((long long)(iy + dj) + GH, (long long)GH - 1,
                                      (long long)coarse_ny + GH);
      return (int)(pj * cw + pi);
    };
    idx[4 * m] = at(0.0, 0.0);
    idx[4 * m + 1] = at(1.0, 0.0);
    idx[4 * m + 2] = at(0.0, 1.0);
    idx[4 * m + 3] = at(1.0, 1.0);
    w[4 * m] = (1.0 - wx) * (1.0 - wy);
    w[4 * m + 1] = wx * (1.0 - wy);
    w[4 * m + 2] = (1.0 - wx) * wy;
    w[4 * m + 3] = wx * wy;
  }
  auto sm = std::make_unique<wc_sampler>();
  sm->device = zone->s->device;
  sm->n = (long long)n;
  sm->cnx = (int)coarse_nx;
  sm->cny = (int)coarse_ny;
  HIP_OK(hipSetDevice(sm->device));
  struct Undo {
    std::unique_ptr<wc_sampler>& sm;
    ~Undo() {
      if (sm)
        for (void* p : sm->owned) (void)hipFree(p);
    }
  } undo{sm};
  auto get = [&](void** p, size_t bytes) -> int {
    HIP_OK(hipMalloc(p, std::max<size_t>(bytes, 8)));
    sm->owned.push_back(*p);
    return WC_OK;
  };
  TRY(get((void**)&sm->idx, idx.size() * sizeof(int)));
  TRY(get((void**)&sm->w, w.size() * sizeof(double)));
  for (int q = 0; q < 2; ++q) TRY(get((void**)&sm->smp[q], 3 * n * sizeof(double)));
  TRY(get((void**)&sm->targets, 3 * n * sizeof(double)));
  HIP_OK(hipMemcpy(sm->idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(sm->w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  *out = sm.release();
  return WC_OK;
}

void wc_sampler_free(wc_sampler* sm) {
  if (!sm) return;
  (void)hipSetDevice(sm->device);
  for (void* p : sm->owned) (void)hipFree(p);
  delete sm;
}

// Sampler::sample (lib.rs:1592-1614) of the coarse solver's current state into slot 0 (the
// state before the coarse step that brackets the fine one) or 1 (after it), at its time.
int wc_sampler_sample(wc_sampler* sm, wc_solver* coarse, int32_t slot) {
  if (!sm || !coarse) return fail(WC_EINVAL, "null argument");
  if (slot != 0 && slot != 1) return fail(WC_EINVAL, "slot must be 0 or 1");
  if (coarse->geo.nx != sm->cnx || coarse->geo.ny != sm->cny)
    return fail(WC_EINVAL, "the coarse solver's grid is not the sampler's");
  if (coarse->device != sm->device) return fail(WC_EINVAL, "the sampler is on another device");
  TRY(set_device(coarse));
  TRY(join_state(coarse));
  Geo g = coarse->geo;
  g.periodic = 0;  // fill_boundaries(grid, false): walls
  if (sm->n > 0)
    TRY(launch(coarse, F_RELAX, k_sample, over(sm->n), dim3(NT), (long long)sm->n,
               (const int*)sm->idx, (const double*)sm->w, g, (const double*)coarse->h,
               (const double*)coarse->hu, (const double*)coarse->hv,
               (const double*)coarse->b, sm->smp[slot]));
  sm->t[slot] = coarse->time;
  HIP_OK(hipStreamSynchronize(coarse->stream));
  return WC_OK;
}

// The nested run's relaxation after a fine step of dt ending at `end` (lib.rs:1866-1897): the
// targets interpolated between the two samples at w = (end - t0) / (t1 - t0) clamped to [0, 1],
// then Relaxation::apply on the fine solver.
int wc_sampler_apply(wc_sampler* sm, wc_relaxation* zone, double end, double dt) {
  if (!sm || !zone) return fail(WC_EINVAL, "null argument");
  if (!zone->s) return fail(WC_EINVAL, "the relaxation's solver has been freed");
  if ((long long)zone->ci.size() != sm->n)
    return fail(WC_EINVAL, "the relaxation is not the sampler's zone");
  if (!(std::isfinite(sm->t[0]) && std::isfinite(sm->t[1])))
    return fail(WC_EINVAL, "both samples must be taken before the targets are interpolated");
  wc_solver* s = zone->s;
  TRY(set_device(s));
  const double w = std::clamp((end - sm->t[0]) / (sm->t[1] - sm->t[0]), 0.0, 1.0);
  if (sm->n > 0)
    TRY(launch(s, F_RELAX, k_lerp_targets, over(sm->n), dim3(NT), (long long)sm->n,
               (const double*)sm->smp[0], (const double*)sm->smp[1], w, sm->targets));
  TRY(relax_apply(zone, dt, sm->targets));
  if (!s->prof) HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

// slot 0 / 1: a sample, 2: the last targets; 3 n doubles (eta, u, v) per zone cell, and the
// sample's time.
int wc_sampler_read(const wc_sampler* sm, int32_t slot, double* out, double* time) {
  if (!sm) return fail(WC_EINVAL, "null sampler");
  if (slot < 0 || slot > 2) return fail(WC_EINVAL, "slot must be 0, 1 or 2");
  HIP_OK(hipSetDevice(sm->device));
  HIP_OK(hipDeviceSynchronize());
  const double* src = slot == 2 ? sm->targets : sm->smp[slot];
  if (out && sm->n)
    HIP_OK(hipMemcpy(out, src, 3 * sm->n * sizeof(double), hipMemcpyDeviceToHost));
  if (time) *time = slot == 2 ? NAN : sm->t[slot];
  return WC_OK;
}

int wc_run_steps(wc_solver* s, uint64_t steps, uint64_t* done_steps, double* last_dt) {
  TRY(set_device(s));
  uint64_t n = 0;
  double dt = 0.0;
  const bool f32 = s->precision == WC_PRECISION_FP32;
  for (; n < steps; ++n) {
    TRY(f32 ? stable_dt32(s, &dt) : stable_dt(s, &dt));
    if (!(isfinite(dt) && dt > 0.0)) break;
    TRY(f32 ? do_step32(s, dt) : do_step(s, dt));
  }
  HIP_OK(hipStreamSynchronize(s->stream));
  if (done_steps) *done_steps = n;
  if (last_dt) *last_dt = dt;
  return WC_OK;
}

int wc_limiting_cell(wc_solver* s, int32_t* found, wc_limiting_t* out) {
  TRY(set_device(s));
  TRY(join_state(s));
  const Geo& g = s->geo;
  const long long n = (long long)g.nx * g.ny;
  HIP_OK(hipMemsetAsync(s->red64, 0, sizeof(unsigned long long), s->stream));
  HIP_OK(hipMemsetAsync(s->red32 + 1, 0xff, 2 * sizeof(unsigned int), s->stream));
  TRY(launch(s, F_REDUCE, k_limit_scan, over(n), dim3(NT), g, (const double*)s->h,
             (const double*)s->hu, (const double*)s->hv, s->red64, s->red32 + 1));
  HIP_OK(hipMemcpyAsync(s->red64_host, s->red64, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(s->red32_host, s->red32, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  unsigned int cell = s->red32_host[1];
  if (cell == 0xffffffffu) {
    if (s->red64_host[0] == 0ull) {
      *found = 0;
      return WC_OK;
    }
    TRY(launch(s, F_REDUCE, k_limit_first, over(n), dim3(NT), g, (const double*)s->h,
               (const double*)s->hu, (const double*)s->hv,
               (const unsigned long long*)s->red64, s->red32 + 2));
    HIP_OK(hipMemcpyAsync(s->red32_host, s->red32, 4 * sizeof(unsigned int),
                          hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    cell = s->red32_host[2];
  }
  const size_t i = GH + cell % (unsigned)g.nx, j = GH + cell / (unsigned)g.nx;
  const size_t k = j * (size_t)g.w + i;
  double v[3];
  HIP_OK(hipMemcpyAsync(&v[0], s->h + k, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(&v[1], s->hu + k, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(&v[2], s->hv + k, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  const double u = wc_velocity(v[0], v[1]);
  const double vv = wc_velocity(v[0], v[2]);
  *found = 1;
  out->i = i - GH;
  out->j = j - GH;
  out->depth = v[0];
  out->speed = wc_hypot(u, vv);
  return WC_OK;
}

int wc_breaking_indicator(wc_solver* s, double* out) {
  TRY(set_device(s));
  TRY(join_state(s));
  const Geo& g = s->geo;
  const size_t n = (size_t)g.nx * g.ny;
  TRY(launch(s, F_REDUCE, k_breaking, over((long long)n), dim3(NT), g, (const double*)s->h,
             (const double*)s->b, breaking_of(s), s->lh));
  HIP_OK(hipMemcpyAsync(out, s->lh, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

int wc_dispersion_stats(wc_solver* s, int32_t* has, wc_dispersion_stats_t* out) {
  TRY(set_device(s));
  *has = s->dispersive ? 1 : 0;
  if (!s->dispersive) return WC_OK;
  HIP_OK(hipMemcpyAsync(s->sc_host, s->sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  out->solves = s->sc_host->solves;
  out->iterations = s->sc_host->iterations;
  out->unconverged = s->sc_host->unconverged;
  return WC_OK;
}

static int state_sums(wc_solver* s) {
  TRY(join_state(s));
  const Geo& g = s->geo;
  HIP_OK(hipMemsetAsync(s->red64 + 1, 0, sizeof(unsigned long long), s->stream));
  HIP_OK(hipMemsetAsync(s->red64 + 2, 0xff, sizeof(unsigned long long), s->stream));
  TRY(launch(s, F_REDUCE, k_state_sums, tiles(s), dim3(NT), g, (const double*)s->h,
             (const double*)s->hu, (const double*)s->hv, s->part, s->red64 + 1, s->red64 + 2));
  TRY(launch(s, F_REDUCE, k_sum_tiles, dim3(1), dim3(NT), (const double*)s->part,
             g.ntx * g.nty, s->redf));
  HIP_OK(hipMemcpyAsync(s->red64_host, s->red64, 4 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s->stream));
  HIP_OK(hipMemcpyAsync(s->redf_host, s->redf, sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIP_OK(hipStreamSynchronize(s->stream));
  return WC_OK;
}

int wc_state_volume(wc_solver* s, double* out) {
  TRY(set_device(s));
  TRY(state_sums(s));
  *out = s->redf_host[0] * s->geo.dx * s->geo.dy;
  return WC_OK;
}

int wc_state_max_abs_momentum(wc_solver* s, double* out) {
  TRY(set_device(s));
  TRY(state_sums(s));
  const unsigned long long bits = s->red64_host[1];
  memcpy(out, &bits, sizeof(double));
  return WC_OK;
}

int wc_state_min_depth(wc_solver* s, double* out) {
  TRY(set_device(s));
  TRY(state_sums(s));
  unsigned long long key = s->red64_host[2];
  if (key == 0xffffffffffffffffull) {
    *out = INFINITY;
    return WC_OK;
  }
  const unsigned long long bits = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
  memcpy(out, &bits, sizeof(double));
  return WC_OK;
}

int wc_linear_wave(double omega, double h, int32_t dispersive, wc_linear_wave_t* out) {
  if (!linear_wave(omega, h, dispersive != 0, out))
    return fail(WC_EINVAL, "a wave of this angular frequency is too short for the dispersive "
                           "model in this depth");
  return WC_OK;
}

double wc_manning_factor(double h, double speed, double dt, double n) {
  const double k = WC_G * n * n / pow(h, 4.0 / 3.0);
  return 1.0 / (1.0 + dt * k * speed);
}

int wc_check_shared_division(double y, const double* x, uint64_t n, double* shared,
                              double* plain) {
  if (!x || !shared || !plain) return fail(WC_EINVAL, "null buffer");
  if (n == 0) return WC_OK;
  if (n > (1ull << 26)) return fail(WC_EINVAL, "at most 2^26 values per check");
  double* d = nullptr;
  HIP_OK(hipMalloc(&d, 3 * n * sizeof(double)));
  auto run = [&]() -> int {
    HIP_OK(hipMemcpy(d, x, n * sizeof(double), hipMemcpyHostToDevice));
    const int nn = (int)n;
    hipLaunchKernelGGL(k_division_check, dim3((nn + NT - 1) / NT), dim3(NT), 0, 0, y, d, nn,
                       d + n, d + 2 * n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(shared, d + n, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(plain, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost));
    return WC_OK;
  };
  const int rc = run();
  (void)hipFree(d);
  return rc;
}

int wc_profile_enable(wc_solver* s, int32_t on) {
  TRY(set_device(s));
  TRY(prof_collect(s));
  s->prof = on == 1;
  s->prof_batch = on == 2;
  return WC_OK;
}

int wc_profile_reset(wc_solver* s) {
  TRY(set_device(s));
  TRY(prof_collect(s));
  for (int f = 0; f < F_COUNT; ++f) {
    s->prof_ms[f] = 0.0;
    s->prof_n[f] = 0;
  }
  return WC_OK;
}

int wc_profile_read(wc_solver* s, char* names, uint64_t names_len, double* total_ms,
                    uint64_t* launches, uint64_t max_kernels, uint64_t* n_kernels) {
  TRY(set_device(s));
  TRY(prof_collect(s));
  std::string all;
  for (int f = 0; f < F_COUNT; ++f) {
    if (f) all += ";";
    all += kFamilyNames[f];
    if ((uint64_t)f < max_kernels) {
      total_ms[f] = s->prof_ms[f];
      launches[f] = s->prof_n[f];
    }
  }
  if (names && names_len) {
    strncpy(names, all.c_str(), names_len - 1);
    names[names_len - 1] = '\0';
  }
  *n_kernels = F_COUNT;
  return WC_OK;
}

}  // extern "C"

// ================================================ many instances in one launch (wc_batch_*)
// Every phase of a step is one launch over all instances (kb_* kernels): each instance keeps
// its own arrays, dt and PCG scalars, and its virtual blocks run exactly the single-instance
// bodies, so an instance advances bit for bit as it would alone (FP32 form, one-kernel PCG).
// Per step the host reads the instances' rates once, evaluates their step parameters into one
// pinned buffer, and reads the PCG flags once per batch of iterations.

struct wc_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<wc_solver*> inst;
  std::vector<int> active;  // indices of the instances still stepping
  bool rk2 = true, sw = false, dispersive = false;
  int lanes = 256;  // the instances' one-kernel PCG lanes
  bool march = true;  // and its layout (pcg_march)
  bool recon64 = false;  // the instances' flux reconstruction in FP64
  bool march_flux = true;  // the instances' flux: marching strips (flux_kernel 3) or 64x4 tiles
  int flux_rows = 32;      // the marching flux's band height (tuned unless fixed by hand)
  bool flux_rows_fixed = false;
  std::vector<std::pair<int, double>> tune_rows;  // (band height, us per launch) as measured
  unsigned long long maxit = 0, pcg_batch = 4;
  int grid_nt = 2048, grid_pt = 512;  // persistent widths of the flux and the PCG launches
  int grid_st = 1 << 30;  // and of the other 256-thread launches: a block per virtual block
                          // (their blocks are short and few registers deep), unless set by hand
  bool tuned = false;
  std::vector<std::pair<int, double>> tune_nt, tune_pt;  // (width, us per launch) as measured

  // device: argument blocks of the active instances and the launch maps
  char* d_args = nullptr;
  size_t args_cap = 0;
  BFill* fill[2] = {};
  BSwitch* swa[2] = {};
  BFlux* flux[2] = {};
  Sgn32* sgn[2] = {};
  CgG* cg[2] = {};
  Stage32* stg[2] = {};
  BRate* rate = nullptr;
  BMap m_fill{}, m_sw{}, m_tiles{}, m_rate{}, m_cg{}, m_march{};
  long long* d_off = nullptr;
  double** d_dst = nullptr;
  double* d_stage = nullptr;  // [dt of each active instance | their step parameters]
  size_t stage_cap = 0;
  unsigned long long* d_red = nullptr;
  int* d_done = nullptr;
  unsigned long long* d_it = nullptr;
  // pinned
  double* h_stage = nullptr;
  unsigned long long* h_red = nullptr;
  int* h_done = nullptr;
  unsigned long long* h_it = nullptr;
  std::vector<long long> off;  // step-parameter offsets of the active instances
  std::vector<unsigned long long> prev_it[2];
  std::vector<void*> owned, owned_host;

  // HIP graphs of the fixed stretches of a step (use_graph): [0] stage 1 up to K[0] PCG
  // iterations, [1] the stage-1 update and stage 2 up to K[1] iterations, [2] the stage-2
  // update; rebuilt when the argument blocks move or a K changes
  bool use_graph = true;
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
  unsigned long long graph_k[2] = {~0ull, ~0ull};
  unsigned long long graph_builds = 0;

  // profiling (option "profile"): each PCG iteration launch between a pair of events, the
  // stretches issued directly instead of from the graphs
  bool prof = false;
  double prof_ms = 0.0;
  unsigned long long prof_n = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pending;
  std::vector<hipEvent_t> prof_pool;
};

namespace {

int bfail(const std::string& m) { return fail(WC_EINVAL, m); }

void drop_graphs(wc_batch* b) {
  for (hipGraphExec_t& g : b->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  b->graph_k[0] = b->graph_k[1] = ~0ull;
}

template <class T>
int balloc(wc_batch* b, T** p, size_t n) {
  void* v = nullptr;
  HIP_OK(hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T)));
  b->owned.push_back(v);
  *p = static_cast<T*>(v);
  return WC_OK;
}

template <class T>
int bhost(wc_batch* b, T** p, size_t n) {
  void* v = nullptr;
  HIP_OK(hipHostMalloc(&v, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
  b->owned_host.push_back(v);
  *p = static_cast<T*>(v);
  return WC_OK;
}

// The one-kernel PCG's blocks that have work: 64x16 tiles, or strip units when marching
// (mrch::segments), at most its lanes.
int pcg_blocks(const wc_solver* s) {
  const int L = s->pcg_lanes;
  if (s->pcg_march) {
    const int strips = (s->geo.nx + 63) / 64;
    const int segs = std::min(strips >= L ? 1 : L / strips, s->geo.ny);
    return std::min(strips * segs, L);
  }
  return tile_stride(s);
}

// (Re)builds every argument block and launch map for the active instances, in one upload.
int batch_build(wc_batch* b) {
  // largest instances first: their virtual blocks are the longest (a PCG block walks several
  // tiles of a large grid), so they start first and the short ones fill the tail of a launch
  std::stable_sort(b->active.begin(), b->active.end(), [&](int x, int y) {
    const Geo &gx = b->inst[x]->geo, &gy = b->inst[y]->geo;
    return (long long)gx.nx * gx.ny > (long long)gy.nx * gy.ny;
  });
  const int n = (int)b->active.size();
  std::vector<int> f_fill(n + 1, 0), f_sw(n + 1, 0), f_tiles(n + 1, 0), f_rate(n + 1, 0),
      f_cg(n + 1, 0), f_march(n + 1, 0);
  auto rows_of = [&](const Geo& g) { return std::max(1, std::min(b->flux_rows, g.ny)); };
  b->off.assign(n + 1, 0);
  for (int q = 0; q < n; ++q) {
    const wc_solver* s = b->inst[b->active[q]];
    const Geo& g = s->geo;
    f_fill[q + 1] = f_fill[q] + (int)((6LL * g.w + 6LL * g.ny + NT - 1) / NT);
    f_sw[q + 1] = f_sw[q] + (int)(((long long)g.w * g.rows + NT - 1) / NT);
    f_tiles[q + 1] = f_tiles[q] + g.ntx * g.nty;
    f_rate[q + 1] = f_rate[q] + (int)rate_blocks((long long)g.nx * g.ny).x;
    f_cg[q + 1] = f_cg[q] + pcg_blocks(s);
    f_march[q + 1] = f_march[q] + ((g.nx + MARCH_MW - 1) / MARCH_MW) *
                                      ((g.ny + rows_of(g) - 1) / rows_of(g));
    b->off[q + 1] = b->off[q] + (long long)s->prm_len;
  }
  // host image: the blocks, then the maps and the parameter routes
  std::vector<char> img;
  auto put = [&](const void* p, size_t bytes) {
    size_t at = (img.size() + 255) & ~(size_t)255;
    img.resize(at + bytes);
    memcpy(img.data() + at, p, bytes);
    return at;
  };
  std::vector<BFill> fill[2];
  std::vector<BSwitch> sw[2];
  std::vector<BFlux> flux[2];
  std::vector<Sgn32> sgn[2];
  std::vector<CgG> cg[2];
  std::vector<Stage32> stg[2];
  std::vector<BRate> rate;
  std::vector<double*> dst;
  for (int q = 0; q < n; ++q) {
    wc_solver* s = b->inst[b->active[q]];
    for (int st = 0; st < 2; ++st) {
      BFill f{};
      f.g = s->geo;
      f.h = st == 0 ? s->f_sh : s->f_s1h;
      f.hl = st == 0 ? s->f_slh : s->f_s1hl;
      f.hu = st == 0 ? s->f_shu : s->f_s1hu;
      f.hv = st == 0 ? s->f_shv : s->f_s1hv;
      f.hul = st == 0 ? s->f_slhu : s->f_s1hul;
      f.hvl = st == 0 ? s->f_slhv : s->f_s1hvl;
      fill[st].push_back(f);
      sw[st].push_back(switch_args32(s, st));
      flux[st].push_back(flux_args32(s, st));
      flux[st].back().rows = rows_of(s->geo);
      sgn[st].push_back(sgn32_args(s, st));
      cg[st].push_back(cg1_args(s, st));
      stg[st].push_back(stage_args32(s, 0.0, st));  // dt comes per step from d_stage
    }
    BRate r{};
    r.g = s->geo;
    r.h = s->f_sh;
    r.hl = s->f_slh;
    r.hu = s->f_shu;
    r.hul = s->f_slhu;
    r.hv = s->f_shv;
    r.hvl = s->f_slhv;
    r.out = b->d_red + q;
    rate.push_back(r);
    dst.push_back(s->prm);
  }
  size_t o_fill[2], o_sw[2], o_flux[2], o_sgn[2], o_cg[2], o_stg[2];
  for (int st = 0; st < 2; ++st) {
    o_fill[st] = put(fill[st].data(), n * sizeof(BFill));
    o_sw[st] = put(sw[st].data(), n * sizeof(BSwitch));
    o_flux[st] = put(flux[st].data(), n * sizeof(BFlux));
    o_sgn[st] = put(sgn[st].data(), n * sizeof(Sgn32));
    o_cg[st] = put(cg[st].data(), n * sizeof(CgG));
    o_stg[st] = put(stg[st].data(), n * sizeof(Stage32));
  }
  const size_t o_rate = put(rate.data(), n * sizeof(BRate));
  const size_t o_ffill = put(f_fill.data(), (n + 1) * sizeof(int));
  const size_t o_fsw = put(f_sw.data(), (n + 1) * sizeof(int));
  const size_t o_ftiles = put(f_tiles.data(), (n + 1) * sizeof(int));
  const size_t o_frate = put(f_rate.data(), (n + 1) * sizeof(int));
  const size_t o_fcg = put(f_cg.data(), (n + 1) * sizeof(int));
  const size_t o_fmarch = put(f_march.data(), (n + 1) * sizeof(int));
  const size_t o_off = put(b->off.data(), (n + 1) * sizeof(long long));
  const size_t o_dst = put(dst.data(), n * sizeof(double*));
  if (img.size() > b->args_cap) {
    if (b->d_args) {
      HIP_OK(hipStreamSynchronize(b->stream));
      b->owned.erase(std::find(b->owned.begin(), b->owned.end(), (void*)b->d_args));
      HIP_OK(hipFree(b->d_args));
      b->d_args = nullptr;
    }
    TRY(balloc(b, &b->d_args, img.size()));
    b->args_cap = img.size();
  }
  HIP_OK(hipMemcpyAsync(b->d_args, img.data(), img.size(), hipMemcpyHostToDevice, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  char* d = b->d_args;
  for (int st = 0; st < 2; ++st) {
    b->fill[st] = reinterpret_cast<BFill*>(d + o_fill[st]);
    b->swa[st] = reinterpret_cast<BSwitch*>(d + o_sw[st]);
    b->flux[st] = reinterpret_cast<BFlux*>(d + o_flux[st]);
    b->sgn[st] = reinterpret_cast<Sgn32*>(d + o_sgn[st]);
    b->cg[st] = reinterpret_cast<CgG*>(d + o_cg[st]);
    b->stg[st] = reinterpret_cast<Stage32*>(d + o_stg[st]);
  }
  b->rate = reinterpret_cast<BRate*>(d + o_rate);
  b->m_fill = BMap{reinterpret_cast<const int*>(d + o_ffill), n, f_fill[n]};
  b->m_sw = BMap{reinterpret_cast<const int*>(d + o_fsw), n, f_sw[n]};
  b->m_tiles = BMap{reinterpret_cast<const int*>(d + o_ftiles), n, f_tiles[n]};
  b->m_rate = BMap{reinterpret_cast<const int*>(d + o_frate), n, f_rate[n]};
  b->m_cg = BMap{reinterpret_cast<const int*>(d + o_fcg), n, f_cg[n]};
  b->m_march = BMap{reinterpret_cast<const int*>(d + o_fmarch), n, f_march[n]};
  b->d_off = reinterpret_cast<long long*>(d + o_off);
  b->d_dst = reinterpret_cast<double**>(d + o_dst);
  drop_graphs(b);  // the captured launches point at the old blocks
  const size_t need = (size_t)n + (size_t)b->off[n];
  if (need > b->stage_cap) {
    // the pinned and device staging are only ever reused after a synchronisation
    TRY(bhost(b, &b->h_stage, need));
    TRY(balloc(b, &b->d_stage, need));
    b->stage_cap = need;
  }
  for (int st = 0; st < 2; ++st) {
    std::vector<unsigned long long> keep(n, 0);
    for (int q = 0; q < n; ++q) keep[q] = b->inst[b->active[q]]->prev_it[st];
    b->prev_it[st] = keep;
  }
  return WC_OK;
}

dim3 bgrid(int total, int width) { return dim3((unsigned)std::max(1, std::min(total, width))); }

// The launches of one stretch of a step, in order (captured into a graph or issued directly).
void enq_fill(wc_batch* b, int st) {
  hipLaunchKernelGGL(kb_fill32, bgrid(b->m_fill.total, b->grid_st), dim3(NT), 0, b->stream,
                     b->fill[st], b->m_fill);
}

// The rates of every instance: the marching flux (one wave per strip and band, a virtual block
// each) or the 64x4 tile flux at its persistent width.
void enq_flux(wc_batch* b, int st) {
  if (b->march_flux)
    hipLaunchKernelGGL(b->recon64 ? kb_flux_march32r : kb_flux_march32,
                       bgrid(b->m_march.total, 1 << 30), dim3(TX), 0, b->stream, b->flux[st],
                       b->m_march);
  else
    hipLaunchKernelGGL(b->recon64 ? kb_flux32r : kb_flux32, bgrid(b->m_tiles.total, b->grid_nt),
                       dim3(NT), 0, b->stream, b->flux[st], b->m_tiles);
}

void enq_stage_head(wc_batch* b, int st) {  // the switch, the rates, the SGN set-up
  if (b->sw)
    hipLaunchKernelGGL(kb_switch32, bgrid(b->m_sw.total, b->grid_st), dim3(NT), 0, b->stream,
                       b->swa[st], b->m_sw);
  enq_flux(b, st);
  if (b->dispersive)
    hipLaunchKernelGGL(kb_sgn_setup32, bgrid(b->m_tiles.total, b->grid_st), dim3(NT), 0,
                       b->stream, b->sgn[st], b->m_tiles);
}

// A PCG iteration launch, timed between two events when the batch profiles.
template <class Launch>
void timed_iter(wc_batch* b, Launch fn) {
  hipEvent_t e[2] = {nullptr, nullptr};
  if (b->prof)
    for (hipEvent_t& x : e) {
      if (!b->prof_pool.empty()) {
        x = b->prof_pool.back();
        b->prof_pool.pop_back();
      } else if (hipEventCreate(&x) != hipSuccess) {
        x = nullptr;
      }
    }
  if (e[0] && e[1]) (void)hipEventRecord(e[0], b->stream);
  fn();
  if (e[0] && e[1]) {
    (void)hipEventRecord(e[1], b->stream);
    b->prof_pending.push_back({e[0], e[1]});
  } else {
    for (hipEvent_t x : e)
      if (x) b->prof_pool.push_back(x);
  }
}

int batch_prof_collect(wc_batch* b) {
  if (b->prof_pending.empty()) return WC_OK;
  HIP_OK(hipStreamSynchronize(b->stream));
  for (auto& pe : b->prof_pending) {
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, pe.first, pe.second));
    b->prof_ms += ms;
    b->prof_n += 1;
    b->prof_pool.push_back(pe.first);
    b->prof_pool.push_back(pe.second);
  }
  b->prof_pending.clear();
  return WC_OK;
}

template <int L>
void enq_pcg_head_l(wc_batch* b, int st, unsigned long long k) {
  const dim3 gp = bgrid(b->m_cg.total, b->grid_pt);
  if (b->march) {
    hipLaunchKernelGGL(kbm_cg1_init<L>, gp, dim3(1024), 0, b->stream, b->cg[st], b->m_cg);
    hipLaunchKernelGGL(kbm_cg1_start<L>, gp, dim3(1024), 0, b->stream, b->cg[st], b->m_cg);
    for (unsigned long long it = 1; it <= k; ++it)
      timed_iter(b, [&] {
        hipLaunchKernelGGL(kbm_cg1_iter<L>, gp, dim3(1024), 0, b->stream, b->cg[st], it,
                           b->m_cg);
      });
    return;
  }
  hipLaunchKernelGGL(kb_cg1_init<L>, gp, dim3(1024), 0, b->stream, b->cg[st], b->m_cg);
  hipLaunchKernelGGL(kb_cg1_start<L>, gp, dim3(1024), 0, b->stream, b->cg[st], b->m_cg);
  for (unsigned long long it = 1; it <= k; ++it)
    timed_iter(b, [&] {
      hipLaunchKernelGGL(kb_cg1_iter<L>, gp, dim3(1024), 0, b->stream, b->cg[st], it, b->m_cg);
    });
}

void enq_pcg_head(wc_batch* b, int st, unsigned long long k) {  // init, start, iterations 1..k
  if (b->lanes == 512)
    enq_pcg_head_l<512>(b, st, k);
  else
    enq_pcg_head_l<256>(b, st, k);
}

void enq_iter_untimed(wc_batch* b, int st, unsigned long long it) {
  const dim3 gp = bgrid(b->m_cg.total, b->grid_pt);
  if (b->march) {
    if (b->lanes == 512)
      hipLaunchKernelGGL(kbm_cg1_iter<512>, gp, dim3(1024), 0, b->stream, b->cg[st], it, b->m_cg);
    else
      hipLaunchKernelGGL(kbm_cg1_iter<256>, gp, dim3(1024), 0, b->stream, b->cg[st], it, b->m_cg);
    return;
  }
  if (b->lanes == 512)
    hipLaunchKernelGGL(kb_cg1_iter<512>, gp, dim3(1024), 0, b->stream, b->cg[st], it, b->m_cg);
  else
    hipLaunchKernelGGL(kb_cg1_iter<256>, gp, dim3(1024), 0, b->stream, b->cg[st], it, b->m_cg);
}

void enq_iter(wc_batch* b, int st, unsigned long long it) {
  timed_iter(b, [&] { enq_iter_untimed(b, st, it); });
}

void enq_check(wc_batch* b, int st, unsigned long long it, int stats) {
  const int n = (int)b->active.size();
  if (b->lanes == 512)
    hipLaunchKernelGGL(kb_cg1_check<512>, dim3(n), dim3(64), 0, b->stream, b->cg[st], it, stats,
                       b->d_done, b->d_it);
  else
    hipLaunchKernelGGL(kb_cg1_check<256>, dim3(n), dim3(64), 0, b->stream, b->cg[st], it, stats,
                       b->d_done, b->d_it);
}

void enq_update(wc_batch* b, int st, int mode) {
  hipLaunchKernelGGL(kb_stage32, bgrid(b->m_tiles.total, b->grid_st), dim3(NT), 0, b->stream,
                     b->stg[st], (const double*)b->d_stage, mode, b->m_tiles);
}

// The rest of stage st's solves after iterations 1..k: first up to the largest count any
// instance needed last time, then batches of pcg_batch iterations, one look at all flags per
// batch; then the solves are counted. An instance that has converged leaves the launches after
// reading its flag, so queueing past its count changes none of its numbers.
int batch_pcg_tail(wc_batch* b, int st, unsigned long long k) {
  const int n = (int)b->active.size();
  unsigned long long it = k + 1;
  // the largest count of the instances' last solves at this stage, leaving out solves that ran
  // out of iterations (one instance at max_iterations would otherwise queue that many unchecked
  // launches for all), and at most pcg_batch past the smallest
  unsigned long long last = 0, least = ~0ull;
  for (int q = 0; q < n; ++q) {
    const unsigned long long v = b->prev_it[st][q];
    least = std::min(least, v);
    if (v < b->maxit) last = std::max(last, v);
  }
  if (least != ~0ull) last = std::min(last, least + b->pcg_batch);
  unsigned long long batch = last > k ? last - k : b->pcg_batch;
  bool look = k > 0 && last <= k;  // nothing to queue before the first look
  while (true) {
    if (!look) {
      const unsigned long long stop = std::min(it + batch, b->maxit + 1);
      for (; it < stop; ++it) enq_iter(b, st, it);
      batch = b->pcg_batch;
    }
    look = false;
    enq_check(b, st, it, 0);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(b->h_done, b->d_done, n * sizeof(int), hipMemcpyDeviceToHost,
                          b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    bool all = true;
    for (int q = 0; q < n; ++q) all = all && b->h_done[q];
    if (all || it > b->maxit) break;
  }
  enq_check(b, st, it, 1);
  hipLaunchKernelGGL(kb_cg1_flush, bgrid(b->m_tiles.total, b->grid_st), dim3(NT), 0, b->stream,
                     b->cg[st], b->m_tiles);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(b->h_it, b->d_it, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  for (int q = 0; q < n; ++q) {
    b->prev_it[st][q] = b->h_it[q];
    b->inst[b->active[q]]->prev_it[st] = b->h_it[q];
  }
  return WC_OK;
}

// The iterations a stage's graph queues: a multiple of 4 below every instance's last count,
// so that a graph is rebuilt only when the counts drift.
unsigned long long graph_iterations(const wc_batch* b, int st) {
  unsigned long long first = ~0ull;
  for (unsigned long long v : b->prev_it[st]) first = std::min(first, v);
  if (first == ~0ull || first < 2) return 0;
  return ((first - 1) / 4) * 4;
}

// Stretch `seg` of a step (see wc_batch::graph), issued directly.
void enq_segment(wc_batch* b, int seg, const unsigned long long* k) {
  if (seg == 0) {
    enq_fill(b, 0);
    enq_stage_head(b, 0);
    if (b->dispersive) enq_pcg_head(b, 0, k[0]);
  } else if (seg == 1) {
    enq_update(b, 0, b->rk2 ? MODE_S1 : MODE_EULER);
    if (!b->rk2) return;
    enq_fill(b, 1);
    enq_stage_head(b, 1);
    if (b->dispersive) enq_pcg_head(b, 1, k[1]);
  } else {
    enq_update(b, 1, MODE_FINAL);
  }
}

int run_segment(wc_batch* b, int seg, const unsigned long long* k) {
  if (!b->use_graph || b->prof) {
    enq_segment(b, seg, k);
    HIP_OK(hipGetLastError());
    return WC_OK;
  }
  const bool stale = !b->graph[seg] || (seg == 0 && b->graph_k[0] != k[0]) ||
                     (seg == 1 && b->graph_k[1] != k[1]);
  if (stale) {
    if (b->graph[seg]) HIP_OK(hipGraphExecDestroy(b->graph[seg]));
    b->graph[seg] = nullptr;
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enq_segment(b, seg, k);
    const hipError_t launched = hipGetLastError();
    HIP_OK(hipStreamEndCapture(b->stream, &g));
    HIP_OK(launched);
    const hipError_t inst = hipGraphInstantiate(&b->graph[seg], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(inst);
    if (seg < 2) b->graph_k[seg] = k[seg];
    b->graph_builds += 1;
  }
  HIP_OK(hipGraphLaunch(b->graph[seg], b->stream));
  return WC_OK;
}

// One step of every active instance: three stretches with the PCG tails between them.
int batch_step(wc_batch* b) {
  const unsigned long long k[2] = {b->dispersive ? graph_iterations(b, 0) : 0,
                                   b->dispersive ? graph_iterations(b, 1) : 0};
  TRY(run_segment(b, 0, k));
  if (b->dispersive) TRY(batch_pcg_tail(b, 0, k[0]));
  TRY(run_segment(b, 1, k));
  if (b->rk2) {
    if (b->dispersive) TRY(batch_pcg_tail(b, 1, k[1]));
    TRY(run_segment(b, 2, k));
  }
  return WC_OK;
}

// The autotuner, each choice the fastest of a few candidates timed on launches that change no
// state: the marching flux's band height, the tile shape of its waves (58 x rows cells; timed
// on the flux launch, which only writes rates), or with the 64x4 tile flux the persistent width
// of its launch; and the persistent width of the 1024-thread PCG launches (timed on
// kb_cg1_start, whose outputs the next solve's start rewrites). No choice changes a bit.
int batch_tune(wc_batch* b) {
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  auto time = [&](auto launch_fn) -> double {
    launch_fn();  // warm
    HIP_OK(hipEventRecord(e0, b->stream));
    for (int r = 0; r < 3; ++r) launch_fn();
    HIP_OK(hipEventRecord(e1, b->stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    return 1e3 * ms / 3.0;
  };
  b->tune_nt.clear();
  b->tune_pt.clear();
  int best_nt = b->grid_nt, best_pt = b->grid_pt;
  double t_nt = 1e30, t_pt = 1e30;
  if (b->march_flux && !b->flux_rows_fixed) {
    b->tune_rows.clear();
    int best_rows = b->flux_rows;
    double t_rows = 1e30;
    for (int r : {8, 16, 24, 32, 48, 64, 96, 128}) {
      b->flux_rows = r;
      TRY(batch_build(b));  // the map of strips and bands for this height
      const double us = time([&] { enq_flux(b, 0); });
      HIP_OK(hipGetLastError());
      b->tune_rows.push_back({r, us});
      if (us < t_rows) t_rows = us, best_rows = r;
    }
    b->flux_rows = best_rows;
    TRY(batch_build(b));
  }
  for (int w : {512, 1024, 2048, 4096, 8192, 1 << 30}) {
    if (b->march_flux) break;  // the tile flux's width only
    const int width = std::min(w, b->m_tiles.total);
    if (!b->tune_nt.empty() && b->tune_nt.back().first == width) continue;
    double us = 0.0;
    const int rc = [&]() -> int {
      us = time([&] {
        hipLaunchKernelGGL(b->recon64 ? kb_flux32r : kb_flux32, bgrid(b->m_tiles.total, width),
                           dim3(NT), 0, b->stream, b->flux[0], b->m_tiles);
      });
      return WC_OK;
    }();
    if (rc) break;
    HIP_OK(hipGetLastError());
    b->tune_nt.push_back({width, us});
    if (us < t_nt) t_nt = us, best_nt = width;
  }
  if (b->dispersive) {
    for (int w : {256, 512, 1024, 2048, 1 << 30}) {
      const int width = std::min(w, b->m_cg.total);
      if (!b->tune_pt.empty() && b->tune_pt.back().first == width) continue;
      double us = 0.0;
      const int rc = [&]() -> int {
        us = time([&] {
          if (b->march && b->lanes == 512)
            hipLaunchKernelGGL(kbm_cg1_start<512>, bgrid(b->m_cg.total, width), dim3(1024), 0,
                               b->stream, b->cg[0], b->m_cg);
          else if (b->march)
            hipLaunchKernelGGL(kbm_cg1_start<256>, bgrid(b->m_cg.total, width), dim3(1024), 0,
                               b->stream, b->cg[0], b->m_cg);
          else if (b->lanes == 512)
            hipLaunchKernelGGL(kb_cg1_start<512>, bgrid(b->m_cg.total, width), dim3(1024), 0,
                               b->stream, b->cg[0], b->m_cg);
          else
            hipLaunchKernelGGL(kb_cg1_start<256>, bgrid(b->m_cg.total, width), dim3(1024), 0,
                               b->stream, b->cg[0], b->m_cg);
        });
        return WC_OK;
      }();
      if (rc) break;
      HIP_OK(hipGetLastError());
      b->tune_pt.push_back({width, us});
      if (us < t_pt) t_pt = us, best_pt = width;
    }
  }
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  b->grid_nt = best_nt;
  b->grid_pt = best_pt;
  b->tuned = true;
  return WC_OK;
}

}  // namespace

extern "C" {

// Whether instance s can step in a batch whose settings are those of s0 (at creation, s0 is the
// first instance; before every run, each instance against its settings at creation, which the
// batch's launches were built from).
int batch_compatible(const wc_solver* s, const wc_solver* s0) {
  if (s->precision != WC_PRECISION_FP32)
    return bfail("a batch runs the FP32 form (wc_solver_set_precision)");
  if (s->precond != WC_PRECOND_JACOBI || s->capture || !(s->pcg_kernels == 0 || s->pcg_kernels == 4))
    return bfail("a batch runs the one-kernel Jacobi PCG, without capture");
  if (s->device != s0->device) return bfail("a batch's instances share one device");
  if (s->pcg_lanes != s0->pcg_lanes || s->pcg_march != s0->pcg_march ||
      s->recon64 != s0->recon64 || (s->flux_kernel == 3) != (s0->flux_kernel == 3))
    return bfail("a batch's instances share pcg_lanes, pcg_march, flux_reconstruction and "
                 "the flux layout (flux_kernel 3 or not)");
  if ((s->order == WC_ORDER_FIRST) != (s0->order == WC_ORDER_FIRST) ||
      switch_on(s) != switch_on(s0) || s->dispersive != s0->dispersive)
    return bfail("a batch's instances share the time stepping, the switch and the dispersion");
  return WC_OK;
}

int wc_batch_new(wc_solver* const* solvers, uint64_t n, wc_batch** out) {
  if (!solvers || !out || n == 0) return fail(WC_EINVAL, "null argument");
  *out = nullptr;
  wc_solver* s0 = solvers[0];
  for (uint64_t q = 0; q < n; ++q) {
    wc_solver* s = solvers[q];
    if (!s) return bfail("null solver");
    TRY(batch_compatible(s, s0));
    for (uint64_t p = 0; p < q; ++p)
      if (solvers[p] == s) return bfail("an instance appears twice in the batch");
  }
  auto b = std::make_unique<wc_batch>();
  b->device = s0->device;
  HIP_OK(hipSetDevice(b->device));
  HIP_OK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  struct Undo {  // a failed construction leaves nothing behind
    std::unique_ptr<wc_batch>& b;
    ~Undo() {
      if (!b) return;
      (void)hipStreamSynchronize(b->stream);
      for (void* p : b->owned) (void)hipFree(p);
      for (void* p : b->owned_host) (void)hipHostFree(p);
      (void)hipStreamDestroy(b->stream);
    }
  } undo{b};
  for (uint64_t q = 0; q < n; ++q) {
    wc_solver* s = solvers[q];
    HIP_OK(hipStreamSynchronize(s->stream));
    b->inst.push_back(s);
    b->active.push_back((int)q);
    b->maxit = std::max<unsigned long long>(b->maxit, s->disp.max_iterations);
    // lanes past an instance's blocks must read zero
    HIP_OK(hipMemsetAsync(s->lanes, 0, 8 * 512 * sizeof(double), b->stream));
  }
  b->rk2 = s0->order != WC_ORDER_FIRST;
  b->lanes = s0->pcg_lanes;
  b->march = s0->pcg_march != 0;
  b->recon64 = s0->recon64 != 0;
  b->march_flux = s0->flux_kernel == 3;
  b->sw = switch_on(s0);
  b->dispersive = s0->dispersive;
  b->pcg_batch = s0->pcg_batch;
  TRY(balloc(b.get(), &b->d_red, n));
  TRY(balloc(b.get(), &b->d_done, n));
  TRY(balloc(b.get(), &b->d_it, n));
  TRY(bhost(b.get(), &b->h_red, n));
  TRY(bhost(b.get(), &b->h_done, n));
  TRY(bhost(b.get(), &b->h_it, n));
  TRY(batch_build(b.get()));
  *out = b.release();
  return WC_OK;
}

void wc_batch_free(wc_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (auto& pe : b->prof_pending) {
    (void)hipEventDestroy(pe.first);
    (void)hipEventDestroy(pe.second);
  }
  for (hipEvent_t e : b->prof_pool) (void)hipEventDestroy(e);
  for (void* p : b->owned) (void)hipFree(p);
  for (void* p : b->owned_host) (void)hipHostFree(p);
  (void)hipStreamDestroy(b->stream);
  delete b;
}

int wc_batch_set_option(wc_batch* b, const char* key, int64_t value) {
  if (!b || !key) return fail(WC_EINVAL, "null argument");
  const std::string k = key;
  if (k == "grid_nt" || k == "grid_pt") {
    if (value < 1) return fail(WC_EINVAL, "a launch width must be positive");
    HIP_OK(hipStreamSynchronize(b->stream));
    drop_graphs(b);  // the widths are part of the captured launches
    (k == "grid_nt" ? b->grid_nt : b->grid_pt) = (int)std::min<int64_t>(value, 1 << 30);
    if (k == "grid_nt") b->grid_st = b->grid_nt;
    b->tuned = true;  // set by hand: no tuning
    return WC_OK;
  }
  if (k == "flux_rows") {  // the marching flux's band height: 0 = tuned at the next run
    if (value < 0 || value > (1 << 30)) return fail(WC_EINVAL, "flux_rows must be >= 0");
    HIP_OK(hipStreamSynchronize(b->stream));
    b->flux_rows_fixed = value > 0;
    if (value > 0) b->flux_rows = (int)value;
    b->tune_rows.clear();
    if (value == 0) b->tuned = false;
    return batch_build(b);
  }
  if (k == "tune") {
    b->tuned = value == 0;
    if (value) b->grid_st = 1 << 30;
    HIP_OK(hipStreamSynchronize(b->stream));
    drop_graphs(b);
    return WC_OK;
  }
  if (k == "pcg_batch") {
    if (value < 1) return fail(WC_EINVAL, "pcg_batch must be >= 1");
    b->pcg_batch = (unsigned long long)value;
    return WC_OK;
  }
  if (k == "graph") {
    HIP_OK(hipStreamSynchronize(b->stream));
    drop_graphs(b);
    b->use_graph = value != 0;
    return WC_OK;
  }
  if (k == "profile") {  // 1: time every PCG iteration launch (counters reset), 0: stop
    if (value != 0 && value != 1) return fail(WC_EINVAL, "profile must be 0 or 1");
    TRY(batch_prof_collect(b));
    b->prof = value == 1;
    if (b->prof) {
      b->prof_ms = 0.0;
      b->prof_n = 0;
    }
    return WC_OK;
  }
  return fail(WC_EINVAL, "unknown batch option: " + k);
}

int wc_batch_profile_read(wc_batch* b, double* total_ms, uint64_t* launches) {
  if (!b || !total_ms || !launches) return fail(WC_EINVAL, "null argument");
  HIP_OK(hipSetDevice(b->device));
  TRY(batch_prof_collect(b));
  *total_ms = b->prof_ms;
  *launches = b->prof_n;
  return WC_OK;
}

int wc_batch_flux_rows(const wc_batch* b, int32_t* rows, int32_t* tried_rows, double* tried_us,
                       int32_t capacity, int32_t* tried) {
  if (!b) return fail(WC_EINVAL, "null batch");
  if (capacity < 0) return fail(WC_EINVAL, "capacity must be >= 0");
  if (rows) *rows = b->march_flux ? b->flux_rows : 0;
  const int n = (int)b->tune_rows.size();
  for (int q = 0; q < std::min(n, (int)capacity); ++q) {
    if (tried_rows) tried_rows[q] = b->tune_rows[q].first;
    if (tried_us) tried_us[q] = b->tune_rows[q].second;
  }
  if (tried) *tried = n;
  return WC_OK;
}

int wc_batch_launch_widths(const wc_batch* b, int32_t* grid_nt, int32_t* grid_pt) {
  if (!b) return fail(WC_EINVAL, "null batch");
  if (grid_nt) *grid_nt = b->grid_nt;
  if (grid_pt) *grid_pt = b->grid_pt;
  return WC_OK;
}

int wc_batch_run_steps(wc_batch* b, uint64_t steps, uint64_t* done_steps, double* last_dt) {
  if (!b) return fail(WC_EINVAL, "null batch");
  HIP_OK(hipSetDevice(b->device));
  const size_t all = b->inst.size();
  if (done_steps)
    for (size_t q = 0; q < all; ++q) done_steps[q] = 0;
  // The instances may have been used alone since the last run: their settings must still be
  // the batch's, and their PCG lanes past the blocks the batch gives them must read zero again
  // (a solve alone with other lanes or blocks may have written them).
  for (size_t q = 0; q < all; ++q) {
    const wc_solver* s = b->inst[q];
    if (s->pcg_lanes != b->lanes || (s->pcg_march != 0) != b->march ||
        (s->recon64 != 0) != b->recon64 || (s->flux_kernel == 3) != b->march_flux ||
        (s->order != WC_ORDER_FIRST) != b->rk2 ||
        switch_on(s) != b->sw || s->dispersive != b->dispersive)
      return bfail("an instance's settings changed since the batch was made");
    TRY(batch_compatible(s, b->inst[0]));
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemsetAsync(s->lanes, 0, 8 * 512 * sizeof(double), b->stream));
  }
  if (!b->tuned && !b->active.empty()) {
    TRY(batch_tune(b));
    drop_graphs(b);
  }
  for (uint64_t step = 0; step < steps && !b->active.empty(); ++step) {
    int n = (int)b->active.size();
    // stable_dt of every instance (solver.rs:434-450), one read
    HIP_OK(hipMemsetAsync(b->d_red, 0, n * sizeof(unsigned long long), b->stream));
    hipLaunchKernelGGL(kb_max_rate32, bgrid(b->m_rate.total, b->grid_st), dim3(NT), 0, b->stream,
                       b->rate, b->m_rate);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(b->h_red, b->d_red, n * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    std::vector<double> dts(n);
    std::vector<int> keep;
    for (int q = 0; q < n; ++q) {
      wc_solver* s = b->inst[b->active[q]];
      double rate;
      memcpy(&rate, &b->h_red[q], sizeof rate);
      const double dt = std::isnan(rate) ? NAN : rate > 0.0 ? s->cfl / rate : INFINITY;
      if (last_dt) last_dt[b->active[q]] = dt;
      if (std::isfinite(dt) && dt > 0.0) {
        dts[keep.size()] = dt;
        keep.push_back(b->active[q]);
      }
    }
    if ((int)keep.size() != n) {  // an instance whose step is not finite stops, as wc_run_steps
      b->active = keep;
      n = (int)keep.size();
      if (n == 0) break;
      TRY(batch_build(b));
    }
    // the step parameters of every instance (glibc, as the reference), one upload
    for (int q = 0; q < n; ++q) {
      const wc_solver* s = b->inst[b->active[q]];
      b->h_stage[q] = dts[q];
      fill_step_params(s, s->time, dts[q], b->h_stage + n + b->off[q]);
    }
    HIP_OK(hipMemcpyAsync(b->d_stage, b->h_stage, (n + b->off[n]) * sizeof(double),
                          hipMemcpyHostToDevice, b->stream));
    if (b->off[n] > 0)
      hipLaunchKernelGGL(kb_scatter, dim3(64), dim3(NT), 0, b->stream,
                         (const double*)(b->d_stage + n), (const long long*)b->d_off,
                         (double* const*)b->d_dst, n);
    TRY(batch_step(b));
    for (int q = 0; q < n; ++q) {
      wc_solver* s = b->inst[b->active[q]];
      s->time += dts[q];
      if (done_steps) done_steps[b->active[q]] += 1;
    }
  }
  HIP_OK(hipStreamSynchronize(b->stream));
  return WC_OK;
}

}  // extern "C"
