// Host side of libwavecore_hip: what the C ABI's source files share. wc_api.hip holds the
// solver (its set-up, step and PCG drivers, diagnostics, profiling), wc_stats_host.hip the run
// statistics, frames and the stage step, wc_nested_host.hip the relaxation zones and the
// coarse-to-fine sampler, wc_batch_host.hip the batched instances.
#pragma once

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
void detach(wc_stats* st);  // releases the device memory; the object stays, inert
void detach(wc_relaxation* r);
}  // namespace wc

namespace wc {
namespace host {

constexpr double TAU = 6.283185307179586476925286766559;
constexpr double PI = 3.14159265358979323846264338327950;
constexpr int PRM_SLOTS = 3;
constexpr int MARCH_MW = TX - 6;  // interior columns of a marching flux strip

// Records msg as the calling thread's last error (wc_last_error) and returns code.
int fail(int code, const std::string& msg);
const char* last_error();

#define HIP_OK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? WC_ENOMEM : WC_EDEVICE,                     \
                  std::string(#expr) + " failed: " + hipGetErrorString(e_));              \
  } while (0)

#define TRY(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != WC_OK) return rc_; \
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
extern const char* kFamilyNames[F_COUNT];

struct HostSponge {
  uint64_t width;
  double strength, level;
  bool west, east, south, north;
  long long offset, len;
};

bool linear_wave(double omega, double h, bool dispersive, wc_linear_wave_t* out);

}  // namespace host
}  // namespace wc

using namespace wc;
using namespace wc::host;

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

namespace wc {
namespace host {

inline int set_device(wc_solver* s) {
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

inline hipEvent_t take_event(wc_solver* s) {
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

dim3 tiles(const wc_solver* s);
dim3 over(long long n);
dim3 rate_blocks(long long n);
Breaking breaking_of(const wc_solver* s);
void fill_step_params(const wc_solver* s, double t, double dt, double* p);
int join_state(wc_solver* s);
Sgn32 sgn32_args(wc_solver* s, int stage);
int tile_stride(const wc_solver* s);
CgG cg1_args(wc_solver* s, int stage);
BSwitch switch_args32(wc_solver* s, int stage);
BFlux flux_args32(wc_solver* s, int stage);
Stage32 stage_args32(wc_solver* s, double dt, int stage);
bool switch_on(const wc_solver* s);

}  // namespace host
}  // namespace wc
