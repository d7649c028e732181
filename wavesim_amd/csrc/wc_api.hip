// Host side of libwavecore_hip: the C ABI of include/wavecore_hip.h over the gfx950 kernels.
//
// A wc_solver is the reference's Solver (crates/wavecore/src/solver.rs:260-280) plus the one
// State it steps, resident on one GPU. Everything a step needs lives in device memory; the host
// only evaluates what depends on nothing but the time and dt (the wave-maker's ramp and phase,
// the sponges' exp(-rate dt)), with the same libm calls the reference makes, and ships those few
// numbers each step.

#include "wc_host.h"

namespace wc {
namespace host {

namespace {
thread_local std::string g_err;
}  // namespace

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

const char* last_error() { return g_err.c_str(); }

const char* kFamilyNames[F_COUNT] = {"fill_state", "switch",    "flux",      "finish_euler",
                                     "sgn_setup",  "cg_init",   "cg_matvec", "cg_update",
                                     "cg_pupdate", "cg_stats",  "sgn_finish", "reduce",
                                     "cg_reduce",  "relax",     "cg_iter",   "net",
                                     "frame",      "cg_flush",  "cg_pair"};


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

}  // namespace host
}  // namespace wc

namespace wc {
namespace host {






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
  const dim3 setup_tiles((unsigned)(s->geo.ntx * ((s->geo.ny + SETUP_ROWS - 1) / SETUP_ROWS)));
  TRY(launch(s, F_SGN_SETUP, k_sgn_setup, setup_tiles, dim3(TX * SETUP_ROWS), sgn_args(s, stage)));
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
    const dim3 setup_tiles((unsigned)(s->geo.ntx * ((s->geo.ny + SETUP32_ROWS - 1) / SETUP32_ROWS)));
    TRY(launch(s, F_SGN_SETUP, k_sgn_setup32, setup_tiles, dim3(TX * SETUP32_ROWS),
               sgn32_args(s, stage)));
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

}  // namespace host
}  // namespace wc

// ================================================================== the C ABI
namespace wc {
// The error setter for the other sources of the library (wc_unet.hip).
int api_fail(int code, const char* msg) { return fail(code, msg); }
}  // namespace wc

extern "C" {

const char* wc_last_error(void) { return last_error(); }

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
