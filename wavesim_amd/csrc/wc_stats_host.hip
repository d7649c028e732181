// The run's statistics, frames and blocks of the run directory, the step watch and the stage
// step: the C ABI over wc_stats.hip and wc_run.hip.

#include "wc_host.h"

extern "C" {

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

}  // namespace wc
