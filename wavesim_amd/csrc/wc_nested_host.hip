// Relaxation zones (forcing.rs:227-284) and the coarse-to-fine sampler of nested runs
// (crates/wavesim/src/lib.rs:1548-1616): the C ABI over wc_relax.hip.

#include "wc_host.h"

extern "C" {

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

}  // extern "C"
