// Many solvers stepped together, one launch per phase (wc_batch_*): the C ABI over the kb_*
// kernels.

#include "wc_host.h"

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
  int setup_rows = SETUP32_ROWS;  // the SGN set-up's tile height, 4, 8 or 16 (tuned unless fixed)
  bool setup_rows_fixed = false;
  std::vector<std::pair<int, double>> tune_setup;  // (tile height, us per launch) as measured
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
  BMap m_fill{}, m_sw{}, m_tiles{}, m_setup{}, m_rate{}, m_cg{}, m_march{};
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
  std::vector<int> f_fill(n + 1, 0), f_sw(n + 1, 0), f_tiles(n + 1, 0), f_setup(n + 1, 0),
      f_rate(n + 1, 0), f_cg(n + 1, 0), f_march(n + 1, 0);
  auto rows_of = [&](const Geo& g) { return std::max(1, std::min(b->flux_rows, g.ny)); };
  b->off.assign(n + 1, 0);
  for (int q = 0; q < n; ++q) {
    const wc_solver* s = b->inst[b->active[q]];
    const Geo& g = s->geo;
    f_fill[q + 1] = f_fill[q] + (int)((6LL * g.w + 6LL * g.ny + NT - 1) / NT);
    f_sw[q + 1] = f_sw[q] + (int)(((long long)g.w * g.rows + NT - 1) / NT);
    f_tiles[q + 1] = f_tiles[q] + g.ntx * g.nty;
    f_setup[q + 1] = f_setup[q] + g.ntx * ((g.ny + b->setup_rows - 1) / b->setup_rows);
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
  const size_t o_fsetup = put(f_setup.data(), (n + 1) * sizeof(int));
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
  b->m_setup = BMap{reinterpret_cast<const int*>(d + o_fsetup), n, f_setup[n]};
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

// The SGN set-up of every instance on its 64 x setup_rows tiles, a block per tile.
void enq_setup(wc_batch* b, int st) {
  const dim3 grid = bgrid(b->m_setup.total, 1 << 30);
  if (b->setup_rows == 4)
    hipLaunchKernelGGL(kb_sgn_setup32<4>, grid, dim3(TX * 4), 0, b->stream, b->sgn[st],
                       b->m_setup);
  else if (b->setup_rows == 8)
    hipLaunchKernelGGL(kb_sgn_setup32<8>, grid, dim3(TX * 8), 0, b->stream, b->sgn[st],
                       b->m_setup);
  else
    hipLaunchKernelGGL(kb_sgn_setup32<16>, grid, dim3(TX * 16), 0, b->stream, b->sgn[st],
                       b->m_setup);
}

void enq_stage_head(wc_batch* b, int st) {  // the switch, the rates, the SGN set-up
  if (b->sw)
    hipLaunchKernelGGL(kb_switch32, bgrid(b->m_sw.total, b->grid_st), dim3(NT), 0, b->stream,
                       b->swa[st], b->m_sw);
  enq_flux(b, st);
  if (b->dispersive) enq_setup(b, st);
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
// of its launch; the SGN set-up's tile height (64 x 4, 8 or 16 cells; timed on the set-up,
// whose coefficients and right-hand side are a function of the state it reads); and the
// persistent width of the 1024-thread PCG launches (timed on kb_cg1_start, whose outputs the
// next solve's start rewrites). No choice changes a bit.
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
  if (b->dispersive && !b->setup_rows_fixed) {
    b->tune_setup.clear();
    int best = b->setup_rows;
    double t_best = 1e30;
    for (int r : {4, 8, 16}) {
      b->setup_rows = r;
      TRY(batch_build(b));  // the map of the instances' tiles for this height
      const double us = time([&] { enq_setup(b, 0); });
      HIP_OK(hipGetLastError());
      b->tune_setup.push_back({r, us});
      if (us < t_best) t_best = us, best = r;
    }
    b->setup_rows = best;
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
  if (k == "setup_rows") {  // the SGN set-up's tile height, 4, 8 or 16: 0 = tuned at the next run
    if (value != 0 && value != 4 && value != 8 && value != 16)
      return fail(WC_EINVAL, "setup_rows must be 0 (tuned), 4, 8 or 16");
    HIP_OK(hipStreamSynchronize(b->stream));
    b->setup_rows_fixed = value > 0;
    if (value > 0) b->setup_rows = (int)value;
    b->tune_setup.clear();
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

int wc_batch_setup_rows(const wc_batch* b, int32_t* rows, int32_t* tried_rows, double* tried_us,
                        int32_t capacity, int32_t* tried) {
  if (!b) return fail(WC_EINVAL, "null batch");
  if (capacity < 0) return fail(WC_EINVAL, "capacity must be >= 0");
  if (rows) *rows = b->setup_rows;
  const int n = (int)b->tune_setup.size();
  for (int q = 0; q < std::min(n, (int)capacity); ++q) {
    if (tried_rows) tried_rows[q] = b->tune_setup[q].first;
    if (tried_us) tried_us[q] = b->tune_setup[q].second;
  }
  if (tried) *tried = n;
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
