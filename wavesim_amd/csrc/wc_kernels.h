// Kernel interfaces of the FP64 step: argument blocks and declarations (kernels in
// wc_kernels.hip, launches in the host sources: wc_api.hip, wc_*_host.hip).
#pragma once

#include "wc_device.h"

namespace wc {

enum { MODE_S1 = 0, MODE_FINAL = 1, MODE_RATES = 2, MODE_EULER = 3 };
enum { REDUCE_INIT = 0, REDUCE_ALPHA = 1, REDUCE_BETA = 2 };

struct Breaking {
  double er_lo, er_hi, sl_lo, sl_hi;
  // with the eddy viscosity on (Breaking::viscosity), the breaking memory of the state:
  // where it is not zero the switch stays off (EddyViscosity::keeps_off, eddy.rs:116-118)
  const double* memory;
};

// The switch's "off" for the thresholds br (wc_math.h: wc_switch_off, the portable hypot
// skipped where its last bits cannot matter).
__device__ __forceinline__ double switch_off(const Breaking& br, double ratio, double ex,
                                             double ey) {
  return wc_switch_off(br.er_lo, br.er_hi, br.sl_lo, br.sl_hi, ratio, ex, ey);
}

// breaking_into's indicator at interior cell (i, j), padded index k (dispersion.rs:183-201):
// 1 - phi with phi the switch of the thresholds br, 0 where the water is shallower than
// H_DISPERSION_MIN.
__device__ __forceinline__ double breaking_at(const double* h, const double* b, const Geo& g,
                                              const Breaking& br, int i, int j, int k) {
  const double hc = h[k];
  if (hc < WC_H_DISPERSION_MIN) return 0.0;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const double ex = ((ghosted(h, g, i + 1, j) + b[k + 1]) - (ghosted(h, g, i - 1, j) + b[k - 1])) * ix;
  const double ey =
      ((ghosted(h, g, i, j + 1) + b[k + g.w]) - (ghosted(h, g, i, j - 1) + b[k - g.w])) * iy;
  const double ratio = wc_max((hc + b[k]) / hc, 0.0);
  double off = switch_off(br, ratio, ex, ey);
  if (br.memory && br.memory[k] != 0.0) off = 1.0;
  return 1.0 - (1.0 - off);
}

// One sponge as the kernels use it: its table of rates and per-step factors, indexed by the
// distance in cells to the nearest active edge (forcing.rs:181-206).
struct SpongeDev {
  long long offset, len;
  double level;
  int west, east, south, north;
};

constexpr int MAX_SPONGES = 16;

// Scalars of one PCG solve and the DispersionStats counters, in device memory.
struct CgScalars {
  double rz, target, alpha, beta;
  double rzs[2];  // r.z of the previous iteration, by iteration parity (wc_pcg.hip)
  double alphas[2];  // alpha of the previous iteration, by parity (wc_pcg1.hip)
  // the flexible PCG with the network (wc_fpcg.hip)
  double rr_prev, sigma;
  int stall, fallback, restart, net_off;
  unsigned long long fallbacks;
  unsigned long long it;
  int converged, done;
  unsigned long long solves, iterations, unconverged;
};

struct StepArgs {
  Geo g;
  int order;
  double dt;
  const double* b;
  const double* phi;
  // stage input (s for stage 1, s1 for stage 2)
  const double *in_h, *in_hu, *in_hv;
  double *s_h, *s_hu, *s_hv;     // the state s (read and written by MODE_FINAL)
  double *s1_h, *s1_hu, *s1_hv;  // written by MODE_S1
  double *lh, *lhu, *lhv;        // rates
  // sources: the wave-maker columns and the host-evaluated per-step values
  int has_drive;
  int src_i0, src_i1;
  const double* src_col;
  const double* prm;
  int ramp_prm, sin_prm, sponge_prm;
  int n_sponges;
  SpongeDev sponges[MAX_SPONGES];
  const double* sponge_rate;
  double manning;
};

// Inputs of the tiled flux kernels (wc_flux.hip); R is the storage type of the stage state.
template <typename R>
struct FluxTileArgs {
  Geo g;
  int order;
  const R *h, *h_lo, *hu, *hv;  // h_lo: low part of the FP32 depth pair, null in FP64
  const R *hu_lo, *hv_lo;       // low parts of the FP32 momentum pairs, null in FP64
  const double* b;
  const R* phi;
  int has_drive;
  int src_i0, src_i1;
  const double* src_col;
  const double* prm;
  int ramp_prm, sin_prm;
};

// The extrapolated warm start of the SGN solves (the FP32 form's default; an option in FP64). The reference starts each solve
// from the previous solution x_{n-1} (dispersion.rs:220-222, 488-490). Solves come in a fixed
// rhythm of `period` per step (2 for SSP-RK2, 1 for Euler), so solve n's solution is a smooth
// function of time, sampled every `period` solves for each kind (RK stage), and
//   x_n ~ x_{n-1} + sum_{j=1..K} c_j d_{n - j period},   d_n = x_n - x_{n-1}.
// The weights c are exact for steps d that are polynomials of degree < m in time (each kind its
// own: then the solutions are a polynomial of degree <= m plus, per kind, one of degree < m) and,
// when the window K exceeds m, of least norm among those: sum c_j^2 sets how much of the
// solutions' own noise (the FP32 rates' rounding, the solves' tolerance) reaches the start.
// K = m gives the binomial weights of the m-th backward difference set to zero (3, -3, 1 for
// m = 3). The ring holds the last WS_SLOTS steps d, slot n % WS_SLOTS for solve n, indexed by
// the solve counter in CgScalars (so replays of captured graphs and batched instances find
// their own slots). The set-up stores the extrapolation x0 - x_{n-1} into slot n and starts the
// solve at x0; the stage update, after the solve, adds the solve's correction, so slot n holds
// d_n = x_n - x_{n-1}. Masked cells start from zero, as the reference's, and record d = 0. The
// start never changes what a solve converges to (same residual test, same tolerance), only how
// many iterations it takes.
constexpr int WS_SLOTS = 32;
constexpr int WS_MAX_ORDER = 5;
constexpr int WS_MAX_WINDOW = 12;  // WS_MAX_WINDOW * period < WS_SLOTS
struct WarmStart {
  float2* ring;                    // WS_SLOTS x stride, interleaved (x, y); null: the reference's
  double2* ring64;                 // the FP64 path's ring (its slot n holds x_{n-1} during solve
                                   // n, as the oracle's), or null
  unsigned long long stride;       // cells per slot
  const CgScalars* sc;             // sc->solves: solves completed so far
  int terms, period;               // K, solves per step
  double coef[WS_MAX_WINDOW];      // c_1 .. c_K
  float guard;                     // cells whose switch reads below it start unextrapolated
};

struct SgnArgs {
  Geo g;
  const double *h, *hu, *hv, *phi;
  const double *lh, *lhu, *lhv;
  double *c, *dgx, *dgy, *bx, *by, *wx, *wy, *rx, *ry, *px, *py, *ax, *ay;
  double *zx, *zy;  // z = r / diag, kept between the two kernels of a fused iteration
  double2* dg2;     // set: the diagonal is stored interleaved here instead (wc_pcg2.hip)
  double* part;
  CgScalars* sc;
  int nonlinear;
  double tolerance;
  unsigned long long max_iterations;
  WarmStart ws;  // ring64 set: the FP64 path's extrapolated warm start
};

// The 256-lane PCG (wc_pcg.hip), on FP64 (V = double) or FP32 (V = float) storage.
template <typename V>
struct CgV {
  Geo g;
  const V *h, *c;
  const V *dgx, *dgy;  // the Jacobi diagonal as the set-up stored it
  const double2* dg2;  // wc_pcg2.hip: the diagonal, r and p as interleaved (x, y) pairs
  double2* r2;
  const double *bx, *by;
  const double *solx, *soly;  // FP32 form: the solve's FP64 start x0 (w its correction); FP64: null
  V *wx, *wy, *rx, *ry, *px, *py, *zx, *zy, *ax, *ay;
  double *acc_init, *acc_pap, *acc_update;  // 3 x 256, 256, 2 x 256 lane sums
  CgScalars* sc;
  double tolerance;
  unsigned long long max_iterations;
};

// The one-kernel-per-iteration PCG of wc_pcg1.hip (FP32 form): r, q (= A u) and s (= A p)
// alternate between two buffers by iteration parity, p and w are updated in place, di holds the
// inverse Jacobi diagonal; r, q, s, p and di are interleaved (x, y) pairs.
struct CgG {
  Geo g;
  const float *h, *c;
  const double *bx, *by;
  const double *solx, *soly;  // the solve's start x0 (FP64, see Sgn32)
  float *wx, *wy;             // the correction x - x0 (FP32)
  float2 *p, *di;
  float2* r[2];
  float2* q[2];
  float2* s[2];
  double* lanes_b;  // 256 lanes of b.b
  double* lanes;    // 2 x 3 x 256 lanes of r.u, q.u, r.r, by parity
  CgScalars* sc;
  double tolerance;
  unsigned long long max_iterations;
  // w += alpha p is taken two iterations at a time, in the odd ones (w += alpha_{i-1} p_{i-1}
  // + alpha_i p_i, p_{i-1} being read there anyway), which halves the traffic of w; a solve
  // that ends after an odd number of iterations adds its last step in k_cg1_flush.
  int defer;
  int stride;  // 64x16-tile layout: the launch's blocks; block b takes tiles b, b + stride, ...
};
// The flexible PCG with the learned preconditioner (wc_fpcg.hip), FP32 form. Vectors are
// interleaved (x, y) pairs in grid layout; lanes: [0] b.b, [1] r.r, [2] r.z, [3] r.z_old,
// [4] p.Ap, 256 each.
struct FcgArgs {
  Geo g;
  const float *h, *c;
  const double *bx, *by;
  const double *solx, *soly;  // the solve's start x0 (FP64, see Sgn32)
  float *wx, *wy;             // the correction x - x0 (FP32)
  float2 *r, *q, *z, *di;
  float2* p[2];
  const float *bgx, *bgy;  // bed-gradient features
  __bf16* x0;              // network input, NHWC 8 channels, one-pixel border, width nw
  const __bf16* y0;        // network's last hidden layer, 16 channels
  int nw;
  const float* wo;         // output layer: 16 x 2 weights, 2 biases
  double hs;               // 1 / (4 dx): depth feature and output scale
  double* lanes;
  CgScalars* sc;
  double tolerance;
  unsigned long long max_iterations;
  int stall_limit;            // 0: never fall back
  double min_ratio;           // an iteration that keeps more than this of |r| stalls
  double* history;            // r.r before each iteration of the solve (may be null)
  unsigned long long history_len;
};
__global__ void k_fcg_init(FcgArgs a);
__global__ void k_fcg_in(FcgArgs a, unsigned long long n);
__global__ void k_fcg_out(FcgArgs a, unsigned long long n);
__global__ void k_fcg_mv(FcgArgs a, unsigned long long n);
__global__ void k_fcg_up(FcgArgs a, unsigned long long n);
__global__ void k_fcg_check(FcgArgs a, unsigned long long n);
__global__ void k_bed_features(Geo g, const double* b, double scale, float* bgx, float* bgy);

// L: the blocks (and lanes) of one instance's launch, 256 or 512
template <int L>
__global__ void k_cg1_init(CgG a);
template <int L>
__global__ void k_cg1_start(CgG a);
template <int L>
__global__ void k_cg1_iter(CgG a, unsigned long long n);
template <int L>
__global__ void k_cg1_check(CgG a, unsigned long long n);
__global__ void k_cg1_flush(CgG a);  // 64x4 tiles
// the same, blocks marching down column strips (wc_march.h)
template <int L>
__global__ void k_cg1m_init(CgG a);
template <int L>
__global__ void k_cg1m_start(CgG a);
template <int L>
__global__ void k_cg1m_iter(CgG a, unsigned long long n);

template <typename V>
__global__ void k_cg3_init(CgV<V> a);
template <typename V>
__global__ void k_cg3_matvec(CgV<V> a, unsigned long long n, const V* psx, const V* psy, V* pdx,
                             V* pdy);
template <typename V>
__global__ void k_cg3_update(CgV<V> a, unsigned long long n, const V* pcx, const V* pcy);
template <typename V>
__global__ void k_cg3_check(CgV<V> a, unsigned long long n);

// The FP64 path's default PCG (wc_pcg2.hip): 64x16 tiles, 256 blocks, the oracle's dot order 2.
__global__ void k_cg2_init(CgV<double> a);
__global__ void k_cg2_matvec(CgV<double> a, unsigned long long n, const double2* __restrict__ po,
                             double2* __restrict__ pn);
template <int PAIR>
__global__ void k_cg2_update(CgV<double> a, unsigned long long n, const double2* __restrict__ pc,
                             const double2* __restrict__ pp);
__global__ void k_cg2_flush(CgV<double> a, unsigned long long it, const double2* __restrict__ p);
__global__ void k_cg2_check(CgV<double> a, unsigned long long n);

// The FP32 chosen form (wc_fp32.hip).

// The SGN solution of the FP32 form is kept in FP64 (solx, soly: the warm start, shared with the
// FP64 path's), and each solve iterates on an FP32 correction w = x - x0 from its start x0: the
// bulk PCG traffic stays FP32 while the solution, which the warm start carries from solve to
// solve, keeps FP64 accuracy (FP32 rounding of x itself would be amplified by the extrapolation
// and cost iterations). The set-up writes x0 and w = 0, the iterations update w, the stage
// update folds x = x0 + w and uses it.
struct Sgn32 {
  Geo g;
  double *solx, *soly;
  const float *h, *hu, *hv, *phi;  // stage state (high parts) and the switch
  const float *hl, *hul, *hvl;     // and its low parts
  const double* lh;                // mass rate (FP64: the faces' fluxes cancel exactly)
  const double *lhu, *lhv;         // momentum rates (FP64: their rounding would reach b)
  float* c;
  double *bx, *by;                 // the right-hand side (FP64, as the rates)
  float *wx, *wy, *rx, *ry, *px, *py, *ax, *ay, *zx, *zy;
  double* part;
  CgScalars* sc;
  int nonlinear;
  double tolerance;
  unsigned long long max_iterations;
  WarmStart ws;
};

struct Stage32 {
  Geo g;
  double dt;
  const double* b;
  float *sh, *shu, *shv, *sl_h, *sl_hu, *sl_hv;  // the state as FP32 pairs (hi, lo)
  float *s1h, *s1hl, *s1hu, *s1hv;               // the stage-1 state as FP32 pairs:
  float *s1hul, *s1hvl;                          // hi s1h, s1hu, s1hv; lo s1hl, s1hul, s1hvl
  double* l1h;                                   // the stage-1 rates (mass in FP64)
  float *l1hu, *l1hv;
  const double* lh;                              // this stage's shallow-water rates
  const double *lhu, *lhv;
  const float *wx, *wy;                          // the SGN solution's FP32 correction
  double *solx, *soly;                           // its FP64 start, then the solution
  int dispersive;
  StepArgs damp;                                 // sponges and friction (dt, tables, manning)
  WarmStart ws;                                  // records this stage's solution step
};

__global__ void k_split32(long long n, const double* a, float* hi, float* lo);
__global__ void k_join32(long long n, const float* hi, const float* lo, double* a);
__global__ void k_fill32(Geo g, float* h, float* hl, float* hu, float* hv, float* hul,
                         float* hvl);
__global__ void k_switch32(Geo g, const float* h, const float* hl, const double* b, Breaking br,
                           const uint8_t* held, float* phi);
__global__ void k_sgn_setup32(Sgn32 a);
__global__ void k_cg_init32(Sgn32 a);
__global__ void k_cg_matvec32(Sgn32 a, int first, const float* psx, const float* psy,
                              float* pdx, float* pdy);
__global__ void k_cg_update32(Sgn32 a, const float* pcx, const float* pcy);
__global__ void k_cg_reduce32(Sgn32 a, int kind);
__global__ void k_stage32(Stage32 a, int mode);
// Rows of the tiles of the SGN set-up (k_sgn_setup, k_sgn_setup32: 64 x ROWS threads a block;
// the batch's kb_sgn_setup32 takes 4, 8 or 16 rows, tuned or set).
constexpr int SETUP_ROWS = 16;
constexpr int SETUP32_ROWS = 16;
__global__ void k_division_check(double y, const double* x, int n, double* shared,
                                 double* plain);
__global__ void k_max_rate32(Geo g, const float* h, const float* hl, const float* hu,
                             const float* hul, const float* hv, const float* hvl,
                             unsigned long long* out);

__global__ void k_fill_state(Geo g, double* h, double* hu, double* hv, double* br, double* ah,
                             double* ahu, double* ahv);
__global__ void k_switch(Geo g, const double* h, const double* b, Breaking br,
                         const uint8_t* held, double* phi);
__global__ void k_breaking(Geo g, const double* h, const double* b, Breaking br, double* out);
__global__ void k_flux(StepArgs a, int mode);
__global__ void k_finish_euler(StepArgs a);
__global__ void k_sgn_setup(SgnArgs a);
__global__ void k_cg_init(SgnArgs a);
__global__ void k_cg_matvec(SgnArgs a);
__global__ void k_cg_update(SgnArgs a);
__global__ void k_cg_pupdate(SgnArgs a);
// Counts a finished solve; `earlier` iterations of it ran in an earlier phase (the network's).
__global__ void k_cg_stats(CgScalars* sc, unsigned long long earlier);
__global__ void k_cg_matvec2(SgnArgs a, int first, const double* psx, const double* psy,
                             double* pdx, double* pdy);
__global__ void k_cg_update2(SgnArgs a, const double* pcx, const double* pcy);
__global__ void k_sgn_finish(StepArgs a, const double* wx, const double* wy, int mode,
                             WarmStart ws);
__global__ void k_flux_tile64(FluxTileArgs<double> f, StepArgs a, int mode);
__global__ void k_flux_tile32(FluxTileArgs<float> f, double* lh, double* lhu, double* lhv);
__global__ void k_flux_tile32r(FluxTileArgs<float> f, double* lh, double* lhu, double* lhv);
// One wave per strip of 58 columns walking down a band of `rows` rows (wc_flux.hip).
__global__ void k_flux_march64(FluxTileArgs<double> f, StepArgs a, int mode, int rows);
__global__ void k_flux_march32(FluxTileArgs<float> f, double* lh, double* lhu, double* lhv,
                               int rows);
__global__ void k_flux_march32r(FluxTileArgs<float> f, double* lh, double* lhu, double* lhv,
                                int rows);
__global__ void k_max_rate(Geo g, const double* h, const double* hu, const double* hv,
                           unsigned long long* out);
__global__ void k_limit_scan(Geo g, const double* h, const double* hu, const double* hv,
                             unsigned long long* maxbits, unsigned int* first_nan);
__global__ void k_limit_first(Geo g, const double* h, const double* hu, const double* hv,
                              const unsigned long long* maxbits, unsigned int* first);
__global__ void k_state_sums(Geo g, const double* h, const double* hu, const double* hv,
                             double* part, unsigned long long* maxmom,
                             unsigned long long* minh_key);
__global__ void k_sum_tiles(const double* part, int np, double* out);
__global__ void k_cg_reduce(SgnArgs a, int kind);

// Stage::record (crates/wavesim/src/lib.rs:1276-1300) for one step: per interior cell n, the
// surface eta = h + elevation joins the extremes, the wave in progress and the first-harmonic
// sums, and the cell's breaking seconds grow by `inside` where the indicator is >= 0.8. With
// close_wave the wave in progress is closed first (lib.rs:918-932): its crest-to-trough height
// (0 if it saw no sample) is added to wave_sum and kept as the last wave (and the first).
struct RecordArgs {
  Geo g;
  const double *h, *b;
  Breaking br;
  const float* elevation;  // nx*ny, the caller's bed elevation; null: the solver's bed
  double inside, w0, w1, w2;
  int close_wave, first_wave;
  double *breaking_seconds, *eta_max, *eta_min, *wave_max, *wave_min, *harm;
  double *wave_sum, *wave_first, *wave_last;
};
__global__ void k_record(RecordArgs a);
// Stats::close_wave alone (lib.rs:918-932), for the end of a run (lib.rs:1451-1453).
__global__ void k_close_wave(RecordArgs a);

// Stage::frame_if_due's frame (crates/wavesim/src/lib.rs:1348-1380) for interior cell n:
// eta[n] = (h + elevation[n]) as f32, breaking[n] = the indicator as f32; rise_bits gets the
// largest (eta as f64) - tide over the cells deeper than 1e-3 m (as the bits of a non-negative
// double, from 0), nonfinite is set where an eta is not finite.
struct FrameArgs {
  Geo g;
  const double *h, *b;
  Breaking br;
  const float* elevation;
  double tide;
  float *eta, *breaking;
  unsigned long long* rise_bits;
  unsigned int* nonfinite;
};
__global__ void k_frame(FrameArgs a);

// Relaxation::apply (forcing.rs:259-283): cell m of the zone list relaxes by the factor
// factor[rate[m]] (exp(-rate dt), evaluated on the host per distinct rate) towards
// target[3m .. 3m+2] = (eta, u, v).
struct RelaxArgs {
  long long n;
  const int* cell;
  const int* rate;
  const double* factor;
  const double* target;
  const double* b;
};
__global__ void k_relax64(RelaxArgs a, double* h, double* hu, double* hv);
__global__ void k_sample(long long n, const int* idx, const double* w, Geo g, const double* h,
                         const double* hu, const double* hv, const double* b, double* out);
__global__ void k_lerp_targets(long long n, const double* from, const double* to, double w,
                               double* out);
__global__ void k_relax32(RelaxArgs a, float* h, float* hl, float* hu, float* hul, float* hv,
                          float* hvl);

// ---- many instances in one launch (wc_fp32.hip, wc_flux.hip, wc_pcg1.hip: kb_*) -------------
// Instance q owns virtual blocks [first[q], first[q+1]) of `total`; first lives in device
// memory. Argument blocks are arrays in device memory, one per instance.
struct BMap {
  const int* first;
  int n, total;
};
__device__ __forceinline__ int batch_instance(const BMap& m, int vb) {
  int lo = 0, hi = m.n;  // first[lo] <= vb < first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (m.first[mid] <= vb)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
struct BFill {
  Geo g;
  float *h, *hl, *hu, *hv, *hul, *hvl;
};
struct BSwitch {
  Geo g;
  const float *h, *hl;
  const double* b;
  Breaking br;
  const uint8_t* held;
  float* phi;
};
struct BFlux {
  FluxTileArgs<float> f;
  double *lh, *lhu, *lhv;
  int recon64;  // reconstruct in FP64 (k_flux_tile32r, kb_flux32r) rather than FP32
  int rows;     // the marching flux's band height (kb_flux_march32*)
};
struct BRate {
  Geo g;
  const float *h, *hl, *hu, *hul, *hv, *hvl;
  unsigned long long* out;
};
__global__ void kb_fill32(const BFill* args, BMap m);
__global__ void kb_switch32(const BSwitch* args, BMap m);
__global__ void kb_flux32(const BFlux* args, BMap m);
__global__ void kb_flux_march32(const BFlux* args, BMap m);
__global__ void kb_flux_march32r(const BFlux* args, BMap m);
__global__ void kb_flux32r(const BFlux* args, BMap m);  // reconstruction in FP64
template <int ROWS>
__global__ void kb_sgn_setup32(const Sgn32* args, BMap m);
__global__ void kb_stage32(const Stage32* args, const double* dt, int mode, BMap m);
__global__ void kb_max_rate32(const BRate* args, BMap m);
__global__ void kb_scatter(const double* src, const long long* off, double* const* dst, int n);
template <int L>
__global__ void kb_cg1_init(const CgG* args, BMap m);
template <int L>
__global__ void kb_cg1_start(const CgG* args, BMap m);
template <int L>
__global__ void kb_cg1_iter(const CgG* args, unsigned long long n, BMap m);
__global__ void kb_cg1_flush(const CgG* args, BMap m);  // m: the instances' 64x4 tiles
template <int L>
__global__ void kbm_cg1_init(const CgG* args, BMap m);
template <int L>
__global__ void kbm_cg1_start(const CgG* args, BMap m);
template <int L>
__global__ void kbm_cg1_iter(const CgG* args, unsigned long long n, BMap m);
template <int L>
__global__ void kb_cg1_check(const CgG* args, unsigned long long n, int stats, int* done,
                             unsigned long long* it);

// ---- breaking fronts and their eddy viscosity (eddy.rs; wc_eddy.hip) ---------------------
struct EddyP {
  double initial, settled, duration, mixing, linger;
};
// EddyViscosity::follow's two passes and diffuse's sub-step (eddy.rs:163-302)
__global__ void k_eddy_rise(Geo g, const double* hu, const double* hv, double* rise);
__global__ void k_eddy_follow(Geo g, EddyP e, const double* h, const double* old,
                              const double* phi, const double* rise, double dt, double* memory,
                              double* nu);
__global__ void k_eddy_top(Geo g, const double* nu, unsigned long long* top);
__global__ void k_eddy_diffuse(Geo g, const double* h, const double* hu, const double* hv,
                               const double* nu, double sub, double* out_hu, double* out_hv);
__global__ void k_clean_dry(Geo g, double* h, double* hu, double* hv);

}  // namespace wc
