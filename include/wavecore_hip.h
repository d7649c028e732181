/* wavecore_hip — wavecore's time step (FV + SGN) on MI355X (gfx950), behind a C ABI.
 *
 * Each entry point replaces one item of wavecore's public API (crates/wavecore/src/lib.rs:60-66)
 * and keeps its argument meaning. A Rust crate, wavecore-hip, would bind these one to one;
 * the Python mirror wavesim_amd and the C++ header wavecore_hip.hpp wrap them too.
 *
 * Ownership and threading. A wc_solver owns, on one GPU, the bed, the settings, the current
 * state, every scratch array and the dispersive solve's warm start (the reference keeps the
 * warm start and the switch inside Solver through RefCell, solver.rs:272-279). The state stays
 * on the device between steps; wc_state_set / wc_state_get are the only host<->device copies.
 * Calls on one solver must not overlap (the reference's Solver is !Sync); different solvers
 * are independent. Every call is synchronous unless its comment says otherwise.
 *
 * Arrays are padded like the reference's (grid.rs:6-58): (nx+6)*(ny+6) doubles, row-major with
 * y outer, interior cells at [3, nx+3) x [3, ny+3). The library needs nx >= 3 and ny >= 3.
 *
 * Errors. Every int-returning call returns WC_OK or an error code and leaves a message for
 * wc_last_error() (thread-local) where the reference would panic (grid.rs:21-27,
 * forcing.rs:63-66,89-90, solver.rs:377-381,424-427). Numerical failure is not an error:
 * wc_stable_dt reports NaN, 0 or +inf exactly as Solver::stable_dt does (solver.rs:431-450).
 */
#ifndef WAVECORE_HIP_H
#define WAVECORE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WC_OK 0
#define WC_EINVAL 1  /* invalid setup: where the reference panics */
#define WC_EDEVICE 2 /* a HIP call failed, or no gfx950 device */
#define WC_ENOMEM 3  /* device memory exhausted */

/* Order (solver.rs:34-50). */
#define WC_ORDER_FIRST 0
#define WC_ORDER_SECOND 1
#define WC_ORDER_FIFTH 2

/* Arithmetic of the step. */
#define WC_PRECISION_FP64 0 /* the plain path: FP64 throughout, matches the FP64 oracle */
/* The chosen form: FP32 storage and reconstruction, the state as FP32 (hi, lo) pairs with a
 * compensated update, FP64 inner products, norms, PCG scalars and surface-minus-bed differences.
 * Differs from FP64 by rounding; in this mode the ghosts returned by wc_state_get are refilled
 * from the final interior. */
#define WC_PRECISION_FP32 1

typedef struct wc_solver wc_solver;

/* WaveMaker (forcing.rs:18-33). */
typedef struct {
  double x, amplitude, period, angle, sigma, ramp_periods;
} wc_wavemaker_t;

/* Sponge (forcing.rs:146-178); the edges are booleans. */
typedef struct {
  uint64_t width;
  double strength, level;
  int32_t west, east, south, north;
} wc_sponge_t;

/* Breaking (dispersion.rs:44-74); its `viscosity` is wc_solver_with_eddy_viscosity. */
typedef struct {
  double eta_ratio_lo, eta_ratio_hi, slope_lo, slope_hi;
} wc_breaking_t;

/* EddyViscosity (eddy.rs:63-89): Kennedy et al.'s defaults are 0.65, 0.15, 5, 1.2, 5. */
typedef struct {
  double initial, settled, duration, mixing, linger;
} wc_eddy_t;

/* Dispersion (dispersion.rs:76-97). */
typedef struct {
  wc_breaking_t breaking;
  int32_t nonlinear;
  int32_t reserved; /* must be 0 */
  double tolerance;
  uint64_t max_iterations;
} wc_dispersion_t;

/* DispersionStats (dispersion.rs:99-106). */
typedef struct {
  uint64_t solves, iterations, unconverged;
} wc_dispersion_stats_t;

/* Limiting (solver.rs:248-258); i, j count interior cells from 0. */
typedef struct {
  uint64_t i, j;
  double depth, speed;
} wc_limiting_t;

/* LinearWave (forcing.rs:35-44). */
typedef struct {
  double k, phase_speed, group_speed;
} wc_linear_wave_t;

/* The message of the last failed call on this thread. */
const char* wc_last_error(void);

/* Solver::new(Grid::new(nx, ny, dx, dy), Bathymetry { b }) (solver.rs:283-299, grid.rs:20-27).
 * bed is a padded array, ghosts already filled as Bathymetry builds them. Defaults as the
 * reference: cfl 0.4, Order::Second, no friction, walls. Runs on the current HIP device. */
int wc_solver_new(uint64_t nx, uint64_t ny, double dx, double dy, const double* bed,
                  wc_solver** out);
void wc_solver_free(wc_solver* s);

/* Builder (solver.rs:301-419). The `cfl` field (solver.rs:263) has a setter and a getter. */
int wc_solver_set_cfl(wc_solver* s, double cfl);
int wc_solver_cfl(const wc_solver* s, double* cfl);
int wc_solver_with_order(wc_solver* s, int32_t order);                /* solver.rs:302 */
int wc_solver_with_manning(wc_solver* s, double n);                   /* solver.rs:308 */
int wc_solver_with_sponge(wc_solver* s, const wc_sponge_t* sponge);   /* solver.rs:314 */
int wc_solver_with_wavemaker(wc_solver* s, const wc_wavemaker_t* m); /* solver.rs:321 */
int wc_solver_with_dispersion(wc_solver* s, const wc_dispersion_t* d); /* solver.rs:330 */
/* Breaking::viscosity = Some(*e) (NULL: None), after wc_solver_with_dispersion (which resets
 * it): the breaking fronts are followed in the state's breaking array after every step, keep
 * the switch off where they are and behind them, and get their eddy viscosity
 * (solver.rs:512-532, eddy.rs). In the FP32 form this part runs on the FP64 state. */
int wc_solver_with_eddy_viscosity(wc_solver* s, const wc_eddy_t* e);
int wc_solver_with_periodic_y(wc_solver* s);                          /* solver.rs:414 */
/* solver.rs:376: one byte per interior cell, row-major with y outer, nonzero = held. */
int wc_solver_with_shallow_cells(wc_solver* s, const uint8_t* cells, uint64_t n);
/* The bed as the solver holds it (with_periodic_y wraps its ghost rows; solver.rs:416). */
int wc_solver_bed(const wc_solver* s, double* out);
/* Implementation choices (no counterpart in the reference). In FP64 they change speed, and
 * pcg_kernels the summation order of the PCG's dot products (1-3: the oracle's order 1, 5: its
 * order 2), nothing else; in the FP32 form the PCG layout and the warm start also change the
 * rounding, and the warm start the iteration count (never the tolerance a solve meets):
 *   "pcg_kernels" 1 = three kernels and two reductions per PCG iteration; 2 = two kernels and
 *                 two reductions; 3 = two kernels, reductions folded into them, 64x4 tiles;
 *                 4 = one kernel per iteration, Chronopoulos-Gear recurrences (FP32 default;
 *                 FP64 runs 5 instead); 5 = FP64: two kernels over 64x16 tiles, the update
 *                 recomputing A p (FP64 default);
 *   "pcg_lanes"   blocks of the one-kernel PCG: 256 (default) or 512;
 *   "pcg_march"   one-kernel PCG layout: 0 = 64x16 tiles (default), 1 = blocks marching down
 *                 64-column strips with circular LDS row buffers;
 *   "pcg_tiles_per_block" n, 1..64: the one-kernel PCG on 64x16 tiles runs ceil(tiles / n)
 *                 blocks, at most pcg_lanes, so a block takes about n tiles (at most n unless
 *                 the lanes cap the block count; then more); default 1. Larger values give a
 *                 small grid fewer, longer blocks, which pays when many instances share a
 *                 launch. Ignored by pcg_march 1, pcg_kernels 1-3 and the network's PCG;
 *   "pcg_batch"   PCG iterations queued between looks at the convergence flag (default 4);
 *   "flux_kernel" 1 = one thread per cell, as the reference's loop (FP64; the FP32 form runs
 *                 2 instead); 2 = 64x4 LDS tiles, each cell reconstructed and each face fluxed
 *                 once; 3 = one wave per 58-column strip walking down a band of rows, the y
 *                 stencil in registers and the x stencil in the wave's LDS row (default);
 *   "flux_rows"   flux_kernel 3's band height, rows per wave (each wave's tile is 58 x rows
 *                 cells; the bits do not depend on it): 0 = chosen by the tuner at the first
 *                 flux launch, timing a few heights (default); n > 0 = n rows;
 *   "flux_reconstruction" FP32 form: 32 = reconstruct in FP32, 64 = in FP64 (default);
 *   "warm_start"  FP32 form: the SGN solves start from an extrapolation of the past solutions,
 *                 exact for solution steps of degree < m (m = 0..5; 0 = the previous
 *                 solution, as the reference; default 3);
 *   "warm_start_window" the past steps weighed, K = max(window, m) (default 6): the
 *                 least-norm weights exact to that degree (wc_warm_start_weights). */
int wc_solver_set_option(wc_solver* s, const char* key, int64_t value);
/* flux_kernel 3's band height in use (0 before the first flux launch) into *rows, and the
 * tuner's measurements: up to `capacity` (height, microseconds per launch) pairs into
 * tried_rows / tried_us, their count into *tried (0 when the height was fixed or not timed).
 * No counterpart in the reference (its flux is one loop over the cells, solver.rs:625-718). */
int wc_solver_flux_rows(const wc_solver* s, int32_t* rows, int32_t* tried_rows, double* tried_us,
                        int32_t capacity, int32_t* tried);
/* The warm start's weights c_1..c_K (host only, no GPU): x_n ~ x_{n-1} + sum c_j d_{n - j P},
 * d the steps between successive solutions of one RK stage, exact for steps of degree < order
 * and of least norm among such weights. K = max(window, order) values into out. No counterpart
 * in the reference (its solves start from the previous solution, dispersion.rs:220-222). */
int wc_warm_start_weights(int32_t order, int32_t window, double* out);
/* Selects the arithmetic (WC_PRECISION_*); FP64 by default. Switching converts the held state. */
int wc_solver_set_precision(wc_solver* s, int32_t precision);

/* The state the solver steps (State, state.rs:7-17): padded h, hu, hv, breaking and the time.
 * breaking may be NULL on set (zeros). Any output pointer may be NULL on get. */
int wc_state_set(wc_solver* s, const double* h, const double* hu, const double* hv,
                 const double* breaking, double time);
int wc_state_get(wc_solver* s, double* h, double* hu, double* hv, double* breaking,
                 double* time);
/* Sets only the time of the state held by the solver. */
int wc_state_set_time(wc_solver* s, double time);

/* Solver::stable_dt (solver.rs:434-450): cfl / max rate, +inf if nothing moves, NaN if any
 * cell is NaN, 0 if any moves infinitely fast. */
int wc_stable_dt(wc_solver* s, double* dt);
/* wc_stable_dt's value for a padded FP64 state (h, hu, hv: (nx+6)(ny+6) doubles each) that the
 * caller keeps in device memory of the current device, written to dt_out (device memory) by
 * work enqueued on `stream` (a hipStream_t, NULL for the default stream); the call does not
 * wait for it. No counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid it), it allocates nothing and keeps
 * its partial results in a fixed 128 KiB device array per GPU, held from its first call on that
 * GPU until the process ends. Limit: nx * ny <= 67,108,864 interior cells (WC_EINVAL beyond).
 * Calls that overlap on one GPU (other streams, other threads) race for that array. */
int wc_stable_dt_noalloc(const double* h, const double* hu, const double* hv, uint64_t nx,
                         uint64_t ny, double dx, double dy, double cfl, double* dt_out,
                         void* stream);
/* Solver::step (solver.rs:481-507): one SSP-RK2 step (Euler for Order::First) of size dt,
 * with sponges and friction, advancing the time by dt. */
int wc_step(wc_solver* s, double dt);
/* Solver::limiting_cell (solver.rs:455-477); *found = 0 where the reference returns None. */
int wc_limiting_cell(wc_solver* s, int32_t* found, wc_limiting_t* out);
/* Solver::breaking_indicator (solver.rs:341-349): nx*ny values, row-major with y outer. */
int wc_breaking_indicator(wc_solver* s, double* out);
/* Solver::dispersion_stats (solver.rs:392-396); *has = 0 for shallow water (None). */
int wc_dispersion_stats(wc_solver* s, int32_t* has, wc_dispersion_stats_t* out);
/* State::volume / max_abs_momentum / min_depth (state.rs:79-102) of the held state. The
 * volume sums the interior in the GPU's canonical order (64x4 tiles, wavefront butterflies),
 * not serially: it agrees with the reference's serial sum to rounding. */
int wc_state_volume(wc_solver* s, double* out);
int wc_state_max_abs_momentum(wc_solver* s, double* out);
int wc_state_min_depth(wc_solver* s, double* out);

/* Free functions (forcing.rs:53-75, 294-297). wc_linear_wave returns WC_EINVAL where the
 * reference panics (a wave too short for the dispersive model). */
int wc_linear_wave(double omega, double h, int32_t dispersive, wc_linear_wave_t* out);
double wc_manning_factor(double h, double speed, double dt, double n);

/* Runs `steps` iterations of { dt = stable_dt; step(dt) } on the device without returning to
 * the caller in between (the loop of crates/wavecore/tests/common/mod.rs:53-66 and
 * wavesim's Stage::step, lib.rs:1385-1434). Stops early, with WC_OK, if dt is not finite and
 * positive; *done_steps reports how many steps ran and *last_dt the last dt. */
int wc_run_steps(wc_solver* s, uint64_t steps, uint64_t* done_steps, double* last_dt);

/* The dispersive solve's preconditioner (no counterpart in the reference, whose PCG uses the
 * Jacobi diagonal, dispersion.rs:568-589). WC_PRECOND_NETWORK: the learned network of
 * include/wavecore_unet.h (params: wcu_param_count() FP32 values, host memory) inside a
 * flexible PCG with the same tolerance, falling back to Jacobi for the rest of a solve whose
 * residual stops decreasing; used by the FP32 form (the FP64 path keeps Jacobi, to stay equal
 * to the oracle). Features: depth / (4 dx), WC_BED_FEATURE_SCALE times the bed's centred
 * gradient, the residual over its RMS; the output times RMS / (4 dx) is the correction.
 * wc_solver_precond_fallbacks: solves that fell back, since the solver was made. */
#define WC_PRECOND_JACOBI 0
#define WC_PRECOND_NETWORK 1
#define WC_BED_FEATURE_SCALE 10.0
int wc_solver_set_preconditioner(wc_solver* s, int32_t kind, const float* params, uint64_t n);
int wc_solver_precond_fallbacks(wc_solver* s, uint64_t* out);
/* Training pairs from the solver's own solves (FP32 form, Jacobi PCG): with the option
 * "capture" set to 1, each solve records its depth and, before its iterations 0, 2, 5 and 10
 * (slots 0..3), the residual r_n = b - A w_n and the iterate w_n. wc_solver_capture copies the
 * last solve's slot into host arrays of (nx+6)*(ny+6) floats (any may be NULL), the error as
 * w* - w_n with w* the solve's solution; *valid = 0 if the solve ended before that iteration.
 * Options "precond_stall" and "precond_ratio": the network falls back to Jacobi after
 * precond_stall iterations in a row (default 2, 0 = never) that each keep more than
 * precond_ratio percent (default 70) of the residual norm. wc_solver_residual_history: r.r before each iteration of
 * the last network-preconditioned solve (*n values, at most max). */
int wc_solver_capture(wc_solver* s, int32_t slot, float* h, float* rx, float* ry, float* ex,
                      float* ey, int32_t* valid);
int wc_solver_residual_history(wc_solver* s, double* out, uint64_t max, uint64_t* n);

/* The run's per-step statistics (wavesim's Stage::record and Stats, crates/wavesim/src/
 * lib.rs:849-937, 1099-1130, 1276-1300), accumulated on the GPU for the solver's state:
 *   wc_stats_new     recording from time `from`, waves of `period` s (the first ends at
 *                    from + period); elevation: the nx*ny bed elevation (f32, wavesim's
 *                    Bed::elevation) that eta = h + elevation uses, or NULL for the solver's bed.
 *   wc_stats_record  Stage::record after a step of dt has ended (the solver's time is its end).
 *   wc_stats_read    seconds recorded, waves closed, and per cell (nx*ny, any NULL): breaking
 *                    seconds (indicator >= 0.8), eta max and min, the mean, first and last wave
 *                    heights (crest to trough per period).
 *   wc_stats_harmonic FirstHarmonic::amplitudes: (a, -b) per cell, *ok = 0 for None.
 *   wc_stats_finish  the end of a run (Stage::finish, lib.rs:1451-1453): closes the wave in
 *                    progress if its period has ended (to within 1e-9 s).
 * Freeing the solver first leaves its statistics detached: their calls then fail with
 * WC_EINVAL and wc_stats_free stays safe.
 * StepWatch (lib.rs:849-896): wc_watch_init with the still sea's step; wc_watch_collapsed counts
 * a stable step and sets *ratio > 0 at the end of a block of WC_WATCH_STEPS whose mean step was
 * more than WC_COLLAPSE times shorter. wc_stage_step is Stage::next_dt + advance + record
 * (lib.rs:1385-1434) without frames: *status WC_STAGE_OK (stepped by *dt), _UNSTABLE (stable
 * dt NaN or <= 0), _NO_WATER (+inf), _COLLAPSED (*ratio set); st and w may be NULL. */
typedef struct wc_stats wc_stats;
typedef struct {
  double reference;
  uint32_t steps;
  double seconds;
} wc_watch_t;
#define WC_WATCH_STEPS 1000
#define WC_COLLAPSE 5.0
#define WC_STAGE_OK 0
#define WC_STAGE_UNSTABLE 1
#define WC_STAGE_NO_WATER 2
#define WC_STAGE_COLLAPSED 3
int wc_stats_new(wc_solver* s, double from, double period, const float* elevation,
                 wc_stats** out);
void wc_stats_free(wc_stats* st);
int wc_stats_record(wc_stats* st, double dt);
int wc_stats_read(wc_stats* st, double* seconds, uint64_t* waves, double* breaking_seconds,
                  double* eta_max, double* eta_min, double* wave_mean, double* wave_first,
                  double* wave_last);
int wc_stats_harmonic(wc_stats* st, int32_t* ok, double* a, double* b);
int wc_stats_finish(wc_stats* st);
int wc_watch_init(wc_watch_t* w, double reference);
int wc_watch_collapsed(wc_watch_t* w, double stable, double* ratio);
int wc_stage_step(wc_solver* s, wc_stats* st, wc_watch_t* w, double until, double* dt,
                  int32_t* status, double* ratio);

/* Relaxation zones (forcing.rs:227-284): the water in the zones is pulled towards targets the
 * caller supplies at every step (one-way nesting, crates/wavesim/src/lib.rs:1894-1897, and
 * the lab's wave paddle, crates/wavecore/tests/lab.rs:253-271). A relaxation belongs to the
 * solver it was made for and acts on that solver's resident state.
 *   wc_relaxation_new   Relaxation::new (forcing.rs:239-248): every interior cell where one of
 *                       the zones has a rate, at the largest of their rates, in the reference's
 *                       row order. Zones are Sponges (rate only; their level is not used).
 *   wc_relaxation_len / wc_relaxation_cells
 *                       Relaxation::cells (forcing.rs:252-254): padded (i, j) of each cell, in
 *                       the order apply takes the targets.
 *   wc_relaxation_apply Relaxation::apply (forcing.rs:259-283) over dt: targets holds
 *                       (eta, u, v) per cell, interleaved (3 * len doubles), in host memory;
 *                       wc_relaxation_apply_device takes them in device memory of the
 *                       solver's GPU (a coarse run's output, say) and does not copy them. */
typedef struct wc_relaxation wc_relaxation;
int wc_relaxation_new(wc_solver* s, const wc_sponge_t* zones, uint64_t n_zones,
                      wc_relaxation** out);
void wc_relaxation_free(wc_relaxation* r);
uint64_t wc_relaxation_len(const wc_relaxation* r);
int wc_relaxation_cells(const wc_relaxation* r, uint64_t* i, uint64_t* j);
int wc_relaxation_apply(wc_relaxation* r, double dt, const double* targets);
int wc_relaxation_apply_device(wc_relaxation* r, double dt, const double* targets);

/* The coarse-to-fine sampler of nested runs (wavesim's Sampler, crates/wavesim/src/
 * lib.rs:1548-1616, and its use after each fine step, lib.rs:1866-1897), for a relaxation zone
 * of the fine solver:
 *   wc_sampler_new    Sampler::new: the zone's cells, the fine grid's first column fine_x0 in
 *                     the bed file and its cell size, the coarse grid (cropped by `crop`
 *                     columns) and its cell size; four coarse corners and bilinear weights per
 *                     zone cell, clamped to one ghost layer.
 *   wc_sampler_sample Sampler::sample of the coarse solver's current state (ghosts as walls)
 *                     into slot 0 (before the coarse step that brackets the fine step's end)
 *                     or 1 (after it), with the state's time.
 *   wc_sampler_apply  the fine step's relaxation: targets interpolated at (end - t0) / (t1 - t0)
 *                     clamped to [0, 1] (the nearer sample where one is dry), then
 *                     Relaxation::apply with dt on the zone's solver.
 *   wc_sampler_read   a sample (slot 0, 1) or the last targets (2): 3 doubles per zone cell. */
typedef struct wc_sampler wc_sampler;
int wc_sampler_new(const wc_relaxation* zone, uint64_t fine_x0, double fine_dx, double fine_dy,
                   uint64_t coarse_nx, uint64_t coarse_ny, double coarse_dx, double coarse_dy,
                   uint64_t crop, wc_sampler** out);
void wc_sampler_free(wc_sampler* sm);
int wc_sampler_sample(wc_sampler* sm, wc_solver* coarse, int32_t slot);
int wc_sampler_apply(wc_sampler* sm, wc_relaxation* zone, double end, double dt);
int wc_sampler_read(const wc_sampler* sm, int32_t slot, double* out, double* time);

/* Many instances in one launch (the batched sweep of independent runs: beds, swell angles,
 * tides, grids of uneven size). A batch groups existing solvers of one device, all in the FP32
 * form with the one-kernel Jacobi PCG, sharing the time stepping (Euler or RK2), the switch and
 * the dispersion; each keeps its own grid, bed, settings, state and dt. wc_batch_run_steps
 * advances every instance `steps` times with its own stable dt (as wc_run_steps does for one):
 * every phase of a step is one launch over all instances, whose blocks walk each instance's
 * tiles, so an instance ends bit for bit where it would alone. An instance whose dt is not
 * finite and positive stops there (done_steps[q], last_dt[q] per instance, both may be NULL).
 * While a batch steps its solvers, they must not be used on their own; afterwards their states
 * are read with wc_state_get as usual. The instances share the flux layout too: with
 * flux_kernel 3 (the default) the flux marches one wave per 58-column strip and band of rows.
 * The first run tunes, by timing candidates, the marching flux's band height (the tile shape of
 * its waves, 58 x rows cells; wc_batch_flux_rows reports it and the timings,
 * wc_batch_set_option("flux_rows", r) fixes it, 0 tunes it again), the SGN set-up's tile height
 * (wc_batch_setup_rows, option "setup_rows") and the launch widths
 * (persistent grids of the 64x4 tile flux and the PCG launches; the other launches run a block
 * per virtual block); wc_batch_launch_widths reports them, wc_batch_set_option("grid_nt" |
 * "grid_pt", w) fixes them ("grid_nt" then for every 256-thread launch), ("tune", 1) tunes
 * again, ("pcg_batch", k) sets the iterations queued between looks at the PCG flags. None of
 * these choices changes a bit of the results. */
typedef struct wc_batch wc_batch;
int wc_batch_new(wc_solver* const* solvers, uint64_t n, wc_batch** out);
void wc_batch_free(wc_batch* b);
int wc_batch_set_option(wc_batch* b, const char* key, int64_t value);
int wc_batch_launch_widths(const wc_batch* b, int32_t* grid_nt, int32_t* grid_pt);
/* The marching flux's band height in use (0 with the tile flux) and the tuner's (height,
 * microseconds per launch) measurements, as wc_solver_flux_rows. */
int wc_batch_flux_rows(const wc_batch* b, int32_t* rows, int32_t* tried_rows, double* tried_us,
                       int32_t capacity, int32_t* tried);
/* The SGN set-up's tile height in use (64 x rows cells a block: 4, 8 or 16, tuned at the first
 * run unless wc_batch_set_option("setup_rows", r) fixed it; 0 tunes it again) and the tuner's
 * (height, microseconds per launch) measurements, as wc_batch_flux_rows. */
int wc_batch_setup_rows(const wc_batch* b, int32_t* rows, int32_t* tried_rows, double* tried_us,
                        int32_t capacity, int32_t* tried);
int wc_batch_run_steps(wc_batch* b, uint64_t steps, uint64_t* done_steps, double* last_dt);
/* Time spent in the batch's PCG iteration launches since wc_batch_set_option("profile", 1),
 * each launch between its own pair of events (while profiling, the stretches of a step are
 * issued directly, not from the graphs). Measurement only; no counterpart in the reference. */
int wc_batch_profile_read(wc_batch* b, double* total_ms, uint64_t* launches);

/* Self-check of the flux kernels' division by a divisor every cell shares (the 60 of the
 * fifth-order formula, dx, dy; a reciprocal computed once, three instructions per quotient where
 * the division's scaling and fix-up steps would change nothing): shared[i] = x[i] / y that way and
 * plain[i] = x[i] / y by the device's own division, on the current device, for n <= 2^26
 * values. The two must agree bit for bit. No counterpart in the reference. */
int wc_check_shared_division(double y, const double* x, uint64_t n, double* shared,
                             double* plain);
/* No counterpart in the reference: what plain streams reach on this GPU's HBM, the yardstick
 * for the stencil kernels' bytes. Four device arrays of `bytes` each (a multiple of 16, at most
 * 16 GiB) are allocated for the call; each test runs once untimed and `reps` times timed with
 * `grid` blocks of 256 threads looping over the arrays 16 B a thread a step. tbs[8] receives
 * TB/s moved: copy, copy with nontemporal loads and stores, read, read nontemporal, write,
 * write nontemporal, three reads and a write, the same nontemporal. */
int wc_stream_ceilings(uint64_t bytes, int32_t grid, int32_t reps, double* tbs);

/* Measurement hooks for benchmarks: device time in milliseconds of each kernel family over the
 * steps since the last reset, measured with HIP events on the solver's stream, and the kernel
 * launch counts. names receives a ';'-separated list. on = 1 puts a pair of events around every
 * launch; on = 2 times only the PCG iterations (the FP32 form's one kernel, "cg_iter"; the FP64
 * path's matvec and update, "cg_pair"), one pair of events around each batch of back-to-back
 * launches, which leaves a timed run nearly undisturbed; 0 (the default) times nothing. */
int wc_profile_enable(wc_solver* s, int32_t on);
int wc_profile_reset(wc_solver* s);
int wc_profile_read(wc_solver* s, char* names, uint64_t names_len, double* total_ms,
                    uint64_t* launches, uint64_t max_kernels, uint64_t* n_kernels);

#ifdef __cplusplus
}
#endif

#endif
