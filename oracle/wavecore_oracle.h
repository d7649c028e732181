/* TEST INFRASTRUCTURE ONLY — the CPU FP64 oracle for wavecore's time step.
 *
 * A from-scratch C++ restatement of the reference's step (crates/wavecore/src/solver.rs,
 * dispersion.rs, forcing.rs, grid.rs, state.rs, bathymetry.rs, par.rs), built with
 * -ffp-contract=off so that, like Rust, it never fuses a*b+c. Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it, and only as the checker or as the timed
 * CPU baseline: the product path never calls it.
 *
 * Two switches select how the oracle evaluates the two things the GPU cannot reproduce from
 * the reference's text alone:
 *   - dot order (wco_set_dot_order): 0 = the reference's row order (each row summed serially,
 *     then the rows in order from 0.0, par.rs:90-98); 1 = the GPU's canonical order (64x4-cell
 *     tiles, a 64-lane xor butterfly per tile row adding lane l^1 first and l^32 last, rows
 *     paired, tiles summed by 256 strided lanes and a final butterfly), which the GPU kernels
 *     follow bit for bit; 2 = the order of the GPU's default FP64 PCG (64x16-cell tiles taken
 *     by 256 strided lanes, a running sum per cell position over the lane's tiles, each row of
 *     64 by the butterfly, the 16 rows in order, the lanes as in 1). The volume sums in order 1
 *     whenever the dot order is not 0;
 *   - libm (wco_set_math): 0 = glibc's hypot and pow (what the Rust reference calls on Linux);
 *     1 = the portable IEEE-only hypot and pow(h, 4/3) that the GPU evaluates.
 * exp, sin and cos only ever see host-known arguments and always come from glibc, in the
 * oracle and in the product alike.
 *
 * Same entry points and argument meaning as the product's C ABI (include/wavecore_hip.h) with
 * the prefix wco_ instead of wc_, plus white-box entry points for the unit tests.
 */
#ifndef WAVECORE_ORACLE_H
#define WAVECORE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wco_solver wco_solver;

typedef struct {
  double x, amplitude, period, angle, sigma, ramp_periods;
} wco_wavemaker_t;

typedef struct {
  uint64_t width;
  double strength, level;
  int32_t west, east, south, north;
} wco_sponge_t;

typedef struct {
  double eta_ratio_lo, eta_ratio_hi, slope_lo, slope_hi;
} wco_breaking_t;

typedef struct {
  wco_breaking_t breaking;
  int32_t nonlinear;
  int32_t reserved;
  double tolerance;
  uint64_t max_iterations;
} wco_dispersion_t;

typedef struct {
  uint64_t solves, iterations, unconverged;
} wco_dispersion_stats_t;

/* EddyViscosity (eddy.rs:63-89). */
typedef struct {
  double initial, settled, duration, mixing, linger;
} wco_eddy_t;

typedef struct {
  uint64_t i, j;
  double depth, speed;
} wco_limiting_t;

typedef struct {
  double k, phase_speed, group_speed;
} wco_linear_wave_t;

const char* wco_last_error(void);

int wco_solver_new(uint64_t nx, uint64_t ny, double dx, double dy, const double* bed_padded,
                   wco_solver** out);
void wco_solver_free(wco_solver* s);
int wco_set_cfl(wco_solver* s, double cfl);
int wco_with_order(wco_solver* s, int32_t order);
int wco_with_manning(wco_solver* s, double n);
int wco_with_sponge(wco_solver* s, const wco_sponge_t* sp);
int wco_with_wavemaker(wco_solver* s, const wco_wavemaker_t* wm);
int wco_with_dispersion(wco_solver* s, const wco_dispersion_t* d);
int wco_with_periodic_y(wco_solver* s);
int wco_with_shallow_cells(wco_solver* s, const uint8_t* cells, uint64_t n);
int wco_bed(const wco_solver* s, double* out);

/* Oracle switches (see above) and OpenMP threads for the row loops (1 = serial). */
int wco_set_dot_order(wco_solver* s, int32_t order);
int wco_set_math(wco_solver* s, int32_t mode);
// The FP32 form's extrapolated warm start (order 0: the reference's previous solution): weights
// exact for solution steps of degree < order, of least norm over max(order, window) past steps;
// set it after the order and the dispersion. wco_ws_weights: those weights. wco_ws_extrapolate:
// the start of solve n from past[i] = x_{n-1-i}, i < n_past (one cell), as the solver computes it.
int wco_set_warm_start(wco_solver* s, int32_t order, int32_t window);
int wco_ws_weights(int32_t order, int32_t window, double* out);
int wco_ws_extrapolate(int32_t order, int32_t window, int32_t period, uint64_t n_past,
                       const double* past, double* out);
int wco_set_threads(wco_solver* s, int32_t threads);

int wco_state_set(wco_solver* s, const double* h, const double* hu, const double* hv,
                  const double* breaking, double time);
int wco_state_get(const wco_solver* s, double* h, double* hu, double* hv, double* breaking,
                  double* time);
int wco_stable_dt(wco_solver* s, double* dt);
int wco_step(wco_solver* s, double dt);
int wco_limiting_cell(wco_solver* s, int32_t* found, wco_limiting_t* out);
/* Breaking::viscosity = Some(e) (NULL: None), after wco_with_dispersion (solver.rs:512-532). */
int wco_with_eddy_viscosity(wco_solver* s, const wco_eddy_t* e);
/* EddyViscosity::follow and diffuse on padded arrays, for the unit tests (eddy.rs:163-302). */
int wco_eddy_follow(uint64_t nx, uint64_t ny, double dx, double dy, const wco_eddy_t* e,
                    double* h, double* hu, double* hv, double* breaking, const double* phi,
                    double dt, double* nu_out);
int wco_eddy_diffuse(uint64_t nx, uint64_t ny, double dx, double dy, double* h, double* hu,
                     double* hv, const double* nu, double dt, uint64_t* steps);
int wco_breaking_indicator(wco_solver* s, double* out);
int wco_dispersion_stats(const wco_solver* s, int32_t* has, wco_dispersion_stats_t* out);
int wco_state_volume(wco_solver* s, double* out);
int wco_state_max_abs_momentum(wco_solver* s, double* out);
int wco_state_min_depth(wco_solver* s, double* out);
int wco_fill_boundaries(wco_solver* s);

/* Relaxation (forcing.rs:227-284), acting on the oracle solver's state. */
typedef struct wco_relaxation wco_relaxation;
int wco_relaxation_new(wco_solver* s, const wco_sponge_t* zones, uint64_t n_zones,
                       wco_relaxation** out);
void wco_relaxation_free(wco_relaxation* r);
uint64_t wco_relaxation_len(const wco_relaxation* r);
int wco_relaxation_cells(const wco_relaxation* r, uint64_t* i, uint64_t* j);
int wco_relaxation_apply(wco_relaxation* r, double dt, const double* targets);

/* wavesim's per-step statistics (crates/wavesim/src/lib.rs:899-932, 1099-1160, 1276-1300) on
 * the oracle solver's state; the product's wc_stats_* with the prefix wco_. */
typedef struct wco_stats wco_stats;
int wco_stats_new(wco_solver* s, double from, double period, const float* elevation,
                  wco_stats** out);
void wco_stats_free(wco_stats* st);
int wco_stats_record(wco_stats* st, double dt);
int wco_stats_read(wco_stats* st, double* seconds, uint64_t* waves, double* breaking_seconds,
                   double* eta_max, double* eta_min, double* wave_mean, double* wave_first,
                   double* wave_last);
int wco_stats_harmonic(wco_stats* st, int32_t* ok, double* a, double* b);
// Stage::finish (lib.rs:1451-1453): the wave in progress closes if its period has ended.
int wco_stats_finish(wco_stats* st);

int wco_linear_wave(double omega, double h, int32_t dispersive, wco_linear_wave_t* out);
double wco_manning_factor(double h, double speed, double dt, double n);

/* White-box entry points for the unit tests. */
void wco_face_values(const double v[5], double blend, double* lo, double* hi);
double wco_smoothstep(double lo, double hi, double x);
double wco_hypot(double a, double b, int32_t portable);
double wco_pow43(double h, int32_t portable);
/* The GPU switch's shortcut and the full expression it replaces (er_lo, er_hi, sl_lo, sl_hi in
 * thresholds), cell by cell. */
int wco_switch_off(const double* thresholds, const double* ratio, const double* ex,
                   const double* ey, uint64_t n, double* shortcut, double* full);
/* A p for the SGN operator on a grid, with the given heff, mask and c (padded arrays);
 * p is copied, its ghosts filled, and the result written to ax, ay (interior). */
int wco_matvec(uint64_t nx, uint64_t ny, double dx, double dy, int32_t periodic,
               const double* heff, const double* mask, const double* c, const double* px,
               const double* py, double* ax, double* ay);
double wco_dot(uint64_t nx, uint64_t ny, int32_t order, const double* ax, const double* ay,
               const double* bx, const double* by);
/* Dispersion::apply on a given state, phi and rates, with fresh scratch (warm start zero). */
int wco_dispersion_apply(uint64_t nx, uint64_t ny, double dx, double dy, int32_t periodic,
                         const wco_dispersion_t* d, const double* h, const double* hu,
                         const double* hv, const double* phi, const double* lh, double* lhu,
                         double* lhv, int32_t dot_order);

#ifdef __cplusplus
}
#endif

#endif
