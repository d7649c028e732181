// TEST INFRASTRUCTURE ONLY — CPU FP64 oracle for wavecore's time step. See wavecore_oracle.h.
//
// Every function below restates one function of the reference in C++, in the reference's
// order of floating-point operations, and cites it. Loops that the reference runs through
// par.rs row helpers run here as OpenMP row loops (bit-identical at any thread count, as
// par.rs:1-6 promises for rayon).
//
// Parity status of each part (how it is pinned without the reference binary) is listed in the
// test files under tests/ that exercise it.

#include "wavecore_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../wavesim_amd/csrc/wc_math.h"

namespace {

constexpr size_t GH = 3;
constexpr double G = 9.81;
constexpr double H_DRY = 1e-8;
constexpr double H_SECOND = 1e-3;
constexpr double H_FACE_MIN = 1e-4;
constexpr double H_DISP_MIN = 0.05;
constexpr double H_FRIC_MIN = 1e-3;
constexpr double TAU = 6.283185307179586476925286766559;
constexpr double PI = 3.14159265358979323846264338327950;
constexpr size_t MIN_CELLS = 40000;  // par.rs:13

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

// grid.rs:12-58
struct Grid {
  size_t nx = 0, ny = 0;
  double dx = 0, dy = 0;
  size_t width() const { return nx + 2 * GH; }
  size_t rows() const { return ny + 2 * GH; }
  size_t cells() const { return width() * rows(); }
  size_t idx(size_t i, size_t j) const { return j * width() + i; }
  double cx(size_t i) const { return ((double)(i - GH) + 0.5) * dx; }
  double cy(size_t j) const { return ((double)(j - GH) + 0.5) * dy; }
};

// Which math the oracle uses where the platform libm enters (see the header).
struct Mathlib {
  bool portable = false;
  double hypot(double a, double b) const { return portable ? wc_hypot(a, b) : ::hypot(a, b); }
  double pow43(double h) const { return portable ? wc_pow43(h) : ::pow(h, 4.0 / 3.0); }
};

// Row-loop threading, as par.rs: threaded only at >= MIN_CELLS padded cells.
struct Par {
  int threads = 1;
  bool on(size_t cells) const { return threads > 1 && cells >= MIN_CELLS; }
};

// grid.rs:61-75
void mirror_ghosts(const Grid& g, double* a) {
  const size_t nx = g.nx, ny = g.ny, w = g.width();
  for (size_t j = GH; j < GH + ny; ++j) {
    for (size_t k = 0; k < GH; ++k) {
      a[g.idx(GH - 1 - k, j)] = a[g.idx(GH + k, j)];
      a[g.idx(GH + nx + k, j)] = a[g.idx(GH + nx - 1 - k, j)];
    }
  }
  for (size_t i = 0; i < w; ++i) {
    for (size_t k = 0; k < GH; ++k) {
      a[g.idx(i, GH - 1 - k)] = a[g.idx(i, GH + k)];
      a[g.idx(i, GH + ny + k)] = a[g.idx(i, GH + ny - 1 - k)];
    }
  }
}

// grid.rs:79-87
void wrap_ghosts_y(const Grid& g, double* a) {
  const size_t ny = g.ny, w = g.width();
  for (size_t i = 0; i < w; ++i) {
    for (size_t k = 0; k < GH; ++k) {
      a[g.idx(i, GH - 1 - k)] = a[g.idx(i, GH + ny - 1 - k)];
      a[g.idx(i, GH + ny + k)] = a[g.idx(i, GH + k)];
    }
  }
}

// Negate the x component in the x ghost columns and the y component in the y ghost rows
// (state.rs:50-63, grid.rs:95-109).
void flip_normals(const Grid& g, double* vx, double* vy) {
  const size_t nx = g.nx, ny = g.ny, w = g.width();
  for (size_t j = 0; j < ny + 2 * GH; ++j) {
    for (size_t k = 0; k < GH; ++k) {
      const size_t l = g.idx(GH - 1 - k, j), r = g.idx(GH + nx + k, j);
      vx[l] = -vx[l];
      vx[r] = -vx[r];
    }
  }
  for (size_t i = 0; i < w; ++i) {
    for (size_t k = 0; k < GH; ++k) {
      const size_t lo = g.idx(i, GH - 1 - k), hi = g.idx(i, GH + ny + k);
      vy[lo] = -vy[lo];
      vy[hi] = -vy[hi];
    }
  }
}

// grid.rs:92-114
void fill_vector_ghosts(const Grid& g, double* vx, double* vy, bool periodic) {
  mirror_ghosts(g, vx);
  mirror_ghosts(g, vy);
  flip_normals(g, vx, vy);
  if (periodic) {
    wrap_ghosts_y(g, vx);
    wrap_ghosts_y(g, vy);
  }
}

// state.rs:7-17
struct State {
  std::vector<double> h, hu, hv, breaking;
  double time = 0.0;
  void zeros(const Grid& g) {
    h.assign(g.cells(), 0.0);
    hu.assign(g.cells(), 0.0);
    hv.assign(g.cells(), 0.0);
    breaking.assign(g.cells(), 0.0);
    time = 0.0;
  }
  // state.rs:44-64
  void fill_walls(const Grid& g) {
    mirror_ghosts(g, h.data());
    mirror_ghosts(g, hu.data());
    mirror_ghosts(g, hv.data());
    mirror_ghosts(g, breaking.data());
    flip_normals(g, hu.data(), hv.data());
  }
  // state.rs:68-76
  void fill_boundaries(const Grid& g, bool periodic) {
    fill_walls(g);
    if (periodic) {
      wrap_ghosts_y(g, h.data());
      wrap_ghosts_y(g, hu.data());
      wrap_ghosts_y(g, hv.data());
      wrap_ghosts_y(g, breaking.data());
    }
  }
};

// EddyViscosity (eddy.rs:63-89) and its threshold (eddy.rs:108-131)
struct Eddy {
  double initial = 0.65, settled = 0.15, duration = 5.0, mixing = 1.2, linger = 5.0;
  double span(double h) const { return duration * sqrt(h / G); }
  double threshold(double age, double h) const {
    const double sp = span(h);
    if (age >= sp) return settled;
    return initial + (settled - initial) * age / sp;
  }
};

struct Breaking {
  double er_lo = 0.30, er_hi = 0.55, sl_lo = 0.40, sl_hi = 0.65;  // dispersion.rs:67-73
  bool viscosity = false;  // Breaking::viscosity is Some(eddy)
  Eddy eddy;
};

struct DispersionCfg {
  Breaking breaking;
  bool nonlinear = true;
  double tolerance = 1e-4;
  uint64_t max_iterations = 200;  // dispersion.rs:88-97
};

// forcing.rs:146-207
struct Sponge {
  uint64_t width = 0;
  double strength = 0, level = 0;
  bool west = false, east = false, south = false, north = false;
  double ramp(uint64_t d) const {
    if (d < width) {
      const double w = (double)width;
      const double f = (w - (double)d) / w;
      return f * f;
    }
    return 0.0;
  }
  double rate(const Grid& g, size_t i, size_t j) const {
    const size_t a = i - GH, b = j - GH;
    double r = 0.0;
    if (west) r = wc_max(r, ramp(a));
    if (east) r = wc_max(r, ramp(g.nx - 1 - a));
    if (south) r = wc_max(r, ramp(b));
    if (north) r = wc_max(r, ramp(g.ny - 1 - b));
    return strength * r;
  }
};

struct WaveMaker {
  double x, amplitude, period, angle, sigma, ramp_periods;
};

// forcing.rs:53-75
bool linear_wave(double omega, double h, bool dispersive, wco_linear_wave_t* out) {
  if (!dispersive) {
    const double c = sqrt(G * h);
    out->k = omega / c;
    out->phase_speed = c;
    out->group_speed = c;
    return true;
  }
  const double nu = omega * omega * h / G;
  if (!(nu < 3.0)) return false;
  const double kh = sqrt(nu / (1.0 - nu / 3.0));
  const double k = kh / h;
  const double phase = omega / k;
  out->k = k;
  out->phase_speed = phase;
  out->group_speed = phase / (1.0 + kh * kh / 3.0);
  return true;
}

// forcing.rs:77-139
struct Drive {
  WaveMaker spec{};
  double omega = 0, ky = 0, strength = 0;
  bool init(const WaveMaker& s, const Grid& g, double still_depth, bool dispersive) {
    if (!(still_depth > 0.0)) return fail("wave-maker must sit in water") == 0;
    if (!(s.period > 0.0 && s.sigma > 0.0 && s.amplitude >= 0.0))
      return fail("wave-maker needs period > 0, sigma > 0 and amplitude >= 0") == 0;
    spec = s;
    omega = TAU / s.period;
    wco_linear_wave_t lw;
    if (!linear_wave(omega, still_depth, dispersive, &lw))
      return fail("the wave is too short for the dispersive model at the wave-maker") == 0;
    const double k = lw.k, c = lw.group_speed;
    // angle.sin() and angle.cos(): one sincos call, as LLVM lowers the pair for glibc (its
    // last bit can differ from separate sin and cos calls)
    double sa, ca;
    sincos(s.angle, &sa, &ca);
    ky = k * sa;
    const double kx = k * ca;
    const double rolloff = exp(0.5 * kx * kx * s.sigma * s.sigma);
    const double flux = 2.0 * c * s.amplitude * ca * rolloff;
    double norm = -0.0;  // Rust's float Sum folds from -0.0
    for (size_t i = 0; i < g.nx; ++i) {
      const double x = ((double)i + 0.5) * g.dx;
      const double d = x - s.x;
      norm = norm + exp(-(d * d) / (2.0 * s.sigma * s.sigma)) * g.dx;
    }
    strength = flux / norm;
    return true;
  }
  double source(double x, double y, double t) const {
    const double d = x - spec.x;
    const double profile = exp(-(d * d) / (2.0 * spec.sigma * spec.sigma));
    if (profile < 1e-12) return 0.0;
    const double ramp_time = spec.ramp_periods * spec.period;
    const double ramp = t >= ramp_time ? 1.0 : 0.5 * (1.0 - cos(PI * t / ramp_time));
    return strength * profile * ramp * sin(omega * t - ky * y);
  }
};

// ---------------------------------------------------------------- reconstruction and flux
enum { ORDER_FIRST = 0, ORDER_SECOND = 1, ORDER_FIFTH = 2 };

struct Cell {
  double h, un, ut, b, phi;
};
struct Side {
  double h, un, ut, z;
};
struct Face {
  double flux[3];
  double corr_left, corr_right;
};

// solver.rs:102-116
double minmod3(double a, double b, double c) {
  if (a > 0.0 && b > 0.0 && c > 0.0) return wc_min(wc_min(a, b), c);
  if (a < 0.0 && b < 0.0 && c < 0.0) return wc_max(wc_max(a, b), c);
  return 0.0;
}
double limited_slope(double dl, double dh) { return minmod3(2.0 * dl, 0.5 * (dl + dh), 2.0 * dh); }

// solver.rs:125-138
void face_values(const double v[5], double blend, double* lo, double* hi) {
  const double low2 = v[0], low = v[1], c = v[2], high = v[3], high2 = v[4];
  const double s = limited_slope(c - low, high - c);
  const double mc_lo = c - 0.5 * s, mc_hi = c + 0.5 * s;
  if (blend <= 0.0) {
    *lo = mc_lo;
    *hi = mc_hi;
    return;
  }
  const double fifth_hi = (2.0 * low2 - 13.0 * low + 47.0 * c + 27.0 * high - 3.0 * high2) / 60.0;
  const double fifth_lo = (2.0 * high2 - 13.0 * high + 47.0 * c + 27.0 * low - 3.0 * low2) / 60.0;
  *lo = mc_lo + blend * (fifth_lo - mc_lo);
  *hi = mc_hi + blend * (fifth_hi - mc_hi);
}

// solver.rs:146-183
void reconstruct(const Cell* cells, int order, Side* lo_out, Side* hi_out) {
  const Cell &low2 = cells[0], &low = cells[1], &c = cells[2], &high = cells[3],
             &high2 = cells[4];
  const Side first{c.h, c.un, c.ut, c.b};
  if (order == ORDER_FIRST || low.h < H_SECOND || c.h < H_SECOND || high.h < H_SECOND) {
    *lo_out = first;
    *hi_out = first;
    return;
  }
  double blend = 0.0;
  if (order == ORDER_FIFTH && low2.h >= H_SECOND && high2.h >= H_SECOND) {
    blend = 1.0;
    for (int m = 0; m < 5; ++m) blend = wc_min(blend, cells[m].phi);
  }
  double v[5], eta_lo, eta_hi, un_lo, un_hi, ut_lo, ut_hi;
  for (int m = 0; m < 5; ++m) v[m] = cells[m].h + cells[m].b;
  face_values(v, blend, &eta_lo, &eta_hi);
  for (int m = 0; m < 5; ++m) v[m] = cells[m].un;
  face_values(v, blend, &un_lo, &un_hi);
  for (int m = 0; m < 5; ++m) v[m] = cells[m].ut;
  face_values(v, blend, &ut_lo, &ut_hi);
  const double s_z = 0.25 * (high.b - low.b);
  Side l, r;
  l.z = c.b + -1.0 * s_z;
  l.h = eta_lo - l.z;
  l.un = un_lo;
  l.ut = ut_lo;
  r.z = c.b + 1.0 * s_z;
  r.h = eta_hi - r.z;
  r.un = un_hi;
  r.ut = ut_hi;
  if (l.h < H_FACE_MIN || r.h < H_FACE_MIN) {
    *lo_out = first;
    *hi_out = first;
    return;
  }
  *lo_out = l;
  *hi_out = r;
}

// solver.rs:185-226
Face face(const Side& l, const Side& r) {
  const double z_star = wc_max(l.z, r.z);
  const double hl = wc_max(l.h + l.z - z_star, 0.0);
  const double hr = wc_max(r.h + r.z - z_star, 0.0);
  const double cl = sqrt(G * hl);
  const double cr = sqrt(G * hr);
  const double s_l = wc_min(l.un - cl, r.un - cr);
  const double s_r = wc_max(l.un + cl, r.un + cr);
  const double f_l[3] = {hl * l.un, hl * l.un * l.un + 0.5 * G * hl * hl, hl * l.un * l.ut};
  const double f_r[3] = {hr * r.un, hr * r.un * r.un + 0.5 * G * hr * hr, hr * r.un * r.ut};
  const double u_l[3] = {hl, hl * l.un, hl * l.ut};
  const double u_r[3] = {hr, hr * r.un, hr * r.ut};
  Face f;
  if (hl <= H_DRY && hr <= H_DRY) {
    f.flux[0] = f.flux[1] = f.flux[2] = 0.0;
  } else if (s_l >= 0.0) {
    for (int k = 0; k < 3; ++k) f.flux[k] = f_l[k];
  } else if (s_r <= 0.0) {
    for (int k = 0; k < 3; ++k) f.flux[k] = f_r[k];
  } else {
    const double inv = 1.0 / (s_r - s_l);
    for (int k = 0; k < 3; ++k)
      f.flux[k] = (s_r * f_l[k] - s_l * f_r[k] + s_l * s_r * (u_r[k] - u_l[k])) * inv;
  }
  f.corr_left = 0.5 * G * (l.h * l.h - hl * hl);
  f.corr_right = 0.5 * G * (r.h * r.h - hr * hr);
  return f;
}

struct Axis {
  Face low, high;
  double source;
};

// solver.rs:229-239
Axis axis_terms(const Cell* st, double spacing, int order) {
  Side a_lo, a_hi, lo, hi, b_lo, b_hi;
  reconstruct(st + 0, order, &a_lo, &a_hi);
  reconstruct(st + 1, order, &lo, &hi);
  reconstruct(st + 2, order, &b_lo, &b_hi);
  Axis ax;
  ax.low = face(a_hi, lo);
  ax.high = face(hi, b_lo);
  ax.source = -G * 0.5 * (lo.h + hi.h) * (hi.z - lo.z) / spacing;
  return ax;
}

// ---------------------------------------------------------------- reductions
// Lane 0's value of a 64-lane xor butterfly that adds lane l^1 first, then l^2, ..., l^32.
void butterfly(double* v) {
  for (size_t off = 1; off < 64; off *= 2)
    for (size_t l = 0; l < 64; l += 2 * off) v[l] = v[l] + v[l + off];
}

// The GPU's canonical summation order for an interior reduction: see the header.
template <class Term>
double canonical_sum(const Grid& g, const Par& par, Term term) {
  const size_t nx = g.nx, ny = g.ny;
  const size_t ntx = (nx + 63) / 64, nty = (ny + 3) / 4, np = ntx * nty;
  std::vector<double> part(np);
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long p = 0; p < (long)np; ++p) {
    const size_t tx = (size_t)p % ntx, ty = (size_t)p / ntx;
    double rs[4];
    for (size_t r = 0; r < 4; ++r) {
      double v[64];
      for (size_t l = 0; l < 64; ++l) {
        const size_t i = tx * 64 + l, j = ty * 4 + r;
        v[l] = (i < nx && j < ny) ? term(GH + i, GH + j) : 0.0;
      }
      butterfly(v);
      rs[r] = v[0];
    }
    part[p] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
  }
  double acc[256];
  for (size_t t = 0; t < 256; ++t) {
    acc[t] = 0.0;
    for (size_t p = t; p < np; p += 256) acc[t] = acc[t] + part[p];
  }
  double ws[4];
  for (size_t wv = 0; wv < 4; ++wv) {
    double* v = acc + 64 * wv;
    butterfly(v);
    ws[wv] = v[0];
  }
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// The FP64 PCG kernels' order (wc_pcg2.hip): 64x16-cell tiles t = ty * ntx + tx; lane b of 256
// takes tiles b, b+256, ... and each of its 1024 threads (one cell of a tile) adds its cell's
// term tile by tile from 0.0 (0.0 for cells outside the grid); each row of 64 running sums goes
// through the butterfly, the 16 row sums are added in row order from 0.0, and the 256 lanes
// end as in canonical_sum.
template <class Term>
double canonical16_sum(const Grid& g, const Par& par, Term term) {
  const size_t nx = g.nx, ny = g.ny;
  const size_t ntx = (nx + 63) / 64, nty = (ny + 15) / 16, np = ntx * nty;
  double acc[256];
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long b = 0; b < 256; ++b) {
    double run[16][64];
    for (size_t r = 0; r < 16; ++r)
      for (size_t l = 0; l < 64; ++l) run[r][l] = 0.0;
    for (size_t p = (size_t)b; p < np; p += 256) {
      const size_t tx = p % ntx, ty = p / ntx;
      for (size_t r = 0; r < 16; ++r)
        for (size_t l = 0; l < 64; ++l) {
          const size_t i = tx * 64 + l, j = ty * 16 + r;
          run[r][l] = run[r][l] + ((i < nx && j < ny) ? term(GH + i, GH + j) : 0.0);
        }
    }
    double t = 0.0;
    for (size_t r = 0; r < 16; ++r) {
      butterfly(run[r]);
      t = t + run[r][0];
    }
    acc[b] = t;
  }
  double ws[4];
  for (size_t wv = 0; wv < 4; ++wv) {
    double* v = acc + 64 * wv;
    butterfly(v);
    ws[wv] = v[0];
  }
  return (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// par.rs:90-98 with dispersion.rs:319-329: rows summed serially, then in row order from 0.0.
template <class Term>
double row_sum(const Grid& g, const Par& par, Term term) {
  std::vector<double> rs(g.ny);
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long jj = 0; jj < (long)g.ny; ++jj) {
    const size_t j = GH + (size_t)jj;
    double s = 0.0;
    for (size_t i = GH; i < GH + g.nx; ++i) s += term(i, j);
    rs[(size_t)jj] = s;
  }
  double total = 0.0;
  for (double v : rs) total = total + v;
  return total;
}

double dot(const Grid& g, const Par& par, int order, const double* ax, const double* ay,
           const double* bx, const double* by) {
  const size_t w = g.width();
  auto term = [&](size_t i, size_t j) {
    const size_t k = j * w + i;
    return ax[k] * bx[k] + ay[k] * by[k];
  };
  if (order == 1) return canonical_sum(g, par, term);
  if (order == 2) return canonical16_sum(g, par, term);
  return row_sum(g, par, term);
}

// ---------------------------------------------------------------- SGN dispersion
// The FP32 form's extrapolated warm start (not the reference's: it starts each solve from the
// previous solution, dispersion.rs:220-222). Solves come `period` per step; with d_n = x_n -
// x_{n-1} the start of solve n is x_{n-1} + sum_{j=1..K} c_j d_{n - j period}, the weights exact
// for steps d of degree < m in time (each kind of solve, RK stage, its own) and of least norm
// among such weights when K > m; K = m gives the m-th backward difference (stride `period`) of
// the steps set to zero, c_j = (-1)^(j+1) C(m, j). Restated here from the mathematics: the
// least-norm weights solve the normal equations (V V^T) y = e_0, c = V^T y, V[r][j] = (-j)^r.
constexpr int WS_SLOTS = 32;
constexpr int WS_MAX_WINDOW = 12;
void ws_weights(int m, int K, double* c) {
  if (K == m) {
    double cf = 1.0;
    for (int j = 1; j <= m; ++j) {
      cf = cf * (double)(m - j + 1) / (double)j;
      c[j - 1] = (j & 1) ? cf : -cf;
    }
    return;
  }
  std::vector<std::vector<double>> a(m, std::vector<double>(m + 1, 0.0));
  auto v = [](int r, int j) { return pow(-(double)j, (double)r); };
  for (int r = 0; r < m; ++r) {
    for (int q = 0; q < m; ++q)
      for (int j = 1; j <= K; ++j) a[r][q] += v(r, j) * v(q, j);
    a[r][m] = r == 0 ? 1.0 : 0.0;
  }
  for (int col = 0; col < m; ++col) {  // Gauss-Jordan with partial pivoting
    int piv = col;
    for (int r = col + 1; r < m; ++r)
      if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
    std::swap(a[col], a[piv]);
    for (int r = 0; r < m; ++r) {
      if (r == col) continue;
      const double f = a[r][col] / a[col][col];
      for (int q = col; q <= m; ++q) a[r][q] -= f * a[col][q];
    }
  }
  for (int j = 1; j <= K; ++j) {
    double t = 0.0;
    for (int r = 0; r < m; ++r) t += v(r, j) * (a[r][m] / a[r][r]);
    c[j - 1] = t;
  }
}

template <class D>
double ws_extrapolate(double x1, int K, const double* c, D step) {
  double x = x1;
  for (int j = 1; j <= K; ++j) x += c[j - 1] * step(j);
  return x;
}

struct Work {
  std::vector<double> u, v, mask, um, vm, div, c, f, heff, diag_x, diag_y, wx, wy, bx, by, rx, ry,
      zx, zy, px, py, ax, ay;
  wco_dispersion_stats_t stats{0, 0, 0};
  // the extrapolated warm start (order 0: the reference's): steps d of the last WS_SLOTS solves,
  // slot n % WS_SLOTS for solve n (which holds x_{n-1} while solve n runs); K weights
  int ws_order = 0, ws_period = 2, ws_terms = 0;
  double ws_c[WS_MAX_WINDOW] = {};
  std::vector<double> ws_x[WS_SLOTS], ws_y[WS_SLOTS];
  void init(const Grid& g) {
    const size_t n = g.cells();
    for (auto* a : {&u, &v, &mask, &um, &vm, &div, &c, &f, &heff, &diag_x, &diag_y, &wx, &wy,
                    &bx, &by, &rx, &ry, &zx, &zy, &px, &py, &ax, &ay})
      a->assign(n, 0.0);
    stats = {0, 0, 0};
  }
};

// dispersion.rs:274-317
void matvec(const Grid& g, const Par& par, bool periodic, const double* heff, const double* mask,
            const double* c, double* f, double* px, double* py, double* ax, double* ay) {
  fill_vector_ghosts(g, px, py, periodic);
  const size_t w = g.width(), nx = g.nx, ny = g.ny;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long jj = (long)GH - 1; jj <= (long)(ny + GH); ++jj) {
    const size_t j = (size_t)jj;
    for (size_t i = GH - 1; i < nx + GH + 1; ++i) {
      const size_t k = j * w + i;
      f[k] = c[k] * ((px[k + 1] - px[k - 1]) * ix + (py[k + w] - py[k - w]) * iy);
    }
  }
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long jj = (long)GH; jj < (long)(ny + GH); ++jj) {
    const size_t j = (size_t)jj;
    for (size_t i = GH; i < nx + GH; ++i) {
      const size_t k = j * w + i;
      ax[k] = mask[k] * (heff[k] * px[k] - (f[k + 1] - f[k - 1]) * ix);
      ay[k] = mask[k] * (heff[k] * py[k] - (f[k + w] - f[k - w]) * iy);
    }
  }
}

// The interior-only row loops of dispersion.rs:566-635 and the p update at 527-539.
template <class Fn>
void interior(const Grid& g, const Par& par, Fn fn) {
  const size_t w = g.width();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long jj = (long)GH; jj < (long)(g.ny + GH); ++jj) {
    const size_t j = (size_t)jj;
    for (size_t i = GH; i < g.nx + GH; ++i) fn(j * w + i);
  }
}

// dispersion.rs:335-563
void dispersion_apply(const Grid& g, const Par& par, int dot_order, const DispersionCfg& d,
                      bool periodic, const State& s, const double* lh, double* lhu, double* lhv,
                      const double* phi, Work& wk) {
  const size_t w = g.width(), nx = g.nx, ny = g.ny, rows = ny + 2 * GH;
  const double dx = g.dx, dy = g.dy;
  const double ix = 0.5 / dx, iy = 0.5 / dy;
  const double *h = s.h.data(), *hu = s.hu.data(), *hv = s.hv.data();
  const bool par_on = par.on(g.cells());
  double *u = wk.u.data(), *v = wk.v.data(), *mask = wk.mask.data(), *um = wk.um.data(),
         *vm = wk.vm.data(), *div = wk.div.data(), *c = wk.c.data(), *f = wk.f.data(),
         *heff = wk.heff.data(), *diag_x = wk.diag_x.data(), *diag_y = wk.diag_y.data(),
         *wx = wk.wx.data(), *wy = wk.wy.data(), *bx = wk.bx.data(), *by = wk.by.data(),
         *rx = wk.rx.data(), *ry = wk.ry.data(), *zx = wk.zx.data(), *zy = wk.zy.data(),
         *px = wk.px.data(), *py = wk.py.data(), *ax = wk.ax.data(), *ay = wk.ay.data();

#pragma omp parallel for schedule(static) num_threads(par.threads) if (par_on)
  for (long jj = 0; jj < (long)rows; ++jj) {
    for (size_t i = 0; i < w; ++i) {
      const size_t k = (size_t)jj * w + i;
      u[k] = wc_velocity(h[k], hu[k]);
      v[k] = wc_velocity(h[k], hv[k]);
    }
  }
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par_on)
  for (long jj = 0; jj < (long)rows; ++jj) {
    for (size_t i = 0; i < w; ++i) {
      const size_t k = (size_t)jj * w + i;
      const double m = h[k] >= H_DISP_MIN ? 1.0 : 0.0;
      mask[k] = m;
      um[k] = m * u[k];
      vm[k] = m * v[k];
    }
  }
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par_on)
  for (long jj = 1; jj < (long)rows - 1; ++jj) {
    for (size_t i = 1; i < w - 1; ++i) {
      const size_t k = (size_t)jj * w + i;
      div[k] = (um[k + 1] - um[k - 1]) * ix + (vm[k + w] - vm[k - w]) * iy;
    }
  }
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par_on)
  for (long jj = (long)GH - 1; jj <= (long)(ny + GH); ++jj) {
    for (size_t i = GH - 1; i < nx + GH + 1; ++i) {
      const size_t k = (size_t)jj * w + i;
      const double depth = h[k];
      const double coeff = depth < H_DISP_MIN ? 0.0 : phi[k] * depth * depth * depth / 3.0;
      c[k] = coeff;
      if (d.nonlinear && coeff > 0.0) {
        const double n = um[k] * (div[k + 1] - div[k - 1]) * ix +
                         vm[k] * (div[k + w] - div[k - w]) * iy - div[k] * div[k];
        f[k] = coeff * n;
      } else {
        f[k] = 0.0;
      }
    }
  }
  interior(g, par, [&](size_t k) {
    heff[k] = wc_max(h[k], H_DRY);
    if (mask[k] == 0.0) {
      diag_x[k] = 1.0;
      diag_y[k] = 1.0;
    } else {
      diag_x[k] = heff[k] + (c[k + 1] + c[k - 1]) * 0.25 / (dx * dx);
      diag_y[k] = heff[k] + (c[k + w] + c[k - w]) * 0.25 / (dy * dy);
    }
  });
  interior(g, par, [&](size_t k) {
    if (mask[k] == 0.0) {
      bx[k] = 0.0;
      by[k] = 0.0;
    } else {
      bx[k] = lhu[k] - u[k] * lh[k] + (f[k + 1] - f[k - 1]) * ix;
      by[k] = lhv[k] - v[k] * lh[k] + (f[k + w] - f[k - w]) * iy;
    }
  });
  if (wk.ws_order > 0) {
    const uint64_t n = wk.stats.solves;
    for (int m = 0; m < WS_SLOTS; ++m)
      if (wk.ws_x[m].size() != g.cells()) {
        wk.ws_x[m].assign(g.cells(), 0.0);
        wk.ws_y[m].assign(g.cells(), 0.0);
      }
    double *ex = wk.ws_x[n % WS_SLOTS].data(), *ey = wk.ws_y[n % WS_SLOTS].data();
    auto slot = [&](int j) { return (n - (uint64_t)(j * wk.ws_period)) % WS_SLOTS; };
    interior(g, par, [&](size_t k) {
      if (mask[k] == 0.0) {
        wx[k] = 0.0;
        wy[k] = 0.0;
        ex[k] = 0.0;
        ey[k] = 0.0;
      } else {
        ex[k] = wx[k];
        ey[k] = wy[k];
        wx[k] = ws_extrapolate(wx[k], wk.ws_terms, wk.ws_c,
                               [&](int j) { return wk.ws_x[slot(j)][k]; });
        wy[k] = ws_extrapolate(wy[k], wk.ws_terms, wk.ws_c,
                               [&](int j) { return wk.ws_y[slot(j)][k]; });
      }
    });
  } else {
    interior(g, par, [&](size_t k) {
      if (mask[k] == 0.0) {
        wx[k] = 0.0;
        wy[k] = 0.0;
      }
    });
  }

  matvec(g, par, periodic, heff, mask, c, f, wx, wy, ax, ay);
  interior(g, par, [&](size_t k) {
    rx[k] = bx[k] - ax[k];
    ry[k] = by[k] - ay[k];
  });
  interior(g, par, [&](size_t k) {
    zx[k] = rx[k] / diag_x[k];
    zy[k] = ry[k] / diag_y[k];
  });
  interior(g, par, [&](size_t k) {
    px[k] = zx[k];
    py[k] = zy[k];
  });
  const double bnorm = sqrt(dot(g, par, dot_order, bx, by, bx, by));
  const double target = d.tolerance * bnorm + 1e-10;
  double rz = dot(g, par, dot_order, rx, ry, zx, zy);
  uint64_t iterations = 0;
  bool converged = sqrt(dot(g, par, dot_order, rx, ry, rx, ry)) <= target;
  while (!converged && iterations < d.max_iterations) {
    matvec(g, par, periodic, heff, mask, c, f, px, py, ax, ay);
    const double pap = dot(g, par, dot_order, px, py, ax, ay);
    const double alpha = rz / pap;
    interior(g, par, [&](size_t k) {
      wx[k] += alpha * px[k];
      wy[k] += alpha * py[k];
    });
    const double na = -alpha;
    interior(g, par, [&](size_t k) {
      rx[k] += na * ax[k];
      ry[k] += na * ay[k];
    });
    interior(g, par, [&](size_t k) {
      zx[k] = rx[k] / diag_x[k];
      zy[k] = ry[k] / diag_y[k];
    });
    iterations += 1;
    converged = sqrt(dot(g, par, dot_order, rx, ry, rx, ry)) <= target;
    const double rz_next = dot(g, par, dot_order, rx, ry, zx, zy);
    const double beta = rz_next / rz;
    rz = rz_next;
    interior(g, par, [&](size_t k) {
      px[k] = zx[k] + beta * px[k];
      py[k] = zy[k] + beta * py[k];
    });
  }
  wk.stats.solves += 1;
  wk.stats.iterations += iterations;
  if (!converged) wk.stats.unconverged += 1;

  interior(g, par, [&](size_t k) {
    if (mask[k] > 0.0) {
      lhu[k] = h[k] * wx[k] + u[k] * lh[k];
      lhv[k] = h[k] * wy[k] + v[k] * lh[k];
    }
  });
  if (wk.ws_order > 0) {  // slot n: x_{n-1} -> d_n = x_n - x_{n-1}
    double *ex = wk.ws_x[(wk.stats.solves - 1) % WS_SLOTS].data(),
           *ey = wk.ws_y[(wk.stats.solves - 1) % WS_SLOTS].data();
    interior(g, par, [&](size_t k) {
      ex[k] = wx[k] - ex[k];
      ey[k] = wy[k] - ey[k];
    });
  }
}

// dispersion.rs:120-160; with the eddy viscosity on, a cell with breaking memory stays off.
void switch_phi(const Grid& g, const Par& par, const Mathlib& m, const double* bed,
                const State& s, const Breaking& br, bool periodic, double* phi) {
  const size_t w = g.width();
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const double* h = s.h.data();
  auto eta = [&](size_t k) { return h[k] + bed[k]; };
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
  for (long jj = (long)GH; jj < (long)(g.ny + GH); ++jj) {
    const size_t j = (size_t)jj;
    for (size_t i = GH; i < g.nx + GH; ++i) {
      const size_t k = j * w + i;
      if (h[k] < H_DISP_MIN) {
        phi[k] = 0.0;
        continue;
      }
      const double ex = (eta(k + 1) - eta(k - 1)) * ix;
      const double ey = (eta(k + w) - eta(k - w)) * iy;
      const double ratio = wc_max(eta(k) / h[k], 0.0);
      double off = wc_max(wc_smoothstep(br.er_lo, br.er_hi, ratio),
                          wc_smoothstep(br.sl_lo, br.sl_hi, m.hypot(ex, ey)));
      if (br.viscosity && s.breaking[k] != 0.0) off = 1.0;  // EddyViscosity::keeps_off
      phi[k] = 1.0 - off;
    }
  }
  mirror_ghosts(g, phi);
  if (periodic) wrap_ghosts_y(g, phi);
}

// ---- breaking fronts and their eddy viscosity (eddy.rs) --------------------------------------
constexpr double H_BREAKING_MIN = 0.01;  // eddy.rs:55
constexpr double BREAKING_ONSET = 0.8;   // eddy.rs:59

struct EddyWork {  // eddy.rs:93-100
  std::vector<double> rise, nu, memory, hu, hv;
  void init(const Grid& g) {
    rise.assign(g.cells(), 0.0);
    nu.assign(g.cells(), 0.0);
    memory.assign(g.cells(), 0.0);
    hu.assign(g.cells(), 0.0);
    hv.assign(g.cells(), 0.0);
  }
};

// EddyViscosity::follow (eddy.rs:163-221): the fronts into s.breaking, their viscosity into nu.
void eddy_follow(const Grid& g, const Eddy& e, State& s, const double* phi, double dt,
                 EddyWork& work) {
  const size_t w = g.width(), nx = g.nx, ny = g.ny;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const double *h = s.h.data(), *hu = s.hu.data(), *hv = s.hv.data();
  double* rise = work.rise.data();
  std::fill(work.rise.begin(), work.rise.end(), 0.0);
  for (size_t j = GH; j < ny + GH; ++j)
    for (size_t i = GH; i < nx + GH; ++i) {
      const size_t k = j * w + i;
      rise[k] = -((hu[k + 1] - hu[k - 1]) * ix + (hv[k + w] - hv[k - w]) * iy);
    }
  const double* old = s.breaking.data();
  std::fill(work.memory.begin(), work.memory.end(), 0.0);
  std::fill(work.nu.begin(), work.nu.end(), 0.0);
  double *rm = work.memory.data(), *rn = work.nu.data();
  for (size_t j = GH; j < ny + GH; ++j)
    for (size_t i = GH; i < nx + GH; ++i) {
      const size_t k = j * w + i;
      const double depth = h[k];
      if (depth < H_BREAKING_MIN) continue;
      const double c = sqrt(G * depth);
      double front = 0.0;
      for (size_t m : {k, k - 1, k + 1, k - w, k + w}) front = wc_max(front, old[m]);
      const bool starts = depth >= H_DISP_MIN && 1.0 - phi[k] >= BREAKING_ONSET && rise[k] > 0.0;
      const bool goes_on = front > 0.0 && rise[k] > e.threshold(front, depth) * c;
      const double age = (front > 0.0 && (starts || goes_on)) ? front + dt : starts ? dt : 0.0;
      if (age > 0.0) {
        double b = rise[k] / (e.threshold(age, depth) * c) - 1.0;
        b = b < 0.0 ? 0.0 : (b > 1.0 ? 1.0 : b);  // f64::clamp(0, 1)
        rm[k] = age;
        rn[k] = b * e.mixing * e.mixing * depth * rise[k];
      } else if (old[k] > 0.0) {
        rm[k] = -e.linger * sqrt(depth / G);
      } else if (old[k] < 0.0) {
        rm[k] = wc_min(old[k] + dt, 0.0);
      }
    }
  std::swap(s.breaking, work.memory);
}

// diffuse (eddy.rs:225-302): nu applied to the momentum in explicit sub-steps; returns them.
template <class Fill>
size_t eddy_diffuse(const Grid& g, State& s, double dt, EddyWork& work, Fill fill) {
  const size_t w = g.width(), nx = g.nx, ny = g.ny;
  const double* nu = work.nu.data();
  double top = 0.0;
  for (size_t j = GH; j < GH + ny; ++j)
    for (size_t i = GH; i < GH + nx; ++i) top = wc_max(top, nu[j * w + i]);
  if (top <= 0.0) return 0;
  const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
  const double limit = 0.5 / (top * wc_max(2.0 * idx2 + idy2, idx2 + 2.0 * idy2));
  const size_t steps = (size_t)wc_max(ceil(dt / limit), 1.0);
  const double sub = dt / (double)steps;
  const double cross = 0.25 / (g.dx * g.dy);
  for (size_t step = 0; step < steps; ++step) {
    if (step > 0) fill(s);
    const double *h = s.h.data(), *hu = s.hu.data(), *hv = s.hv.data();
    auto link = [&](size_t a, size_t b) {
      return (h[a] < H_BREAKING_MIN || h[b] < H_BREAKING_MIN) ? 0.0 : 0.5 * (nu[a] + nu[b]);
    };
    auto wet = [&](size_t a) { return h[a] < H_BREAKING_MIN ? 0.0 : nu[a]; };
    double *ru = work.hu.data(), *rv = work.hv.data();
    for (size_t j = GH; j < ny + GH; ++j)
      for (size_t i = GH; i < nx + GH; ++i) {
        const size_t k = j * w + i;
        if (h[k] < H_BREAKING_MIN) {
          ru[k] = hu[k];
          rv[k] = hv[k];
          continue;
        }
        const size_t e = k + 1, west = k - 1, north = k + w, south = k - w;
        auto xx = [&](const double* q) {
          return (link(k, e) * (q[e] - q[k]) - link(k, west) * (q[k] - q[west])) * idx2;
        };
        auto yy = [&](const double* q) {
          return (link(k, north) * (q[north] - q[k]) - link(k, south) * (q[k] - q[south])) * idy2;
        };
        auto yx = [&](const double* q) {
          return (wet(north) * (q[north + 1] - q[north - 1]) -
                  wet(south) * (q[south + 1] - q[south - 1])) * cross;
        };
        auto xy = [&](const double* q) {
          return (wet(e) * (q[e + w] - q[e - w]) - wet(west) * (q[west + w] - q[west - w])) *
                 cross;
        };
        ru[k] = hu[k] + sub * (xx(hu) + 0.5 * (yy(hu) + yx(hv)));
        rv[k] = hv[k] + sub * (yy(hv) + 0.5 * (xx(hv) + xy(hu)));
      }
    std::swap(s.hu, work.hu);
    std::swap(s.hv, work.hv);
  }
  return steps;
}

}  // namespace

// ---------------------------------------------------------------- the solver
struct wco_solver {
  Grid g;
  std::vector<double> bed;
  double cfl = 0.4;
  int order = ORDER_SECOND;
  double manning = 0.0;
  bool periodic = false;
  bool has_maker = false;
  WaveMaker maker{};
  bool has_drive = false;
  Drive drive;
  std::vector<Sponge> sponges;
  bool dispersive = false;
  DispersionCfg disp;
  Work work;
  EddyWork eddy_work;
  std::vector<double> phi;
  bool has_shallow = false;
  std::vector<uint8_t> shallow;
  int dot_order = 0;
  Mathlib math;
  Par par;
  State s;

  Breaking breaking_config() const { return dispersive ? disp.breaking : Breaking(); }

  // solver.rs:421-429
  bool check_periodic() const {
    if (periodic && has_drive) {
      const double turns = drive.ky * (double)g.ny * g.dy / TAU;
      if (!(fabs(turns - round(turns)) < 1e-6)) {
        char buf[200];
        snprintf(buf, sizeof buf,
                 "wave-maker does not fit the periodic domain: k_y * Ly / 2pi = %.17g, must be "
                 "a whole number",
                 turns);
        fail(buf);
        return false;
      }
    }
    return true;
  }

  // solver.rs:400-409
  bool rebuild_drive() {
    if (!has_maker) return true;
    double fi = round(maker.x / g.dx - 0.5);
    fi = wc_max(fi, 0.0);
    size_t i = fi >= (double)SIZE_MAX ? SIZE_MAX : (size_t)fi;
    if (i > g.nx - 1) i = g.nx - 1;
    i += GH;
    const double depth = -bed[g.idx(i, GH + g.ny / 2)];
    Drive d;
    if (!d.init(maker, g, depth, dispersive)) return false;
    drive = d;
    has_drive = true;
    return check_periodic();
  }

  // solver.rs:577-591
  void clean_dry(State& st) const {
    const size_t w = g.width();
    double *h = st.h.data(), *hu = st.hu.data(), *hv = st.hv.data();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
    for (long jj = (long)GH; jj < (long)(g.ny + GH); ++jj) {
      for (size_t i = GH; i < g.nx + GH; ++i) {
        const size_t k = (size_t)jj * w + i;
        if (h[k] < H_DRY) {
          h[k] = wc_max(h[k], 0.0);
          hu[k] = 0.0;
          hv[k] = 0.0;
        }
      }
    }
  }

  // solver.rs:594-616
  State advance(const State& st, const std::vector<double>* rhs, double dt) const {
    State next = st;
    const size_t w = g.width();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
    for (long jj = (long)GH; jj < (long)(g.ny + GH); ++jj) {
      for (size_t i = GH; i < g.nx + GH; ++i) {
        const size_t k = (size_t)jj * w + i;
        next.h[k] += dt * rhs[0][k];
        next.hu[k] += dt * rhs[1][k];
        next.hv[k] += dt * rhs[2][k];
      }
    }
    clean_dry(next);
    return next;
  }

  // solver.rs:625-733
  void rhs(const State& st, double t, std::vector<double>* out) {
    const size_t n = g.cells(), w = g.width();
    if (order == ORDER_FIFTH || dispersive) {
      switch_phi(g, par, math, bed.data(), st, breaking_config(), periodic, phi.data());
      if (has_shallow) {
        size_t m = 0;
        for (size_t j = GH; j < GH + g.ny; ++j)
          for (size_t i = GH; i < GH + g.nx; ++i, ++m)
            if (shallow[m]) phi[g.idx(i, j)] = 0.0;
        mirror_ghosts(g, phi.data());
        if (periodic) wrap_ghosts_y(g, phi.data());
      }
    }
    for (int c = 0; c < 3; ++c) out[c].assign(n, 0.0);
    const double *h = st.h.data(), *hu = st.hu.data(), *hv = st.hv.data(), *b = bed.data(),
                 *ph = phi.data();
    auto cell = [&](size_t k, bool normal_is_x) {
      const double un = normal_is_x ? hu[k] : hv[k];
      const double ut = normal_is_x ? hv[k] : hu[k];
      return Cell{h[k], wc_velocity(h[k], un), wc_velocity(h[k], ut), b[k], ph[k]};
    };
    double *rh = out[0].data(), *rhu = out[1].data(), *rhv = out[2].data();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
    for (long jj = (long)GH; jj < (long)(g.ny + GH); ++jj) {
      const size_t j = (size_t)jj;
      for (size_t i = GH; i < g.nx + GH; ++i) {
        Cell sx[7], sy[7];
        for (size_t m = 0; m < 7; ++m) {
          sx[m] = cell(g.idx(i + m - 3, j), true);
          sy[m] = cell(g.idx(i, j + m - 3), false);
        }
        const Axis x = axis_terms(sx, g.dx, order);
        const Axis y = axis_terms(sy, g.dy, order);
        const size_t k = j * w + i;
        rh[k] = -(x.high.flux[0] - x.low.flux[0]) / g.dx - (y.high.flux[0] - y.low.flux[0]) / g.dy;
        rhu[k] = -((x.high.flux[1] + x.high.corr_left) - (x.low.flux[1] + x.low.corr_right)) / g.dx +
                 x.source - (y.high.flux[2] - y.low.flux[2]) / g.dy;
        rhv[k] = -((y.high.flux[1] + y.high.corr_left) - (y.low.flux[1] + y.low.corr_right)) / g.dy +
                 y.source - (x.high.flux[2] - x.low.flux[2]) / g.dx;
        if (has_drive) rh[k] += drive.source(g.cx(i), g.cy(j), t);
      }
    }
    if (dispersive)
      dispersion_apply(g, par, dot_order, disp, periodic, st, rh, rhu, rhv, phi.data(), work);
  }

  // solver.rs:535-574
  void damp(State& st, double dt) const {
    if (sponges.empty() && manning == 0.0) return;
    const size_t w = g.width();
    double *h = st.h.data(), *hu = st.hu.data(), *hv = st.hv.data();
    const double* b = bed.data();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
    for (long jj = (long)GH; jj < (long)(g.ny + GH); ++jj) {
      const size_t j = (size_t)jj;
      for (size_t i = GH; i < g.nx + GH; ++i) {
        const size_t k = j * w + i;
        for (const Sponge& sp : sponges) {
          const double rate = sp.rate(g, i, j);
          if (rate > 0.0) {
            const double f = exp(-rate * dt);
            // forcing.rs:213-225 with target (level, 0, 0)
            const double ht = wc_max(sp.level - b[k], 0.0);
            h[k] = f * h[k] + (1.0 - f) * ht;
            hu[k] = f * hu[k] + (1.0 - f) * ht * 0.0;
            hv[k] = f * hv[k] + (1.0 - f) * ht * 0.0;
          }
        }
        if (manning > 0.0 && h[k] > H_FRIC_MIN) {
          const double speed = math.hypot(hu[k], hv[k]) / h[k];
          // forcing.rs:294-297
          const double kk = G * manning * manning / math.pow43(h[k]);
          const double fac = 1.0 / (1.0 + dt * kk * speed);
          hu[k] *= fac;
          hv[k] *= fac;
        }
      }
    }
    clean_dry(st);
  }

  // solver.rs:481-507 (follow_breaking is a no-op: viscosity is never on here)
  void step(double dt) {
    const double t = s.time;
    s.fill_boundaries(g, periodic);
    std::vector<double> r[3];
    if (order == ORDER_FIRST) {
      rhs(s, t, r);
      const double keep = s.time;
      s = advance(s, r, dt);
      s.time = keep;
    } else {
      rhs(s, t, r);
      State s1 = advance(s, r, dt);
      s1.fill_boundaries(g, periodic);
      rhs(s1, t + dt, r);
      State s2 = advance(s1, r, dt);
      const size_t n = g.cells(), w = g.width();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(n))
      for (long jj = 0; jj < (long)g.rows(); ++jj) {
        for (size_t i = 0; i < w; ++i) {
          const size_t k = (size_t)jj * w + i;
          s.h[k] = 0.5 * (s.h[k] + s2.h[k]);
          s.hu[k] = 0.5 * (s.hu[k] + s2.hu[k]);
          s.hv[k] = 0.5 * (s.hv[k] + s2.hv[k]);
        }
      }
      clean_dry(s);
    }
    damp(s, dt);
    follow_breaking(s, dt);
    s.time = t + dt;
  }

  // solver.rs:512-532
  void follow_breaking(State& st, double dt) {
    if (!(dispersive && disp.breaking.viscosity)) return;
    st.fill_boundaries(g, periodic);
    Breaking fresh = breaking_config();
    fresh.viscosity = false;  // the switch as the surface alone reads it
    switch_phi(g, par, math, bed.data(), st, fresh, periodic, phi.data());
    const Eddy& e = disp.breaking.eddy;
    eddy_follow(g, e, st, phi.data(), dt, eddy_work);
    if (e.mixing > 0.0)
      eddy_diffuse(g, st, dt, eddy_work, [&](State& x) { x.fill_boundaries(g, periodic); });
    clean_dry(st);
  }

  // solver.rs:95-100
  double rate(size_t k) const {
    const double c = sqrt(G * s.h[k]);
    const double u = fabs(wc_velocity(s.h[k], s.hu[k]));
    const double v = fabs(wc_velocity(s.h[k], s.hv[k]));
    return (u + c) / g.dx + (v + c) / g.dy;
  }

  // solver.rs:434-450
  double stable_dt() const {
    std::vector<double> rows(g.ny);
    const size_t w = g.width();
#pragma omp parallel for schedule(static) num_threads(par.threads) if (par.on(g.cells()))
    for (long jj = 0; jj < (long)g.ny; ++jj) {
      const size_t j = GH + (size_t)jj;
      double m = 0.0;
      for (size_t i = GH; i < GH + g.nx; ++i) m = wc_max_or_nan(m, rate(j * w + i));
      rows[(size_t)jj] = m;
    }
    double r = 0.0;
    for (double v : rows) r = wc_max_or_nan(r, v);
    if (isnan(r)) return NAN;
    if (r > 0.0) return cfl / r;
    return INFINITY;
  }
};

// ---------------------------------------------------------------- C ABI
extern "C" {

const char* wco_last_error(void) { return g_err.c_str(); }

int wco_solver_new(uint64_t nx, uint64_t ny, double dx, double dy, const double* bed,
                   wco_solver** out) {
  if (!(nx > 0 && ny > 0)) return fail("grid must have at least one interior cell");
  if (!(dx > 0.0 && dy > 0.0)) return fail("cell size must be positive");
  if (!bed || !out) return fail("null argument");
  auto* s = new wco_solver();
  s->g.nx = nx;
  s->g.ny = ny;
  s->g.dx = dx;
  s->g.dy = dy;
  const size_t n = s->g.cells();
  s->bed.assign(bed, bed + n);
  s->phi.assign(n, 0.0);
  s->s.zeros(s->g);
  *out = s;
  return 0;
}

void wco_solver_free(wco_solver* s) { delete s; }

int wco_set_cfl(wco_solver* s, double cfl) {
  s->cfl = cfl;
  return 0;
}

int wco_with_order(wco_solver* s, int32_t order) {
  if (order < 0 || order > 2) return fail("order must be 0 (first), 1 (second) or 2 (fifth)");
  s->order = order;
  s->work.ws_period = order == ORDER_FIRST ? 1 : 2;
  return 0;
}

int wco_with_manning(wco_solver* s, double n) {
  s->manning = n;
  return 0;
}

int wco_with_sponge(wco_solver* s, const wco_sponge_t* sp) {
  Sponge x;
  x.width = sp->width;
  x.strength = sp->strength;
  x.level = sp->level;
  x.west = sp->west != 0;
  x.east = sp->east != 0;
  x.south = sp->south != 0;
  x.north = sp->north != 0;
  s->sponges.push_back(x);
  return 0;
}

// forcing.rs:227-284
struct wco_relaxation {
  wco_solver* s = nullptr;
  std::vector<uint64_t> i, j;
  std::vector<double> rate;
};

// Relaxation::new (forcing.rs:239-248)
int wco_relaxation_new(wco_solver* s, const wco_sponge_t* zones, uint64_t n_zones,
                       wco_relaxation** out) {
  auto* r = new wco_relaxation;
  r->s = s;
  std::vector<Sponge> zs;
  for (uint64_t z = 0; z < n_zones; ++z) {
    Sponge x;
    x.width = zones[z].width;
    x.strength = zones[z].strength;
    x.level = zones[z].level;
    x.west = zones[z].west != 0;
    x.east = zones[z].east != 0;
    x.south = zones[z].south != 0;
    x.north = zones[z].north != 0;
    zs.push_back(x);
  }
  const Grid& g = s->g;
  for (size_t j = GH; j < GH + g.ny; ++j)
    for (size_t i = GH; i < GH + g.nx; ++i) {
      double rate = 0.0;
      for (const Sponge& z : zs) rate = wc_max(rate, z.rate(g, i, j));
      if (rate > 0.0) {
        r->i.push_back(i);
        r->j.push_back(j);
        r->rate.push_back(rate);
      }
    }
  *out = r;
  return 0;
}

void wco_relaxation_free(wco_relaxation* r) { delete r; }

uint64_t wco_relaxation_len(const wco_relaxation* r) { return r->i.size(); }

int wco_relaxation_cells(const wco_relaxation* r, uint64_t* i, uint64_t* j) {
  for (size_t m = 0; m < r->i.size(); ++m) {
    if (i) i[m] = r->i[m];
    if (j) j[m] = r->j[m];
  }
  return 0;
}

// Relaxation::apply (forcing.rs:259-283) with relax_cell (213-225)
int wco_relaxation_apply(wco_relaxation* r, double dt, const double* targets) {
  State& st = r->s->s;
  const Grid& g = r->s->g;
  const double* b = r->s->bed.data();
  for (size_t m = 0; m < r->i.size(); ++m) {
    const size_t k = g.idx(r->i[m], r->j[m]);
    const double f = exp(-r->rate[m] * dt);
    const double eta = targets[3 * m], u = targets[3 * m + 1], v = targets[3 * m + 2];
    const double ht = wc_max(eta - b[k], 0.0);
    st.h[k] = f * st.h[k] + (1.0 - f) * ht;
    st.hu[k] = f * st.hu[k] + (1.0 - f) * ht * u;
    st.hv[k] = f * st.hv[k] + (1.0 - f) * ht * v;
    if (st.h[k] < H_DRY) {
      st.h[k] = wc_max(st.h[k], 0.0);
      st.hu[k] = 0.0;
      st.hv[k] = 0.0;
    }
  }
  return 0;
}

int wco_with_wavemaker(wco_solver* s, const wco_wavemaker_t* wm) {
  s->maker = WaveMaker{wm->x, wm->amplitude, wm->period, wm->angle, wm->sigma, wm->ramp_periods};
  s->has_maker = true;
  return s->rebuild_drive() ? 0 : 1;
}

int wco_with_dispersion(wco_solver* s, const wco_dispersion_t* d) {
  s->disp.breaking.er_lo = d->breaking.eta_ratio_lo;
  s->disp.breaking.er_hi = d->breaking.eta_ratio_hi;
  s->disp.breaking.sl_lo = d->breaking.slope_lo;
  s->disp.breaking.sl_hi = d->breaking.slope_hi;
  s->disp.nonlinear = d->nonlinear != 0;
  s->disp.tolerance = d->tolerance;
  s->disp.max_iterations = d->max_iterations;
  s->disp.breaking.viscosity = false;
  s->dispersive = true;
  s->work.init(s->g);
  return s->rebuild_drive() ? 0 : 1;
}

int wco_with_eddy_viscosity(wco_solver* s, const wco_eddy_t* e) {
  if (!s->dispersive) return fail("the eddy viscosity is part of the dispersion (with_dispersion first)");
  if (!e) {
    s->disp.breaking.viscosity = false;
    return 0;
  }
  s->disp.breaking.viscosity = true;
  s->disp.breaking.eddy = Eddy{e->initial, e->settled, e->duration, e->mixing, e->linger};
  s->eddy_work.init(s->g);
  return 0;
}

// Test hooks: EddyViscosity::follow and diffuse on given arrays (eddy.rs:163-302).
int wco_eddy_follow(uint64_t nx, uint64_t ny, double dx, double dy, const wco_eddy_t* e,
                    double* h, double* hu, double* hv, double* breaking, const double* phi,
                    double dt, double* nu_out) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.dx = dx;
  g.dy = dy;
  State s;
  s.h.assign(h, h + g.cells());
  s.hu.assign(hu, hu + g.cells());
  s.hv.assign(hv, hv + g.cells());
  s.breaking.assign(breaking, breaking + g.cells());
  EddyWork work;
  work.init(g);
  eddy_follow(g, Eddy{e->initial, e->settled, e->duration, e->mixing, e->linger}, s, phi, dt,
              work);
  std::copy(s.breaking.begin(), s.breaking.end(), breaking);
  if (nu_out) std::copy(work.nu.begin(), work.nu.end(), nu_out);
  return 0;
}

int wco_eddy_diffuse(uint64_t nx, uint64_t ny, double dx, double dy, double* h, double* hu,
                     double* hv, const double* nu, double dt, uint64_t* steps) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.dx = dx;
  g.dy = dy;
  State s;
  s.h.assign(h, h + g.cells());
  s.hu.assign(hu, hu + g.cells());
  s.hv.assign(hv, hv + g.cells());
  s.breaking.assign(g.cells(), 0.0);
  EddyWork work;
  work.init(g);
  std::copy(nu, nu + g.cells(), work.nu.begin());
  const size_t n = eddy_diffuse(g, s, dt, work, [&](State& x) { x.fill_walls(g); });
  std::copy(s.hu.begin(), s.hu.end(), hu);
  std::copy(s.hv.begin(), s.hv.end(), hv);
  if (steps) *steps = n;
  return 0;
}

int wco_with_periodic_y(wco_solver* s) {
  s->periodic = true;
  wrap_ghosts_y(s->g, s->bed.data());
  return s->check_periodic() ? 0 : 1;
}

int wco_with_shallow_cells(wco_solver* s, const uint8_t* cells, uint64_t n) {
  if (n != s->g.nx * s->g.ny) return fail("one entry per interior cell");
  s->shallow.assign(cells, cells + n);
  s->has_shallow = true;
  return 0;
}

int wco_bed(const wco_solver* s, double* out) {
  memcpy(out, s->bed.data(), s->bed.size() * sizeof(double));
  return 0;
}

int wco_set_dot_order(wco_solver* s, int32_t order) {
  if (order < 0 || order > 2)
    return fail("dot order must be 0 (reference), 1 (canonical, 64x4 tiles) or 2 (64x16 tiles)");
  s->dot_order = order;
  return 0;
}

int wco_set_math(wco_solver* s, int32_t mode) {
  if (mode != 0 && mode != 1) return fail("math must be 0 (glibc) or 1 (portable)");
  s->math.portable = mode == 1;
  return 0;
}

int wco_set_warm_start(wco_solver* s, int32_t order, int32_t window) {
  if (order < 0 || order > 5 || window < 0 || window > WS_MAX_WINDOW)
    return fail("warm start order must be in [0, 5], window in [0, 12]");
  s->work.ws_order = order;
  s->work.ws_terms = order > 0 ? std::max(order, window) : 0;
  if (order > 0) ws_weights(order, s->work.ws_terms, s->work.ws_c);
  s->work.ws_period = s->order == ORDER_FIRST ? 1 : 2;
  for (int m = 0; m < WS_SLOTS; ++m) {
    s->work.ws_x[m].clear();
    s->work.ws_y[m].clear();
  }
  return 0;
}

int wco_ws_weights(int32_t order, int32_t window, double* out) {
  if (order < 1 || order > 5 || window < 0 || window > WS_MAX_WINDOW)
    return fail("order must be in [1, 5], window in [0, 12]");
  ws_weights(order, std::max(order, window), out);
  return 0;
}

int wco_ws_extrapolate(int32_t order, int32_t window, int32_t period, uint64_t n_past,
                       const double* past, double* out) {
  if (order < 0 || order > 5 || window < 0 || window > WS_MAX_WINDOW || period < 1 ||
      period > 2)
    return fail("order must be in [0, 5], window in [0, 12] and period 1 or 2");
  if (n_past < 1) return fail("need x_{n-1}");
  double c[WS_MAX_WINDOW] = {};
  const int K = order > 0 ? std::max(order, window) : 0;
  if (order > 0) ws_weights(order, K, c);
  // past[i] = x_{n-1-i}; d_{n-m} = past[m-1] - past[m], zero where the history runs out
  *out = ws_extrapolate(past[0], K, c, [&](int j) {
    const uint64_t m = (uint64_t)(j * period);
    return m < n_past ? past[m - 1] - past[m] : 0.0;
  });
  return 0;
}

int wco_set_threads(wco_solver* s, int32_t threads) {
  s->par.threads = threads < 1 ? 1 : threads;
  return 0;
}

int wco_state_set(wco_solver* s, const double* h, const double* hu, const double* hv,
                  const double* breaking, double time) {
  const size_t n = s->g.cells();
  s->s.h.assign(h, h + n);
  s->s.hu.assign(hu, hu + n);
  s->s.hv.assign(hv, hv + n);
  if (breaking)
    s->s.breaking.assign(breaking, breaking + n);
  else
    s->s.breaking.assign(n, 0.0);
  s->s.time = time;
  return 0;
}

int wco_state_get(const wco_solver* s, double* h, double* hu, double* hv, double* breaking,
                  double* time) {
  const size_t n = s->g.cells() * sizeof(double);
  if (h) memcpy(h, s->s.h.data(), n);
  if (hu) memcpy(hu, s->s.hu.data(), n);
  if (hv) memcpy(hv, s->s.hv.data(), n);
  if (breaking) memcpy(breaking, s->s.breaking.data(), n);
  if (time) *time = s->s.time;
  return 0;
}

int wco_stable_dt(wco_solver* s, double* dt) {
  *dt = s->stable_dt();
  return 0;
}

int wco_step(wco_solver* s, double dt) {
  s->step(dt);
  return 0;
}

// solver.rs:455-477
int wco_limiting_cell(wco_solver* s, int32_t* found, wco_limiting_t* out) {
  const Grid& g = s->g;
  bool have = false;
  double worst = 0.0;
  size_t wi = 0, wj = 0;
  for (size_t j = GH; j < GH + g.ny; ++j) {
    bool stop = false;
    for (size_t i = GH; i < GH + g.nx; ++i) {
      const double r = s->rate(g.idx(i, j));
      if (isnan(r)) {
        have = true;
        worst = r;
        wi = i;
        wj = j;
        stop = true;
        break;
      }
      if (r > (have ? worst : 0.0)) {
        have = true;
        worst = r;
        wi = i;
        wj = j;
      }
    }
    if (stop) break;
  }
  *found = have ? 1 : 0;
  if (!have) return 0;
  const size_t k = g.idx(wi, wj);
  const double u = wc_velocity(s->s.h[k], s->s.hu[k]);
  const double v = wc_velocity(s->s.h[k], s->s.hv[k]);
  out->i = wi - GH;
  out->j = wj - GH;
  out->depth = s->s.h[k];
  out->speed = s->math.hypot(u, v);
  return 0;
}

// dispersion.rs:166-201 (on a copy: the state's own ghosts stay as they are)
int wco_breaking_indicator(wco_solver* s, double* out) {
  const Grid& g = s->g;
  State filled = s->s;
  filled.fill_boundaries(g, s->periodic);
  std::vector<double> phi(g.cells(), 0.0);
  switch_phi(g, s->par, s->math, s->bed.data(), filled, s->breaking_config(), s->periodic,
             phi.data());
  size_t n = 0;
  for (size_t j = GH; j < GH + g.ny; ++j)
    for (size_t i = GH; i < GH + g.nx; ++i, ++n) {
      const size_t k = g.idx(i, j);
      out[n] = filled.h[k] < H_DISP_MIN ? 0.0 : 1.0 - phi[k];
    }
  return 0;
}

// wavesim's Stats and Stage::record (crates/wavesim/src/lib.rs:899-932, 1099-1130, 1276-1300)
struct wco_stats {
  wco_solver* s = nullptr;
  double from = 0, period = 0, next_wave = 0, seconds = 0, omega = 0;
  double moments[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> breaking_seconds, eta_max, eta_min, wave_max, wave_min;
  std::vector<std::vector<double>> waves;
  std::vector<double> sums;  // 3 per cell
  std::vector<float> elevation;
};

int wco_stats_new(wco_solver* s, double from, double period, const float* elevation,
                  wco_stats** out) {
  auto* st = new wco_stats;
  const size_t n = s->g.nx * s->g.ny;
  st->s = s;
  st->from = from;
  st->period = period;
  st->next_wave = from + period;
  st->omega = TAU / period;
  st->breaking_seconds.assign(n, 0.0);
  st->eta_max.assign(n, -INFINITY);
  st->eta_min.assign(n, INFINITY);
  st->wave_max.assign(n, -INFINITY);
  st->wave_min.assign(n, INFINITY);
  st->sums.assign(3 * n, 0.0);
  if (elevation) st->elevation.assign(elevation, elevation + n);
  *out = st;
  return 0;
}

void wco_stats_free(wco_stats* st) { delete st; }

namespace {
// Stats::close_wave (lib.rs:918-932)
void close_wave(wco_stats* st) {
  std::vector<double> heights(st->wave_max.size());
  for (size_t n = 0; n < heights.size(); ++n)
    heights[n] = std::isfinite(st->wave_max[n]) ? st->wave_max[n] - st->wave_min[n] : 0.0;
  st->waves.push_back(heights);
  std::fill(st->wave_max.begin(), st->wave_max.end(), -INFINITY);
  std::fill(st->wave_min.begin(), st->wave_min.end(), INFINITY);
  st->next_wave += st->period;
}
}  // namespace

int wco_stats_finish(wco_stats* st) {
  if (st->s->s.time >= st->next_wave - 1e-9) close_wave(st);
  return 0;
}

int wco_stats_record(wco_stats* st, double dt) {
  wco_solver* s = st->s;
  const Grid& g = s->g;
  const double inside = std::min(dt, s->s.time - st->from);
  if (inside <= 0.0) return 0;
  std::vector<double> now(g.nx * g.ny);
  wco_breaking_indicator(s, now.data());
  if (s->s.time > st->next_wave + 1e-9) close_wave(st);
  double sn, cs;  // (omega t).sin_cos(), one sincos call as above
  sincos(st->omega * s->s.time, &sn, &cs);
  const double v[6] = {1.0, cs, sn, cs * cs, cs * sn, sn * sn};
  for (int m = 0; m < 6; ++m) st->moments[m] += inside * v[m];
  const double w[3] = {inside, inside * cs, inside * sn};
  size_t n = 0;
  for (size_t j = GH; j < GH + g.ny; ++j)
    for (size_t i = GH; i < GH + g.nx; ++i, ++n) {
      const size_t k = g.idx(i, j);
      const double eta =
          s->s.h[k] + (st->elevation.empty() ? s->bed[k] : (double)st->elevation[n]);
      for (int m = 0; m < 3; ++m) st->sums[3 * n + m] += w[m] * eta;
      st->eta_max[n] = wc_max(st->eta_max[n], eta);
      st->eta_min[n] = wc_min(st->eta_min[n], eta);
      st->wave_max[n] = wc_max(st->wave_max[n], eta);
      st->wave_min[n] = wc_min(st->wave_min[n], eta);
      if (now[n] >= 0.8) st->breaking_seconds[n] += inside;
    }
  st->seconds += inside;
  return 0;
}

int wco_stats_read(wco_stats* st, double* seconds, uint64_t* waves, double* breaking_seconds,
                   double* eta_max, double* eta_min, double* wave_mean, double* wave_first,
                   double* wave_last) {
  const size_t n = st->eta_max.size();
  if (seconds) *seconds = st->seconds;
  if (waves) *waves = st->waves.size();
  for (size_t q = 0; q < n; ++q) {
    if (breaking_seconds) breaking_seconds[q] = st->breaking_seconds[q];
    if (eta_max) eta_max[q] = st->eta_max[q];
    if (eta_min) eta_min[q] = st->eta_min[q];
    double sum = 0.0;
    for (const auto& wv : st->waves) sum += wv[q];
    if (wave_mean) wave_mean[q] = st->waves.empty() ? 0.0 : sum / (double)st->waves.size();
    if (wave_first) wave_first[q] = st->waves.empty() ? 0.0 : st->waves.front()[q];
    if (wave_last) wave_last[q] = st->waves.empty() ? 0.0 : st->waves.back()[q];
  }
  return 0;
}

// FirstHarmonic::amplitudes (lib.rs:1136-1160)
int wco_stats_harmonic(wco_stats* st, int32_t* ok, double* a, double* b) {
  const double* M = st->moments;
  const double m[3][3] = {{M[0], M[1], M[2]}, {M[1], M[3], M[4]}, {M[2], M[4], M[5]}};
  auto cof = [&](int r, int k) {
    const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    return m[r1][k1] * m[r2][k2] - m[r1][k2] * m[r2][k1];
  };
  double det = 0.0;
  for (int k = 0; k < 3; ++k) det += m[0][k] * cof(0, k);
  if (std::isnan(det) || det <= 1e-3 * M[0] * M[0] * M[0] / 4.0) {
    *ok = 0;
    return 0;
  }
  *ok = 1;
  const size_t n = st->eta_max.size();
  for (size_t q = 0; q < n; ++q) {
    auto row = [&](int r) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += cof(k, r) / det * st->sums[3 * q + k];
      return t;
    };
    if (a) a[q] = row(1);
    if (b) b[q] = -row(2);
  }
  return 0;
}

int wco_dispersion_stats(const wco_solver* s, int32_t* has, wco_dispersion_stats_t* out) {
  *has = s->dispersive ? 1 : 0;
  if (s->dispersive) *out = s->work.stats;
  return 0;
}

// state.rs:79-85 (order 0: Rust's serial sum from -0.0; otherwise the canonical tree of the
// GPU's state sums, 64x4 tiles, whichever order its PCG takes)
int wco_state_volume(wco_solver* s, double* out) {
  const Grid& g = s->g;
  double sum;
  if (s->dot_order != 0) {
    const double* h = s->s.h.data();
    sum = canonical_sum(g, s->par, [&](size_t i, size_t j) { return h[g.idx(i, j)]; });
  } else {
    sum = -0.0;
    for (size_t j = GH; j < GH + g.ny; ++j)
      for (size_t i = GH; i < GH + g.nx; ++i) sum = sum + s->s.h[g.idx(i, j)];
  }
  *out = sum * g.dx * g.dy;
  return 0;
}

// state.rs:88-95
int wco_state_max_abs_momentum(wco_solver* s, double* out) {
  const Grid& g = s->g;
  double m = 0.0;
  for (size_t j = GH; j < GH + g.ny; ++j)
    for (size_t i = GH; i < GH + g.nx; ++i) {
      const size_t k = g.idx(i, j);
      m = wc_max(m, wc_max(fabs(s->s.hu[k]), fabs(s->s.hv[k])));
    }
  *out = m;
  return 0;
}

// state.rs:98-102
int wco_state_min_depth(wco_solver* s, double* out) {
  const Grid& g = s->g;
  double m = INFINITY;
  for (size_t j = GH; j < GH + g.ny; ++j)
    for (size_t i = GH; i < GH + g.nx; ++i) m = wc_min(m, s->s.h[g.idx(i, j)]);
  *out = m;
  return 0;
}

int wco_fill_boundaries(wco_solver* s) {
  s->s.fill_boundaries(s->g, s->periodic);
  return 0;
}

int wco_linear_wave(double omega, double h, int32_t dispersive, wco_linear_wave_t* out) {
  if (!linear_wave(omega, h, dispersive != 0, out))
    return fail("a wave of this angular frequency is too short for the dispersive model");
  return 0;
}

double wco_manning_factor(double h, double speed, double dt, double n) {
  const double k = G * n * n / pow(h, 4.0 / 3.0);
  return 1.0 / (1.0 + dt * k * speed);
}

void wco_face_values(const double v[5], double blend, double* lo, double* hi) {
  face_values(v, blend, lo, hi);
}

double wco_smoothstep(double lo, double hi, double x) { return wc_smoothstep(lo, hi, x); }

double wco_hypot(double a, double b, int32_t portable) {
  return portable ? wc_hypot(a, b) : hypot(a, b);
}

double wco_pow43(double h, int32_t portable) { return portable ? wc_pow43(h) : pow(h, 4.0 / 3.0); }

// The GPU switch's shortcut (wc_math.h: wc_switch_off, which skips the portable hypot where its
// bits cannot matter) and the full expression, over n cells, for the tests that pin the
// shortcut against the expression it replaces.
int wco_switch_off(const double* thresholds, const double* ratio, const double* ex,
                   const double* ey, uint64_t n, double* shortcut, double* full) {
  const double er_lo = thresholds[0], er_hi = thresholds[1], sl_lo = thresholds[2],
               sl_hi = thresholds[3];
  for (uint64_t k = 0; k < n; ++k) {
    shortcut[k] = wc_switch_off(er_lo, er_hi, sl_lo, sl_hi, ratio[k], ex[k], ey[k]);
    full[k] = wc_switch_off_full(er_lo, er_hi, sl_lo, sl_hi, ratio[k], ex[k], ey[k]);
  }
  return 0;
}

int wco_matvec(uint64_t nx, uint64_t ny, double dx, double dy, int32_t periodic,
               const double* heff, const double* mask, const double* c, const double* px,
               const double* py, double* ax, double* ay) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.dx = dx;
  g.dy = dy;
  const size_t n = g.cells();
  std::vector<double> f(n, 0.0), x(px, px + n), y(py, py + n);
  Par par;
  matvec(g, par, periodic != 0, heff, mask, c, f.data(), x.data(), y.data(), ax, ay);
  return 0;
}

double wco_dot(uint64_t nx, uint64_t ny, int32_t order, const double* ax, const double* ay,
               const double* bx, const double* by) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.dx = 1.0;
  g.dy = 1.0;
  Par par;
  return dot(g, par, order, ax, ay, bx, by);
}

int wco_dispersion_apply(uint64_t nx, uint64_t ny, double dx, double dy, int32_t periodic,
                         const wco_dispersion_t* d, const double* h, const double* hu,
                         const double* hv, const double* phi, const double* lh, double* lhu,
                         double* lhv, int32_t dot_order) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.dx = dx;
  g.dy = dy;
  const size_t n = g.cells();
  State st;
  st.h.assign(h, h + n);
  st.hu.assign(hu, hu + n);
  st.hv.assign(hv, hv + n);
  st.breaking.assign(n, 0.0);
  DispersionCfg cfg;
  cfg.breaking.er_lo = d->breaking.eta_ratio_lo;
  cfg.breaking.er_hi = d->breaking.eta_ratio_hi;
  cfg.breaking.sl_lo = d->breaking.slope_lo;
  cfg.breaking.sl_hi = d->breaking.slope_hi;
  cfg.nonlinear = d->nonlinear != 0;
  cfg.tolerance = d->tolerance;
  cfg.max_iterations = d->max_iterations;
  Work wk;
  wk.init(g);
  Par par;
  dispersion_apply(g, par, dot_order, cfg, periodic != 0, st, lh, lhu, lhv, phi, wk);
  return 0;
}

}  // extern "C"
