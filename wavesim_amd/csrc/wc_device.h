// Device-side building blocks of wavecore's step for gfx950: the ghost map, the canonical
// wavefront reduction, and the per-cell finite-volume arithmetic.
//
// Every floating-point expression keeps the reference's order of operations (Rust evaluates
// left to right and never contracts a*b+c; this file is compiled with -ffp-contract=off), so
// a kernel built from these pieces reproduces the CPU oracle bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wc_math.h"

namespace wc {

constexpr int GH = WC_GHOST;
constexpr int TX = 64;  // tile width: one wavefront along x, reads and writes coalesced
constexpr int TY = 4;   // tile height: four wavefronts per 256-thread block
constexpr int NT = TX * TY;
constexpr int NXCD = 8;  // MI355X: 8 XCDs, each with its own L2
// Cells per block of the largest-rate reductions (stable_dt): one atomic per block, not per
// wave, so that 8 M cells take 2048 atomics on the one address instead of 131,072.
constexpr int RATE_CELLS = 16 * NT;

// Grid geometry as kernels see it.
struct Geo {
  int nx, ny;      // interior cells
  int w, rows;     // padded width and rows
  double dx, dy;
  int periodic;    // south/north edges wrap
  int ntx, nty;    // 64x4 tiles over the interior
};

// ---- the ghost map ------------------------------------------------------------------------
// What the reference's fills (grid.rs:61-114, state.rs:44-76) leave in a ghost cell, as a
// pure function of the interior: mirror across the wall, or wrap in y when periodic. Valid
// for nx, ny >= 3, which the host enforces.
__device__ __forceinline__ int map_x(int i, int nx) {
  if (i < GH) return 2 * GH - 1 - i;
  if (i >= nx + GH) return 2 * (nx + GH) - 1 - i;
  return i;
}
__device__ __forceinline__ int map_y(int j, int ny, int periodic) {
  if (j < GH) return periodic ? j + ny : 2 * GH - 1 - j;
  if (j >= ny + GH) return periodic ? j - ny : 2 * (ny + GH) - 1 - j;
  return j;
}
__device__ __forceinline__ bool is_ghost_x(int i, int nx) { return i < GH || i >= nx + GH; }
__device__ __forceinline__ bool is_ghost_y(int j, int ny) { return j < GH || j >= ny + GH; }

// ---- division by a divisor every cell shares ------------------------------------------------
// hipcc expands an FP64 division x / y into v_div_scale (of y, then of x), v_rcp_f64 of the
// scaled y, two Newton steps on that reciprocal (four FMAs), the product m = x r, the remainder
// e = fma(-y, m, x), v_div_fmas (fma(e, r, m), scaled back when v_div_scale scaled) and
// v_div_fixup (zeros, infinities, NaN, overflow and underflow): eleven instructions, one of them
// quarter rate, with the correctly rounded quotient. When y is the same for every cell (the 60
// of the fifth-order formula, dx, dy), the reciprocal and its refinement are computed once
// (quot_of), and where neither v_div_scale nor v_div_fixup changes anything (0 < |x| within
// [2^-900, 2^700], 2^-60 <= |y| <= 2^60) the quotient takes the three instructions the expansion
// runs there (div_q), so its bits are the division's; other x take the full division.
struct Quot {
  double y, r;
  bool fast;
};

// The same double in every lane, held in scalar registers.
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// y must be the same in every lane.
__device__ __forceinline__ Quot quot_of(double y) {
  double v = y;
  asm volatile("" : "+v"(v));  // a run-time value: the reciprocal is the instruction's, not folded
  const double rcp = __builtin_amdgcn_rcp(v);
  const double e0 = fma(-v, rcp, 1.0);
  const double r1 = fma(rcp, e0, rcp);
  const double e1 = fma(-v, r1, 1.0);
  const double ay = fabs(y);
  return {y, wave_uniform(fma(r1, e1, r1)), ay >= 0x1p-60 && ay <= 0x1p60};
}

__device__ __forceinline__ double div_q(double x, const Quot& d) {
  const double ax = fabs(x);
  if (d.fast && ax >= 0x1p-900 && ax <= 0x1p700) {
    const double m = x * d.r;
    return fma(fma(-d.y, m, x), d.r, m);
  }
  return x / d.y;
}

// ---- stores of streamed outputs ------------------------------------------------------------
// A pass's outputs at its own cells, which no other block of the launch reads and the next launch
// reads from HBM anyway (a 4096x2048 grid's arrays are 67 MB each), stored with the
// nontemporal hint.
template <typename T>
__device__ __forceinline__ void store_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void store_nt(double2* p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
__device__ __forceinline__ void store_nt(float2* p, float2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}

// A scalar field at any padded cell, from its interior.
__device__ __forceinline__ double ghosted(const double* a, const Geo& g, int i, int j) {
  return a[map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx)];
}
// The x and y components of a vector field at any padded cell: the normal component flips at
// a wall, nothing flips at a periodic seam.
__device__ __forceinline__ double ghosted_vx(const double* a, const Geo& g, int i, int j) {
  const double v = ghosted(a, g, i, j);
  return is_ghost_x(i, g.nx) ? -v : v;
}
__device__ __forceinline__ double ghosted_vy(const double* a, const Geo& g, int i, int j) {
  const double v = ghosted(a, g, i, j);
  return (!g.periodic && is_ghost_y(j, g.ny)) ? -v : v;
}

// ---- XCD-aware tile order -------------------------------------------------------------------
// Hardware block b runs on XCD b % 8. Give each XCD a contiguous band of tile rows, so the
// halo rows two vertically adjacent tiles share sit in the same L2.
__device__ __forceinline__ int tile_of_block(int b, int nb) {
  const int x = b % NXCD, slot = b / NXCD;
  const int q = nb / NXCD, r = nb % NXCD;
  return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// ---- the canonical reduction --------------------------------------------------------------
// Interior reductions sum in one fixed order, which the oracle's canonical mode repeats:
// within a 64x4 tile, a 64-lane butterfly per row (lane l adds lane l^1, then l^2, l^4, ...,
// l^32), rows paired as (r0 + r1) + (r2 + r3); then the tile sums, by 256 lanes that each add
// tiles lane, lane+256, ... from 0.0, and a final butterfly per wave with the waves paired.
//
// The butterfly runs without the LDS crossbar: the first two steps are quad permutations, and
// once each quad (then each half row) holds one value, the l^4 and l^8 partners are the same
// values the half-row and row mirrors bring; the l^16 and l^32 steps add the four row sums,
// read from lanes 0, 16, 32, 48, as (R0 + R1) + (R2 + R3). Every lane computes its own sum
// and its partner's with the same two operands, so all 64 lanes end with the same bits, which
// are those of the plain butterfly. All 64 lanes of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v = v + dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]: lane l^1
  v = v + dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]: lane l^2
  v = v + dpp_move<0x141>(v);  // row_half_mirror: the other quad of the half row, l^4
  v = v + dpp_move<0x140>(v);  // row_mirror: the other half row, l^8
  const double r0 = lane_value(v, 0), r1 = lane_value(v, 16);
  const double r2 = lane_value(v, 32), r3 = lane_value(v, 48);
  return (r0 + r1) + (r2 + r3);
}

// The largest of the block's values (IEEE bit patterns of non-negative doubles, or NaN, whose
// patterns exceed every number's, so NaN wins as the reference's max_or_nan) into *out: a
// wave butterfly, the block's waves through LDS, one atomic per block. Maxima do not depend on
// the order they are taken in, so any grouping gives the same bits.
__device__ __forceinline__ void block_max_to(unsigned long long bits, unsigned long long* out) {
  __shared__ unsigned long long lds_max[NT / 64];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_xor(bits, off, 64);
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0) lds_max[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; ++w) bits = lds_max[w] > bits ? lds_max[w] : bits;
    if (bits != 0ull) atomicMax(out, bits);
  }
  __syncthreads();
}

template <int N>
struct Partials {
  double v[N];
};

// Reduce N per-thread terms over the 64x4 tile; thread 0 returns the tile sums.
template <int N>
__device__ __forceinline__ void tile_sums(double (&v)[N], double (&out)[N]) {
  __shared__ double lds[N][TY];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const double s = wave_sum(v[n]);
    if (lane == 0) lds[n][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < N; ++n) out[n] = (lds[n][0] + lds[n][1]) + (lds[n][2] + lds[n][3]);
}

// The second stage, run by one whole 256-thread block over np tile sums (stride ntiles between
// the N reductions); every thread gets the results.
template <int N>
__device__ __forceinline__ void final_sums(const double* part, int np, double (&out)[N]) {
  __shared__ double lds[N][TY];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double acc = 0.0;
    for (int p = threadIdx.x; p < np; p += NT) acc = acc + part[n * np + p];
    const double s = wave_sum(acc);
    if (lane == 0) lds[n][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < N; ++n) out[n] = (lds[n][0] + lds[n][1]) + (lds[n][2] + lds[n][3]);
  __syncthreads();
}

// Writes this tile's N sums to part[n * np + tile] with plain stores; a separate one-block
// kernel adds them up once every tile has finished (the kernel boundary makes them visible on
// every XCD, which a device-scope fence per block would do at the price of an L2 write-back).
template <int N>
__device__ __forceinline__ void store_tile(double (&v)[N], double* part, int np, int tile) {
  double sums[N];
  tile_sums<N>(v, sums);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int n = 0; n < N; ++n) part[n * np + tile] = sums[n];
  }
}

// ---- finite volumes (solver.rs:102-239) -------------------------------------------------------
struct Cell {
  double h, un, ut, b, phi;
};
struct Side {
  double h, un, ut, z;
};
struct Face {
  double f0, f1, f2, cl, cr;
};

__device__ __forceinline__ double minmod3(double a, double b, double c) {
  if (a > 0.0 && b > 0.0 && c > 0.0) return wc_min(wc_min(a, b), c);
  if (a < 0.0 && b < 0.0 && c < 0.0) return wc_max(wc_max(a, b), c);
  return 0.0;
}

__device__ __forceinline__ void face_values(double low2, double low, double c, double high,
                                            double high2, double blend, double& lo, double& hi) {
  const double s = minmod3(2.0 * (c - low), 0.5 * ((c - low) + (high - c)), 2.0 * (high - c));
  const double mc_lo = c - 0.5 * s, mc_hi = c + 0.5 * s;
  if (blend <= 0.0) {
    lo = mc_lo;
    hi = mc_hi;
    return;
  }
  const double fifth_hi = (2.0 * low2 - 13.0 * low + 47.0 * c + 27.0 * high - 3.0 * high2) / 60.0;
  const double fifth_lo = (2.0 * high2 - 13.0 * high + 47.0 * c + 27.0 * low - 3.0 * low2) / 60.0;
  lo = mc_lo + blend * (fifth_lo - mc_lo);
  hi = mc_hi + blend * (fifth_hi - mc_hi);
}

__device__ __forceinline__ void reconstruct(const Cell* cs, int order, Side& lo_out,
                                            Side& hi_out) {
  const Cell &low2 = cs[0], &low = cs[1], &c = cs[2], &high = cs[3], &high2 = cs[4];
  const Side first{c.h, c.un, c.ut, c.b};
  if (order == 0 || low.h < WC_H_SECOND || c.h < WC_H_SECOND || high.h < WC_H_SECOND) {
    lo_out = first;
    hi_out = first;
    return;
  }
  double blend = 0.0;
  if (order == 2 && low2.h >= WC_H_SECOND && high2.h >= WC_H_SECOND) {
    blend = 1.0;
#pragma unroll
    for (int m = 0; m < 5; ++m) blend = wc_min(blend, cs[m].phi);
  }
  double eta_lo, eta_hi, un_lo, un_hi, ut_lo, ut_hi;
  face_values(low2.h + low2.b, low.h + low.b, c.h + c.b, high.h + high.b, high2.h + high2.b,
              blend, eta_lo, eta_hi);
  face_values(low2.un, low.un, c.un, high.un, high2.un, blend, un_lo, un_hi);
  face_values(low2.ut, low.ut, c.ut, high.ut, high2.ut, blend, ut_lo, ut_hi);
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
  if (l.h < WC_H_FACE_MIN || r.h < WC_H_FACE_MIN) {
    lo_out = first;
    hi_out = first;
    return;
  }
  lo_out = l;
  hi_out = r;
}

__device__ __forceinline__ Face face(const Side& l, const Side& r) {
  const double z_star = wc_max(l.z, r.z);
  const double hl = wc_max(l.h + l.z - z_star, 0.0);
  const double hr = wc_max(r.h + r.z - z_star, 0.0);
  const double cl = sqrt(WC_G * hl);
  const double cr = sqrt(WC_G * hr);
  const double s_l = wc_min(l.un - cl, r.un - cr);
  const double s_r = wc_max(l.un + cl, r.un + cr);
  const double fl0 = hl * l.un, fl1 = hl * l.un * l.un + 0.5 * WC_G * hl * hl,
               fl2 = hl * l.un * l.ut;
  const double fr0 = hr * r.un, fr1 = hr * r.un * r.un + 0.5 * WC_G * hr * hr,
               fr2 = hr * r.un * r.ut;
  Face f;
  if (hl <= WC_H_DRY && hr <= WC_H_DRY) {
    f.f0 = f.f1 = f.f2 = 0.0;
  } else if (s_l >= 0.0) {
    f.f0 = fl0;
    f.f1 = fl1;
    f.f2 = fl2;
  } else if (s_r <= 0.0) {
    f.f0 = fr0;
    f.f1 = fr1;
    f.f2 = fr2;
  } else {
    const double ul0 = hl, ul1 = hl * l.un, ul2 = hl * l.ut;
    const double ur0 = hr, ur1 = hr * r.un, ur2 = hr * r.ut;
    const double inv = 1.0 / (s_r - s_l);
    f.f0 = (s_r * fl0 - s_l * fr0 + s_l * s_r * (ur0 - ul0)) * inv;
    f.f1 = (s_r * fl1 - s_l * fr1 + s_l * s_r * (ur1 - ul1)) * inv;
    f.f2 = (s_r * fl2 - s_l * fr2 + s_l * s_r * (ur2 - ul2)) * inv;
  }
  f.cl = 0.5 * WC_G * (l.h * l.h - hl * hl);
  f.cr = 0.5 * WC_G * (r.h * r.h - hr * hr);
  return f;
}

// Fluxes through the two faces of the middle of seven cells and the centred bed source.
__device__ __forceinline__ void axis_terms(const Cell* st, double spacing, int order, Face& low,
                                           Face& high, double& source) {
  Side a_lo, a_hi, lo, hi, b_lo, b_hi;
  reconstruct(st + 0, order, a_lo, a_hi);
  reconstruct(st + 1, order, lo, hi);
  reconstruct(st + 2, order, b_lo, b_hi);
  low = face(a_hi, lo);
  high = face(hi, b_lo);
  source = -WC_G * 0.5 * (lo.h + hi.h) * (hi.z - lo.z) / spacing;
}

}  // namespace wc
