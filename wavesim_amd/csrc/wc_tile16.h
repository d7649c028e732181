// 64x16-cell tiles for the PCG variants that do not follow the oracle's summation order
// (wc_pcg1.hip, wc_fpcg.hip): one 1024-thread block per tile, a cell per thread, the vector
// operand staged with its two-cell halo in LDS; 256 blocks each add their tiles' sums into one
// lane, and every wave of the next kernel reduces the 256 lanes itself.
#pragma once

#include "wc_kernels.h"

namespace wc {
namespace t16 {

constexpr int LANES = 256;
constexpr int BX = 64, BY = 16;  // tile
constexpr int PT = BX * BY;      // threads per block, a cell each
constexpr int S2X = BX + 4, S2Y = BY + 4;
constexpr int S1X = BX + 2, S1Y = BY + 2;
constexpr int H2 = S2X * S2Y, H1 = S1X * S1Y;

template <int N, int L = LANES>
__device__ __forceinline__ void lane_sums(const double* acc, double (&out)[N]) {
  static_assert(L == 256 || L == 512, "256 or 512 lanes");
  const int lane = threadIdx.x & 63;
  constexpr int G = L / 64;
  double v[N][G];
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int w = 0; w < G; ++w) v[n][w] = acc[n * L + 64 * w + lane];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const double lo = (wave_sum(v[n][0]) + wave_sum(v[n][1])) + (wave_sum(v[n][2]) + wave_sum(v[n][3]));
    if constexpr (G == 4) {
      out[n] = lo;
    } else {
      const double hi =
          (wave_sum(v[n][4]) + wave_sum(v[n][5])) + (wave_sum(v[n][6]) + wave_sum(v[n][7]));
      out[n] = lo + hi;
    }
  }
}

__device__ __forceinline__ int tiles_x(const Geo& g) { return (g.nx + BX - 1) / BX; }
__device__ __forceinline__ int tiles_all(const Geo& g) {
  return tiles_x(g) * ((g.ny + BY - 1) / BY);
}

// The summation lane this block carries, in a launch of `stride` blocks where lane b takes
// tiles b, b + stride, ... and adds their sums in that order whichever block runs it, so the
// choice changes no bit. Blocks run on XCD blockIdx.x % 8. When the tile columns are a multiple
// of 8 and divide the stride, each XCD takes the lanes of its own eighth of the tile columns (at
// 64 tile columns and 256 lanes: 8 columns by the 4 tile rows of each pass), so a tile's left
// and right neighbours, whose halo columns share its 128-byte lines, run on the same XCD and
// meet in its L2, as the tiles above and below do. Otherwise lane b runs on block b. Block 0
// carries lane 0 either way.
__device__ __forceinline__ int lane_of_block(const Geo& g, int stride = LANES) {
  const int tx = tiles_x(g);
  const int b = blockIdx.x;
  if ((int)gridDim.x != stride || (stride & 7) != 0 || tx < 8 || (tx & 7) != 0 ||
      stride % tx != 0)
    return b;
  const int cols = tx / 8;           // tile columns per XCD
  const int x = b & 7, s = b >> 3;  // XCD, slot (0 .. stride / 8 - 1)
  return (s / cols) * tx + x * cols + s % cols;
}

// Adds this tile's terms into the thread's own running sums (its cells of the block's tiles, in
// tile order), and keeps the next tile's LDS staging from overwriting what this one still reads.
// The block reduction runs once, after the last tile (rows_total), not once per tile.
template <int N>
__device__ __forceinline__ void add_rows(double (&v)[N], double (*)[BY][N], int,
                                         double (&acc)[N]) {
#pragma unroll
  for (int n = 0; n < N; ++n) acc[n] = acc[n] + v[n];
  __syncthreads();
}

// The block's total of the threads' running sums into thread 0's acc: each row of 64 threads by
// the wave butterfly, then the rows in order.
template <int N>
__device__ __forceinline__ void rows_total(double (&acc)[N], double (*sums)[BY][N]) {
  const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const double t = wave_sum(acc[n]);
    if (lane == 0) sums[0][row][n] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double t = 0.0;
      for (int y = 0; y < BY; ++y) t = t + sums[0][y][n];
      acc[n] = t;
    }
}

// c (grad . u) on the tile + 1 from the staged u, then a barrier.
__device__ __forceinline__ void stencil_f(const Geo& g, const float* c, int tile,
                                          double (*s_u)[S2Y][S2X], double (*s_f)[S1X]) {
  const int tx = tiles_x(g);
  const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < H1; t += PT) {
    const int lx = t % S1X, ly = t / S1X;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    double f = 0.0;
    if (i <= g.nx + GH && j <= g.ny + GH) {
      const int x = lx + 1, y = ly + 1;
      f = (double)c[j * g.w + i] * ((s_u[0][y][x + 1] - s_u[0][y][x - 1]) * ix +
                                    (s_u[1][y + 1][x] - s_u[1][y - 1][x]) * iy);
    }
    s_f[ly][lx] = f;
  }
  __syncthreads();
}

// c at the thread's (at most two) tile + 1 points, zero outside the used rows and columns (as
// stencil_f has it), fetched a tile ahead for stencil_f_staged.
__device__ __forceinline__ void prefetch_c(const Geo& g, const float* c, int tile, float (&v)[2]) {
  const int tx = tiles_x(g);
  const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = threadIdx.x + e * PT;
    v[e] = 0.f;
    if (t < H1) {
      const int i = i0 - 1 + t % S1X, j = j0 - 1 + t / S1X;
      if (i <= g.nx + GH && j <= g.ny + GH) v[e] = c[j * g.w + i];
    }
  }
}

// stencil_f with c staged in LDS (s_c, the tile + 1 in rows of S1X) instead of read from memory
// after the barrier. Where c is staged as zero (outside the used rows and columns) f is a zero,
// possibly of either sign, which no cell of the tile reads.
__device__ __forceinline__ void stencil_f_staged(const Geo& g, double (*s_u)[S2Y][S2X],
                                                 const float* s_c, double (*s_f)[S1X]) {
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < H1; t += PT) {
    const int lx = t % S1X, ly = t / S1X;
    const int x = lx + 1, y = ly + 1;
    s_f[ly][lx] = (double)s_c[t] * ((s_u[0][y][x + 1] - s_u[0][y][x - 1]) * ix +
                                    (s_u[1][y + 1][x] - s_u[1][y - 1][x]) * iy);
  }
  __syncthreads();
}

// A u at the thread's cell (and u there) from the staged u and f.
__device__ __forceinline__ void stencil_own_h(const Geo& g, float h_own, int tile,
                                              double (*s_u)[S2Y][S2X], double (*s_f)[S1X],
                                              bool& own, int& k, double& ux, double& uy,
                                              double& ax, double& ay);

__device__ __forceinline__ void stencil_own(const Geo& g, const float* hh, int tile,
                                            double (*s_u)[S2Y][S2X], double (*s_f)[S1X],
                                            bool& own, int& k, double& ux, double& uy,
                                            double& ax, double& ay) {
  const int t0 = threadIdx.x;
  const int tx = tiles_x(g);
  const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const int lx = t0 & 63, ly = t0 >> 6;
  const int i = i0 + lx, j = j0 + ly;
  own = i < g.nx + GH && j < g.ny + GH;
  k = j * g.w + i;
  ux = s_u[0][ly + 2][lx + 2];
  uy = s_u[1][ly + 2][lx + 2];
  ax = ay = 0.0;
  if (own) {
    const double h = (double)hh[k];
    const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
    const double heff = wc_max(h, WC_H_DRY);
    const int x1 = lx + 1, y1 = ly + 1;
    ax = mask * (heff * ux - (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix);
    ay = mask * (heff * uy - (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy);
  }
}

// stencil_own with the cell's depth h_own fetched ahead (any value outside the interior).
__device__ __forceinline__ void stencil_own_h(const Geo& g, float h_own, int tile,
                                              double (*s_u)[S2Y][S2X], double (*s_f)[S1X],
                                              bool& own, int& k, double& ux, double& uy,
                                              double& ax, double& ay) {
  const int t0 = threadIdx.x;
  const int tx = tiles_x(g);
  const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const int lx = t0 & 63, ly = t0 >> 6;
  const int i = i0 + lx, j = j0 + ly;
  own = i < g.nx + GH && j < g.ny + GH;
  k = j * g.w + i;
  ux = s_u[0][ly + 2][lx + 2];
  uy = s_u[1][ly + 2][lx + 2];
  ax = ay = 0.0;
  if (own) {
    const double h = (double)h_own;
    const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
    const double heff = wc_max(h, WC_H_DRY);
    const int x1 = lx + 1, y1 = ly + 1;
    ax = mask * (heff * ux - (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix);
    ay = mask * (heff * uy - (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy);
  }
}

// Stages the vector u (given per halo cell by fn(m, vx, vy), m the mapped interior index; the
// wall flips are applied here) on the tile + 2 and c (grad . u) on the tile + 1; returns A u
// at the thread's cell (and u there).
template <class Fn>
__device__ __forceinline__ void stencil(const Geo& g, const float* c, const float* hh, int tile,
                                        double (*s_u)[S2Y][S2X], double (*s_f)[S1X], Fn fn,
                                        bool& own, int& k, double& ux, double& uy, double& ax,
                                        double& ay) {
  const int t0 = threadIdx.x;
  const int tx = tiles_x(g);
  const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
  for (int t = t0; t < H2; t += PT) {
    const int lx = t % S2X, ly = t / S2X;
    const int i = i0 - 2 + lx, j = j0 - 2 + ly;
    double vx = 0.0, vy = 0.0;
    if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
      const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
      fn(m, vx, vy);
      if (is_ghost_x(i, g.nx)) vx = -vx;
      if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
    }
    s_u[0][ly][lx] = vx;
    s_u[1][ly][lx] = vy;
  }
  __syncthreads();
  stencil_f(g, c, tile, s_u, s_f);
  stencil_own(g, hh, tile, s_u, s_f, own, k, ux, uy, ax, ay);
}

// The staged vector's inputs at a tile's halo points (at most two per thread), fetched into
// registers ahead of their use, so that the next tile's loads are in flight while this tile
// computes. NV float2 values per point.
template <int NV>
struct Prefetch {
  float2 v[2][NV];
  int flags[2];  // 1: inside the grid's two rings; 2, 4: flip x, y (wall ghosts)
};

template <int NV, class Fetch>
__device__ __forceinline__ void prefetch(const Geo& g, int tile, Fetch fetch, Prefetch<NV>& P) {
  const int tx = tiles_x(g);
  const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = threadIdx.x + e * PT;
    P.flags[e] = 0;
    if (t < H2) {
      const int lx = t % S2X, ly = t / S2X;
      const int i = i0 - 2 + lx, j = j0 - 2 + ly;
      if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
        const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
        fetch(m, P.v[e]);
        P.flags[e] = 1 | (is_ghost_x(i, g.nx) ? 2 : 0) |
                     (!g.periodic && is_ghost_y(j, g.ny) ? 4 : 0);
      }
    }
  }
}

// u from the prefetched inputs by make(v, vx, vy), into LDS on the tile + 2; the inputs of the
// tile's own cells also into `own` (NV x 64x16), where the cell epilogue reads them instead of
// loading them again (by then L2 may have let them go).
template <int NV, class Make>
__device__ __forceinline__ void stage(const Prefetch<NV>& P, double (*s_u)[S2Y][S2X],
                                      float2 (*own)[BY][BX], Make make) {
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = threadIdx.x + e * PT;
    if (t < H2) {
      const int lx = t % S2X, ly = t / S2X;
      double vx = 0.0, vy = 0.0;
      if (P.flags[e] & 1) {
        make(P.v[e], vx, vy);
        if (P.flags[e] & 2) vx = -vx;
        if (P.flags[e] & 4) vy = -vy;
      }
      s_u[0][ly][lx] = vx;
      s_u[1][ly][lx] = vy;
      if (lx >= 2 && lx < BX + 2 && ly >= 2 && ly < BY + 2)
#pragma unroll
        for (int q = 0; q < NV; ++q) own[q][ly - 2][lx - 2] = P.v[e][q];
    }
  }
}

}  // namespace t16
}  // namespace wc
