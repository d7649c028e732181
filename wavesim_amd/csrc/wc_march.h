// Column-marching stencil for the FP32 form's one-kernel PCG (wc_pcg1.hip).
//
// The SGN operator A u needs u on the cell's two-cell neighbourhood and f = c div u on its
// one-cell neighbourhood (dispersion.rs:274-317). Instead of staging a separate halo for every
// 64x16 tile (a third of the loads again), each 1024-thread block marches down a strip 64
// cells wide, 16 rows at a time, keeping u and f in circular LDS buffers: a chunk loads only
// its 16 new rows of u (plus the two side columns each way), computes 16 new rows of f, and
// uses the rows the previous chunk left behind. The vertical halo is then loaded once per strip
// segment, and a chunk needs two barriers: after staging u and after computing f; the next
// chunk's staging overlaps this chunk's cell updates (the buffers are deep enough that it never
// overwrites a row still being read).
//
// Work: strips of 64 columns; when there are fewer strips than blocks each strip is cut into
// `segs` row segments (segment boundaries cost one extra 4-row prologue each). Block b takes
// units b, b + L, ... Each thread adds its cells' terms into its own running sums in the order
// it meets them, and the block's total is taken once at the end (t16::rows_total), so the
// result does not depend on timing.
#pragma once

#include "wc_tile16.h"

namespace wc {
namespace mrch {

constexpr int BX = 64;        // strip width: one wavefront per row, coalesced along x
constexpr int R = 16;         // rows per chunk: 16 wavefronts, one row each
constexpr int PT = BX * R;    // 1024 threads
constexpr int UX = BX + 4;    // u columns staged: the strip and two each side
constexpr int FX = BX + 2;    // f columns: the strip and one each side
constexpr int NBU = 36;       // u rows kept: a chunk reads rows [j, j+16) while the next
                              // writes [j+18, j+34), 36 apart at most
constexpr int NBF = 18;       // f rows kept: a chunk reads [j-1, j+17)

using LdsU = double[2][NBU][UX];
using LdsF = double[NBF][FX];

struct Unit {
  int i0, y0, y1;  // first column, rows [y0, y1) (padded indices)
};

__device__ __forceinline__ int strips(const Geo& g) { return (g.nx + BX - 1) / BX; }

// Row segments per strip for L blocks.
__device__ __forceinline__ int segments(const Geo& g, int L) {
  const int s = strips(g);
  int segs = s >= L ? 1 : L / s;
  return segs > g.ny ? g.ny : segs;
}

__device__ __forceinline__ Unit unit_of(const Geo& g, int u, int segs) {
  const int s = strips(g);
  const int strip = u % s, seg = u / s;
  Unit r;
  r.i0 = GH + strip * BX;
  r.y0 = GH + (int)((long long)seg * g.ny / segs);
  r.y1 = GH + (int)((long long)(seg + 1) * g.ny / segs);
  return r;
}

// March block `blk` (of L) over its units. load(m, vx, vy): the vector u at mapped interior
// index m (the wall flips are applied here). cell(own, k, ux, uy, ax, ay, v): the cell's
// epilogue, with A u at the cell; its N terms are added into acc (zero when not own).
template <int N, class Load, class Cell>
__device__ __forceinline__ void march(const Geo& g, const float* c, const float* hh, int blk,
                                      int L, LdsU& s_u, LdsF& s_f, Load load, Cell cell,
                                      double (&acc)[N]) {
  const int segs = segments(g, L);
  const int units = strips(g) * segs;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const int tid = threadIdx.x;
  for (int un = blk; un < units; un += L) {
    const Unit U = unit_of(g, un, segs);
    for (int j = U.y0; j < U.y1; j += R) {
      // u on rows [ua, j + R + 2): the first chunk of a unit also loads its two rows above
      const int ua = j == U.y0 ? j - 2 : j + 2, ub = j + R + 2;
      const int nu = (ub - ua) * UX;
      for (int t = tid; t < nu; t += PT) {
        const int ly = t / UX, lx = t - ly * UX;
        const int row = ua + ly, col = U.i0 - 2 + lx;
        double vx = 0.0, vy = 0.0;
        if (col <= g.nx + GH + 1 && row <= g.ny + GH + 1) {
          const int m = map_y(row, g.ny, g.periodic) * g.w + map_x(col, g.nx);
          load(m, vx, vy);
          if (is_ghost_x(col, g.nx)) vx = -vx;
          if (!g.periodic && is_ghost_y(row, g.ny)) vy = -vy;
        }
        const int r = (row - U.y0 + 2) % NBU;
        s_u[0][r][lx] = vx;
        s_u[1][r][lx] = vy;
      }
      __syncthreads();
      // f = c div u on rows [fa, j + R + 1)
      const int fa = j == U.y0 ? j - 1 : j + 1, fb = j + R + 1;
      const int nf = (fb - fa) * FX;
      for (int t = tid; t < nf; t += PT) {
        const int ly = t / FX, lx = t - ly * FX;
        const int row = fa + ly, col = U.i0 - 1 + lx;
        double f = 0.0;
        if (col <= g.nx + GH && row <= g.ny + GH) {
          const int r = (row - U.y0 + 2) % NBU;
          const int rm = (row - U.y0 + 1) % NBU, rp = (row - U.y0 + 3) % NBU;
          f = (double)c[row * g.w + col] * ((s_u[0][r][lx + 2] - s_u[0][r][lx]) * ix +
                                            (s_u[1][rp][lx + 1] - s_u[1][rm][lx + 1]) * iy);
        }
        s_f[(row - U.y0 + 1) % NBF][lx] = f;
      }
      __syncthreads();
      // the cells of rows [j, j + R)
      const int lx = tid & (BX - 1), ly = tid / BX;
      const int row = j + ly, col = U.i0 + lx;
      const bool own = row < U.y1 && col < g.nx + GH;
      const int k = row * g.w + col;
      const int ru = (row - U.y0 + 2) % NBU;
      const double ux = s_u[0][ru][lx + 2], uy = s_u[1][ru][lx + 2];
      double ax = 0.0, ay = 0.0;
      if (own) {
        const double h = (double)hh[k];
        const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
        const double heff = wc_max(h, WC_H_DRY);
        const int rf = (row - U.y0 + 1) % NBF;
        const int rfm = (row - U.y0) % NBF, rfp = (row - U.y0 + 2) % NBF;
        ax = mask * (heff * ux - (s_f[rf][lx + 2] - s_f[rf][lx]) * ix);
        ay = mask * (heff * uy - (s_f[rfp][lx + 1] - s_f[rfm][lx + 1]) * iy);
      }
      double v[N];
#pragma unroll
      for (int n = 0; n < N; ++n) v[n] = 0.0;
      cell(own, k, ux, uy, ax, ay, v);
#pragma unroll
      for (int n = 0; n < N; ++n) acc[n] = acc[n] + v[n];
    }
    __syncthreads();  // the next unit's first chunk restages every buffer row
  }
}

}  // namespace mrch
}  // namespace wc
