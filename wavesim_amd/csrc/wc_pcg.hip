// PCG (dispersion.rs:493-545) in two kernels per iteration and no separate reduction kernel.
//
// The grid is exactly 256 blocks (the lanes of the canonical summation's second stage) of
// 1024 threads: four groups of 256 threads, each group one 64x4 tile at a time. Block b owns
// tiles b, b+256, b+512, ... and takes eight of them per round (group q has tiles 8r+q and
// 8r+q+4 of its list); after each round its thread 0 adds the eight tile sums in tile order to
// a register. That register is, bit for bit, lane b of the canonical second stage (lane t adds
// tiles t, t+256, ... from 0.0), so the consumer of a reduction only has to read 256 numbers
// and do the final 64-lane butterflies. Every wave of the next kernel does that by itself (four
// butterflies, one per 64 lanes, then (w0 + w1) + (w2 + w3)), so the scalars (alpha, beta, the
// convergence test) need no barrier and are identical in every wave; block 0 records them.
// Each kernel issues its loads for the first round before it waits for those scalars, so the
// two latencies overlap. The iteration's two kernels are
//   M_n: beta from U_{n-1}'s lanes (or the start from I's), p = z + beta p on the tile + 2,
//        A p, lanes of p.Ap;
//   U_n: alpha = rz / p.Ap from M_n's lanes; w += alpha p; r -= alpha Ap; z = r / diag;
//        lanes of r.r and r.z.
// A one-wave check kernel with M_n's prologue ends a batch, so the host sees convergence.
//
// Templated on the storage type V: double is the FP64 path (same arithmetic, same order, as
// wc_kernels.hip's kernels and the reference), float the FP32 form (FP32 storage, FP64
// arithmetic and reductions).

#include "wc_kernels.h"

namespace wc {

namespace {

constexpr int LANES = 256;
constexpr int PG = 4;          // groups (tiles in flight) per block
constexpr int SL_OP = 1;       // tiles per group per round, kernels with the operator
constexpr int SL_UP = 2;       // tiles per group per round, the update
constexpr int PT = NT * PG;    // threads per block
constexpr int S2X = TX + 4, S2Y = TY + 4;
constexpr int S1X = TX + 2, S1Y = TY + 2;
constexpr int H2 = S2X * S2Y, E2 = (H2 + NT - 1) / NT;  // tile + 2 ring, per thread
constexpr int H1 = S1X * S1Y, E1 = (H1 + NT - 1) / NT;  // tile + 1 ring, per thread

// The canonical final stage over the 256 lanes in acc[n * 256 + lane], by one wave alone.
template <int N>
__device__ __forceinline__ void lane_sums(const double* acc, double (&out)[N]) {
  const int lane = threadIdx.x & 63;
  double v[N][4];
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int w = 0; w < 4; ++w) v[n][w] = acc[n * LANES + 64 * w + lane];
#pragma unroll
  for (int n = 0; n < N; ++n)
    out[n] = (wave_sum(v[n][0]) + wave_sum(v[n][1])) + (wave_sum(v[n][2]) + wave_sum(v[n][3]));
}

// The tile a thread's slot s works on in round r (-1 if the block has no such tile), with SL
// tiles per group per round.
template <int SL>
__device__ __forceinline__ int slot_tile(int r, int s, int q, int np) {
  const int tile = blockIdx.x + LANES * (PG * SL * r + PG * s + q);
  return tile < np ? tile : -1;
}

// Row sums of each slot's tile into sums[r & 1][s][q][n][row]; after the barrier thread 0
// adds the tiles in tile order (slot-major: tile index 4 SL r + 4s + q) to acc.
template <int N, int SL>
__device__ __forceinline__ void add_tiles(double (&v)[SL][N], double (*sums)[SL][PG][N][TY],
                                          int r, int np, double (&acc)[N]) {
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  const int lane = local & 63, row = local >> 6;
#pragma unroll
  for (int s = 0; s < SL; ++s)
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const double t = wave_sum(v[s][n]);
      if (lane == 0) sums[r & 1][s][q][n][row] = t;
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int s = 0; s < SL; ++s)
      for (int g2 = 0; g2 < PG; ++g2) {
        if (slot_tile<SL>(r, s, g2, np) < 0) continue;
        const double* t = &sums[r & 1][s][g2][0][0];
        for (int n = 0; n < N; ++n)
          acc[n] = acc[n] + ((t[n * TY + 0] + t[n * TY + 1]) + (t[n * TY + 2] + t[n * TY + 3]));
      }
  }
}

// The Jacobi diagonal (dispersion.rs:466-480) from the cell's depth and its four neighbours' c.
__device__ __forceinline__ void diag_of(const Geo& g, double h, double cxm, double cxp,
                                        double cym, double cyp, double& dgx, double& dgy) {
  if (!(h >= WC_H_DISPERSION_MIN)) {
    dgx = 1.0;
    dgy = 1.0;
    return;
  }
  const double heff = wc_max(h, WC_H_DRY);
  dgx = heff + (cxp + cxm) * 0.25 / (g.dx * g.dx);
  dgy = heff + (cyp + cym) * 0.25 / (g.dy * g.dy);
}

// The prologue of M_n (and of the check kernel): the scalars iteration n starts from.
// Returns false when the solve is over. Every wave computes the same numbers.
template <typename V>
__device__ __forceinline__ bool start_iteration(const CgV<V>& a, unsigned long long n,
                                                double& beta) {
  CgScalars* sc = a.sc;
  bool done;
  double rz;
  if (n == 1) {
    double s[3];
    lane_sums<3>(a.acc_init, s);
    const double target = a.tolerance * sqrt(s[0]) + 1e-10;
    rz = s[1];
    const int conv = sqrt(s[2]) <= target ? 1 : 0;
    done = conv || a.max_iterations == 0;
    beta = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->target = target;
      sc->converged = conv;
      sc->it = 0;
    }
  } else {
    // A solve that ended in an earlier kernel of the batch stays as it ended: the later
    // kernels write nothing. (Block 0 of this kernel may already have set the flag; a wave
    // that reads it would have reached the same verdict.)
    if (sc->done) return false;
    double s[2];
    lane_sums<2>(a.acc_update, s);
    const int conv = sqrt(s[0]) <= sc->target ? 1 : 0;
    const unsigned long long it = n - 1;
    done = conv || it >= a.max_iterations;
    beta = s[1] / sc->rzs[(n - 1) & 1];
    rz = s[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->converged = conv;
      sc->it = it;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rzs[n & 1] = rz;
    sc->done = done ? 1 : 0;
  }
  return !done;
}

// The operator on one round of tiles (dispersion.rs:274-317): the caller has put the vector
// (ghosts signed) on each slot's tile + 2 in s_x, s_y; this fills s_f on the tile + 1 from the
// preloaded c, and returns A p at the thread's cell of each slot (and p there).
template <typename V>
struct Op {
  static constexpr int SL = SL_OP;
  V (*s_x)[PG][S2Y][S2X];
  V (*s_y)[PG][S2Y][S2X];
  double (*s_f)[PG][S1Y][S1X];
};

template <typename V>
__device__ __forceinline__ void op_flux(const Geo& g, const Op<V>& o, const int (&tile)[SL_OP],
                                        const double (&cr)[SL_OP][E1]) {
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
#pragma unroll
  for (int s = 0; s < SL_OP; ++s)
#pragma unroll
    for (int e = 0; e < E1; ++e) {
      const int t = local + NT * e;
      if (t >= H1) continue;
      const int lx = t % S1X, ly = t / S1X;
      double f = 0.0;
      if (tile[s] >= 0) {
        const int x = lx + 1, y = ly + 1;
        f = cr[s][e] * (((double)o.s_x[s][q][y][x + 1] - (double)o.s_x[s][q][y][x - 1]) * ix +
                        ((double)o.s_y[s][q][y + 1][x] - (double)o.s_y[s][q][y - 1][x]) * iy);
      }
      o.s_f[s][q][ly][lx] = f;
    }
}

// c on the tile + 1 ring (0 outside the padded grid's used part), for op_flux.
template <typename V>
__device__ __forceinline__ void load_c(const Geo& g, const V* c, const int (&tile)[SL_OP],
                                       double (&cr)[SL_OP][E1]) {
  const int local = threadIdx.x % NT;
#pragma unroll
  for (int s = 0; s < SL_OP; ++s) {
    const int i0 = GH + (tile[s] >= 0 ? tile[s] % g.ntx : 0) * TX;
    const int j0 = GH + (tile[s] >= 0 ? tile[s] / g.ntx : 0) * TY;
#pragma unroll
    for (int e = 0; e < E1; ++e) {
      const int t = local + NT * e;
      cr[s][e] = 0.0;
      if (t >= H1 || tile[s] < 0) continue;
      const int i = i0 - 1 + t % S1X, j = j0 - 1 + t / S1X;
      if (i <= g.nx + GH && j <= g.ny + GH) cr[s][e] = (double)c[j * g.w + i];
    }
  }
}

// A p at the thread's own cell of slot s, from the staged p and f.
template <typename V>
__device__ __forceinline__ void op_apply(const Geo& g, const Op<V>& o, int s, double h,
                                         double& px, double& py, double& ax, double& ay) {
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  const int lx = local & 63, ly = local >> 6;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
  const double heff = wc_max(h, WC_H_DRY);
  const int x1 = lx + 1, y1 = ly + 1;
  px = (double)o.s_x[s][q][ly + 2][lx + 2];
  py = (double)o.s_y[s][q][ly + 2][lx + 2];
  ax = mask * (heff * px - (o.s_f[s][q][y1][x1 + 1] - o.s_f[s][q][y1][x1 - 1]) * ix);
  ay = mask * (heff * py - (o.s_f[s][q][y1 + 1][x1] - o.s_f[s][q][y1 - 1][x1]) * iy);
}

// The padded index of ring element t of a tile + 2 halo and whether it is inside the grid's
// used part; also whether its x or y component flips (a wall ghost).
__device__ __forceinline__ bool ring2(const Geo& g, int tile, int t, int& m, bool& fx, bool& fy) {
  const int i0 = GH + (tile % g.ntx) * TX, j0 = GH + (tile / g.ntx) * TY;
  const int i = i0 - 2 + t % S2X, j = j0 - 2 + t / S2X;
  if (i > g.nx + GH + 1 || j > g.ny + GH + 1) return false;
  m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
  fx = is_ghost_x(i, g.nx);
  fy = !g.periodic && is_ghost_y(j, g.ny);
  return true;
}

}  // namespace

// r = b - A w, p = z = r / diag, lanes of b.b, r.z, r.r (dispersion.rs:493-514).
template <typename V>
__global__ void __launch_bounds__(PT) k_cg3_init(CgV<V> a) {
  constexpr int SL = SL_OP, TPR = PG * SL;
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  __shared__ V s_x[SL][PG][S2Y][S2X], s_y[SL][PG][S2Y][S2X];
  __shared__ double s_f[SL][PG][S1Y][S1X];
  __shared__ double sums[2][SL][PG][3][TY];
  const Op<V> o{s_x, s_y, s_f};
  double acc[3] = {0.0, 0.0, 0.0};
  for (int r = 0; blockIdx.x + LANES * TPR * r < np; ++r) {
    int tile[SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) tile[s] = slot_tile<SL>(r, s, q, np);
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
      for (int e = 0; e < E2; ++e) {
        const int t = local + NT * e;
        if (t >= H2) continue;
        int m;
        bool fx, fy;
        double vx = 0.0, vy = 0.0;
        if (tile[s] >= 0 && ring2(g, tile[s], t, m, fx, fy)) {
          if (a.solx) {  // the FP32 form's iterate x0 + w
            vx = a.solx[m] + (double)a.wx[m];
            vy = a.soly[m] + (double)a.wy[m];
          } else {
            vx = (double)a.wx[m];
            vy = (double)a.wy[m];
          }
          if (fx) vx = -vx;
          if (fy) vy = -vy;
        }
        s_x[s][q][t / S2X][t % S2X] = (V)vx;
        s_y[s][q][t / S2X][t % S2X] = (V)vy;
      }
    double cr[SL][E1];
    load_c(g, a.c, tile, cr);
    __syncthreads();
    op_flux(g, o, tile, cr);
    __syncthreads();
    double v[SL][3];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      v[s][0] = v[s][1] = v[s][2] = 0.0;
      if (tile[s] < 0) continue;
      const int i = GH + (tile[s] % g.ntx) * TX + (local & 63);
      const int j = GH + (tile[s] / g.ntx) * TY + (local >> 6);
      if (i >= g.nx + GH || j >= g.ny + GH) continue;
      const int k = j * g.w + i;
      const double h = (double)a.h[k];
      double px, py, ax, ay;
      op_apply(g, o, s, h, px, py, ax, ay);
      const double bx = (double)a.bx[k], by = (double)a.by[k];
      const double rx = (double)(V)(bx - ax), ry = (double)(V)(by - ay);
      double dgx, dgy;
      diag_of(g, h, (double)a.c[k - 1], (double)a.c[k + 1], (double)a.c[k - g.w],
              (double)a.c[k + g.w], dgx, dgy);
      const double zx = (double)(V)(rx / dgx), zy = (double)(V)(ry / dgy);
      a.rx[k] = (V)rx;
      a.ry[k] = (V)ry;
      a.px[k] = (V)zx;
      a.py[k] = (V)zy;
      v[s][0] = bx * bx + by * by;
      v[s][1] = rx * zx + ry * zy;
      v[s][2] = rx * rx + ry * ry;
    }
    add_tiles<3, SL>(v, sums, r, np, acc);
  }
  if (threadIdx.x == 0)
    for (int n = 0; n < 3; ++n) a.acc_init[n * LANES + blockIdx.x] = acc[n];
}

// The matvec stages p_n = z + beta p_{n-1} on the tile + 2 straight from the loads (the scalars
// first): measured faster here than holding the loads in registers across the prologue.
template <typename V>
__global__ void __launch_bounds__(PT) k_cg3_matvec(CgV<V> a, unsigned long long n,
                                                   const V* psx, const V* psy, V* pdx, V* pdy) {
  constexpr int SL = SL_OP;
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  double beta = 0.0;
  if (blockIdx.x < np && !start_iteration<V>(a, n, beta)) return;
  const bool first = n == 1;
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  __shared__ double s_px[PG][S2Y][S2X], s_py[PG][S2Y][S2X], s_f[PG][S1Y][S1X];
  __shared__ double sums[2][SL][PG][1][TY];
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  double acc[1] = {0.0};
  for (int r = 0; blockIdx.x + LANES * PG * r < np; ++r) {
    const int tile = slot_tile<SL>(r, 0, q, np);
    const bool have = tile >= 0;
    const int i0 = GH + (have ? tile % g.ntx : 0) * TX, j0 = GH + (have ? tile / g.ntx : 0) * TY;
    for (int t = local; t < H2; t += NT) {
      const int lx = t % S2X, ly = t / S2X;
      const int i = i0 - 2 + lx, j = j0 - 2 + ly;
      double vx = 0.0, vy = 0.0;
      if (have && i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
        const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
        if (first) {
          vx = (double)psx[m];
          vy = (double)psy[m];
        } else {
          vx = (double)(V)((double)a.zx[m] + beta * (double)psx[m]);
          vy = (double)(V)((double)a.zy[m] + beta * (double)psy[m]);
        }
        if (is_ghost_x(i, g.nx)) vx = -vx;
        if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
      }
      s_px[q][ly][lx] = vx;
      s_py[q][ly][lx] = vy;
    }
    __syncthreads();
    for (int t = local; t < H1; t += NT) {
      const int lx = t % S1X, ly = t / S1X;
      const int i = i0 - 1 + lx, j = j0 - 1 + ly;
      double f = 0.0;
      if (have && i <= g.nx + GH && j <= g.ny + GH) {
        const int x = lx + 1, y = ly + 1;
        f = (double)a.c[j * g.w + i] *
            ((s_px[q][y][x + 1] - s_px[q][y][x - 1]) * ix +
             (s_py[q][y + 1][x] - s_py[q][y - 1][x]) * iy);
      }
      s_f[q][ly][lx] = f;
    }
    __syncthreads();
    double v[SL][1] = {{0.0}};
    const int lx = local & 63, ly = local >> 6;
    const int i = i0 + lx, j = j0 + ly;
    if (have && i < g.nx + GH && j < g.ny + GH) {
      const int k = j * g.w + i;
      const double h = (double)a.h[k];
      const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
      const double heff = wc_max(h, WC_H_DRY);
      const int x1 = lx + 1, y1 = ly + 1;
      const double px = s_px[q][ly + 2][lx + 2], py = s_py[q][ly + 2][lx + 2];
      const double ax = mask * (heff * px - (s_f[q][y1][x1 + 1] - s_f[q][y1][x1 - 1]) * ix);
      const double ay = mask * (heff * py - (s_f[q][y1 + 1][x1] - s_f[q][y1 - 1][x1]) * iy);
      if (!first) {
        pdx[k] = (V)px;
        pdy[k] = (V)py;
      }
      a.ax[k] = (V)ax;
      a.ay[k] = (V)ay;
      v[0][0] = px * ax + py * ay;
    }
    add_tiles<1, SL>(v, sums, r, np, acc);
  }
  if (threadIdx.x == 0) a.acc_pap[blockIdx.x] = acc[0];
}

template <typename V>
__global__ void __launch_bounds__(PT) k_cg3_update(CgV<V> a, unsigned long long n, const V* pcx,
                                                   const V* pcy) {
  constexpr int SL = SL_UP, TPR = PG * SL;
  CgScalars* sc = a.sc;
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  __shared__ double sums[2][SL][PG][2][TY];
  double acc[2] = {0.0, 0.0};
  double alpha = 0.0;
  for (int r = 0; blockIdx.x + LANES * TPR * r < np; ++r) {
    int tile[SL], kk[SL];
    V lp[SL][2], la[SL][2], lw[SL][2], lr[SL][2], lc[SL][4], lh[SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      tile[s] = slot_tile<SL>(r, s, q, np);
      kk[s] = -1;
      if (tile[s] < 0) continue;
      const int i = GH + (tile[s] % g.ntx) * TX + (local & 63);
      const int j = GH + (tile[s] / g.ntx) * TY + (local >> 6);
      if (i >= g.nx + GH || j >= g.ny + GH) continue;
      const int k = j * g.w + i;
      kk[s] = k;
      lp[s][0] = pcx[k];
      lp[s][1] = pcy[k];
      la[s][0] = a.ax[k];
      la[s][1] = a.ay[k];
      lw[s][0] = a.wx[k];
      lw[s][1] = a.wy[k];
      lr[s][0] = a.rx[k];
      lr[s][1] = a.ry[k];
      lh[s] = a.h[k];
      lc[s][0] = a.c[k - 1];
      lc[s][1] = a.c[k + 1];
      lc[s][2] = a.c[k - g.w];
      lc[s][3] = a.c[k + g.w];
    }
    if (r == 0) {
      if (sc->done) return;
      double s1[1];
      lane_sums<1>(a.acc_pap, s1);
      alpha = sc->rzs[n & 1] / s1[0];
      if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha = alpha;
    }
    const double na = -alpha;
    double v[SL][2];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      v[s][0] = v[s][1] = 0.0;
      const int k = kk[s];
      if (k < 0) continue;
      a.wx[k] = (V)((double)lw[s][0] + alpha * (double)lp[s][0]);
      a.wy[k] = (V)((double)lw[s][1] + alpha * (double)lp[s][1]);
      const double rx = (double)(V)((double)lr[s][0] + na * (double)la[s][0]);
      const double ry = (double)(V)((double)lr[s][1] + na * (double)la[s][1]);
      a.rx[k] = (V)rx;
      a.ry[k] = (V)ry;
      double dgx, dgy;
      diag_of(g, (double)lh[s], (double)lc[s][0], (double)lc[s][1], (double)lc[s][2],
              (double)lc[s][3], dgx, dgy);
      const double zx = (double)(V)(rx / dgx), zy = (double)(V)(ry / dgy);
      a.zx[k] = (V)zx;
      a.zy[k] = (V)zy;
      v[s][0] = rx * rx + ry * ry;
      v[s][1] = rx * zx + ry * zy;
    }
    add_tiles<2, SL>(v, sums, r, np, acc);
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < 2; ++m) a.acc_update[m * LANES + blockIdx.x] = acc[m];
}

// M_n's prologue alone, one wave: records whether the solve is over after iteration n - 1.
template <typename V>
__global__ void __launch_bounds__(64) k_cg3_check(CgV<V> a, unsigned long long n) {
  double beta;
  start_iteration<V>(a, n, beta);
}

template __global__ void k_cg3_init<double>(CgV<double>);
template __global__ void k_cg3_init<float>(CgV<float>);
template __global__ void k_cg3_matvec<double>(CgV<double>, unsigned long long, const double*,
                                              const double*, double*, double*);
template __global__ void k_cg3_matvec<float>(CgV<float>, unsigned long long, const float*,
                                             const float*, float*, float*);
template __global__ void k_cg3_update<double>(CgV<double>, unsigned long long, const double*,
                                              const double*);
template __global__ void k_cg3_update<float>(CgV<float>, unsigned long long, const float*,
                                             const float*);
template __global__ void k_cg3_check<double>(CgV<double>, unsigned long long);
template __global__ void k_cg3_check<float>(CgV<float>, unsigned long long);

}  // namespace wc
