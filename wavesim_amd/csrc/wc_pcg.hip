// PCG (dispersion.rs:493-545) in two kernels per iteration and no separate reduction kernel.
//
// The grid is exactly 256 blocks (the lanes of the canonical summation's second stage) of
// 1024 threads. Block b owns tiles b, b+256, b+512, ... and processes four of them at a time;
// after each round its thread 0 adds the four tile sums in tile order to a register. That
// register is, bit for bit, lane b of the canonical second stage (lane t adds tiles t, t+256,
// ... from 0.0), so the consumer of a reduction only has to read 256 numbers and do the final
// 64-lane butterflies, which every block of the next kernel does in its prologue: the scalars
// (alpha, beta, the convergence test) are computed redundantly and identically by every block,
// and block 0 records them. The iteration's two kernels are
//   M_n: beta from U_{n-1}'s lanes (or the start from I's), p = z + beta p on the tile + 2,
//        A p, lanes of p.Ap;
//   U_n: alpha = rz / p.Ap from M_n's lanes; w += alpha p; r -= alpha Ap; z = r / diag;
//        lanes of r.r and r.z.
// A one-block check kernel with M_n's prologue ends a batch, so the host sees convergence.
//
// Templated on the storage type V: double is the FP64 path (same arithmetic, same order, as
// wc_kernels.hip's kernels and the reference), float the FP32 form (FP32 storage, FP64
// arithmetic and reductions).

#include "wc_kernels.h"

namespace wc {

namespace {

constexpr int LANES = 256;
constexpr int PG = 4;          // tiles in flight per block
constexpr int PT = NT * PG;    // threads per block
constexpr int S2X = TX + 4, S2Y = TY + 4;
constexpr int S1X = TX + 2, S1Y = TY + 2;

// The canonical final stage over the 256 lanes in acc[n * 256 + lane]; every thread gets it.
template <int N>
__device__ __forceinline__ void lane_sums(const double* acc, double (&out)[N]) {
  __shared__ double lds[N][4];
  const int t = threadIdx.x;
  if (t < LANES) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const double s = wave_sum(acc[n * LANES + t]);
      if ((t & 63) == 0) lds[n][t >> 6] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < N; ++n) out[n] = (lds[n][0] + lds[n][1]) + (lds[n][2] + lds[n][3]);
  __syncthreads();
}

// The canonical tile sum of group q's 256 threads (one 64x4 tile); result in gsum[q][n].
template <int N>
__device__ __forceinline__ void group_tile_sums(double (&v)[N], double (*gsum)[N]) {
  __shared__ double lds[PG][N][TY];
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  const int lane = local & 63, wave = local >> 6;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const double s = wave_sum(v[n]);
    if (lane == 0) lds[q][n][wave] = s;
  }
  __syncthreads();
  if (local == 0) {
#pragma unroll
    for (int n = 0; n < N; ++n)
      gsum[q][n] = (lds[q][n][0] + lds[q][n][1]) + (lds[q][n][2] + lds[q][n][3]);
  }
  __syncthreads();
}

template <typename V>
__device__ __forceinline__ void diag_v(const CgV<V>& a, int k, double h, double& dgx,
                                       double& dgy) {
  if (!(h >= WC_H_DISPERSION_MIN)) {
    dgx = 1.0;
    dgy = 1.0;
    return;
  }
  const double heff = wc_max(h, WC_H_DRY);
  const int w = a.g.w;
  dgx = heff + ((double)a.c[k + 1] + (double)a.c[k - 1]) * 0.25 / (a.g.dx * a.g.dx);
  dgy = heff + ((double)a.c[k + w] + (double)a.c[k - w]) * 0.25 / (a.g.dy * a.g.dy);
}

// The prologue of M_n (and of the check kernel): the scalars iteration n starts from.
// Returns false when the solve is over. Every block computes the same numbers.
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
    // kernels write nothing. (Block 0 of this kernel may already have set the flag; a block
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

}  // namespace

// r = b - A w, p = z = r / diag, lanes of b.b, r.z, r.r (dispersion.rs:493-514).
template <typename V>
__global__ void __launch_bounds__(PT) k_cg3_init(CgV<V> a) {
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  __shared__ double s_px[PG][S2Y][S2X], s_py[PG][S2Y][S2X], s_f[PG][S1Y][S1X];
  __shared__ double gsum[PG][3];
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int base = blockIdx.x; base < np; base += LANES * PG) {
    const int tile = base + LANES * q;
    const bool have = tile < np;
    const int i0 = GH + (have ? tile % g.ntx : 0) * TX, j0 = GH + (have ? tile / g.ntx : 0) * TY;
    for (int t = local; t < S2X * S2Y; t += NT) {
      const int lx = t % S2X, ly = t / S2X;
      const int i = i0 - 2 + lx, j = j0 - 2 + ly;
      double vx = 0.0, vy = 0.0;
      if (have && i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
        const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
        vx = (double)a.wx[m];
        vy = (double)a.wy[m];
        if (is_ghost_x(i, g.nx)) vx = -vx;
        if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
      }
      s_px[q][ly][lx] = vx;
      s_py[q][ly][lx] = vy;
    }
    __syncthreads();
    for (int t = local; t < S1X * S1Y; t += NT) {
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
    double v[3] = {0.0, 0.0, 0.0};
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
      const double bx = (double)a.bx[k], by = (double)a.by[k];
      const double rx = (double)(V)(bx - ax), ry = (double)(V)(by - ay);
      double dgx, dgy;
      diag_v(a, k, h, dgx, dgy);
      const double zx = (double)(V)(rx / dgx), zy = (double)(V)(ry / dgy);
      a.rx[k] = (V)rx;
      a.ry[k] = (V)ry;
      a.px[k] = (V)zx;
      a.py[k] = (V)zy;
      v[0] = bx * bx + by * by;
      v[1] = rx * zx + ry * zy;
      v[2] = rx * rx + ry * ry;
    }
    group_tile_sums<3>(v, gsum);
    if (threadIdx.x == 0) {
      for (int g2 = 0; g2 < PG; ++g2)
        if (base + LANES * g2 < np)
          for (int n = 0; n < 3; ++n) acc[n] = acc[n] + gsum[g2][n];
    }
  }
  if (threadIdx.x == 0)
    for (int n = 0; n < 3; ++n) a.acc_init[n * LANES + blockIdx.x] = acc[n];
}

template <typename V>
__global__ void __launch_bounds__(PT) k_cg3_matvec(CgV<V> a, unsigned long long n,
                                                   const V* psx, const V* psy, V* pdx, V* pdy) {
  double beta;
  if (!start_iteration<V>(a, n, beta)) return;
  const bool first = n == 1;
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  __shared__ double s_px[PG][S2Y][S2X], s_py[PG][S2Y][S2X], s_f[PG][S1Y][S1X];
  __shared__ double gsum[PG][1];
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  double acc = 0.0;
  for (int base = blockIdx.x; base < np; base += LANES * PG) {
    const int tile = base + LANES * q;
    const bool have = tile < np;
    const int i0 = GH + (have ? tile % g.ntx : 0) * TX, j0 = GH + (have ? tile / g.ntx : 0) * TY;
    for (int t = local; t < S2X * S2Y; t += NT) {
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
    for (int t = local; t < S1X * S1Y; t += NT) {
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
    double v[1] = {0.0};
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
      v[0] = px * ax + py * ay;
    }
    group_tile_sums<1>(v, gsum);
    if (threadIdx.x == 0) {
      for (int g2 = 0; g2 < PG; ++g2)
        if (base + LANES * g2 < np) acc = acc + gsum[g2][0];
    }
  }
  if (threadIdx.x == 0) a.acc_pap[blockIdx.x] = acc;
}

template <typename V>
__global__ void __launch_bounds__(PT) k_cg3_update(CgV<V> a, unsigned long long n, const V* pcx,
                                                   const V* pcy) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  double s[1];
  lane_sums<1>(a.acc_pap, s);
  const double alpha = sc->rzs[n & 1] / s[0];
  const double na = -alpha;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha = alpha;
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  __shared__ double gsum[PG][2];
  double acc[2] = {0.0, 0.0};
  for (int base = blockIdx.x; base < np; base += LANES * PG) {
    const int tile = base + LANES * q;
    const bool have = tile < np;
    double v[2] = {0.0, 0.0};
    if (have) {
      const int i = GH + (tile % g.ntx) * TX + (local & 63);
      const int j = GH + (tile / g.ntx) * TY + (local >> 6);
      if (i < g.nx + GH && j < g.ny + GH) {
        const int k = j * g.w + i;
        a.wx[k] = (V)((double)a.wx[k] + alpha * (double)pcx[k]);
        a.wy[k] = (V)((double)a.wy[k] + alpha * (double)pcy[k]);
        const double rx = (double)(V)((double)a.rx[k] + na * (double)a.ax[k]);
        const double ry = (double)(V)((double)a.ry[k] + na * (double)a.ay[k]);
        a.rx[k] = (V)rx;
        a.ry[k] = (V)ry;
        double dgx, dgy;
        diag_v(a, k, (double)a.h[k], dgx, dgy);
        const double zx = (double)(V)(rx / dgx), zy = (double)(V)(ry / dgy);
        a.zx[k] = (V)zx;
        a.zy[k] = (V)zy;
        v[0] = rx * rx + ry * ry;
        v[1] = rx * zx + ry * zy;
      }
    }
    group_tile_sums<2>(v, gsum);
    if (threadIdx.x == 0) {
      for (int g2 = 0; g2 < PG; ++g2)
        if (base + LANES * g2 < np)
          for (int m = 0; m < 2; ++m) acc[m] = acc[m] + gsum[g2][m];
    }
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < 2; ++m) a.acc_update[m * LANES + blockIdx.x] = acc[m];
}

// M_n's prologue alone, one block: records whether the solve is over after iteration n - 1.
template <typename V>
__global__ void __launch_bounds__(LANES) k_cg3_check(CgV<V> a, unsigned long long n) {
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
