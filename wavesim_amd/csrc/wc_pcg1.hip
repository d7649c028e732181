// PCG in one kernel per iteration (Chronopoulos and Gear's formulation), for the FP32 form.
//
// The reference's PCG (dispersion.rs:493-545) needs two grid-wide reductions per iteration,
// p.Ap before the update and r.z after it, hence two kernels. Chronopoulos and Gear (1989)
// compute A u instead of A p and get A p by the recurrence s = A u + beta s, so all three inner
// products of an iteration, gamma = r.u, delta = (A u).u and r.r, are taken at one point:
//   beta_i  = gamma_i / gamma_{i-1}                     (beta_0 = 0)
//   alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1})   (alpha_0 = gamma_0 / delta_0)
//   p_i = u_i + beta_i p_{i-1};  s_i = q_i + beta_i s_{i-1};  w += alpha_i p_i;
//   r_{i+1} = r_i - alpha_i s_i;  u_{i+1} = D^-1 r_{i+1};  q_{i+1} = A u_{i+1}.
// In exact arithmetic the iterates are those of the reference's method (same Jacobi
// preconditioner, same convergence test on the recursively updated residual, same iteration
// count); in floating point they differ by rounding, which is why the FP64 path, held to the
// oracle bit for bit, keeps the reference's two-kernel form.
//
// Kernel K_n does iteration i = n - 1. A u_{i+1} needs u on the tile + 2, so the tile recomputes
// s_i, r_{i+1} and u_{i+1} there from r_i, q_i and s_{i-1}, which K_{n-1} wrote: r, q and s
// alternate between two buffers by iteration parity (neighbouring tiles read the old values
// while this one writes the new), p and w are updated in place. The lane sums follow the
// canonical order of wc_pcg.hip (256 blocks, block b owns tiles b, b+256, ...), in two buffers
// by parity; every wave of K_{n+1} reduces them itself.

#include "wc_kernels.h"

namespace wc {

namespace {

constexpr int LANES = 256;
constexpr int PG = 4;
constexpr int PT = NT * PG;
constexpr int S2X = TX + 4, S2Y = TY + 4;
constexpr int S1X = TX + 2, S1Y = TY + 2;
constexpr int H2 = S2X * S2Y, H1 = S1X * S1Y;

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

// Tile sums of the round's four tiles (one per 256-thread group), added in tile order.
template <int N>
__device__ __forceinline__ void add_round(double (&v)[N], double (*sums)[PG][N][TY], int r,
                                          int np, double (&acc)[N]) {
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  const int lane = local & 63, row = local >> 6;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const double t = wave_sum(v[n]);
    if (lane == 0) sums[r & 1][q][n][row] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int g2 = 0; g2 < PG; ++g2) {
      if ((int)blockIdx.x + LANES * (PG * r + g2) >= np) continue;
      const double* t = &sums[r & 1][g2][0][0];
      for (int n = 0; n < N; ++n)
        acc[n] = acc[n] + ((t[n * TY + 0] + t[n * TY + 1]) + (t[n * TY + 2] + t[n * TY + 3]));
    }
  }
}

// Stages the vector u (given per halo cell by fn(m, vx, vy), m the mapped interior index; the
// wall flips are applied here) on the tile + 2 and c (grad . u) on the tile + 1; returns A u
// at the thread's cell (and u there).
template <typename V, class Fn>
__device__ __forceinline__ void stencil(const Geo& g, const V* c, const V* hh, int tile,
                                        double (*s_ux)[S2Y][S2X], double (*s_uy)[S2Y][S2X],
                                        double (*s_f)[S1Y][S1X], Fn fn, bool& own, int& k,
                                        double& ux, double& uy, double& ax, double& ay) {
  const int q = threadIdx.x / NT, local = threadIdx.x % NT;
  const bool have = tile >= 0;
  const int i0 = GH + (have ? tile % g.ntx : 0) * TX, j0 = GH + (have ? tile / g.ntx : 0) * TY;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = local; t < H2; t += NT) {
    const int lx = t % S2X, ly = t / S2X;
    const int i = i0 - 2 + lx, j = j0 - 2 + ly;
    double vx = 0.0, vy = 0.0;
    if (have && i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
      const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
      fn(m, vx, vy);
      if (is_ghost_x(i, g.nx)) vx = -vx;
      if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
    }
    s_ux[q][ly][lx] = vx;
    s_uy[q][ly][lx] = vy;
  }
  __syncthreads();
  for (int t = local; t < H1; t += NT) {
    const int lx = t % S1X, ly = t / S1X;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    double f = 0.0;
    if (have && i <= g.nx + GH && j <= g.ny + GH) {
      const int x = lx + 1, y = ly + 1;
      f = (double)c[j * g.w + i] * ((s_ux[q][y][x + 1] - s_ux[q][y][x - 1]) * ix +
                                    (s_uy[q][y + 1][x] - s_uy[q][y - 1][x]) * iy);
    }
    s_f[q][ly][lx] = f;
  }
  __syncthreads();
  const int lx = local & 63, ly = local >> 6;
  const int i = i0 + lx, j = j0 + ly;
  own = have && i < g.nx + GH && j < g.ny + GH;
  k = j * g.w + i;
  ux = s_ux[q][ly + 2][lx + 2];
  uy = s_uy[q][ly + 2][lx + 2];
  ax = ay = 0.0;
  if (own) {
    const double h = (double)hh[k];
    const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
    const double heff = wc_max(h, WC_H_DRY);
    const int x1 = lx + 1, y1 = ly + 1;
    ax = mask * (heff * ux - (s_f[q][y1][x1 + 1] - s_f[q][y1][x1 - 1]) * ix);
    ay = mask * (heff * uy - (s_f[q][y1 + 1][x1] - s_f[q][y1 - 1][x1]) * iy);
  }
}

// The scalars of iteration i = n - 1, from K_{n-1}'s lanes (or the start's); false when the
// solve is over. Every wave computes the same numbers; block 0 records them.
template <typename V>
__device__ __forceinline__ bool start(const CgG<V>& a, unsigned long long n, double& alpha,
                                      double& beta) {
  CgScalars* sc = a.sc;
  const unsigned long long i = n - 1;
  double target;
  if (n == 1) {
    double b[1];
    lane_sums<1>(a.lanes_b, b);
    target = a.tolerance * sqrt(b[0]) + 1e-10;
  } else {
    if (sc->done) return false;
    target = sc->target;
  }
  double s[3];
  lane_sums<3>(a.lanes + 3 * LANES * (i & 1), s);
  const double gamma = s[0], delta = s[1];
  const int conv = sqrt(s[2]) <= target ? 1 : 0;
  const bool done = conv || i >= a.max_iterations;
  if (i == 0) {
    beta = 0.0;
    alpha = gamma / delta;
  } else {
    beta = gamma / sc->rzs[(i - 1) & 1];
    alpha = gamma / (delta - beta * gamma / sc->alphas[(i - 1) & 1]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (n == 1) sc->target = target;
    sc->converged = conv;
    sc->it = i;
    sc->rzs[i & 1] = gamma;
    sc->alphas[i & 1] = alpha;
    sc->done = done ? 1 : 0;
  }
  return !done;
}

}  // namespace

// r_0 = b - A w (w zero where masked, as set up), the inverse Jacobi diagonal
// (dispersion.rs:466-480), lanes of b.b.
template <typename V>
__global__ void __launch_bounds__(PT) k_cg1_init(CgG<V> a) {
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT;
  __shared__ double s_ux[PG][S2Y][S2X], s_uy[PG][S2Y][S2X], s_f[PG][S1Y][S1X];
  __shared__ double sums[2][PG][1][TY];
  double acc[1] = {0.0};
  for (int r = 0; (int)blockIdx.x + LANES * PG * r < np; ++r) {
    const int tile0 = blockIdx.x + LANES * (PG * r + q);
    const int tile = tile0 < np ? tile0 : -1;
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil<V>(g, a.c, a.h, tile, s_ux, s_uy, s_f,
               [&](int m, double& vx, double& vy) {
                 vx = (double)a.wx[m];
                 vy = (double)a.wy[m];
               },
               own, k, ux, uy, ax, ay);
    double v[1] = {0.0};
    if (own) {
      const double h = (double)a.h[k];
      double dgx = 1.0, dgy = 1.0;
      if (h >= WC_H_DISPERSION_MIN) {
        const double heff = wc_max(h, WC_H_DRY);
        dgx = heff + ((double)a.c[k + 1] + (double)a.c[k - 1]) * 0.25 / (g.dx * g.dx);
        dgy = heff + ((double)a.c[k + g.w] + (double)a.c[k - g.w]) * 0.25 / (g.dy * g.dy);
      }
      const double bx = (double)a.bx[k], by = (double)a.by[k];
      a.r[0][0][k] = (V)(bx - ax);
      a.r[0][1][k] = (V)(by - ay);
      a.dix[k] = (V)(1.0 / dgx);
      a.diy[k] = (V)(1.0 / dgy);
      a.px[k] = (V)0;
      a.py[k] = (V)0;
      a.s[1][0][k] = (V)0;
      a.s[1][1][k] = (V)0;
      v[0] = bx * bx + by * by;
    }
    add_round<1>(v, sums, r, np, acc);
  }
  if (threadIdx.x == 0) a.lanes_b[blockIdx.x] = acc[0];
}

// u_0 = D^-1 r_0, q_0 = A u_0; lanes of gamma_0 = r.u, delta_0 = q.u, r.r (slot 0).
template <typename V>
__global__ void __launch_bounds__(PT) k_cg1_start(CgG<V> a) {
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT;
  __shared__ double s_ux[PG][S2Y][S2X], s_uy[PG][S2Y][S2X], s_f[PG][S1Y][S1X];
  __shared__ double sums[2][PG][3][TY];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int r = 0; (int)blockIdx.x + LANES * PG * r < np; ++r) {
    const int tile0 = blockIdx.x + LANES * (PG * r + q);
    const int tile = tile0 < np ? tile0 : -1;
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil<V>(g, a.c, a.h, tile, s_ux, s_uy, s_f,
               [&](int m, double& vx, double& vy) {
                 vx = (double)a.r[0][0][m] * (double)a.dix[m];
                 vy = (double)a.r[0][1][m] * (double)a.diy[m];
               },
               own, k, ux, uy, ax, ay);
    double v[3] = {0.0, 0.0, 0.0};
    if (own) {
      const double qx = (double)(V)ax, qy = (double)(V)ay;
      a.q[0][0][k] = (V)qx;
      a.q[0][1][k] = (V)qy;
      const double rx = (double)a.r[0][0][k], ry = (double)a.r[0][1][k];
      v[0] = rx * ux + ry * uy;
      v[1] = qx * ux + qy * uy;
      v[2] = rx * rx + ry * ry;
    }
    add_round<3>(v, sums, r, np, acc);
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) a.lanes[m * LANES + blockIdx.x] = acc[m];
}

// Iteration i = n - 1 (see the top of the file).
template <typename V>
__global__ void __launch_bounds__(PT) k_cg1_iter(CgG<V> a, unsigned long long n) {
  const Geo& g = a.g;
  const int np = g.ntx * g.nty;
  const int q = threadIdx.x / NT;
  double alpha = 0.0, beta = 0.0;
  if ((int)blockIdx.x < np && !start<V>(a, n, alpha, beta)) return;
  const unsigned long long i = n - 1;
  const int cur = (int)(i & 1), nxt = cur ^ 1;
  const V *rx0 = a.r[cur][0], *ry0 = a.r[cur][1], *qx0 = a.q[cur][0], *qy0 = a.q[cur][1];
  const V *sx0 = a.s[nxt][0], *sy0 = a.s[nxt][1];  // s_{i-1}: written by K_{n-1} into nxt
  const double na = -alpha;
  __shared__ double s_ux[PG][S2Y][S2X], s_uy[PG][S2Y][S2X], s_f[PG][S1Y][S1X];
  __shared__ double sums[2][PG][3][TY];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int r = 0; (int)blockIdx.x + LANES * PG * r < np; ++r) {
    const int tile0 = blockIdx.x + LANES * (PG * r + q);
    const int tile = tile0 < np ? tile0 : -1;
    bool own;
    int k;
    double ux, uy, ax, ay;
    // u_{i+1} = D^-1 (r_i - alpha (q_i + beta s_{i-1})) on the tile + 2
    stencil<V>(g, a.c, a.h, tile, s_ux, s_uy, s_f,
               [&](int m, double& vx, double& vy) {
                 const double sx = (double)(V)((double)qx0[m] + beta * (double)sx0[m]);
                 const double sy = (double)(V)((double)qy0[m] + beta * (double)sy0[m]);
                 vx = (double)(V)((double)rx0[m] + na * sx) * (double)a.dix[m];
                 vy = (double)(V)((double)ry0[m] + na * sy) * (double)a.diy[m];
               },
               own, k, ux, uy, ax, ay);
    double v[3] = {0.0, 0.0, 0.0};
    if (own) {
      const double dix = (double)a.dix[k], diy = (double)a.diy[k];
      const double rx = (double)rx0[k], ry = (double)ry0[k];
      const double sx = (double)(V)((double)qx0[k] + beta * (double)sx0[k]);
      const double sy = (double)(V)((double)qy0[k] + beta * (double)sy0[k]);
      const double px = (double)(V)(rx * dix + beta * (double)a.px[k]);
      const double py = (double)(V)(ry * diy + beta * (double)a.py[k]);
      a.px[k] = (V)px;
      a.py[k] = (V)py;
      a.wx[k] = (V)((double)a.wx[k] + alpha * px);
      a.wy[k] = (V)((double)a.wy[k] + alpha * py);
      a.s[cur][0][k] = (V)sx;
      a.s[cur][1][k] = (V)sy;
      const double nrx = (double)(V)(rx + na * sx), nry = (double)(V)(ry + na * sy);
      a.r[nxt][0][k] = (V)nrx;
      a.r[nxt][1][k] = (V)nry;
      const double qx = (double)(V)ax, qy = (double)(V)ay;
      a.q[nxt][0][k] = (V)qx;
      a.q[nxt][1][k] = (V)qy;
      v[0] = nrx * ux + nry * uy;
      v[1] = qx * ux + qy * uy;
      v[2] = nrx * nrx + nry * nry;
    }
    add_round<3>(v, sums, r, np, acc);
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) a.lanes[3 * LANES * (n & 1) + m * LANES + blockIdx.x] = acc[m];
}

// K_n's prologue alone, one wave: whether the solve is over before iteration n - 1.
template <typename V>
__global__ void __launch_bounds__(64) k_cg1_check(CgG<V> a, unsigned long long n) {
  double alpha, beta;
  start<V>(a, n, alpha, beta);
}

template __global__ void k_cg1_init<float>(CgG<float>);
template __global__ void k_cg1_start<float>(CgG<float>);
template __global__ void k_cg1_iter<float>(CgG<float>, unsigned long long);
template __global__ void k_cg1_check<float>(CgG<float>, unsigned long long);

}  // namespace wc
