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
// while this one writes the new), p and w are updated in place.
//
// Nothing here has to repeat the oracle's summation order, so the layout is this kernel's own:
// tiles of 64x16 cells (one 1024-thread block, a cell per thread; the halo is a third of the
// tile rather than the 64x4 tile's 113%), and r, q, s, p and the inverse diagonal stored as
// interleaved (x, y) pairs, one 8-byte load per cell and vector. The 256 blocks own tiles b,
// b+256, ...; each adds its tiles' sums (rows in order) into one lane, and every wave of the
// next kernel reduces the 256 lanes itself, so no separate reduction kernel runs.

#include "wc_march.h"
#include "wc_tile16.h"

namespace wc {

namespace {

using namespace t16;

// The scalars of iteration i = n - 1, from K_{n-1}'s lanes (or the start's); false when the
// solve is over. Every wave computes the same numbers; block 0 records them.
template <int L>
__device__ __forceinline__ bool start(const CgG& a, unsigned long long n, int blk, double& alpha,
                                      double& beta, double& alpha_prev) {
  CgScalars* sc = a.sc;
  const unsigned long long i = n - 1;
  double target;
  if (n == 1) {
    double b[1];
    lane_sums<1, L>(a.lanes_b, b);
    target = a.tolerance * sqrt(b[0]) + 1e-10;
  } else {
    if (sc->done) return false;
    target = sc->target;
  }
  double s[3];
  lane_sums<3, L>(a.lanes + 3 * L * (i & 1), s);
  const double gamma = s[0], delta = s[1];
  const int conv = sqrt(s[2]) <= target ? 1 : 0;
  const bool done = conv || i >= a.max_iterations;
  if (i == 0) {
    beta = 0.0;
    alpha = gamma / delta;
    alpha_prev = 0.0;
  } else {
    alpha_prev = sc->alphas[(i - 1) & 1];
    beta = gamma / sc->rzs[(i - 1) & 1];
    alpha = gamma / (delta - beta * gamma / alpha_prev);
  }
  if (blk == 0 && threadIdx.x == 0) {
    if (n == 1) sc->target = target;
    sc->converged = conv;
    sc->it = i;
    sc->rzs[i & 1] = gamma;
    sc->alphas[i & 1] = alpha;
    sc->done = done ? 1 : 0;
  }
  return !done;
}

__device__ __forceinline__ float2 f2(double x, double y) { return make_float2((float)x, (float)y); }

}  // namespace

// The tile's own cells: the thread's cell and whether it lies in the interior.
__device__ __forceinline__ int own_index(const Geo& g, int tile, bool& inside) {
  const int tx = tiles_x(g);
  const int i = GH + (tile % tx) * BX + (threadIdx.x & 63);
  const int j = GH + (tile / tx) * BY + (threadIdx.x >> 6);
  inside = i < g.nx + GH && j < g.ny + GH;
  return j * g.w + i;
}

// r_0 = b - A w (w zero where masked, as set up), the inverse Jacobi diagonal
// (dispersion.rs:466-480), p = s_{-1} = 0, lanes of b.b. AHEAD tiles' loads in flight (as in
// the iteration: 2 alone, 1 in the batched kernel).
template <int L, int AHEAD>
__device__ __forceinline__ void cg1_init_body(const CgG& a, int blk) {
  const Geo& g = a.g;
  const int nt = blk < a.stride ? tiles_all(g) : 0;  // blocks past the stride: zero lanes
  __shared__ double s_u[2][S2Y][S2X], s_f[S1Y][S1X];
  __shared__ double sums[2][BY][1];
  __shared__ float s_c[H1];
  double acc[1] = {0.0};
  // the iterate x0 + w at the tile + 2 points (its parts, added once loaded), c on the tile + 1,
  // and the cell's depth and right-hand side
  struct Ahead {
    double x0[2][2];
    float2 w[2];
    int flags[2];
    float c[2];
    float h;
    double bx, by;
  };
  auto fetch = [&](int tile, Ahead& A) {
    const int tx = tiles_x(g);
    const int i0 = GH + (tile % tx) * BX, j0 = GH + (tile / tx) * BY;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int t = threadIdx.x + e * PT;
      A.flags[e] = 0;
      if (t < H2) {
        const int i = i0 - 2 + t % S2X, j = j0 - 2 + t / S2X;
        if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
          const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
          A.x0[e][0] = a.solx[m];
          A.x0[e][1] = a.soly[m];
          A.w[e] = make_float2(a.wx[m], a.wy[m]);
          A.flags[e] = 1 | (is_ghost_x(i, g.nx) ? 2 : 0) |
                       (!g.periodic && is_ghost_y(j, g.ny) ? 4 : 0);
        }
      }
    }
    prefetch_c(g, a.c, tile, A.c);
    bool in;
    const int k = own_index(g, tile, in);
    A.h = 0.f;
    A.bx = A.by = 0.0;
    if (in) {
      A.h = a.h[k];
      A.bx = a.bx[k];
      A.by = a.by[k];
    }
  };
  Ahead P[AHEAD];
#pragma unroll
  for (int q = 0; q < AHEAD; ++q)
    if (blk + a.stride * q < nt) fetch(blk + a.stride * q, P[q]);
  auto body = [&](Ahead& Q, int tile, int r) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int t = threadIdx.x + e * PT;
      if (t < H2) {
        double vx = 0.0, vy = 0.0;
        if (Q.flags[e] & 1) {
          vx = Q.x0[e][0] + (double)Q.w[e].x;
          vy = Q.x0[e][1] + (double)Q.w[e].y;
          if (Q.flags[e] & 2) vx = -vx;
          if (Q.flags[e] & 4) vy = -vy;
        }
        s_u[0][t / S2X][t % S2X] = vx;
        s_u[1][t / S2X][t % S2X] = vy;
      }
      if (t < H1) s_c[t] = Q.c[e];
    }
    const float hh = Q.h;
    const double bx = Q.bx, by = Q.by;
    __syncthreads();
    if (tile + AHEAD * a.stride < nt) fetch(tile + AHEAD * a.stride, Q);
    stencil_f_staged(g, s_u, s_c, s_f);
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil_own_h(g, hh, tile, s_u, s_f, own, k, ux, uy, ax, ay);
    double v[1] = {0.0};
    if (own) {
      const double h = (double)hh;
      double dgx = 1.0, dgy = 1.0;
      if (h >= WC_H_DISPERSION_MIN) {
        // c of the cell's neighbours, as staged on the tile + 1
        const int c1 = ((threadIdx.x >> 6) + 1) * S1X + (threadIdx.x & 63) + 1;
        const double heff = wc_max(h, WC_H_DRY);
        dgx = heff + ((double)s_c[c1 + 1] + (double)s_c[c1 - 1]) * 0.25 / (g.dx * g.dx);
        dgy = heff + ((double)s_c[c1 + S1X] + (double)s_c[c1 - S1X]) * 0.25 / (g.dy * g.dy);
      }
      a.r[0][k] = f2(bx - ax, by - ay);
      a.di[k] = f2(1.0 / dgx, 1.0 / dgy);
      a.p[k] = make_float2(0.f, 0.f);
      a.s[1][k] = make_float2(0.f, 0.f);
      v[0] = bx * bx + by * by;
    }
    add_rows<1>(v, sums, r, acc);
  };
  for (int r = 0; blk + a.stride * r < nt; r += AHEAD) {
#pragma unroll
    for (int q = 0; q < AHEAD; ++q)
      if (blk + a.stride * (r + q) < nt) body(P[q], blk + a.stride * (r + q), r + q);
  }
  rows_total<1>(acc, sums);
  if (threadIdx.x == 0) a.lanes_b[blk] = acc[0];
}

template <int L>
__global__ void __launch_bounds__(PT) k_cg1_init(CgG a) {
  cg1_init_body<L, 2>(a, lane_of_block(a.g, a.stride));
}

// u_0 = D^-1 r_0, q_0 = A u_0; lanes of gamma_0 = r.u, delta_0 = q.u, r.r (slot 0).
template <int L, int AHEAD>
__device__ __forceinline__ void cg1_start_body(const CgG& a, int blk) {
  const Geo& g = a.g;
  const int nt = blk < a.stride ? tiles_all(g) : 0;  // blocks past the stride: zero lanes
  __shared__ double s_u[2][S2Y][S2X], s_f[S1Y][S1X];
  __shared__ double sums[2][BY][3];
  __shared__ float2 s_own[2][BY][BX];  // r_0, D^-1 at the tile's cells
  __shared__ float s_c[H1];
  double acc[3] = {0.0, 0.0, 0.0};
  struct Ahead {
    Prefetch<2> halo;  // r_0 and D^-1 at the tile + 2 points
    float c[2];
    float h;
  };
  auto fetch_halo = [&](int m, float2 (&v)[2]) {
    v[0] = a.r[0][m];
    v[1] = a.di[m];
  };
  auto make = [](const float2 (&v)[2], double& vx, double& vy) {
    vx = (double)v[0].x * (double)v[1].x;
    vy = (double)v[0].y * (double)v[1].y;
  };
  auto fetch = [&](int tile, Ahead& A) {
    prefetch(g, tile, fetch_halo, A.halo);
    prefetch_c(g, a.c, tile, A.c);
    bool in;
    const int k = own_index(g, tile, in);
    A.h = in ? a.h[k] : 0.f;
  };
  Ahead P[AHEAD];
#pragma unroll
  for (int q = 0; q < AHEAD; ++q)
    if (blk + a.stride * q < nt) fetch(blk + a.stride * q, P[q]);
  auto body = [&](Ahead& Q, int tile, int r) {
    stage(Q.halo, s_u, s_own, make);
#pragma unroll
    for (int e = 0; e < 2; ++e)
      if (threadIdx.x + e * PT < H1) s_c[threadIdx.x + e * PT] = Q.c[e];
    const float hh = Q.h;
    __syncthreads();
    if (tile + AHEAD * a.stride < nt) fetch(tile + AHEAD * a.stride, Q);
    stencil_f_staged(g, s_u, s_c, s_f);
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil_own_h(g, hh, tile, s_u, s_f, own, k, ux, uy, ax, ay);
    double v[3] = {0.0, 0.0, 0.0};
    if (own) {
      const double qx = (double)(float)ax, qy = (double)(float)ay;
      a.q[0][k] = f2(qx, qy);
      const float2 rr = s_own[0][threadIdx.x >> 6][threadIdx.x & 63];
      const double rx = rr.x, ry = rr.y;
      v[0] = rx * ux + ry * uy;
      v[1] = qx * ux + qy * uy;
      v[2] = rx * rx + ry * ry;
    }
    add_rows<3>(v, sums, r, acc);
  };
  for (int r = 0; blk + a.stride * r < nt; r += AHEAD) {
#pragma unroll
    for (int q = 0; q < AHEAD; ++q)
      if (blk + a.stride * (r + q) < nt) body(P[q], blk + a.stride * (r + q), r + q);
  }
  rows_total<3>(acc, sums);
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) a.lanes[m * L + blk] = acc[m];
}

template <int L>
__global__ void __launch_bounds__(PT) k_cg1_start(CgG a) {
  cg1_start_body<L, 2>(a, lane_of_block(a.g, a.stride));
}

// Iteration i = n - 1 (see the top of the file).
// AHEAD tiles' loads in flight: 2 alone; 1 in the batched kernel, whose argument block and
// instance walk leave no registers for a second set.
template <int L, int AHEAD, bool NTS>
__device__ __forceinline__ void cg1_iter_body(const CgG& a, unsigned long long n, int blk) {
  const Geo& g = a.g;
  const int nt = blk < a.stride ? tiles_all(g) : 0;  // blocks past the stride: zero lanes
  double alpha = 0.0, beta = 0.0, alpha_prev = 0.0;
  if (blk < nt && !start<L>(a, n, blk, alpha, beta, alpha_prev)) return;
  const unsigned long long i = n - 1;
  const int cur = (int)(i & 1), nxt = cur ^ 1;
  const float2 *r0 = a.r[cur], *q0 = a.q[cur];
  const float2* s0 = a.s[nxt];  // s_{i-1}: written by K_{n-1} into nxt
  const double na = -alpha;
  // With two tiles in flight the tiles alternate between two sets of LDS buffers, so a tile's
  // staging never waits for the previous tile's last reads: two barriers per tile, not three.
  constexpr int SETS = AHEAD;
  __shared__ double s_u[SETS][2][S2Y][S2X], s_f[SETS][S1Y][S1X];
  __shared__ double sums[2][BY][3];
  double acc[3] = {0.0, 0.0, 0.0};
  // u_{i+1} = D^-1 (r_i - alpha (q_i + beta s_{i-1})) on the tile + 2, from q_i, s_{i-1}, r_i
  // and D^-1 prefetched a tile ahead
  auto fetch = [&](int m, float2 (&v)[4]) {
    v[0] = q0[m];
    v[1] = s0[m];
    v[2] = r0[m];
    v[3] = a.di[m];
  };
  auto make = [&](const float2 (&v)[4], double& vx, double& vy) {
    const double sx = (double)(float)((double)v[0].x + beta * (double)v[1].x);
    const double sy = (double)(float)((double)v[0].y + beta * (double)v[1].y);
    vx = (double)(float)((double)v[2].x + na * sx) * (double)v[3].x;
    vy = (double)(float)((double)v[2].y + na * sy) * (double)v[3].y;
  };
  __shared__ float2 s_own[SETS][4][BY][BX];  // q_i, s_{i-1}, r_i, D^-1 at the tile's cells
  __shared__ float s_c[SETS][H1];              // c on the tile + 1
  // the iteration reads w at its cells when it updates it (every iteration, or only the odd
  // ones when the step is deferred)
  const bool upd_w = !a.defer || (i & 1);
  // AHEAD tiles' loads in flight (the halo, p and w at the own cell): the tile AHEAD places on
  // is fetched as soon as a tile's halo is staged.
  struct Ahead {
    Prefetch<4> halo;
    float c[2];  // c at the tile + 1 points
    float2 p, w;
    float h;
  };
  auto fetch_own = [&](int tile, Ahead& A) {
    prefetch(g, tile, fetch, A.halo);
    prefetch_c(g, a.c, tile, A.c);
    const int tx = tiles_x(g);
    const int ii = GH + (tile % tx) * BX + (threadIdx.x & 63);
    const int jj = GH + (tile / tx) * BY + (threadIdx.x >> 6);
    A.p = A.w = make_float2(0.f, 0.f);
    A.h = 0.f;
    if (ii < g.nx + GH && jj < g.ny + GH) {
      const int k = jj * g.w + ii;
      A.p = a.p[k];
      A.h = a.h[k];
      if (upd_w) A.w = make_float2(a.wx[k], a.wy[k]);
    }
  };
  Ahead P[AHEAD];
#pragma unroll
  for (int q = 0; q < AHEAD; ++q)
    if (blk + a.stride * q < nt) fetch_own(blk + a.stride * q, P[q]);
  auto body = [&](Ahead& Q, int tile, int r) {
    const int set = r % SETS;
    stage(Q.halo, s_u[set], s_own[set], make);
#pragma unroll
    for (int e = 0; e < 2; ++e)
      if (threadIdx.x + e * PT < H1) s_c[set][threadIdx.x + e * PT] = Q.c[e];
    const float2 pp = Q.p, ww = Q.w;
    const float hh = Q.h;
    __syncthreads();
    if (tile + AHEAD * a.stride < nt) fetch_own(tile + AHEAD * a.stride, Q);
    stencil_f_staged(g, s_u[set], s_c[set], s_f[set]);
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil_own_h(g, hh, tile, s_u[set], s_f[set], own, k, ux, uy, ax, ay);
    double v[3] = {0.0, 0.0, 0.0};
    if (own) {
      const int ox = threadIdx.x & 63, oy = threadIdx.x >> 6;
      const float2 qq = s_own[set][0][oy][ox], ss = s_own[set][1][oy][ox];
      const float2 rr = s_own[set][2][oy][ox], d = s_own[set][3][oy][ox];
      const double sx = (double)(float)((double)qq.x + beta * (double)ss.x);
      const double sy = (double)(float)((double)qq.y + beta * (double)ss.y);
      const double px = (double)(float)((double)rr.x * (double)d.x + beta * (double)pp.x);
      const double py = (double)(float)((double)rr.y * (double)d.y + beta * (double)pp.y);
      // the cell's new values: stored with the nontemporal hint in the batched kernel, plainly
      // alone (where the hint measured slower)
      double nrx, nry, qx, qy;
      if constexpr (NTS) {
        store_nt(&a.p[k], f2(px, py));
        if (!a.defer) {
          store_nt(&a.wx[k], (float)((double)ww.x + alpha * px));
          store_nt(&a.wy[k], (float)((double)ww.y + alpha * py));
        } else if (i & 1) {  // iterations i - 1 and i together (CgG::defer)
          store_nt(&a.wx[k], (float)((double)ww.x + (alpha_prev * (double)pp.x + alpha * px)));
          store_nt(&a.wy[k], (float)((double)ww.y + (alpha_prev * (double)pp.y + alpha * py)));
        }
        store_nt(&a.s[cur][k], f2(sx, sy));
        nrx = (double)(float)((double)rr.x + na * sx);
        nry = (double)(float)((double)rr.y + na * sy);
        store_nt(&a.r[nxt][k], f2(nrx, nry));
        qx = (double)(float)ax;
        qy = (double)(float)ay;
        store_nt(&a.q[nxt][k], f2(qx, qy));
      } else {
        a.p[k] = f2(px, py);
        if (!a.defer) {
          a.wx[k] = (float)((double)ww.x + alpha * px);
          a.wy[k] = (float)((double)ww.y + alpha * py);
        } else if (i & 1) {
          a.wx[k] = (float)((double)ww.x + (alpha_prev * (double)pp.x + alpha * px));
          a.wy[k] = (float)((double)ww.y + (alpha_prev * (double)pp.y + alpha * py));
        }
        a.s[cur][k] = f2(sx, sy);
        nrx = (double)(float)((double)rr.x + na * sx);
        nry = (double)(float)((double)rr.y + na * sy);
        a.r[nxt][k] = f2(nrx, nry);
        qx = (double)(float)ax;
        qy = (double)(float)ay;
        a.q[nxt][k] = f2(qx, qy);
      }
      v[0] = nrx * ux + nry * uy;
      v[1] = qx * ux + qy * uy;
      v[2] = nrx * nrx + nry * nry;
    }
    if constexpr (SETS == 2) {
#pragma unroll
      for (int m = 0; m < 3; ++m) acc[m] = acc[m] + v[m];
    } else {
      add_rows<3>(v, sums, r, acc);
    }
  };
  for (int r = 0; blk + a.stride * r < nt; r += AHEAD) {
#pragma unroll
    for (int q = 0; q < AHEAD; ++q)
      if (blk + a.stride * (r + q) < nt) body(P[q], blk + a.stride * (r + q), r + q);
  }
  rows_total<3>(acc, sums);
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) a.lanes[3 * L * (n & 1) + m * L + blk] = acc[m];
}

template <int L>
__global__ void __launch_bounds__(PT) k_cg1_iter(CgG a, unsigned long long n) {
  cg1_iter_body<L, 2, false>(a, n, lane_of_block(a.g, a.stride));
}

// K_n's prologue alone, one wave: whether the solve is over before iteration n - 1.
template <int L>
__global__ void __launch_bounds__(64) k_cg1_check(CgG a, unsigned long long n) {
  double alpha, beta, alpha_prev;
  start<L>(a, n, 0, alpha, beta, alpha_prev);
}

// The step deferred from the last iteration of a solve that ended after an odd number of
// iterations (CgG::defer): w += alpha_{it-1} p_{it-1}, at the cells of one 64x4 tile.
__device__ __forceinline__ void cg1_flush_body(const CgG& a, int tile) {
  const CgScalars* sc = a.sc;
  const unsigned long long it = sc->it;
  if (!a.defer || !(it & 1)) return;
  const double alpha = sc->alphas[(it - 1) & 1];
  const Geo& g = a.g;
  const int i = GH + (tile % g.ntx) * TX + (threadIdx.x & 63);
  const int j = GH + (tile / g.ntx) * TY + (threadIdx.x >> 6);
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  const int k = j * g.w + i;
  const float2 pp = a.p[k];
  a.wx[k] = (float)((double)a.wx[k] + alpha * (double)pp.x);
  a.wy[k] = (float)((double)a.wy[k] + alpha * (double)pp.y);
}

__global__ void __launch_bounds__(NT) k_cg1_flush(CgG a) { cg1_flush_body(a, blockIdx.x); }

__global__ void __launch_bounds__(NT) kb_cg1_flush(const CgG* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1_flush_body(args[q], vb - m.first[q]);
  }
}

// ---- the same three kernels marching down column strips (wc_march.h) -----------------------
// Same arithmetic per cell as the tile kernels above; only the grouping of cells into blocks,
// and so the order of the lane sums, differs.

template <int L>
__device__ __forceinline__ void cg1m_init_body(const CgG& a, int blk) {
  const Geo& g = a.g;
  __shared__ mrch::LdsU s_u;
  __shared__ mrch::LdsF s_f;
  __shared__ double sums[2][BY][1];
  double acc[1] = {0.0};
  mrch::march<1>(
      g, a.c, a.h, blk, L, s_u, s_f,
      [&](int m, double& vx, double& vy) {  // the iterate x0 + w
        vx = a.solx[m] + (double)a.wx[m];
        vy = a.soly[m] + (double)a.wy[m];
      },
      [&](bool own, int k, double, double, double ax, double ay, double (&v)[1]) {
        if (!own) return;
        const double h = (double)a.h[k];
        double dgx = 1.0, dgy = 1.0;
        if (h >= WC_H_DISPERSION_MIN) {
          const double heff = wc_max(h, WC_H_DRY);
          dgx = heff + ((double)a.c[k + 1] + (double)a.c[k - 1]) * 0.25 / (g.dx * g.dx);
          dgy = heff + ((double)a.c[k + g.w] + (double)a.c[k - g.w]) * 0.25 / (g.dy * g.dy);
        }
        const double bx = (double)a.bx[k], by = (double)a.by[k];
        a.r[0][k] = f2(bx - ax, by - ay);
        a.di[k] = f2(1.0 / dgx, 1.0 / dgy);
        a.p[k] = make_float2(0.f, 0.f);
        a.s[1][k] = make_float2(0.f, 0.f);
        v[0] = bx * bx + by * by;
      },
      acc);
  rows_total<1>(acc, sums);
  if (threadIdx.x == 0) a.lanes_b[blk] = acc[0];
}

template <int L>
__device__ __forceinline__ void cg1m_start_body(const CgG& a, int blk) {
  const Geo& g = a.g;
  __shared__ mrch::LdsU s_u;
  __shared__ mrch::LdsF s_f;
  __shared__ double sums[2][BY][3];
  double acc[3] = {0.0, 0.0, 0.0};
  mrch::march<3>(
      g, a.c, a.h, blk, L, s_u, s_f,
      [&](int m, double& vx, double& vy) {
        const float2 rr = a.r[0][m], d = a.di[m];
        vx = (double)rr.x * (double)d.x;
        vy = (double)rr.y * (double)d.y;
      },
      [&](bool own, int k, double ux, double uy, double ax, double ay, double (&v)[3]) {
        if (!own) return;
        const double qx = (double)(float)ax, qy = (double)(float)ay;
        a.q[0][k] = f2(qx, qy);
        const float2 rr = a.r[0][k];
        const double rx = rr.x, ry = rr.y;
        v[0] = rx * ux + ry * uy;
        v[1] = qx * ux + qy * uy;
        v[2] = rx * rx + ry * ry;
      },
      acc);
  rows_total<3>(acc, sums);
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) a.lanes[m * L + blk] = acc[m];
}

template <int L>
__device__ __forceinline__ void cg1m_iter_body(const CgG& a, unsigned long long n, int blk) {
  const Geo& g = a.g;
  double alpha = 0.0, beta = 0.0, alpha_prev = 0.0;
  if (!start<L>(a, n, blk, alpha, beta, alpha_prev)) return;
  const unsigned long long i = n - 1;
  const int cur = (int)(i & 1), nxt = cur ^ 1;
  const float2 *r0 = a.r[cur], *q0 = a.q[cur];
  const float2* s0 = a.s[nxt];  // s_{i-1}: written by K_{n-1} into nxt
  const double na = -alpha;
  __shared__ mrch::LdsU s_u;
  __shared__ mrch::LdsF s_f;
  __shared__ double sums[2][BY][3];
  double acc[3] = {0.0, 0.0, 0.0};
  mrch::march<3>(
      g, a.c, a.h, blk, L, s_u, s_f,
      [&](int m, double& vx, double& vy) {  // u_{i+1} = D^-1 (r_i - alpha (q_i + beta s_{i-1}))
        const float2 qq = q0[m], ss = s0[m], rr = r0[m], d = a.di[m];
        const double sx = (double)(float)((double)qq.x + beta * (double)ss.x);
        const double sy = (double)(float)((double)qq.y + beta * (double)ss.y);
        vx = (double)(float)((double)rr.x + na * sx) * (double)d.x;
        vy = (double)(float)((double)rr.y + na * sy) * (double)d.y;
      },
      [&](bool own, int k, double ux, double uy, double ax, double ay, double (&v)[3]) {
        if (!own) return;
        const float2 qq = q0[k], ss = s0[k], rr = r0[k], d = a.di[k], pp = a.p[k];
        const double sx = (double)(float)((double)qq.x + beta * (double)ss.x);
        const double sy = (double)(float)((double)qq.y + beta * (double)ss.y);
        const double px = (double)(float)((double)rr.x * (double)d.x + beta * (double)pp.x);
        const double py = (double)(float)((double)rr.y * (double)d.y + beta * (double)pp.y);
        a.p[k] = f2(px, py);
        if (!a.defer) {
          a.wx[k] = (float)((double)a.wx[k] + alpha * px);
          a.wy[k] = (float)((double)a.wy[k] + alpha * py);
        } else if (i & 1) {  // iterations i - 1 and i together (CgG::defer)
          a.wx[k] = (float)((double)a.wx[k] + (alpha_prev * (double)pp.x + alpha * px));
          a.wy[k] = (float)((double)a.wy[k] + (alpha_prev * (double)pp.y + alpha * py));
        }
        a.s[cur][k] = f2(sx, sy);
        const double nrx = (double)(float)((double)rr.x + na * sx);
        const double nry = (double)(float)((double)rr.y + na * sy);
        a.r[nxt][k] = f2(nrx, nry);
        const double qx = (double)(float)ax, qy = (double)(float)ay;
        a.q[nxt][k] = f2(qx, qy);
        v[0] = nrx * ux + nry * uy;
        v[1] = qx * ux + qy * uy;
        v[2] = nrx * nrx + nry * nry;
      },
      acc);
  rows_total<3>(acc, sums);
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; ++m) a.lanes[3 * L * (n & 1) + m * L + blk] = acc[m];
}

template <int L>
__global__ void __launch_bounds__(PT) k_cg1m_init(CgG a) { cg1m_init_body<L>(a, blockIdx.x); }
template <int L>
__global__ void __launch_bounds__(PT) k_cg1m_start(CgG a) { cg1m_start_body<L>(a, blockIdx.x); }
template <int L>
__global__ void __launch_bounds__(PT) k_cg1m_iter(CgG a, unsigned long long n) {
  cg1m_iter_body<L>(a, n, blockIdx.x);
}

// ---- many instances in one launch (see kb_fill32) -------------------------------------------
// Instance q runs min(256, its tiles) virtual blocks; its lanes past those stay zero, as the
// single launch's empty blocks write them, so the sums and every iterate are the same.
template <int L>
__global__ void __launch_bounds__(PT) kb_cg1_init(const CgG* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1_init_body<L, 1>(args[q], vb - m.first[q]);
    __syncthreads();
  }
}

template <int L>
__global__ void __launch_bounds__(PT) kb_cg1_start(const CgG* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1_start_body<L, 1>(args[q], vb - m.first[q]);
    __syncthreads();
  }
}

template <int L>
__global__ void __launch_bounds__(PT) kb_cg1_iter(const CgG* __restrict__ args, unsigned long long n, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1_iter_body<L, 1, true>(args[q], n, vb - m.first[q]);
    __syncthreads();
  }
}

// The marching kernels over many instances: instance q runs min(L, its strip units) virtual
// blocks (mrch::segments), the same blocks its own launch runs work in.
template <int L>
__global__ void __launch_bounds__(PT) kbm_cg1_init(const CgG* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1m_init_body<L>(args[q], vb - m.first[q]);
    __syncthreads();
  }
}

template <int L>
__global__ void __launch_bounds__(PT) kbm_cg1_start(const CgG* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1m_start_body<L>(args[q], vb - m.first[q]);
    __syncthreads();
  }
}

template <int L>
__global__ void __launch_bounds__(PT) kbm_cg1_iter(const CgG* __restrict__ args, unsigned long long n, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    cg1m_iter_body<L>(args[q], n, vb - m.first[q]);
    __syncthreads();
  }
}

// One wave per instance: k_cg1_check, then the instance's flag and iteration into done[q],
// it[q]; with `stats` (the solve is over) DispersionStats counts the solve as k_cg_stats does.
template <int L>
__global__ void __launch_bounds__(64) kb_cg1_check(const CgG* __restrict__ args, unsigned long long n,
                                                   int stats, int* done, unsigned long long* it) {
  const int q = blockIdx.x;
  const CgG& a = args[q];
  double alpha, beta, alpha_prev;
  start<L>(a, n, 0, alpha, beta, alpha_prev);
  if (threadIdx.x == 0) {
    CgScalars* sc = a.sc;
    done[q] = sc->done;
    it[q] = sc->it;
    if (stats) {
      sc->solves += 1;
      sc->iterations += sc->it;
      if (!sc->converged) sc->unconverged += 1;
    }
  }
}

#define WC_CG1_LANES(L)                                                                      \
  template __global__ void k_cg1_init<L>(CgG);                                                 \
  template __global__ void k_cg1_start<L>(CgG);                                                \
  template __global__ void k_cg1_iter<L>(CgG, unsigned long long);                             \
  template __global__ void k_cg1_check<L>(CgG, unsigned long long);                            \
  template __global__ void k_cg1m_init<L>(CgG);                                                \
  template __global__ void k_cg1m_start<L>(CgG);                                               \
  template __global__ void k_cg1m_iter<L>(CgG, unsigned long long);                            \
  template __global__ void kb_cg1_init<L>(const CgG*, BMap);                                   \
  template __global__ void kb_cg1_start<L>(const CgG*, BMap);                                  \
  template __global__ void kb_cg1_iter<L>(const CgG*, unsigned long long, BMap);               \
  template __global__ void kbm_cg1_init<L>(const CgG*, BMap);                                  \
  template __global__ void kbm_cg1_start<L>(const CgG*, BMap);                                 \
  template __global__ void kbm_cg1_iter<L>(const CgG*, unsigned long long, BMap);              \
  template __global__ void kb_cg1_check<L>(const CgG*, unsigned long long, int, int*,          \
                                           unsigned long long*);
WC_CG1_LANES(256)
WC_CG1_LANES(512)
#undef WC_CG1_LANES

}  // namespace wc
