// The FP64 path's PCG (dispersion.rs:493-545): the reference's iterates bit for bit, two
// kernels per iteration over 64x16-cell tiles.
//
// The reference's loop needs two grid-wide reductions per iteration (p.Ap before the update,
// r.r and r.z after it), so an iteration is two launches, each a pass over the grid:
//   M_n: beta_{n-1} from U_{n-1}'s lanes (or the start's); p_{n-1} = z_{n-1} + beta p_{n-2}
//        (p_0 = z_0) on the tile + 2, with z = r / diag taken there from r and the diagonal the
//        set-up stored; A p_{n-1}, lanes of p.Ap; p_{n-1} stored;
//   U_n: alpha_n = r.z / p.Ap; p_{n-1} staged again on the tile + 2 and A p_{n-1} computed
//        again (the same arithmetic, so the same bits) instead of being stored by M_n and read
//        back; w += alpha p; r -= alpha A p; z = r / diag; lanes of r.r and r.z.
// Recomputing A p in U_n saves its store and its load (32 of 208 bytes per cell and
// iteration), and M_n's z from r and the stored diagonal replaces U_n's store of z by a load;
// the halo work runs from LDS and L2. Per cell and iteration the passes move r, p_old, diag
// (read), p (written), c, h in M_n and p, c, h, r (read), r (written) in U_n, and every other
// iteration w and p_(n-2) (read) and w (written), which U_n updates two steps at a time: 168
// bytes on average.
//
// r, p and the diagonal are stored as interleaved (x, y) pairs (double2): a halo point of each
// is one 16-byte load, and a halo row's partial 128-byte lines are shared by both components.
// The kernels are bound by how many loads and L2 requests each CU keeps in flight, not by the
// bytes alone, so halving both for these three vectors shortens both passes.
//
// Layout: 64x16 tiles (one 1024-thread block, a cell per thread; the two-cell halo is a third of
// the tile where a 64x4 tile's is 113%), exactly 256 blocks, lane b taking tiles b, b+256, ...
// The inputs of the next two tiles (their halos, the cells' own values) are loaded into
// registers while the current tile computes. Each thread keeps running sums over its cells, and the block
// reduces them once, at the end: rows by the canonical butterfly, then in row order. That
// register is lane b of the final stage, which every wave of the next kernel finishes itself,
// so no separate reduction kernel, barrier or atomic is needed. The oracle's dot order 2 is this
// order (oracle/wavecore_oracle.cpp: canonical16_sum).
//
// Block b runs on XCD b % 8, and carries the lane lane_of_block gives it: each XCD takes the
// lanes of an eighth of the tile columns, so a tile's neighbours on all four sides share its
// XCD's L2 for their halo lines.

#include "wc_tile16.h"

namespace wc {

namespace {

using namespace t16;

// The scalars iteration n starts from (M_n's prologue and the check kernel); false when the
// solve is over. Every wave computes the same numbers; block 0 records them.
__device__ __forceinline__ bool start(const CgV<double>& a, unsigned long long n, double& beta) {
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
    // A solve that ended in an earlier launch stays as it ended: later launches write nothing.
    // (Block 0 of this launch may already have set the flag; a wave that reads it would have
    // reached the same verdict from the lanes.)
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

// Streamed values of a cell that no other block of the launch reads, loaded with the nontemporal
// hint (stored with it by store_nt).
__device__ __forceinline__ double2 ld_nt(const double2* p) {
  return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}

// Ring element t of a tile + 2 halo: the mapped interior index m and flags (1: inside the
// grid's two rings; 2, 4: the x, y component flips at a wall).
__device__ __forceinline__ int ring2(const Geo& g, int tile, int t, int& m) {
  const int tx = tiles_x(g);
  const int i = GH + (tile % tx) * BX - 2 + t % S2X;
  const int j = GH + (tile / tx) * BY - 2 + t / S2X;
  if (i > g.nx + GH + 1 || j > g.ny + GH + 1) return 0;
  m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
  return 1 | (is_ghost_x(i, g.nx) ? 2 : 0) | (!g.periodic && is_ghost_y(j, g.ny) ? 4 : 0);
}

// Ring element t of a tile + 1 halo: its padded index, or -1 outside the used rows and
// columns (where f is zero).
__device__ __forceinline__ int ring1(const Geo& g, int tile, int t) {
  const int tx = tiles_x(g);
  const int i = GH + (tile % tx) * BX - 1 + t % S1X;
  const int j = GH + (tile / tx) * BY - 1 + t / S1X;
  if (i > g.nx + GH || j > g.ny + GH) return -1;
  return j * g.w + i;
}

// The thread's own cell of a tile: its padded index, or -1 outside the interior.
__device__ __forceinline__ int own_cell(const Geo& g, int tile) {
  const int tx = tiles_x(g);
  const int i = GH + (tile % tx) * BX + (threadIdx.x & 63);
  const int j = GH + (tile / tx) * BY + (threadIdx.x >> 6);
  return (i < g.nx + GH && j < g.ny + GH) ? j * g.w + i : -1;
}

// What a thread loads for one tile, a tile ahead: NV values at each of its (at most two) tile + 2
// halo points, c at its (at most two) tile + 1 points, and NO values at its own cell.
template <int NV, int NO>
struct Pre {
  double v[2][NV];
  int flags[2];
  double c[2];
  double o[NO];
  int k;  // own cell, -1 outside
};

template <int NV, int NO, class Halo, class Own>
__device__ __forceinline__ void prefetch(const CgV<double>& a, int tile, Halo halo, Own own,
                                         Pre<NV, NO>& P) {
  const Geo& g = a.g;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = threadIdx.x + e * PT;
    P.flags[e] = 0;
    int m = 0;
    if (t < H2 && (P.flags[e] = ring2(g, tile, t, m)) != 0) halo(m, P.v[e]);
    P.c[e] = 0.0;
    if (t < H1) {
      const int q = ring1(g, tile, t);
      if (q >= 0) P.c[e] = a.c[q];
    }
  }
  P.k = own_cell(g, tile);
  if (P.k >= 0) own(P.k, P.o);
}

// The vector on the tile + 2 from the prefetched halo values by make(v, vx, vy), the wall flips
// applied; c on the tile + 1.
template <int NV, int NO, class Make>
__device__ __forceinline__ void stage(const Pre<NV, NO>& P, double (*s_u)[S2Y][S2X],
                                      double (*s_c)[S1X], Make make) {
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = threadIdx.x + e * PT;
    if (t < H2) {
      double vx = 0.0, vy = 0.0;
      if (P.flags[e] & 1) {
        make(P.v[e], vx, vy);
        if (P.flags[e] & 2) vx = -vx;
        if (P.flags[e] & 4) vy = -vy;
      }
      s_u[0][t / S2X][t % S2X] = vx;
      s_u[1][t / S2X][t % S2X] = vy;
    }
    if (t < H1) s_c[t / S1X][t % S1X] = P.c[e];
  }
}

// f = c (d/dx px + d/dy py) on the tile + 1 (dispersion.rs:283-296), then a barrier; outside the
// used rows and columns f is zero (c is staged as zero there).
__device__ __forceinline__ void stencil_f(const Geo& g, double (*s_u)[S2Y][S2X],
                                          double (*s_c)[S1X], double (*s_f)[S1X]) {
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < H1; t += PT) {
    const int lx = t % S1X, ly = t / S1X;
    const int x = lx + 1, y = ly + 1;
    s_f[ly][lx] = s_c[ly][lx] * ((s_u[0][y][x + 1] - s_u[0][y][x - 1]) * ix +
                                 (s_u[1][y + 1][x] - s_u[1][y - 1][x]) * iy);
  }
  __syncthreads();
}

// A p at the thread's own cell (dispersion.rs:298-316), p there, from the staged p and f.
__device__ __forceinline__ void apply_own(const Geo& g, double h, double (*s_u)[S2Y][S2X],
                                          double (*s_f)[S1X], double& px, double& py,
                                          double& ax, double& ay) {
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
  const double heff = wc_max(h, WC_H_DRY);
  const int x1 = lx + 1, y1 = ly + 1;
  px = s_u[0][ly + 2][lx + 2];
  py = s_u[1][ly + 2][lx + 2];
  ax = mask * (heff * px - (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix);
  ay = mask * (heff * py - (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy);
}

// The Jacobi diagonal (dispersion.rs:466-480) of a cell from its depth and its neighbours' c.
__device__ __forceinline__ void diag_at(const Geo& g, double h, double cxm, double cxp,
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

// The same at the own cell, c of its neighbours from the staged tile + 1.
__device__ __forceinline__ void own_diag(const Geo& g, double h, double (*s_c)[S1X], double& dgx,
                                         double& dgy) {
  const int x = (threadIdx.x & 63) + 1, y = (threadIdx.x >> 6) + 1;
  diag_at(g, h, s_c[y][x - 1], s_c[y][x + 1], s_c[y - 1][x], s_c[y + 1][x], dgx, dgy);
}

}  // namespace

// r = b - A w; lanes of b.b, r.z, r.r with z = r / diag (dispersion.rs:493-514; p_0 = z_0 is
// staged by M_1 from r).
__global__ void __launch_bounds__(PT) k_cg2_init(CgV<double> a) {
  const Geo& g = a.g;
  const int nt = tiles_all(g);
  __shared__ double s_u[2][S2Y][S2X], s_c[S1Y][S1X], s_f[S1Y][S1X];
  __shared__ double sums[1][BY][3];
  double acc[3] = {0.0, 0.0, 0.0};
  const int lane = lane_of_block(g);
  auto halo = [&](int m, double (&v)[2]) {
    v[0] = a.wx[m];
    v[1] = a.wy[m];
  };
  auto own = [&](int k, double (&o)[3]) {
    o[0] = a.h[k];
    o[1] = a.bx[k];
    o[2] = a.by[k];
  };
  Pre<2, 3> P;
  if (lane < nt) prefetch(a, lane, halo, own, P);
  for (int tile = lane; tile < nt; tile += LANES) {
    stage(P, s_u, s_c, [](const double (&v)[2], double& vx, double& vy) {
      vx = v[0];
      vy = v[1];
    });
    const int k = P.k;
    const double h = P.o[0], bx = P.o[1], by = P.o[2];
    __syncthreads();
    if (tile + LANES < nt) prefetch(a, tile + LANES, halo, own, P);
    stencil_f(g, s_u, s_c, s_f);
    double v[3] = {0.0, 0.0, 0.0};
    if (k >= 0) {
      double wx, wy, ax, ay;
      apply_own(g, h, s_u, s_f, wx, wy, ax, ay);
      const double rx = bx - ax, ry = by - ay;
      double dgx, dgy;
      own_diag(g, h, s_c, dgx, dgy);
      const double zx = rx / dgx, zy = ry / dgy;
      store_nt(&a.r2[k], make_double2(rx, ry));
      v[0] = bx * bx + by * by;
      v[1] = rx * zx + ry * zy;
      v[2] = rx * rx + ry * ry;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = acc[q] + v[q];
    __syncthreads();
  }
  rows_total<3>(acc, sums);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) a.acc_init[q * LANES + lane] = acc[q];
}

// M_n: p_{n-1} = z + beta p_{n-2} (p_0 = z) on the tile + 2, stored at the tile's cells; A p;
// lanes of p.Ap (dispersion.rs:516-518, 537-539).
__global__ void __launch_bounds__(PT) k_cg2_matvec(CgV<double> a, unsigned long long n,
                                                   const double2* __restrict__ po,
                                                   double2* __restrict__ pn) {
  const Geo& g = a.g;
  const int nt = tiles_all(g);
  const bool first = n == 1;
  // z = r / diag at the halo's (mapped) cells from r and the diagonal the set-up stored
  // (k_sgn_setup, dispersion.rs:466-480): no z array
  auto halo = [&](int m, double (&v)[6]) {
    const double2 r = a.r2[m], d = a.dg2[m];
    const double2 p = first ? make_double2(0.0, 0.0) : po[m];
    v[0] = r.x;
    v[1] = r.y;
    v[2] = p.x;
    v[3] = p.y;
    v[4] = d.x;
    v[5] = d.y;
  };
  auto own = [&](int k, double (&o)[1]) { o[0] = a.h[k]; };
  // two tiles' loads in flight: the first two while the scalars are reduced, then each tile's
  // successor-but-one as soon as the tile is staged
  Pre<6, 1> P[2];
  const int t0 = lane_of_block(g);
  if (t0 < nt) prefetch(a, t0, halo, own, P[0]);
  if (t0 + LANES < nt) prefetch(a, t0 + LANES, halo, own, P[1]);
  double beta = 0.0;
  if (t0 < nt && !start(a, n, beta)) return;
  __shared__ double s_u[2][S2Y][S2X], s_c[S1Y][S1X], s_f[S1Y][S1X];
  __shared__ double sums[1][BY][1];
  double acc[1] = {0.0};
  auto body = [&](Pre<6, 1>& Q, int tile) {
    stage(Q, s_u, s_c, [&](const double (&v)[6], double& vx, double& vy) {
      const double zx = v[0] / v[4], zy = v[1] / v[5];
      if (first) {
        vx = zx;
        vy = zy;
      } else {
        vx = zx + beta * v[2];
        vy = zy + beta * v[3];
      }
    });
    const int k = Q.k;
    const double h = Q.o[0];
    __syncthreads();
    if (tile + 2 * LANES < nt) prefetch(a, tile + 2 * LANES, halo, own, Q);
    stencil_f(g, s_u, s_c, s_f);
    double v = 0.0;
    if (k >= 0) {
      double px, py, ax, ay;
      apply_own(g, h, s_u, s_f, px, py, ax, ay);
      store_nt(&pn[k], make_double2(px, py));
      v = px * ax + py * ay;
    }
    acc[0] = acc[0] + v;
    __syncthreads();
  };
  for (int tile = t0; tile < nt; tile += 2 * LANES) {
    body(P[0], tile);
    if (tile + LANES < nt) body(P[1], tile + LANES);
  }
  rows_total<1>(acc, sums);
  if (threadIdx.x == 0) a.acc_pap[t0] = acc[0];
}

// U_n: alpha = r.z / p.Ap; A p again from p_{n-1} on the tile + 2; r -= alpha A p,
// z = r / diag (dispersion.rs:519-526, 568-589); lanes of r.r and r.z. The solution takes two
// iterations' steps in one pass: U_n of odd n leaves w alone, U_n of even n (PAIR) makes
// w = (w + alpha_{n-1} p_{n-2}) + alpha_n p_{n-1}, the reference's two updates in its order, so
// w reaches the same bits as if each U_n had updated it (nothing reads w during the solve; a
// solve that ends after an odd iteration gets its last step from k_cg2_flush). w is read and
// written every other iteration, for one more read of p_{n-2}: 88 instead of 96 bytes per cell
// and iteration on average.
template <int PAIR>
__global__ void __launch_bounds__(PT) k_cg2_update(CgV<double> a, unsigned long long n,
                                                   const double2* __restrict__ pc,
                                                   const double2* __restrict__ pp) {
  constexpr int NO = PAIR ? 7 : 3;
  CgScalars* sc = a.sc;
  const Geo& g = a.g;
  const int nt = tiles_all(g);
  auto halo = [&](int m, double (&v)[2]) {
    const double2 p = pc[m];
    v[0] = p.x;
    v[1] = p.y;
  };
  auto own = [&](int k, double (&o)[NO]) {
    const double2 r = ld_nt(&a.r2[k]);
    o[0] = a.h[k];
    o[1] = r.x;
    o[2] = r.y;
    if constexpr (PAIR) {
      const double2 q = ld_nt(&pp[k]);  // p_{n-2}
      o[3] = __builtin_nontemporal_load(&a.wx[k]);
      o[4] = __builtin_nontemporal_load(&a.wy[k]);
      o[5] = q.x;
      o[6] = q.y;
    }
  };
  Pre<2, NO> P[2];  // two tiles' loads in flight
  const int t0 = lane_of_block(g);
  if (t0 < nt) prefetch(a, t0, halo, own, P[0]);
  if (t0 + LANES < nt) prefetch(a, t0 + LANES, halo, own, P[1]);
  double alpha = 0.0, alpha_prev = 0.0;
  if (t0 < nt) {
    if (sc->done) return;
    double s[1];
    lane_sums<1>(a.acc_pap, s);
    alpha = sc->rzs[n & 1] / s[0];
    if (PAIR) alpha_prev = sc->alphas[(n - 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->alpha = alpha;
      sc->alphas[n & 1] = alpha;
    }
  }
  const double na = -alpha;
  __shared__ double s_u[2][S2Y][S2X], s_c[S1Y][S1X], s_f[S1Y][S1X];
  __shared__ double sums[1][BY][2];
  double acc[2] = {0.0, 0.0};
  auto body = [&](Pre<2, NO>& Q, int tile) {
    stage(Q, s_u, s_c, [](const double (&v)[2], double& vx, double& vy) {
      vx = v[0];
      vy = v[1];
    });
    const int k = Q.k;
    const double h = Q.o[0], rx0 = Q.o[1], ry0 = Q.o[2];
    double wx0 = 0.0, wy0 = 0.0, qx = 0.0, qy = 0.0;
    if constexpr (PAIR) {
      wx0 = Q.o[3];
      wy0 = Q.o[4];
      qx = Q.o[5];
      qy = Q.o[6];
    }
    __syncthreads();
    if (tile + 2 * LANES < nt) prefetch(a, tile + 2 * LANES, halo, own, Q);
    stencil_f(g, s_u, s_c, s_f);
    double v[2] = {0.0, 0.0};
    if (k >= 0) {
      double px, py, ax, ay;
      apply_own(g, h, s_u, s_f, px, py, ax, ay);
      if constexpr (PAIR) {
        store_nt(&a.wx[k], (wx0 + alpha_prev * qx) + alpha * px);
        store_nt(&a.wy[k], (wy0 + alpha_prev * qy) + alpha * py);
      }
      const double rx = rx0 + na * ax, ry = ry0 + na * ay;
      double dgx, dgy;
      own_diag(g, h, s_c, dgx, dgy);
      const double zx = rx / dgx, zy = ry / dgy;
      store_nt(&a.r2[k], make_double2(rx, ry));
      v[0] = rx * rx + ry * ry;
      v[1] = rx * zx + ry * zy;
    }
    acc[0] = acc[0] + v[0];
    acc[1] = acc[1] + v[1];
    __syncthreads();
  };
  for (int tile = t0; tile < nt; tile += 2 * LANES) {
    body(P[0], tile);
    if (tile + LANES < nt) body(P[1], tile + LANES);
  }
  rows_total<2>(acc, sums);
  if (threadIdx.x == 0)
    for (int q = 0; q < 2; ++q) a.acc_update[q * LANES + t0] = acc[q];
}

template __global__ void k_cg2_update<0>(CgV<double>, unsigned long long, const double2*,
                                         const double2*);
template __global__ void k_cg2_update<1>(CgV<double>, unsigned long long, const double2*,
                                         const double2*);

// The step U_it left in a solve that ended after an odd iteration it: w += alpha_it p_{it-1}.
__global__ void __launch_bounds__(NT) k_cg2_flush(CgV<double> a, unsigned long long it,
                                                  const double2* __restrict__ p) {
  const Geo& g = a.g;
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.nx * g.ny) return;
  const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
  const double alpha = a.sc->alphas[it & 1];
  const double2 q = p[k];
  a.wx[k] = a.wx[k] + alpha * q.x;
  a.wy[k] = a.wy[k] + alpha * q.y;
}

// M_n's prologue alone, one wave: records whether the solve is over after iteration n - 1.
__global__ void __launch_bounds__(64) k_cg2_check(CgV<double> a, unsigned long long n) {
  double beta;
  start(a, n, beta);
}

}  // namespace wc
