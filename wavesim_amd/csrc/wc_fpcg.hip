// Flexible PCG with the learned preconditioner, FP32 form (no counterpart in the reference,
// whose PCG is preconditioned by the Jacobi diagonal, dispersion.rs:493-545,568-589).
//
// The network z = N(r) is not a fixed linear operator, so beta is Polak and Ribiere's,
//   beta_n = r_n . (z_n - z_{n-1}) / r_{n-1} . z_{n-1},
// which keeps the search directions conjugate for a preconditioner that varies (Golub and Ye,
// 1999; Notay, 2000). The tolerance and the test on the recursively updated residual are the
// reference's (target = tol ||b|| + 1e-10). When the residual stops decreasing the solve falls
// back to the reference's Jacobi
// preconditioner for the rest of its iterations, restarting the directions. "Stops decreasing":
// stall_limit iterations in a row each cut the residual norm by less than the factor min_ratio.
//
// Per iteration n: F_in (convergence and stall test on r_{n-1}; the network input from r / sigma,
// sigma its RMS), the network (E0 .. D0, skipped while done or fallen back), F_out (z_n: the
// output layer scaled by sigma / (4 dx) on wet cells, or D^-1 r; lanes of r.z_n and r.z_{n-1}),
// F_mv (p_n = z_n + beta p_{n-1}, A p_n, lanes of p.Ap), F_up (w += alpha p, r -= alpha Ap, lanes
// of r.r). The 64x16 tiles and lanes are those of wc_tile16.h.

#include "wc_tile16.h"

namespace wc {

namespace {

using namespace t16;

__device__ __forceinline__ float2 f2(double x, double y) { return make_float2((float)x, (float)y); }

// The scalars F_in works from for iteration n (the residual r_{n-1}); false when the solve is
// over. With update, block 0 also advances the stall counter and the fallback state, which
// only it reads (every block computes done and sigma, which need neither).
__device__ __forceinline__ bool fcg_start(const FcgArgs& a, unsigned long long n, bool update,
                                          double& sigma) {
  CgScalars* sc = a.sc;
  double target;
  if (n == 1) {
    double b[1];
    lane_sums<1>(a.lanes, b);
    target = a.tolerance * sqrt(b[0]) + 1e-10;
  } else {
    if (sc->done) return false;
    target = sc->target;
  }
  double s[1];
  lane_sums<1>(a.lanes + LANES, s);
  const double rr = s[0];
  const unsigned long long it = n - 1;
  const int conv = sqrt(rr) <= target ? 1 : 0;
  const bool done = conv || it >= a.max_iterations;
  sigma = sqrt(rr / ((double)a.g.nx * (double)a.g.ny));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (n == 1) sc->target = target;
    sc->converged = conv;
    sc->it = it;
    sc->done = done ? 1 : 0;
    if (update) {
      // stall counts the iterations in a row whose residual norm fell by less than the factor
      // a.min_ratio (r.r by less than its square): the network has stopped earning its cost
      int stall = 0, fallback = 0;
      if (n > 1) {
        stall = rr > a.min_ratio * a.min_ratio * sc->rr_prev ? sc->stall + 1 : 0;
        fallback = sc->fallback;
      }
      const int fall_now = !fallback && a.stall_limit > 0 && stall >= a.stall_limit;
      if (fall_now) sc->fallbacks += 1;
      fallback = fallback || fall_now;
      sc->stall = stall;
      sc->fallback = fallback;
      sc->restart = fall_now;
      sc->rr_prev = rr;
      sc->sigma = sigma;
      sc->net_off = (done || fallback) ? 1 : 0;
      if (a.history && it < a.history_len) a.history[it] = rr;
    }
  }
  return !done;
}

}  // namespace

// Bed-gradient features: scale * (b[i+1] - b[i-1]) / (2 dx), the same in y, on every cell
// whose neighbours exist (zero on the outer ring).
__global__ void k_bed_features(Geo g, const double* b, double scale, float* bgx, float* bgy) {
  const long long n = (long long)g.w * g.rows;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = (int)(k % g.w), j = (int)(k / g.w);
  float gx = 0.f, gy = 0.f;
  if (i > 0 && i < g.w - 1 && j > 0 && j < g.rows - 1) {
    gx = (float)(scale * (b[k + 1] - b[k - 1]) / (2.0 * g.dx));
    gy = (float)(scale * (b[k + g.w] - b[k - g.w]) / (2.0 * g.dy));
  }
  bgx[k] = gx;
  bgy[k] = gy;
}

// r = b - A w, the inverse Jacobi diagonal (dispersion.rs:466-480), p_0 = z = 0; lanes of b.b
// and r.r.
__global__ void __launch_bounds__(PT) k_fcg_init(FcgArgs a) {
  const Geo& g = a.g;
  const int nt = tiles_all(g);
  __shared__ double s_u[2][S2Y][S2X], s_f[S1Y][S1X];
  __shared__ double sums[2][BY][2];
  double acc[2] = {0.0, 0.0};
  for (int r = 0; (int)blockIdx.x + LANES * r < nt; ++r) {
    const int tile = blockIdx.x + LANES * r;
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil(g, a.c, a.h, tile, s_u, s_f,
            [&](int m, double& vx, double& vy) {  // the iterate x0 + w
              vx = a.solx[m] + (double)a.wx[m];
              vy = a.soly[m] + (double)a.wy[m];
            },
            own, k, ux, uy, ax, ay);
    double v[2] = {0.0, 0.0};
    if (own) {
      const double h = (double)a.h[k];
      double dgx = 1.0, dgy = 1.0;
      if (h >= WC_H_DISPERSION_MIN) {
        const double heff = wc_max(h, WC_H_DRY);
        dgx = heff + ((double)a.c[k + 1] + (double)a.c[k - 1]) * 0.25 / (g.dx * g.dx);
        dgy = heff + ((double)a.c[k + g.w] + (double)a.c[k - g.w]) * 0.25 / (g.dy * g.dy);
      }
      const double bx = (double)a.bx[k], by = (double)a.by[k];
      const double rx = (double)(float)(bx - ax), ry = (double)(float)(by - ay);
      a.r[k] = f2(rx, ry);
      a.di[k] = f2(1.0 / dgx, 1.0 / dgy);
      a.p[0][k] = make_float2(0.f, 0.f);
      a.z[k] = make_float2(0.f, 0.f);
      v[0] = bx * bx + by * by;
      v[1] = rx * rx + ry * ry;
    }
    add_rows<2>(v, sums, r, acc);
  }
  rows_total<2>(acc, sums);
  if (threadIdx.x == 0) {
    a.lanes[blockIdx.x] = acc[0];
    a.lanes[LANES + blockIdx.x] = acc[1];
  }
}

// The network input from r_{n-1}: channels h / (4 dx), the bed gradient, r / sigma, zeros.
__global__ void __launch_bounds__(PT) k_fcg_in(FcgArgs a, unsigned long long n) {
  double sigma;
  if (!fcg_start(a, n, true, sigma)) return;
  // After a fallback the input is not used. A solve starts without one (at n = 1 the flag in
  // memory may still be the previous solve's, which block 0 is resetting); later, a block that
  // reads the flag as it was before block 0 set it only writes an unused input.
  if (n > 1 && a.sc->fallback) return;
  const Geo& g = a.g;
  const double rs = sigma > 0.0 ? 1.0 / sigma : 0.0;
  const long long cells = (long long)g.nx * g.ny;
  for (long long c = (long long)blockIdx.x * PT + threadIdx.x; c < cells;
       c += (long long)gridDim.x * PT) {
    const int x = (int)(c % g.nx), y = (int)(c / g.nx);
    const int k = (y + GH) * g.w + x + GH;
    const float2 r = a.r[k];
    typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
    bf16x8 o;
    o[0] = (__bf16)(float)((double)a.h[k] * a.hs);
    o[1] = (__bf16)a.bgx[k];
    o[2] = (__bf16)a.bgy[k];
    o[3] = (__bf16)(float)((double)r.x * rs);
    o[4] = (__bf16)(float)((double)r.y * rs);
    o[5] = o[6] = o[7] = (__bf16)0.f;
    *reinterpret_cast<bf16x8*>(a.x0 + ((size_t)(y + 1) * (a.nw + 2) + x + 1) * 8) = o;
  }
}

// z_n from the network's last layer (or D^-1 r after a fallback); lanes of r.z_n, r.z_{n-1}.
__global__ void __launch_bounds__(PT) k_fcg_out(FcgArgs a, unsigned long long n) {
  const CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int nt = tiles_all(g), tx = tiles_x(g);
  const bool jacobi = sc->fallback != 0;
  const double zs = sc->sigma * a.hs;
  __shared__ double sums[2][BY][2];
  double acc[2] = {0.0, 0.0};
  typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
  for (int r = 0; (int)blockIdx.x + LANES * r < nt; ++r) {
    const int tile = blockIdx.x + LANES * r;
    const int x = (tile % tx) * BX + (threadIdx.x & 63), y = (tile / tx) * BY + (threadIdx.x >> 6);
    double v[2] = {0.0, 0.0};
    if (x < g.nx && y < g.ny) {
      const int k = (y + GH) * g.w + x + GH;
      const float2 rr = a.r[k], zo = a.z[k];
      double zx, zy;
      if (jacobi) {
        const float2 d = a.di[k];
        zx = (double)rr.x * (double)d.x;
        zy = (double)rr.y * (double)d.y;
      } else {
        const size_t o = ((size_t)(y + 1) * (a.nw + 2) + x + 1) * 16;
        const bf16x8 lo = *reinterpret_cast<const bf16x8*>(a.y0 + o);
        const bf16x8 hi = *reinterpret_cast<const bf16x8*>(a.y0 + o + 8);
        float o0 = a.wo[32], o1 = a.wo[33];
        for (int c = 0; c < 8; ++c) {
          o0 += (float)lo[c] * a.wo[2 * c];
          o1 += (float)lo[c] * a.wo[2 * c + 1];
        }
        for (int c = 0; c < 8; ++c) {
          o0 += (float)hi[c] * a.wo[2 * (c + 8)];
          o1 += (float)hi[c] * a.wo[2 * (c + 8) + 1];
        }
        const double mask = (double)a.h[k] >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
        zx = mask * zs * (double)o0;
        zy = mask * zs * (double)o1;
      }
      if (!isfinite(zx) || !isfinite(zy)) {  // never let the network put NaN into the solve
        const float2 d = a.di[k];
        zx = (double)rr.x * (double)d.x;
        zy = (double)rr.y * (double)d.y;
      }
      zx = (double)(float)zx;
      zy = (double)(float)zy;
      a.z[k] = f2(zx, zy);
      v[0] = (double)rr.x * zx + (double)rr.y * zy;
      v[1] = (double)rr.x * (double)zo.x + (double)rr.y * (double)zo.y;
    }
    add_rows<2>(v, sums, r, acc);
  }
  rows_total<2>(acc, sums);
  if (threadIdx.x == 0) {
    a.lanes[2 * LANES + blockIdx.x] = acc[0];
    a.lanes[3 * LANES + blockIdx.x] = acc[1];
  }
}

// p_n = z_n + beta p_{n-1} on the tile + 2, A p_n; lanes of p.Ap.
__global__ void __launch_bounds__(PT) k_fcg_mv(FcgArgs a, unsigned long long n) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int nt = tiles_all(g);
  double s[2];
  lane_sums<2>(a.lanes + 2 * LANES, s);
  const double rz = s[0];
  // Polak-Ribiere with restarts (beta >= 0, the "PR+" rule of Gilbert and Nocedal, 1992): a
  // fresh direction after a fallback, when the previous r.z was not positive, or beta < 0.
  double beta = 0.0;
  if (n > 1 && !sc->restart) {
    const double prev = sc->rzs[(n - 1) & 1];
    if (prev > 0.0) beta = wc_max((s[0] - s[1]) / prev, 0.0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rzs[n & 1] = rz;
    sc->beta = beta;
  }
  const float2* po = a.p[(n - 1) & 1];
  float2* pn = a.p[n & 1];
  __shared__ double s_u[2][S2Y][S2X], s_f[S1Y][S1X];
  __shared__ double sums[2][BY][1];
  double acc[1] = {0.0};
  for (int r = 0; (int)blockIdx.x + LANES * r < nt; ++r) {
    const int tile = blockIdx.x + LANES * r;
    bool own;
    int k;
    double ux, uy, ax, ay;
    stencil(g, a.c, a.h, tile, s_u, s_f,
            [&](int m, double& vx, double& vy) {
              const float2 z = a.z[m], p = po[m];
              vx = (double)(float)((double)z.x + beta * (double)p.x);
              vy = (double)(float)((double)z.y + beta * (double)p.y);
            },
            own, k, ux, uy, ax, ay);
    double v[1] = {0.0};
    if (own) {
      pn[k] = f2(ux, uy);
      const double qx = (double)(float)ax, qy = (double)(float)ay;
      a.q[k] = f2(qx, qy);
      v[0] = ux * qx + uy * qy;
    }
    add_rows<1>(v, sums, r, acc);
  }
  rows_total<1>(acc, sums);
  if (threadIdx.x == 0) a.lanes[4 * LANES + blockIdx.x] = acc[0];
}

// alpha = r.z / p.Ap; w += alpha p; r -= alpha Ap; lanes of r.r.
__global__ void __launch_bounds__(PT) k_fcg_up(FcgArgs a, unsigned long long n) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int nt = tiles_all(g), tx = tiles_x(g);
  double s[1];
  lane_sums<1>(a.lanes + 4 * LANES, s);
  const double alpha = sc->rzs[n & 1] / s[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha = alpha;
  const float2* p = a.p[n & 1];
  __shared__ double sums[2][BY][1];
  double acc[1] = {0.0};
  for (int r = 0; (int)blockIdx.x + LANES * r < nt; ++r) {
    const int tile = blockIdx.x + LANES * r;
    const int x = (tile % tx) * BX + (threadIdx.x & 63), y = (tile / tx) * BY + (threadIdx.x >> 6);
    double v[1] = {0.0};
    if (x < g.nx && y < g.ny) {
      const int k = (y + GH) * g.w + x + GH;
      const float2 pp = p[k], qq = a.q[k], rr = a.r[k];
      a.wx[k] = (float)((double)a.wx[k] + alpha * (double)pp.x);
      a.wy[k] = (float)((double)a.wy[k] + alpha * (double)pp.y);
      const double rx = (double)(float)((double)rr.x - alpha * (double)qq.x);
      const double ry = (double)(float)((double)rr.y - alpha * (double)qq.y);
      a.r[k] = f2(rx, ry);
      v[0] = rx * rx + ry * ry;
    }
    add_rows<1>(v, sums, r, acc);
  }
  rows_total<1>(acc, sums);
  if (threadIdx.x == 0) a.lanes[LANES + blockIdx.x] = acc[0];
}

// F_in's test alone, one wave, without advancing the stall state: ends a batch.
__global__ void __launch_bounds__(64) k_fcg_check(FcgArgs a, unsigned long long n) {
  double sigma;
  fcg_start(a, n, false, sigma);
}

}  // namespace wc
