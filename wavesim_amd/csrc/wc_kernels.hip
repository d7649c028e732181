// gfx950 kernels of wavecore's FP64 step. Launched by the host sources (wc_api.hip, wc_*_host.hip); declared in wc_kernels.h.
//
// One thread per cell on 64x4-cell tiles (a wavefront per tile row, so every global access of
// a row is coalesced along x), blocks mapped to tiles XCD by XCD (wc::tile_of_block). The
// arithmetic of each kernel is that of the reference function it cites, in the same order.

#include "wc_kernels.h"
#include "wc_stage.h"

namespace wc {

// ---- ghost layers ---------------------------------------------------------------------------
// state.rs:68-76: h, hu, hv and breaking ghosts from the interior. With `avg`, also the ghost
// part of the RK average of solver.rs:493-500: avg = (avg + this) / 2 for h, hu, hv, because
// the second stage's state carries the first stage's ghosts (solver.rs:490-491, 594-595).
__global__ void __launch_bounds__(NT) k_fill_state(Geo g, double* h, double* hu, double* hv,
                                                   double* br, double* ah, double* ahu,
                                                   double* ahv) {
  const int n = 6 * g.w + 6 * g.ny;
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= n) return;
  int i, j;
  if (t < 3 * g.w) {
    j = t / g.w;
    i = t % g.w;
  } else if (t < 6 * g.w) {
    const int u = t - 3 * g.w;
    j = g.ny + GH + u / g.w;
    i = u % g.w;
  } else {
    const int u = t - 6 * g.w;
    j = GH + u / 6;
    const int c = u % 6;
    i = c < 3 ? c : g.nx + GH + (c - 3);
  }
  const int k = j * g.w + i;
  const double vh = ghosted(h, g, i, j);
  const double vu = ghosted_vx(hu, g, i, j);
  const double vv = ghosted_vy(hv, g, i, j);
  h[k] = vh;
  hu[k] = vu;
  hv[k] = vv;
  if (br) br[k] = ghosted(br, g, i, j);
  if (ah) {
    ah[k] = 0.5 * (ah[k] + vh);
    ahu[k] = 0.5 * (ahu[k] + vu);
    ahv[k] = 0.5 * (ahv[k] + vv);
  }
}

// ---- the breaking switch (dispersion.rs:120-160) ----------------------------------------------
// phi on every padded cell: a ghost takes the value of the interior cell it mirrors or wraps,
// as the reference's mirror_ghosts/wrap_ghosts_y leave it. `held` (may be null) forces 0 on
// the cells held to shallow water (solver.rs:640-651). h must have its ghosts filled.
__device__ __forceinline__ double switch_at(const Geo& g, const double* h, const double* b,
                                           const Breaking& br, int i, int j) {
  const int k = j * g.w + i;
  if (h[k] < WC_H_DISPERSION_MIN) return 0.0;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  const double ex = ((h[k + 1] + b[k + 1]) - (h[k - 1] + b[k - 1])) * ix;
  const double ey = ((h[k + g.w] + b[k + g.w]) - (h[k - g.w] + b[k - g.w])) * iy;
  const double ratio = wc_max((h[k] + b[k]) / h[k], 0.0);
  double off = switch_off(br, ratio, ex, ey);
  if (br.memory && br.memory[k] != 0.0) off = 1.0;  // a front, or the water just behind it
  return 1.0 - off;
}

__global__ void __launch_bounds__(NT) k_switch(Geo g, const double* h, const double* b,
                                               Breaking br, const uint8_t* held, double* phi) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.w * g.rows) return;
  const int i = t % g.w, j = t / g.w;
  const int mi = map_x(i, g.nx), mj = map_y(j, g.ny, g.periodic);
  double v;
  if (held && held[(mj - GH) * g.nx + (mi - GH)]) {
    v = 0.0;
  } else {
    v = switch_at(g, h, b, br, mi, mj);
  }
  phi[t] = v;
}

// breaking_into (dispersion.rs:183-201) on a copy of the state whose ghosts are filled: the
// ghost reads of h go through the ghost map, the state's own ghosts are left alone.
__global__ void __launch_bounds__(NT) k_breaking(Geo g, const double* h, const double* b,
                                                 Breaking br, double* out) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.nx * g.ny) return;
  const int i = GH + t % g.nx, j = GH + t / g.nx;
  out[t] = breaking_at(h, b, g, br, i, j, j * g.w + i);
}

// ---- the finite-volume rates (solver.rs:625-718) ----------------------------------------------
__device__ __forceinline__ Cell cell_at(const StepArgs& a, int k, bool normal_is_x) {
  const double h = a.in_h[k];
  const double un = normal_is_x ? a.in_hu[k] : a.in_hv[k];
  const double ut = normal_is_x ? a.in_hv[k] : a.in_hu[k];
  return Cell{h, wc_velocity(h, un), wc_velocity(h, ut), a.b[k], a.phi[k]};
}

__global__ void __launch_bounds__(NT) k_flux(StepArgs a, int mode) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  Cell sx[7], sy[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    sx[m] = cell_at(a, j * g.w + (i + m - 3), true);
    sy[m] = cell_at(a, (j + m - 3) * g.w + i, false);
  }
  Face xl, xh, yl, yh;
  double xs, ys;
  axis_terms(sx, g.dx, a.order, xl, xh, xs);
  axis_terms(sy, g.dy, a.order, yl, yh, ys);
  double rh = -(xh.f0 - xl.f0) / g.dx - (yh.f0 - yl.f0) / g.dy;
  const double rhu = -((xh.f1 + xh.cl) - (xl.f1 + xl.cr)) / g.dx + xs - (yh.f2 - yl.f2) / g.dy;
  const double rhv = -((yh.f1 + yh.cl) - (yl.f1 + yl.cr)) / g.dy + ys - (xh.f2 - xl.f2) / g.dx;
  if (a.has_drive) {
    // Drive::source (forcing.rs:125-138) with the host-evaluated exp, cos and sin.
    const int ii = i - GH;
    double src = 0.0;
    if (ii >= a.src_i0 && ii < a.src_i1) src = a.src_col[ii] * a.prm[a.ramp_prm] * a.prm[a.sin_prm + (j - GH)];
    rh = rh + src;
  }
  stage_out(a, mode, i, j, j * g.w + i, rh, rhu, rhv);
}

// The Euler step of Order::First (solver.rs:485-487): s1 holds clean_dry(s + dt L); copy it to
// the interior of s and damp (the ghosts of s keep their filled values, as advance's clone did).
__global__ void __launch_bounds__(NT) k_finish_euler(StepArgs a) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  const int k = j * g.w + i;
  double h = a.s1_h[k], hu = a.s1_hu[k], hv = a.s1_hv[k];
  damp(a, a.dt, i, j, a.b[k], h, hu, hv);
  a.s_h[k] = h;
  a.s_hu[k] = hu;
  a.s_hv[k] = hv;
}

// ---- SGN dispersion (dispersion.rs:335-563) -----------------------------------------------------
// Setup: u, v, the mask, the masked velocities and their divergence, c = phi h^3/3 and c N,
// the Jacobi diagonal, the right-hand side b, and the warm start zeroed where masked. The tile
// is staged in LDS with three ghost rings: um, vm on the tile + 3, div on + 2, c and f on + 1.
// (The 64x4 tile's rings, for the matrix product's kernels below; the set-up's own tiles are
// 64 x SETUP_ROWS.)
constexpr int SX2 = TX + 4, SY2 = TY + 4;  // tile + 2
constexpr int SX1 = TX + 2, SY1 = TY + 2;  // tile + 1

__global__ void __launch_bounds__(TX * SETUP_ROWS) k_sgn_setup(SgnArgs a) {
  // a tile of 64 x SETUP_ROWS cells, a thread per cell, staged with three rings
  constexpr int ROWS = SETUP_ROWS, NTH = TX * ROWS;
  constexpr int SX3 = TX + 6, SY3 = ROWS + 6, SX2 = TX + 4, SY2 = ROWS + 4;
  constexpr int SX1 = TX + 2, SY1 = ROWS + 2;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i0 = GH + tx * TX, j0 = GH + ty * ROWS;  // padded corner of the tile
  __shared__ double s_h[SY3][SX3], s_um[SY3][SX3], s_vm[SY3][SX3];
  __shared__ double s_div[SY2][SX2];
  __shared__ double s_c[SY1][SX1], s_f[SY1][SX1];
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  // Every load of the kernel is issued before the first is used: the state on the tile + 3
  // (at most two points a thread), phi on the tile + 1 (at most two), the cell's rates.
  constexpr int N3 = SX3 * SY3, R3 = (N3 + NTH - 1) / NTH;
  constexpr int N1 = SX1 * SY1, R1 = (N1 + NTH - 1) / NTH;
  double st[R3][3];
#pragma unroll
  for (int q = 0; q < R3; ++q) {
    const int t = threadIdx.x + q * NTH;
    const int i = i0 - 3 + t % SX3, j = j0 - 3 + t / SX3;
    st[q][0] = st[q][1] = st[q][2] = 0.0;
    if (t < N3 && i < g.w && j < g.rows) {
      const int k = j * g.w + i;
      st[q][0] = a.h[k];
      st[q][1] = a.hu[k];
      st[q][2] = a.hv[k];
    }
  }
  double ph[R1];
#pragma unroll
  for (int q = 0; q < R1; ++q) {
    const int t = threadIdx.x + q * NTH;
    const int i = i0 - 1 + t % SX1, j = j0 - 1 + t / SX1;
    ph[q] = 0.0;
    if (t < N1 && i <= g.nx + GH && j <= g.ny + GH) ph[q] = a.phi[j * g.w + i];
  }
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int i = i0 + lx, j = j0 + ly;
  const bool own = i < g.nx + GH && j < g.ny + GH;
  const int k = j * g.w + i;
  double r_h = 0.0, r_hu = 0.0, r_hv = 0.0;
  if (own) {
    r_h = a.lh[k];
    r_hu = a.lhu[k];
    r_hv = a.lhv[k];
  }
#pragma unroll
  for (int q = 0; q < R3; ++q) {
    const int t = threadIdx.x + q * NTH;
    if (t < N3) {
      const double h = st[q][0];  // zero outside the array: then um = vm = 0
      const double m = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
      s_h[t / SX3][t % SX3] = h;
      s_um[t / SX3][t % SX3] = m * wc_velocity(h, st[q][1]);
      s_vm[t / SX3][t % SX3] = m * wc_velocity(h, st[q][2]);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < SX2 * SY2; t += NTH) {
    const int lx = t % SX2, ly = t / SX2;
    const int x = lx + 1, y = ly + 1;  // in the +3 frame
    s_div[ly][lx] = (s_um[y][x + 1] - s_um[y][x - 1]) * ix + (s_vm[y + 1][x] - s_vm[y - 1][x]) * iy;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < R1; ++q) {
    const int t = threadIdx.x + q * NTH;
    if (t >= N1) break;
    const int lx = t % SX1, ly = t / SX1;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    const int x3 = lx + 2, y3 = ly + 2;  // in the +3 frame
    const int x2 = lx + 1, y2 = ly + 1;  // in the +2 frame
    double coeff = 0.0, f = 0.0;
    if (i <= g.nx + GH && j <= g.ny + GH) {
      const double depth = s_h[y3][x3];
      coeff = depth < WC_H_DISPERSION_MIN ? 0.0 : ph[q] * depth * depth * depth / 3.0;
      if (a.nonlinear && coeff > 0.0) {
        const double n = s_um[y3][x3] * (s_div[y2][x2 + 1] - s_div[y2][x2 - 1]) * ix +
                         s_vm[y3][x3] * (s_div[y2 + 1][x2] - s_div[y2 - 1][x2]) * iy -
                         s_div[y2][x2] * s_div[y2][x2];
        f = coeff * n;
      }
      // c is read by the matrix product on the interior and its first ring.
      const bool inside = i >= GH && i < g.nx + GH && j >= GH && j < g.ny + GH;
      const bool ring1 = i >= GH - 1 && j >= GH - 1;
      const bool mine = (lx >= 1 && lx <= TX && ly >= 1 && ly <= ROWS);
      if ((mine && inside) || (!inside && ring1)) a.c[j * g.w + i] = coeff;
    }
    s_c[ly][lx] = coeff;
    s_f[ly][lx] = f;
  }
  __syncthreads();
  if (!own) return;
  const int x1 = lx + 1, y1 = ly + 1;
  const double h = s_h[ly + 3][lx + 3];
  const double heff = wc_max(h, WC_H_DRY);
  const bool masked = !(h >= WC_H_DISPERSION_MIN);
  double dgx, dgy, bx, by;
  if (a.ws.ring64) {  // the extrapolated start (the oracle's with_warm_start, bit for bit)
    const WarmStart& ws = a.ws;
    const unsigned long long n = ws.sc->solves;
    double2* e = ws.ring64 + (n % WS_SLOTS) * ws.stride;
    if (masked) {
      e[k] = make_double2(0.0, 0.0);
    } else {
      double x = a.wx[k], y = a.wy[k];
      e[k] = make_double2(x, y);
      for (int q = 1; q <= ws.terms; ++q) {
        const unsigned long long m = n - (unsigned long long)(q * ws.period);  // mod 2^64
        const double2 d = ws.ring64[(m % WS_SLOTS) * ws.stride + k];
        x += ws.coef[q - 1] * d.x;
      }
      for (int q = 1; q <= ws.terms; ++q) {
        const unsigned long long m = n - (unsigned long long)(q * ws.period);
        const double2 d = ws.ring64[(m % WS_SLOTS) * ws.stride + k];
        y += ws.coef[q - 1] * d.y;
      }
      a.wx[k] = x;
      a.wy[k] = y;
    }
  }
  if (masked) {
    dgx = 1.0;
    dgy = 1.0;
    bx = 0.0;
    by = 0.0;
    a.wx[k] = 0.0;
    a.wy[k] = 0.0;
  } else {
    dgx = heff + (s_c[y1][x1 + 1] + s_c[y1][x1 - 1]) * 0.25 / (g.dx * g.dx);
    dgy = heff + (s_c[y1 + 1][x1] + s_c[y1 - 1][x1]) * 0.25 / (g.dy * g.dy);
    // the cell's velocities as staged (the mask is 1 here, and 1 u is u)
    const double u = s_um[ly + 3][lx + 3];
    const double v = s_vm[ly + 3][lx + 3];
    bx = r_hu - u * r_h + (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix;
    by = r_hv - v * r_h + (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy;
  }
  if (a.dg2) {
    a.dg2[k] = make_double2(dgx, dgy);
  } else {
    a.dgx[k] = dgx;
    a.dgy[k] = dgy;
  }
  a.bx[k] = bx;
  a.by[k] = by;
}

// A p on the tile (matvec, dispersion.rs:274-317): p on the tile + 2 (ghosts by the vector
// ghost map, as fill_vector_ghosts leaves them), c div p on the tile + 1, then
// A p = mask (heff p - grad(c div p)). Returns this thread's (Apx, Apy, px, py).
struct ApOut {
  double ax, ay, px, py;
  bool valid;
};

__device__ __forceinline__ ApOut tile_matvec(const SgnArgs& a, const double* px, const double* py,
                                             int i0, int j0, double (*s_px)[SX2],
                                             double (*s_py)[SX2], double (*s_f)[SX1]) {
  const Geo& g = a.g;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < SX2 * SY2; t += NT) {
    const int lx = t % SX2, ly = t / SX2;
    const int i = i0 - 2 + lx, j = j0 - 2 + ly;
    double vx = 0.0, vy = 0.0;
    if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
      vx = ghosted_vx(px, g, i, j);
      vy = ghosted_vy(py, g, i, j);
    }
    s_px[ly][lx] = vx;
    s_py[ly][lx] = vy;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < SX1 * SY1; t += NT) {
    const int lx = t % SX1, ly = t / SX1;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    double f = 0.0;
    if (i <= g.nx + GH && j <= g.ny + GH) {
      const int x = lx + 1, y = ly + 1;
      f = a.c[j * g.w + i] *
          ((s_px[y][x + 1] - s_px[y][x - 1]) * ix + (s_py[y + 1][x] - s_py[y - 1][x]) * iy);
    }
    s_f[ly][lx] = f;
  }
  __syncthreads();
  ApOut o;
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int i = i0 + lx, j = j0 + ly;
  o.valid = i < g.nx + GH && j < g.ny + GH;
  o.ax = o.ay = o.px = o.py = 0.0;
  if (o.valid) {
    const int k = j * g.w + i;
    const double h = a.h[k];
    const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
    const double heff = wc_max(h, WC_H_DRY);
    const int x1 = lx + 1, y1 = ly + 1;
    o.px = s_px[ly + 2][lx + 2];
    o.py = s_py[ly + 2][lx + 2];
    o.ax = mask * (heff * o.px - (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix);
    o.ay = mask * (heff * o.py - (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy);
  }
  return o;
}

// PCG start (dispersion.rs:493-514): r = b - A w, z = r / diag, p = z, and the norms.
__global__ void __launch_bounds__(NT) k_cg_init(SgnArgs a) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i0 = GH + tx * TX, j0 = GH + ty * TY;
  __shared__ double s_px[SY2][SX2], s_py[SY2][SX2], s_f[SY1][SX1];
  const ApOut o = tile_matvec(a, a.wx, a.wy, i0, j0, s_px, s_py, s_f);
  double v[3] = {0.0, 0.0, 0.0};
  if (o.valid) {
    const int k = (j0 + (threadIdx.x >> 6)) * g.w + i0 + (threadIdx.x & 63);
    const double bx = a.bx[k], by = a.by[k];
    const double rx = bx - o.ax, ry = by - o.ay;
    const double zx = rx / a.dgx[k], zy = ry / a.dgy[k];
    a.rx[k] = rx;
    a.ry[k] = ry;
    a.px[k] = zx;
    a.py[k] = zy;
    v[0] = bx * bx + by * by;
    v[1] = rx * zx + ry * zy;
    v[2] = rx * rx + ry * ry;
  }
  store_tile<3>(v, a.part, gridDim.x, tile);
}

// One PCG iteration, part 1 (dispersion.rs:516-518): Ap, p.Ap and alpha = rz / p.Ap.
__global__ void __launch_bounds__(NT) k_cg_matvec(SgnArgs a) {
  if (a.sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i0 = GH + tx * TX, j0 = GH + ty * TY;
  __shared__ double s_px[SY2][SX2], s_py[SY2][SX2], s_f[SY1][SX1];
  const ApOut o = tile_matvec(a, a.px, a.py, i0, j0, s_px, s_py, s_f);
  double v[1] = {0.0};
  if (o.valid) {
    const int k = (j0 + (threadIdx.x >> 6)) * g.w + i0 + (threadIdx.x & 63);
    a.ax[k] = o.ax;
    a.ay[k] = o.ay;
    v[0] = o.px * o.ax + o.py * o.ay;
  }
  store_tile<1>(v, a.part, gridDim.x, tile);
}

// Part 2 (dispersion.rs:519-526): w += alpha p, r -= alpha Ap, z = r / diag, the convergence
// test on |r| and beta = rz' / rz.
__global__ void __launch_bounds__(NT) k_cg_update(SgnArgs a) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  const double alpha = sc->alpha;
  const double na = -alpha;
  double v[2] = {0.0, 0.0};
  if (i < g.nx + GH && j < g.ny + GH) {
    const int k = j * g.w + i;
    const double px = a.px[k], py = a.py[k];
    a.wx[k] = a.wx[k] + alpha * px;
    a.wy[k] = a.wy[k] + alpha * py;
    const double rx = a.rx[k] + na * a.ax[k];
    const double ry = a.ry[k] + na * a.ay[k];
    a.rx[k] = rx;
    a.ry[k] = ry;
    const double zx = rx / a.dgx[k], zy = ry / a.dgy[k];
    v[0] = rx * rx + ry * ry;
    v[1] = rx * zx + ry * zy;
  }
  store_tile<2>(v, a.part, gridDim.x, tile);
}

// Part 3 (dispersion.rs:527-539): p = z + beta p. Skipped once the solve is done: p is not
// read again.
__global__ void __launch_bounds__(NT) k_cg_pupdate(SgnArgs a) {
  if (a.sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  const int k = j * g.w + i;
  const double beta = a.sc->beta;
  a.px[k] = a.rx[k] / a.dgx[k] + beta * a.px[k];
  a.py[k] = a.ry[k] / a.dgy[k] + beta * a.py[k];
}

// ---- fused PCG iteration: two kernels, no stored diagonal ----------------------------------
// The Jacobi diagonal (dispersion.rs:442-457) at an interior cell, from c and h as the setup
// computes it.
__device__ __forceinline__ void diag_at(const SgnArgs& a, int k, double h, double& dgx,
                                        double& dgy) {
  if (!(h >= WC_H_DISPERSION_MIN)) {
    dgx = 1.0;
    dgy = 1.0;
    return;
  }
  const double heff = wc_max(h, WC_H_DRY);
  const int w = a.g.w;
  dgx = heff + (a.c[k + 1] + a.c[k - 1]) * 0.25 / (a.g.dx * a.g.dx);
  dgy = heff + (a.c[k + w] + a.c[k - w]) * 0.25 / (a.g.dy * a.g.dy);
}

// Part 1 (dispersion.rs:527-539 of the previous iteration, then 516-518): p = z + beta p on the
// tile + 2 (the first iteration takes p as cg_init stored it), A p, p.Ap and alpha. The new p of
// the tile's own cells goes to pdst for part 2 and the next iteration.
__global__ void __launch_bounds__(NT) k_cg_matvec2(SgnArgs a, int first, const double* psx,
                                                   const double* psy, double* pdx, double* pdy) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i0 = GH + tx * TX, j0 = GH + ty * TY;
  __shared__ double s_px[SY2][SX2], s_py[SY2][SX2], s_f[SY1][SX1];
  const double beta = sc->beta;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < SX2 * SY2; t += NT) {
    const int lx = t % SX2, ly = t / SX2;
    const int i = i0 - 2 + lx, j = j0 - 2 + ly;
    double vx = 0.0, vy = 0.0;
    if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
      const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
      if (first) {
        vx = psx[m];
        vy = psy[m];
      } else {
        vx = a.zx[m] + beta * psx[m];
        vy = a.zy[m] + beta * psy[m];
      }
      if (is_ghost_x(i, g.nx)) vx = -vx;
      if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
    }
    s_px[ly][lx] = vx;
    s_py[ly][lx] = vy;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < SX1 * SY1; t += NT) {
    const int lx = t % SX1, ly = t / SX1;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    double f = 0.0;
    if (i <= g.nx + GH && j <= g.ny + GH) {
      const int x = lx + 1, y = ly + 1;
      f = a.c[j * g.w + i] *
          ((s_px[y][x + 1] - s_px[y][x - 1]) * ix + (s_py[y + 1][x] - s_py[y - 1][x]) * iy);
    }
    s_f[ly][lx] = f;
  }
  __syncthreads();
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int i = i0 + lx, j = j0 + ly;
  double v[1] = {0.0};
  if (i < g.nx + GH && j < g.ny + GH) {
    const int k = j * g.w + i;
    const double h = a.h[k];
    const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
    const double heff = wc_max(h, WC_H_DRY);
    const int x1 = lx + 1, y1 = ly + 1;
    const double px = s_px[ly + 2][lx + 2], py = s_py[ly + 2][lx + 2];
    const double ax = mask * (heff * px - (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix);
    const double ay = mask * (heff * py - (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy);
    if (!first) {
      pdx[k] = px;
      pdy[k] = py;
    }
    a.ax[k] = ax;
    a.ay[k] = ay;
    v[0] = px * ax + py * ay;
  }
  store_tile<1>(v, a.part, gridDim.x, tile);
}

// Part 2 (dispersion.rs:519-526): w += alpha p, r -= alpha Ap, z = r / diag (stored for the
// next iteration's p), |r|, r.z, the convergence test and beta.
__global__ void __launch_bounds__(NT) k_cg_update2(SgnArgs a, const double* pcx,
                                                   const double* pcy) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  const double alpha = sc->alpha;
  const double na = -alpha;
  double v[2] = {0.0, 0.0};
  if (i < g.nx + GH && j < g.ny + GH) {
    const int k = j * g.w + i;
    a.wx[k] = a.wx[k] + alpha * pcx[k];
    a.wy[k] = a.wy[k] + alpha * pcy[k];
    const double rx = a.rx[k] + na * a.ax[k];
    const double ry = a.ry[k] + na * a.ay[k];
    a.rx[k] = rx;
    a.ry[k] = ry;
    double dgx, dgy;
    diag_at(a, k, a.h[k], dgx, dgy);
    const double zx = rx / dgx, zy = ry / dgy;
    a.zx[k] = zx;
    a.zy[k] = zy;
    v[0] = rx * rx + ry * ry;
    v[1] = rx * zx + ry * zy;
  }
  store_tile<2>(v, a.part, gridDim.x, tile);
}

// The second stage of the PCG reductions, one block, and the scalar recurrences they feed:
//   REDUCE_INIT    b.b, r.z, r.r -> target = tol |b| + 1e-10, rz, converged (dispersion.rs:510-514)
//   REDUCE_ALPHA   p.Ap -> alpha = rz / p.Ap (dispersion.rs:517-518)
//   REDUCE_BETA    r.r, r.z -> it + 1, converged, beta = rz' / rz (dispersion.rs:522-526)
__global__ void __launch_bounds__(NT) k_cg_reduce(SgnArgs a, int kind) {
  CgScalars* sc = a.sc;
  const int np = a.g.ntx * a.g.nty;
  if (kind == REDUCE_INIT) {
    double s[3];
    final_sums<3>(a.part, np, s);
    if (threadIdx.x == 0) {
      const double bnorm = sqrt(s[0]);
      sc->target = a.tolerance * bnorm + 1e-10;
      sc->rz = s[1];
      sc->it = 0;
      sc->converged = sqrt(s[2]) <= sc->target ? 1 : 0;
      sc->done = (sc->converged || a.max_iterations == 0) ? 1 : 0;
    }
    return;
  }
  if (sc->done) return;
  if (kind == REDUCE_ALPHA) {
    double s[1];
    final_sums<1>(a.part, np, s);
    if (threadIdx.x == 0) sc->alpha = sc->rz / s[0];
    return;
  }
  double s[2];
  final_sums<2>(a.part, np, s);
  if (threadIdx.x == 0) {
    sc->it += 1;
    sc->converged = sqrt(s[0]) <= sc->target ? 1 : 0;
    sc->beta = s[1] / sc->rz;
    sc->rz = s[1];
    sc->done = (sc->converged || sc->it >= a.max_iterations) ? 1 : 0;
  }
}

// The end of a solve: the counters of DispersionStats (dispersion.rs:541-545).
__global__ void k_cg_stats(CgScalars* sc, unsigned long long earlier) {
  if (threadIdx.x == 0) {
    sc->solves += 1;
    sc->iterations += earlier + sc->it;
    if (!sc->converged) sc->unconverged += 1;
  }
}

// d(hu)/dt = h w + u dh/dt where the cell takes part (dispersion.rs:547-562), then the stage
// update of k_flux.
__global__ void __launch_bounds__(NT) k_sgn_finish(StepArgs a, const double* wx, const double* wy,
                                                   int mode, WarmStart ws) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  const int k = j * g.w + i;
  if (ws.ring64) {  // the solve just counted: its slot from x_{n-1} to d_n = x_n - x_{n-1}
    double2* e = ws.ring64 + ((ws.sc->solves - 1) % WS_SLOTS) * ws.stride;
    const double2 x1 = e[k];
    e[k] = make_double2(wx[k] - x1.x, wy[k] - x1.y);
  }
  const double h = a.in_h[k];
  const double lh = a.lh[k];
  double lhu = a.lhu[k], lhv = a.lhv[k];
  if ((h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0) > 0.0) {
    lhu = h * wx[k] + wc_velocity(h, a.in_hu[k]) * lh;
    lhv = h * wy[k] + wc_velocity(h, a.in_hv[k]) * lh;
  }
  stage_out(a, mode, i, j, k, lh, lhu, lhv);
}

// ---- reductions over the state ---------------------------------------------------------------
// stable_dt (solver.rs:434-450): the largest rate, NaN if any is NaN. Rates are +0, positive,
// +inf or NaN, so their bit patterns order like their values and a NaN of either sign is above
// every number: an unsigned max of the bits is max_or_nan.
__global__ void __launch_bounds__(NT) k_max_rate(Geo g, const double* h, const double* hu,
                                                 const double* hv, unsigned long long* out) {
  const int n = g.nx * g.ny;
  unsigned long long bits = 0ull;
  for (int t = blockIdx.x * RATE_CELLS + threadIdx.x; t < min(n, (int)(blockIdx.x + 1) * RATE_CELLS);
       t += NT) {
    const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
    const double hc = h[k];
    const double c = sqrt(WC_G * hc);
    const double u = fabs(wc_velocity(hc, hu[k]));
    const double v = fabs(wc_velocity(hc, hv[k]));
    const double r = (u + c) / g.dx + (v + c) / g.dy;
    const unsigned long long b = (unsigned long long)__double_as_longlong(r);
    bits = b > bits ? b : bits;
  }
  block_max_to(bits, out);
}

// limiting_cell (solver.rs:455-477), pass 1: the first NaN cell (row-major), and the bits of
// the largest rate.
__global__ void __launch_bounds__(NT) k_limit_scan(Geo g, const double* h, const double* hu,
                                                   const double* hv, unsigned long long* maxbits,
                                                   unsigned int* first_nan) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.nx * g.ny) return;
  const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
  const double hc = h[k];
  const double c = sqrt(WC_G * hc);
  const double u = fabs(wc_velocity(hc, hu[k]));
  const double v = fabs(wc_velocity(hc, hv[k]));
  const double r = (u + c) / g.dx + (v + c) / g.dy;
  if (isnan(r)) {
    atomicMin(first_nan, (unsigned int)t);
  } else if (r > 0.0) {
    atomicMax(maxbits, (unsigned long long)__double_as_longlong(r));
  }
}

// Pass 2: the first cell whose rate equals the largest.
__global__ void __launch_bounds__(NT) k_limit_first(Geo g, const double* h, const double* hu,
                                                    const double* hv,
                                                    const unsigned long long* maxbits,
                                                    unsigned int* first) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.nx * g.ny) return;
  const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
  const double hc = h[k];
  const double c = sqrt(WC_G * hc);
  const double u = fabs(wc_velocity(hc, hu[k]));
  const double v = fabs(wc_velocity(hc, hv[k]));
  const double r = (u + c) / g.dx + (v + c) / g.dy;
  if ((unsigned long long)__double_as_longlong(r) == *maxbits) atomicMin(first, (unsigned int)t);
}

// State::volume in the canonical order (k_sum_h then its last block), State::max_abs_momentum
// and State::min_depth (NaN cells drop out of both, as in the reference's folds).
__global__ void __launch_bounds__(NT) k_state_sums(Geo g, const double* h, const double* hu,
                                                   const double* hv, double* part,
                                                   unsigned long long* maxmom,
                                                   unsigned long long* minh_key) {
  const int tile = blockIdx.x;
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i = GH + tx * TX + (threadIdx.x & 63);
  const int j = GH + ty * TY + (threadIdx.x >> 6);
  double v[1] = {0.0};
  if (i < g.nx + GH && j < g.ny + GH) {
    const int k = j * g.w + i;
    const double hc = h[k];
    v[0] = hc;
    const double m = wc_max(fabs(hu[k]), fabs(hv[k]));
    if (!isnan(m)) atomicMax(maxmom, (unsigned long long)__double_as_longlong(m));
    if (!isnan(hc)) {
      // order-preserving key of a double
      unsigned long long b = (unsigned long long)__double_as_longlong(hc);
      b = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
      atomicMin(minh_key, b);
    }
  }
  store_tile<1>(v, part, gridDim.x, tile);
}

// The second stage of the canonical sum of k_state_sums.
__global__ void __launch_bounds__(NT) k_sum_tiles(const double* part, int np, double* out) {
  double s[1];
  final_sums<1>(part, np, s);
  if (threadIdx.x == 0) *out = s[0];
}

// x / y for n values, both through the shared reciprocal of the flux kernels (div_q) and by the
// device's own division, for a host-side check that the two agree bit for bit.
__global__ void __launch_bounds__(NT) k_division_check(double y, const double* x, int n,
                                                       double* shared, double* plain) {
  const Quot q = quot_of(y);
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  shared[i] = div_q(x[i], q);
  plain[i] = x[i] / y;
}

}  // namespace wc
