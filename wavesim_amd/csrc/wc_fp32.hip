// The FP32 chosen form of wavecore's step: FP32 bulk storage and reconstruction, FP64 where
// accuracy lives.
//
//   - The state is an unevaluated pair of FP32 arrays per field, value = hi + lo (about 48
//     significant bits); each step adds its update to the pair with one FP64 add and an exact
//     split, a compensated accumulation, so 6000 steps of small increments do not lose the
//     low bits FP32 alone would.
//   - The RK2 average (solver.rs:493-500) is taken as s + dt (L1 + L2) / 2 on the pair where no
//     cell was cleaned, so the FP32-rounded stage state only ever feeds the second evaluation of
//     the rates, never the result.
//   - Rates, the switch, c, the SGN right-hand side and every PCG vector are stored in FP32;
//     inner products, residual norms, alpha, beta and the tolerance test are FP64, as are the
//     surface-minus-bed differences of the hydrostatic reconstruction (wc_flux.hip).
//
// Results differ from the FP64 path by rounding; tests/test_gpu_fp32.py measures by how much.

#include "wc_kernels.h"
#include "wc_stage.h"

namespace wc {

namespace {

constexpr int SX2f = TX + 4, SY2f = TY + 4;
constexpr int SX1f = TX + 2, SY1f = TY + 2;

__device__ __forceinline__ void split(double v, float& hi, float& lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

__device__ __forceinline__ int tile_i(const Geo& g, int tile) { return GH + (tile % g.ntx) * TX; }
__device__ __forceinline__ int tile_j(const Geo& g, int tile) { return GH + (tile / g.ntx) * TY; }

}  // namespace

// FP64 state <-> FP32 pairs, on every padded cell.
__global__ void __launch_bounds__(NT) k_split32(long long n, const double* a, float* hi,
                                                float* lo) {
  const long long t = (long long)blockIdx.x * NT + threadIdx.x;
  if (t >= n) return;
  float h, l;
  split(a[t], h, l);
  hi[t] = h;
  lo[t] = l;
}

__global__ void __launch_bounds__(NT) k_join32(long long n, const float* hi, const float* lo,
                                               double* a) {
  const long long t = (long long)blockIdx.x * NT + threadIdx.x;
  if (t >= n) return;
  a[t] = (double)hi[t] + (double)lo[t];
}

// Ghost layers of the FP32 state (both parts of h; the high parts of hu, hv, which is all the
// rates read of them).
__device__ __forceinline__ void fill32_body(const Geo& g, float* h, float* hl, float* hu,
                                            float* hv, float* hul, float* hvl, int blk) {
  const int n = 6 * g.w + 6 * g.ny;
  const int t = blk * NT + threadIdx.x;
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
  const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
  h[k] = h[m];
  if (hl) hl[k] = hl[m];
  const bool fx = is_ghost_x(i, g.nx), fy = !g.periodic && is_ghost_y(j, g.ny);
  hu[k] = fx ? -hu[m] : hu[m];
  hv[k] = fy ? -hv[m] : hv[m];
  if (hul) {  // the momentum pairs' low parts, flipped with them
    hul[k] = fx ? -hul[m] : hul[m];
    hvl[k] = fy ? -hvl[m] : hvl[m];
  }
}

__global__ void __launch_bounds__(NT) k_fill32(Geo g, float* h, float* hl, float* hu, float* hv,
                                               float* hul, float* hvl) {
  fill32_body(g, h, hl, hu, hv, hul, hvl, blockIdx.x);
}

// The switch (dispersion.rs:120-160) in FP64 from the FP32 pair, stored in FP32.
__device__ __forceinline__ void switch32_body(const Geo& g, const float* h, const float* hl,
                                              const double* b, const Breaking& br,
                                              const uint8_t* held, float* phi, int blk) {
  const int t = blk * NT + threadIdx.x;
  if (t >= g.w * g.rows) return;
  const int i = t % g.w, j = t / g.w;
  const int mi = map_x(i, g.nx), mj = map_y(j, g.ny, g.periodic);
  double v = 0.0;
  if (!(held && held[(mj - GH) * g.nx + (mi - GH)])) {
    const int k = mj * g.w + mi;
    auto depth = [&](int q) { return (double)h[q] + (hl ? (double)hl[q] : 0.0); };
    const double hc = depth(k);
    if (hc >= WC_H_DISPERSION_MIN) {
      const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
      const double ex = ((depth(k + 1) + b[k + 1]) - (depth(k - 1) + b[k - 1])) * ix;
      const double ey = ((depth(k + g.w) + b[k + g.w]) - (depth(k - g.w) + b[k - g.w])) * iy;
      const double ratio = wc_max((hc + b[k]) / hc, 0.0);
      double off = switch_off(br, ratio, ex, ey);
      if (br.memory && br.memory[k] != 0.0) off = 1.0;
      v = 1.0 - off;
    }
  }
  phi[t] = (float)v;
}

__global__ void __launch_bounds__(NT) k_switch32(Geo g, const float* h, const float* hl,
                                                 const double* b, Breaking br,
                                                 const uint8_t* held, float* phi) {
  switch32_body(g, h, hl, b, br, held, phi, blockIdx.x);
}

// The start of solve n = sc->solves at interior cell k (Sgn32, WarmStart): the FP32 correction
// w = 0; masked cells x0 = 0 (dispersion.rs:488-490); wet cells x0 = the previous solution, plus
// the extrapolated step when the warm start is on (the step also into slot n).
__device__ __forceinline__ void solve_start(const Sgn32& a, int k, bool wet) {
  a.wx[k] = 0.f;
  a.wy[k] = 0.f;
  const WarmStart& ws = a.ws;
  float2* e = ws.ring ? ws.ring + (ws.sc->solves % WS_SLOTS) * ws.stride : nullptr;
  if (!wet) {
    a.solx[k] = 0.0;
    a.soly[k] = 0.0;
    if (e) e[k] = make_float2(0.f, 0.f);
    return;
  }
  if (!e) return;
  const unsigned long long n = ws.sc->solves;
  double px = 0.0, py = 0.0;
  const int terms = a.phi[k] < ws.guard ? 0 : ws.terms;
  for (int j = 1; j <= terms; ++j) {
    const unsigned long long m = n - (unsigned long long)(j * ws.period);  // mod 2^64
    const float2 d = ws.ring[(m % WS_SLOTS) * ws.stride + k];
    px += ws.coef[j - 1] * (double)d.x;
    py += ws.coef[j - 1] * (double)d.y;
  }
  const float2 step = make_float2((float)px, (float)py);
  e[k] = step;
  a.solx[k] += (double)step.x;
  a.soly[k] += (double)step.y;
}

// After solve n - 1 = sc->solves - 1 (counted by now), at interior cell k: the solution x = x0
// + w into solx, soly (returned), and slot n - 1 from the start's step to d = x - x_{n-2}.
__device__ __forceinline__ void solve_end(const Stage32& a, int k, double& x, double& y) {
  const double wx = (double)a.wx[k], wy = (double)a.wy[k];
  x = a.solx[k] + wx;
  y = a.soly[k] + wy;
  a.solx[k] = x;
  a.soly[k] = y;
  const WarmStart& ws = a.ws;
  if (ws.ring) {
    float2* e = ws.ring + ((ws.sc->solves - 1) % WS_SLOTS) * ws.stride;
    const float2 st = e[k];
    e[k] = make_float2((float)((double)st.x + wx), (float)((double)st.y + wy));
  }
}

// SGN setup (dispersion.rs:385-490) on FP32 storage; arithmetic in FP64. A block takes a tile
// of 64 x ROWS cells, a thread per cell, and stages it with three rings (16 rows: 1.5 staged
// points per cell, against 2.7 on 64x4 tiles). The staging loads six floats per point and is
// bound by the rate the texture data unit returns them, so the taller tile's fewer points per
// cell are what counts.
template <int ROWS>
__device__ __forceinline__ void sgn_setup32_body(const Sgn32& a, int blk, int nblk) {
  constexpr int NTH = TX * ROWS;
  constexpr int SX3 = TX + 6, SY3 = ROWS + 6, SX2 = TX + 4, SY2 = ROWS + 4;
  constexpr int SX1 = TX + 2, SY1 = ROWS + 2;
  const Geo& g = a.g;
  const int tile = tile_of_block(blk, nblk);
  const int i0 = GH + (tile % g.ntx) * TX, j0 = GH + (tile / g.ntx) * ROWS;
  __shared__ double s_h[SY3][SX3], s_um[SY3][SX3], s_vm[SY3][SX3];
  __shared__ double s_div[SY2][SX2];
  __shared__ double s_c[SY1][SX1], s_f[SY1][SX1];
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < SX3 * SY3; t += NTH) {
    const int lx = t % SX3, ly = t / SX3;
    const int i = i0 - 3 + lx, j = j0 - 3 + ly;
    double h = 0.0, um = 0.0, vm = 0.0;
    if (i < g.w && j < g.rows) {
      const int k = j * g.w + i;
      h = (double)a.h[k] + (double)a.hl[k];  // the FP32 pairs
      const double m = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
      um = m * wc_velocity(h, (double)a.hu[k] + (double)a.hul[k]);
      vm = m * wc_velocity(h, (double)a.hv[k] + (double)a.hvl[k]);
    }
    s_h[ly][lx] = h;
    s_um[ly][lx] = um;
    s_vm[ly][lx] = vm;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < SX2 * SY2; t += NTH) {
    const int lx = t % SX2, ly = t / SX2;
    const int x = lx + 1, y = ly + 1;
    s_div[ly][lx] = (s_um[y][x + 1] - s_um[y][x - 1]) * ix + (s_vm[y + 1][x] - s_vm[y - 1][x]) * iy;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < SX1 * SY1; t += NTH) {
    const int lx = t % SX1, ly = t / SX1;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    const int x3 = lx + 2, y3 = ly + 2, x2 = lx + 1, y2 = ly + 1;
    double coeff = 0.0, f = 0.0;
    if (i <= g.nx + GH && j <= g.ny + GH) {
      const double depth = s_h[y3][x3];
      coeff = depth < WC_H_DISPERSION_MIN
                  ? 0.0
                  : (double)a.phi[j * g.w + i] * depth * depth * depth / 3.0;
      if (a.nonlinear && coeff > 0.0) {
        const double n = s_um[y3][x3] * (s_div[y2][x2 + 1] - s_div[y2][x2 - 1]) * ix +
                         s_vm[y3][x3] * (s_div[y2 + 1][x2] - s_div[y2 - 1][x2]) * iy -
                         s_div[y2][x2] * s_div[y2][x2];
        f = coeff * n;
      }
      const bool inside = i >= GH && i < g.nx + GH && j >= GH && j < g.ny + GH;
      const bool ring1 = i >= GH - 1 && j >= GH - 1;
      const bool mine = (lx >= 1 && lx <= TX && ly >= 1 && ly <= ROWS);
      if ((mine && inside) || (!inside && ring1)) a.c[j * g.w + i] = (float)coeff;
    }
    s_c[ly][lx] = (double)(float)coeff;  // the matrix product sees c as stored
    s_f[ly][lx] = f;
  }
  __syncthreads();
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int i = i0 + lx, j = j0 + ly;
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  const int k = j * g.w + i;
  const double h = s_h[ly + 3][lx + 3];
  // which cells take part is read from the depth's high part, as the PCG kernels read it, so
  // that a cell the solve masks always starts (and stays) at zero with a zero right-hand side
  if (!((double)a.h[k] >= WC_H_DISPERSION_MIN)) {
    a.bx[k] = 0.0;
    a.by[k] = 0.0;
    solve_start(a, k, false);
    return;
  }
  solve_start(a, k, true);
  const int x1 = lx + 1, y1 = ly + 1;
  const double u = wc_velocity(h, (double)a.hu[k] + (double)a.hul[k]);
  const double v = wc_velocity(h, (double)a.hv[k] + (double)a.hvl[k]);
  const double lh = a.lh[k];
  a.bx[k] = a.lhu[k] - u * lh + (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix;
  a.by[k] = a.lhv[k] - v * lh + (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy;
}

__device__ __forceinline__ void diag32(const Sgn32& a, int k, double h, double& dgx,
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

// A p on the tile from p given on the tile + 2 in LDS (FP64 arithmetic on FP32 storage).
__device__ __forceinline__ void matvec32_tail(const Sgn32& a, int i0, int j0,
                                              double (*s_px)[SX2f], double (*s_py)[SX2f],
                                              double (*s_f)[SX1f], double& ax, double& ay,
                                              double& px, double& py, bool& valid, int& k) {
  const Geo& g = a.g;
  const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
  for (int t = threadIdx.x; t < SX1f * SY1f; t += NT) {
    const int lx = t % SX1f, ly = t / SX1f;
    const int i = i0 - 1 + lx, j = j0 - 1 + ly;
    double f = 0.0;
    if (i <= g.nx + GH && j <= g.ny + GH) {
      const int x = lx + 1, y = ly + 1;
      f = (double)a.c[j * g.w + i] *
          ((s_px[y][x + 1] - s_px[y][x - 1]) * ix + (s_py[y + 1][x] - s_py[y - 1][x]) * iy);
    }
    s_f[ly][lx] = f;
  }
  __syncthreads();
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int i = i0 + lx, j = j0 + ly;
  valid = i < g.nx + GH && j < g.ny + GH;
  ax = ay = px = py = 0.0;
  k = j * g.w + i;
  if (!valid) return;
  const double h = (double)a.h[k];
  const double mask = h >= WC_H_DISPERSION_MIN ? 1.0 : 0.0;
  const double heff = wc_max(h, WC_H_DRY);
  const int x1 = lx + 1, y1 = ly + 1;
  px = s_px[ly + 2][lx + 2];
  py = s_py[ly + 2][lx + 2];
  ax = mask * (heff * px - (s_f[y1][x1 + 1] - s_f[y1][x1 - 1]) * ix);
  ay = mask * (heff * py - (s_f[y1 + 1][x1] - s_f[y1 - 1][x1]) * iy);
}

__global__ void __launch_bounds__(TX * SETUP32_ROWS) k_sgn_setup32(Sgn32 a) {
  sgn_setup32_body<SETUP32_ROWS>(a, blockIdx.x, gridDim.x);
}

// r = b - A w, p = z = r / diag, and b.b, r.z, r.r (dispersion.rs:493-514).
__global__ void __launch_bounds__(NT) k_cg_init32(Sgn32 a) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int i0 = tile_i(g, tile), j0 = tile_j(g, tile);
  __shared__ double s_px[SY2f][SX2f], s_py[SY2f][SX2f], s_f[SY1f][SX1f];
  for (int t = threadIdx.x; t < SX2f * SY2f; t += NT) {
    const int lx = t % SX2f, ly = t / SX2f;
    const int i = i0 - 2 + lx, j = j0 - 2 + ly;
    double vx = 0.0, vy = 0.0;
    if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
      const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
      vx = a.solx[m] + (double)a.wx[m];
      vy = a.soly[m] + (double)a.wy[m];
      if (is_ghost_x(i, g.nx)) vx = -vx;
      if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
    }
    s_px[ly][lx] = vx;
    s_py[ly][lx] = vy;
  }
  __syncthreads();
  double ax, ay, px, py;
  bool valid;
  int k;
  matvec32_tail(a, i0, j0, s_px, s_py, s_f, ax, ay, px, py, valid, k);
  double v[3] = {0.0, 0.0, 0.0};
  if (valid) {
    const double bx = (double)a.bx[k], by = (double)a.by[k];
    const double rx = bx - ax, ry = by - ay;
    double dgx, dgy;
    diag32(a, k, (double)a.h[k], dgx, dgy);
    const double zx = rx / dgx, zy = ry / dgy;
    a.rx[k] = (float)rx;
    a.ry[k] = (float)ry;
    a.px[k] = (float)zx;
    a.py[k] = (float)zy;
    v[0] = bx * bx + by * by;
    v[1] = rx * zx + ry * zy;
    v[2] = rx * rx + ry * ry;
  }
  store_tile<3>(v, a.part, gridDim.x, tile);
}

// p = z + beta p_old on the tile + 2, A p, p.Ap.
__global__ void __launch_bounds__(NT) k_cg_matvec32(Sgn32 a, int first, const float* psx,
                                                    const float* psy, float* pdx, float* pdy) {
  if (a.sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int i0 = tile_i(g, tile), j0 = tile_j(g, tile);
  __shared__ double s_px[SY2f][SX2f], s_py[SY2f][SX2f], s_f[SY1f][SX1f];
  const double beta = a.sc->beta;
  for (int t = threadIdx.x; t < SX2f * SY2f; t += NT) {
    const int lx = t % SX2f, ly = t / SX2f;
    const int i = i0 - 2 + lx, j = j0 - 2 + ly;
    double vx = 0.0, vy = 0.0;
    if (i <= g.nx + GH + 1 && j <= g.ny + GH + 1) {
      const int m = map_y(j, g.ny, g.periodic) * g.w + map_x(i, g.nx);
      if (first) {
        vx = (double)psx[m];
        vy = (double)psy[m];
      } else {
        vx = (double)(float)((double)a.zx[m] + beta * (double)psx[m]);
        vy = (double)(float)((double)a.zy[m] + beta * (double)psy[m]);
      }
      if (is_ghost_x(i, g.nx)) vx = -vx;
      if (!g.periodic && is_ghost_y(j, g.ny)) vy = -vy;
    }
    s_px[ly][lx] = vx;
    s_py[ly][lx] = vy;
  }
  __syncthreads();
  double ax, ay, px, py;
  bool valid;
  int k;
  matvec32_tail(a, i0, j0, s_px, s_py, s_f, ax, ay, px, py, valid, k);
  double v[1] = {0.0};
  if (valid) {
    if (!first) {
      pdx[k] = (float)px;
      pdy[k] = (float)py;
    }
    a.ax[k] = (float)ax;
    a.ay[k] = (float)ay;
    v[0] = px * ax + py * ay;
  }
  store_tile<1>(v, a.part, gridDim.x, tile);
}

// w += alpha p, r -= alpha Ap, z = r / diag, r.r, r.z.
__global__ void __launch_bounds__(NT) k_cg_update32(Sgn32 a, const float* pcx, const float* pcy) {
  CgScalars* sc = a.sc;
  if (sc->done) return;
  const Geo& g = a.g;
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int i = tile_i(g, tile) + (threadIdx.x & 63);
  const int j = tile_j(g, tile) + (threadIdx.x >> 6);
  const double alpha = sc->alpha;
  double v[2] = {0.0, 0.0};
  if (i < g.nx + GH && j < g.ny + GH) {
    const int k = j * g.w + i;
    a.wx[k] = (float)((double)a.wx[k] + alpha * (double)pcx[k]);
    a.wy[k] = (float)((double)a.wy[k] + alpha * (double)pcy[k]);
    const double rx = (double)(float)((double)a.rx[k] - alpha * (double)a.ax[k]);
    const double ry = (double)(float)((double)a.ry[k] - alpha * (double)a.ay[k]);
    a.rx[k] = (float)rx;
    a.ry[k] = (float)ry;
    double dgx, dgy;
    diag32(a, k, (double)a.h[k], dgx, dgy);
    const double zx = (double)(float)(rx / dgx), zy = (double)(float)(ry / dgy);
    a.zx[k] = (float)zx;
    a.zy[k] = (float)zy;
    v[0] = rx * rx + ry * ry;
    v[1] = rx * zx + ry * zy;
  }
  store_tile<2>(v, a.part, gridDim.x, tile);
}

// The stage update of the FP32 form, with the SGN correction when `dispersive`
// (dispersion.rs:547-562, solver.rs:481-616):
//   MODE_S1     s1 = clean(s + dt L); L1 = L
//   MODE_FINAL  s2 = clean(s1 + dt L); s = s + dt (L1 + L) / 2 where nothing was cleaned, else
//               clean((s + s2) / 2); then damp, on the FP32 pair
//   MODE_EULER  s = damp(clean(s + dt L))
__device__ __forceinline__ void stage32_body(const Stage32& a, double dt, int mode, int blk,
                                             int nblk) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blk, nblk);
  const int i = tile_i(g, tile) + (threadIdx.x & 63);
  const int j = tile_j(g, tile) + (threadIdx.x >> 6);
  if (i >= g.nx + GH || j >= g.ny + GH) return;
  const int k = j * g.w + i;
  // the stage input, as the rates saw it
  const double ih = mode == MODE_FINAL ? (double)a.s1h[k] + (double)a.s1hl[k]
                                       : (double)a.sh[k] + (double)a.sl_h[k];
  const double ihu = mode == MODE_FINAL ? (double)a.s1hu[k] + (double)a.s1hul[k]
                                        : (double)a.shu[k] + (double)a.sl_hu[k];
  const double ihv = mode == MODE_FINAL ? (double)a.s1hv[k] + (double)a.s1hvl[k]
                                        : (double)a.shv[k] + (double)a.sl_hv[k];
  const double lh = a.lh[k];
  double lhu = (double)a.lhu[k], lhv = (double)a.lhv[k];
  if (a.dispersive) {
    double wx, wy;
    solve_end(a, k, wx, wy);
    if (ih >= WC_H_DISPERSION_MIN) {
      lhu = ih * wx + wc_velocity(ih, ihu) * lh;
      lhv = ih * wy + wc_velocity(ih, ihv) * lh;
    }
  }
  // the state s as the pair
  const double sh = (double)a.sh[k] + (double)a.sl_h[k];
  const double shu = (double)a.shu[k] + (double)a.sl_hu[k];
  const double shv = (double)a.shv[k] + (double)a.sl_hv[k];
  if (mode == MODE_S1) {
    double h = sh + dt * lh, hu = shu + dt * lhu, hv = shv + dt * lhv;
    clean_dry(h, hu, hv);
    float hi, lo;
    split(h, hi, lo);
    a.s1h[k] = hi;
    a.s1hl[k] = lo;
    split(hu, hi, lo);
    a.s1hu[k] = hi;
    a.s1hul[k] = lo;
    split(hv, hi, lo);
    a.s1hv[k] = hi;
    a.s1hvl[k] = lo;
    a.l1h[k] = lh;
    a.l1hu[k] = (float)lhu;
    a.l1hv[k] = (float)lhv;
    return;
  }
  double h, hu, hv;
  if (mode == MODE_EULER) {
    h = sh + dt * lh;
    hu = shu + dt * lhu;
    hv = shv + dt * lhv;
    clean_dry(h, hu, hv);
  } else {
    const double s1h = (double)a.s1h[k] + (double)a.s1hl[k];
    const double s1hu = (double)a.s1hu[k] + (double)a.s1hul[k];
    const double s1hv = (double)a.s1hv[k] + (double)a.s1hvl[k];
    double h2 = s1h + dt * lh, hu2 = s1hu + dt * lhu, hv2 = s1hv + dt * lhv;
    const bool cleaned = !(sh + dt * a.l1h[k] >= WC_H_DRY) || !(h2 >= WC_H_DRY);
    clean_dry(h2, hu2, hv2);
    if (!cleaned) {
      h = sh + 0.5 * dt * (a.l1h[k] + lh);
      hu = shu + 0.5 * dt * ((double)a.l1hu[k] + lhu);
      hv = shv + 0.5 * dt * ((double)a.l1hv[k] + lhv);
    } else {
      h = 0.5 * (sh + h2);
      hu = 0.5 * (shu + hu2);
      hv = 0.5 * (shv + hv2);
    }
    clean_dry(h, hu, hv);
  }
  damp(a.damp, dt, i, j, a.b[k], h, hu, hv);
  float hi, lo;
  split(h, hi, lo);
  a.sh[k] = hi;
  a.sl_h[k] = lo;
  split(hu, hi, lo);
  a.shu[k] = hi;
  a.sl_hu[k] = lo;
  split(hv, hi, lo);
  a.shv[k] = hi;
  a.sl_hv[k] = lo;
}

__global__ void __launch_bounds__(NT) k_stage32(Stage32 a, int mode) {
  stage32_body(a, a.dt, mode, blockIdx.x, gridDim.x);
}

// The largest rate (stable_dt) of the FP32 pair state.
__device__ __forceinline__ void max_rate32_body(const Geo& g, const float* h, const float* hl,
                                                const float* hu, const float* hul,
                                                const float* hv, const float* hvl,
                                                unsigned long long* out, int blk) {
  const int n = g.nx * g.ny;
  unsigned long long bits = 0ull;
  for (int t = blk * RATE_CELLS + threadIdx.x; t < min(n, (blk + 1) * RATE_CELLS); t += NT) {
    const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
    const double hc = (double)h[k] + (double)hl[k];
    const double c = sqrt(WC_G * hc);
    const double u = fabs(wc_velocity(hc, (double)hu[k] + (double)hul[k]));
    const double v = fabs(wc_velocity(hc, (double)hv[k] + (double)hvl[k]));
    const double r = (u + c) / g.dx + (v + c) / g.dy;
    const unsigned long long b = (unsigned long long)__double_as_longlong(r);
    bits = b > bits ? b : bits;
  }
  block_max_to(bits, out);
}

__global__ void __launch_bounds__(NT) k_max_rate32(Geo g, const float* h, const float* hl,
                                                   const float* hu, const float* hul,
                                                   const float* hv, const float* hvl,
                                                   unsigned long long* out) {
  max_rate32_body(g, h, hl, hu, hul, hv, hvl, out, blockIdx.x);
}

// The second stage of the PCG reductions for the FP32 form (same recurrences as k_cg_reduce).
__global__ void __launch_bounds__(NT) k_cg_reduce32(Sgn32 a, int kind) {
  CgScalars* sc = a.sc;
  const int np = a.g.ntx * a.g.nty;
  if (kind == REDUCE_INIT) {
    double s[3];
    final_sums<3>(a.part, np, s);
    if (threadIdx.x == 0) {
      sc->target = a.tolerance * sqrt(s[0]) + 1e-10;
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

// ---- many instances in one launch ----------------------------------------------------------
// The grid walks every instance's virtual blocks, gridDim.x at a time; a virtual block runs the
// single-instance body with its instance's arguments, so each instance computes exactly what
// its own launch would.
__global__ void __launch_bounds__(NT) kb_fill32(const BFill* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    const BFill& a = args[q];
    fill32_body(a.g, a.h, a.hl, a.hu, a.hv, a.hul, a.hvl, vb - m.first[q]);
  }
}

__global__ void __launch_bounds__(NT) kb_switch32(const BSwitch* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    const BSwitch& a = args[q];
    switch32_body(a.g, a.h, a.hl, a.b, a.br, a.held, a.phi, vb - m.first[q]);
  }
}

// Virtual blocks: each instance's 64 x ROWS tiles, a block of 64 ROWS threads each (the batch
// tunes ROWS among 4, 8 and 16).
template <int ROWS>
__global__ void __launch_bounds__(TX * ROWS) kb_sgn_setup32(const Sgn32* __restrict__ args,
                                                            BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    sgn_setup32_body<ROWS>(args[q], vb - m.first[q], m.first[q + 1] - m.first[q]);
    __syncthreads();  // the next virtual block reuses the LDS
  }
}
template __global__ void kb_sgn_setup32<4>(const Sgn32*, BMap);
template __global__ void kb_sgn_setup32<8>(const Sgn32*, BMap);
template __global__ void kb_sgn_setup32<16>(const Sgn32*, BMap);

__global__ void __launch_bounds__(NT) kb_stage32(const Stage32* __restrict__ args, const double* dt, int mode,
                                                 BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    stage32_body(args[q], dt[q], mode, vb - m.first[q], m.first[q + 1] - m.first[q]);
  }
}

__global__ void __launch_bounds__(NT) kb_max_rate32(const BRate* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    const BRate& a = args[q];
    max_rate32_body(a.g, a.h, a.hl, a.hu, a.hul, a.hv, a.hvl, a.out, vb - m.first[q]);
  }
}

// Per-step host values of every instance, staged in one buffer: instance q's segment
// [off[q], off[q+1]) goes to dst[q].
__global__ void __launch_bounds__(NT) kb_scatter(const double* src, const long long* off,
                                                 double* const* dst, int n) {
  for (int q = 0; q < n; ++q) {
    const long long len = off[q + 1] - off[q];
    for (long long t = (long long)blockIdx.x * NT + threadIdx.x; t < len;
         t += (long long)gridDim.x * NT)
      dst[q][t] = src[off[q] + t];
  }
}

}  // namespace wc
