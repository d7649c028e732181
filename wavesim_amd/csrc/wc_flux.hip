// The finite-volume rates on a 64x4 tile staged in LDS, each cell reconstructed once and each
// face fluxed once (solver.rs:625-718 reconstructs every cell three times per axis and fluxes
// every face twice; both are pure functions of their inputs, so doing each once gives the same
// numbers).
//
// Template parameter S is the storage type of the state, R the type the reconstruction runs in
// (face_values on the surface and the velocities, solver.rs:125-138). The surface-minus-bed
// differences (face depths, the hydrostatic reconstruction, solver.rs:164-190), the HLL flux and
// the rates stay in FP64 for all: with S = R = double this is the reference's arithmetic bit for
// bit; with S = float (the FP32 form's state pairs) R = float is the FP32 chosen form, whose
// lake at rest stays at rest because eta - b is never formed in FP32, and R = double
// reconstructs the FP32 state in FP64 (same time on this kernel, which is bound by latency and
// LDS occupancy rather than by its bytes, and rates with FP64 rounding instead of FP32's).

#include "wc_kernels.h"
#include "wc_stage.h"

namespace wc {

namespace {

constexpr int BX = TX + 6, BY = TY + 6;  // the tile and three ghost rings
constexpr int RXN = TX + 2;              // x-reconstructed cells per row: the tile +- 1
constexpr int RYN = TY + 2;              // y-reconstructed rows: the tile +- 1

template <typename R>
struct SideT {
  double h, z;
  R un, ut;
};

template <typename R>
__device__ __forceinline__ R minmod3r(R a, R b, R c) {
  const R zero = R(0);
  if (a > zero && b > zero && c > zero) return (R)wc_min(wc_min((double)a, (double)b), (double)c);
  if (a < zero && b < zero && c < zero) return (R)wc_max(wc_max((double)a, (double)b), (double)c);
  return zero;
}

// x / 60 of the fifth-order formula: in FP64 through the shared reciprocal (the division's bits).
template <typename R>
__device__ __forceinline__ R div60(R x, const Quot& q60);
template <>
__device__ __forceinline__ float div60<float>(float x, const Quot&) {
  return x / 60.0f;
}
template <>
__device__ __forceinline__ double div60<double>(double x, const Quot& q60) {
  return div_q(x, q60);
}

// The divisors the rates share, their reciprocals computed once per wave.
struct FluxDivs {
  Quot q60, qdx, qdy;
};
__device__ __forceinline__ FluxDivs flux_divs(const Geo& g) {
  return {quot_of(60.0), quot_of(g.dx), quot_of(g.dy)};
}

// solver.rs:125-138 in R arithmetic.
template <typename R>
__device__ __forceinline__ void face_values_r(R low2, R low, R c, R high, R high2, R blend,
                                              const Quot& q60, R& lo, R& hi) {
  const R s = minmod3r<R>(R(2) * (c - low), R(0.5) * ((c - low) + (high - c)), R(2) * (high - c));
  const R mc_lo = c - R(0.5) * s, mc_hi = c + R(0.5) * s;
  if (blend <= R(0)) {
    lo = mc_lo;
    hi = mc_hi;
    return;
  }
  const R fifth_hi =
      div60<R>(R(2) * low2 - R(13) * low + R(47) * c + R(27) * high - R(3) * high2, q60);
  const R fifth_lo =
      div60<R>(R(2) * high2 - R(13) * high + R(47) * c + R(27) * low - R(3) * low2, q60);
  lo = mc_lo + blend * (fifth_lo - mc_lo);
  hi = mc_hi + blend * (fifth_hi - mc_hi);
}

// solver.rs:146-183 for the cell in the middle of five consecutive cells along a line: e[]
// (surface), d[] (depth, FP64), n[] (normal velocity), t[] (tangential), b[] (bed, FP64), p[]
// (phi).
template <typename R>
__device__ __forceinline__ void reconstruct5(const R (&e)[5], const double (&d)[5],
                                             const R (&n)[5], const R (&t)[5],
                                             const double (&b)[5], const R (&p)[5], int order,
                                             const Quot& q60, SideT<R>& lo_out,
                                             SideT<R>& hi_out) {
  const double h1 = d[1], h2 = d[2], h3 = d[3];
  const SideT<R> first{h2, b[2], n[2], t[2]};
  if (order == 0 || h1 < WC_H_SECOND || h2 < WC_H_SECOND || h3 < WC_H_SECOND) {
    lo_out = first;
    hi_out = first;
    return;
  }
  R blend = R(0);
  if (order == 2 && d[0] >= WC_H_SECOND && d[4] >= WC_H_SECOND) {
    double bl = 1.0;
#pragma unroll
    for (int m = 0; m < 5; ++m) bl = wc_min(bl, (double)p[m]);
    blend = (R)bl;
  }
  R eta_lo, eta_hi, un_lo, un_hi, ut_lo, ut_hi;
  face_values_r<R>(e[0], e[1], e[2], e[3], e[4], blend, q60, eta_lo, eta_hi);
  face_values_r<R>(n[0], n[1], n[2], n[3], n[4], blend, q60, un_lo, un_hi);
  face_values_r<R>(t[0], t[1], t[2], t[3], t[4], blend, q60, ut_lo, ut_hi);
  const double cb = b[2];
  const double s_z = 0.25 * (b[3] - b[1]);
  SideT<R> l, r;
  l.z = cb + -1.0 * s_z;
  l.h = (double)eta_lo - l.z;
  l.un = un_lo;
  l.ut = ut_lo;
  r.z = cb + 1.0 * s_z;
  r.h = (double)eta_hi - r.z;
  r.un = un_hi;
  r.ut = ut_hi;
  if (l.h < WC_H_FACE_MIN || r.h < WC_H_FACE_MIN) {
    lo_out = first;
    hi_out = first;
    return;
  }
  lo_out = l;
  hi_out = r;
}

// The same from a staged line: five values `stride` apart.
template <typename R>
__device__ __forceinline__ void reconstruct_r(const R* e, const double* d, const R* n, const R* t,
                                              const double* b, const R* p, int stride, int order,
                                              const Quot& q60, SideT<R>& lo_out,
                                              SideT<R>& hi_out) {
  R e5[5], n5[5], t5[5], p5[5];
  double d5[5], b5[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    e5[m] = e[m * stride];
    d5[m] = d[m * stride];
    n5[m] = n[m * stride];
    t5[m] = t[m * stride];
    b5[m] = b[m * stride];
    p5[m] = p[m * stride];
  }
  reconstruct5<R>(e5, d5, n5, t5, b5, p5, order, q60, lo_out, hi_out);
}

template <typename R>
__device__ __forceinline__ Face face_r(const SideT<R>& a, const SideT<R>& b) {
  const Side l{a.h, (double)a.un, (double)a.ut, a.z};
  const Side r{b.h, (double)b.un, (double)b.ut, b.z};
  return face(l, r);
}

template <typename R>
struct TileLds {
  union {
    struct {
      R eta[BY][BX], un_x[BY][BX], un_y[BY][BX], phi[BY][BX];
      double h[BY][BX], b[BY][BX];
    } box;
    struct {
      Face fx[TY][TX + 1];
      Face fy[TY + 1][TX];
    } faces;
  } u;
  SideT<R> xlo[TY][RXN], xhi[TY][RXN];
  SideT<R> ylo[RYN][TX], yhi[RYN][TX];
};

}  // namespace

// Rates of every tile cell, then the stage epilogue of `mode` (see stage_out).
// in_lo: the low part of the depth in the FP32 chosen form (null in FP64).
template <typename S, typename R>
__device__ __forceinline__ void flux_tile(const FluxTileArgs<S>& a, int mode, int blk, int nblk,
                                          double& rh, double& rhu, double& rhv, int& i_out,
                                          int& j_out, bool& valid) {
  const Geo& g = a.g;
  const int tile = tile_of_block(blk, nblk);
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int i0 = GH + tx * TX, j0 = GH + ty * TY;
  __shared__ TileLds<R> L;
  const FluxDivs dv = flux_divs(g);
  // stage the tile and three rings: eta = h + b, h, u, v, b, phi
  for (int t = threadIdx.x; t < BX * BY; t += NT) {
    const int lx = t % BX, ly = t / BX;
    const int i = i0 - 3 + lx, j = j0 - 3 + ly;
    double hd = 0.0, bb = 0.0, u = 0.0, v = 0.0;
    R ph = R(0);
    if (i < g.w && j < g.rows) {
      const int k = j * g.w + i;
      hd = (double)a.h[k];
      if (a.h_lo) hd = hd + (double)a.h_lo[k];
      bb = a.b[k];
      double mu = (double)a.hu[k], mv = (double)a.hv[k];
      if (a.hu_lo) {
        mu = mu + (double)a.hu_lo[k];
        mv = mv + (double)a.hv_lo[k];
      }
      u = wc_velocity(hd, mu);
      v = wc_velocity(hd, mv);
      ph = (R)a.phi[k];
    }
    L.u.box.eta[ly][lx] = (R)(hd + bb);
    L.u.box.h[ly][lx] = hd;
    L.u.box.un_x[ly][lx] = (R)u;
    L.u.box.un_y[ly][lx] = (R)v;
    L.u.box.b[ly][lx] = bb;
    L.u.box.phi[ly][lx] = ph;
  }
  __syncthreads();
  // One work list for both axes, so that no pass runs with only a few threads busy:
  // x: cells of tile columns -1..64 of every tile row, stencil along the row; then
  // y: tile columns 0..63 of tile rows -1..4, stencil along the column.
  for (int q = threadIdx.x; q < TY * RXN + RYN * TX; q += NT) {
    if (q < TY * RXN) {
      const int cx = q % RXN, row = q / RXN;
      const int y = row + 3, x0 = cx;  // stencil columns cx..cx+4 = tile cols cx-3..cx+1
      reconstruct_r<R>(&L.u.box.eta[y][x0], &L.u.box.h[y][x0], &L.u.box.un_x[y][x0],
                       &L.u.box.un_y[y][x0], &L.u.box.b[y][x0], &L.u.box.phi[y][x0], 1, a.order,
                       dv.q60, L.xlo[row][cx], L.xhi[row][cx]);
    } else {
      const int qy = q - TY * RXN;
      const int col = qy % TX, ry = qy / TX;
      const int x = col + 3, y0 = ry;
      reconstruct_r<R>(&L.u.box.eta[y0][x], &L.u.box.h[y0][x], &L.u.box.un_y[y0][x],
                       &L.u.box.un_x[y0][x], &L.u.box.b[y0][x], &L.u.box.phi[y0][x], BX,
                       a.order, dv.q60, L.ylo[ry][col], L.yhi[ry][col]);
    }
  }
  __syncthreads();
  // faces, one list again: x-face f of a row sits between tile columns f-1 and f; y-face f of a
  // column between tile rows f-1 and f.
  for (int q = threadIdx.x; q < TY * (TX + 1) + (TY + 1) * TX; q += NT) {
    if (q < TY * (TX + 1)) {
      const int f = q % (TX + 1), row = q / (TX + 1);
      L.u.faces.fx[row][f] = face_r<R>(L.xhi[row][f], L.xlo[row][f + 1]);
    } else {
      const int qy = q - TY * (TX + 1);
      const int col = qy % TX, f = qy / TX;
      L.u.faces.fy[f][col] = face_r<R>(L.yhi[f][col], L.ylo[f + 1][col]);
    }
  }
  __syncthreads();
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int i = i0 + lx, j = j0 + ly;
  i_out = i;
  j_out = j;
  valid = i < g.nx + GH && j < g.ny + GH;
  if (!valid) return;
  const Face& xl = L.u.faces.fx[ly][lx];
  const Face& xh = L.u.faces.fx[ly][lx + 1];
  const Face& yl = L.u.faces.fy[ly][lx];
  const Face& yh = L.u.faces.fy[ly + 1][lx];
  const SideT<R>& xs_lo = L.xlo[ly][lx + 1];
  const SideT<R>& xs_hi = L.xhi[ly][lx + 1];
  const SideT<R>& ys_lo = L.ylo[ly + 1][lx];
  const SideT<R>& ys_hi = L.yhi[ly + 1][lx];
  const double xs = div_q(-WC_G * 0.5 * (xs_lo.h + xs_hi.h) * (xs_hi.z - xs_lo.z), dv.qdx);
  const double ys = div_q(-WC_G * 0.5 * (ys_lo.h + ys_hi.h) * (ys_hi.z - ys_lo.z), dv.qdy);
  rh = div_q(-(xh.f0 - xl.f0), dv.qdx) - div_q(yh.f0 - yl.f0, dv.qdy);
  rhu = div_q(-((xh.f1 + xh.cl) - (xl.f1 + xl.cr)), dv.qdx) + xs - div_q(yh.f2 - yl.f2, dv.qdy);
  rhv = div_q(-((yh.f1 + yh.cl) - (yl.f1 + yl.cr)), dv.qdy) + ys - div_q(xh.f2 - xl.f2, dv.qdx);
  if (a.has_drive) {
    const int ii = i - GH;
    double src = 0.0;
    if (ii >= a.src_i0 && ii < a.src_i1)
      src = a.src_col[ii] * a.prm[a.ramp_prm] * a.prm[a.sin_prm + (j - GH)];
    rh = rh + src;
  }
}

// ---- the marching wave ------------------------------------------------------------------------
// One 64-lane wave per strip of MW = 58 interior columns and a band of rows: lane l holds padded
// column c0 + l (the strip's columns and three more on each side). The wave walks down its band
// one row at a time. Each lane keeps its column's five rows around the row being reconstructed
// along y in registers (the y stencil, the y sides and the y faces need no neighbour); the
// x stencil of a row goes through the wave's own LDS row. Per cell: one reconstruction and one
// face per axis (the 64x4 tile takes 2.5 and 2.3), the inputs read once per band plus six rows,
// and a few KB of LDS per wave instead of 75 KB per 256-thread tile. The arithmetic per cell is
// the tile kernel's (reconstruct5, face_r, the rates), so the results are the same bits.
constexpr int MW = TX - 6;

// What a lane stages per cell: depth (FP64, with the low part of the FP32 pair), the velocities
// and phi in R, the bed.
template <typename R>
struct Cell5 {
  double h, b;
  R u, v, p;
};

template <typename S, typename R>
__device__ __forceinline__ Cell5<R> load_cell(const FluxTileArgs<S>& a, int i, int j) {
  Cell5<R> c{0.0, 0.0, R(0), R(0), R(0)};
  if (i < a.g.w) {
    const int k = j * a.g.w + i;
    double hd = (double)a.h[k];
    if (a.h_lo) hd = hd + (double)a.h_lo[k];
    double mu = (double)a.hu[k], mv = (double)a.hv[k];
    if (a.hu_lo) {
      mu = mu + (double)a.hu_lo[k];
      mv = mv + (double)a.hv_lo[k];
    }
    c.h = hd;
    c.b = a.b[k];
    c.u = (R)wc_velocity(hd, mu);
    c.v = (R)wc_velocity(hd, mv);
    c.p = (R)a.phi[k];
  }
  return c;
}

template <typename R>
struct MarchLds {
  double h[TX], b[TX];
  R u[TX], v[TX], p[TX];
  SideT<R> xhi[TX];
  Face fx[TX];
};

// The rates of the band's cells, handed to out(i, j, k, rh, rhu, rhv) (padded i, j). `rows` is
// the band height; (virtual) block b takes strip b % strips, band b / strips.
template <typename S, typename R, class Out>
__device__ __forceinline__ void flux_march(const FluxTileArgs<S>& a, int rows, int blk, Out out) {
  const Geo& g = a.g;
  __shared__ MarchLds<R> L;
  const int strips = (g.nx + MW - 1) / MW;
  const int strip = blk % strips, band = blk / strips;
  const int l = threadIdx.x;
  const int i = strip * MW + l;  // padded column: interior column strip * MW - 3 + l
  const int j0 = GH + band * rows, j1 = min(j0 + rows, GH + g.ny);
  if (j0 >= j1) return;
  // the five rows around the row being reconstructed along y
  Cell5<R> w[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) w[m] = load_cell<S, R>(a, i, j0 - 3 + m);
  SideT<R> ylo_prev{}, yhi_prev{};
  Face fy_prev{};
  const FluxDivs dv = flux_divs(g);
  for (int jr = j0 - 1; jr <= j1; ++jr) {
    Cell5<R> nxt{};
    if (jr < j1) nxt = load_cell<S, R>(a, i, jr + 3);  // in flight while this row computes
    // y: the sides of row jr, the face between jr - 1 and jr
    SideT<R> ylo, yhi;
    {
      R e[5], n[5], t[5], p[5];
      double d[5], b[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        e[m] = (R)(w[m].h + w[m].b);
        d[m] = w[m].h;
        n[m] = w[m].v;
        t[m] = w[m].u;
        b[m] = w[m].b;
        p[m] = w[m].p;
      }
      reconstruct5<R>(e, d, n, t, b, p, a.order, dv.q60, ylo, yhi);
    }
    Face fy{};
    if (jr >= j0) fy = face_r<R>(yhi_prev, ylo);
    if (jr >= j0 + 1) {
      // x: row jr - 1 (w[1]) through the wave's LDS row
      const int j = jr - 1;
      L.h[l] = w[1].h;
      L.b[l] = w[1].b;
      L.u[l] = w[1].u;
      L.v[l] = w[1].v;
      L.p[l] = w[1].p;
      __syncthreads();
      SideT<R> xlo{}, xhi{};
      if (l >= 2 && l < TX - 2) {
        R e[5], n[5], t[5], p[5];
        double d[5], b[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          const int q = l - 2 + m;
          e[m] = (R)(L.h[q] + L.b[q]);
          d[m] = L.h[q];
          n[m] = L.u[q];
          t[m] = L.v[q];
          b[m] = L.b[q];
          p[m] = L.p[q];
        }
        reconstruct5<R>(e, d, n, t, b, p, a.order, dv.q60, xlo, xhi);
      }
      L.xhi[l] = xhi;
      __syncthreads();
      if (l >= 3 && l < TX - 2) L.fx[l] = face_r<R>(L.xhi[l - 1], xlo);
      __syncthreads();
      const int ii = i;  // padded
      if (l >= 3 && l < TX - 3 && ii < g.nx + GH) {
        const Face& xl = L.fx[l];
        const Face& xh = L.fx[l + 1];
        const Face& yl = fy_prev;
        const Face& yh = fy;
        const double xs = div_q(-WC_G * 0.5 * (xlo.h + xhi.h) * (xhi.z - xlo.z), dv.qdx);
        const double ys =
            div_q(-WC_G * 0.5 * (ylo_prev.h + yhi_prev.h) * (yhi_prev.z - ylo_prev.z), dv.qdy);
        double rh = div_q(-(xh.f0 - xl.f0), dv.qdx) - div_q(yh.f0 - yl.f0, dv.qdy);
        const double rhu = div_q(-((xh.f1 + xh.cl) - (xl.f1 + xl.cr)), dv.qdx) + xs -
                           div_q(yh.f2 - yl.f2, dv.qdy);
        const double rhv = div_q(-((yh.f1 + yh.cl) - (yl.f1 + yl.cr)), dv.qdy) + ys -
                           div_q(xh.f2 - xl.f2, dv.qdx);
        if (a.has_drive) {
          const int ic = ii - GH;
          double src = 0.0;
          if (ic >= a.src_i0 && ic < a.src_i1)
            src = a.src_col[ic] * a.prm[a.ramp_prm] * a.prm[a.sin_prm + (j - GH)];
          rh = rh + src;
        }
        out(ii, j, j * g.w + ii, rh, rhu, rhv);
      }
      __syncthreads();  // the next row's staging overwrites the LDS row
    }
    ylo_prev = ylo;
    yhi_prev = yhi;
    fy_prev = fy;
    if (jr < j1) {
#pragma unroll
      for (int m = 0; m < 4; ++m) w[m] = w[m + 1];
      w[4] = nxt;
    }
  }
}

// FP64: the tile rates with the stage epilogue of k_flux.
__global__ void __launch_bounds__(NT) k_flux_tile64(FluxTileArgs<double> f, StepArgs a,
                                                    int mode) {
  double rh, rhu, rhv;
  int i, j;
  bool valid;
  flux_tile<double, double>(f, mode, blockIdx.x, gridDim.x, rh, rhu, rhv, i, j, valid);
  if (!valid) return;
  stage_out(a, mode, i, j, j * a.g.w + i, rh, rhu, rhv);
}

// The same, marching (flux_kernel 3, the default). Held to 168 VGPRs, three waves per SIMD
// (two spilled registers): 481 -> 419 us at 4096x2048 against 174 VGPRs and two waves; forcing
// four spilled 81 registers.
__global__ void __launch_bounds__(TX, 3) k_flux_march64(FluxTileArgs<double> f, StepArgs a, int mode,
                                                     int rows) {
  flux_march<double, double>(f, rows, blockIdx.x,
                             [&](int i, int j, int k, double rh, double rhu, double rhv) {
                               stage_out(a, mode, i, j, k, rh, rhu, rhv);
                             });
}

// FP32 chosen form: the rates for the stage update that follows, in FP64 (the mass rate so
// that the volume is conserved to FP64 rounding; the momentum rates because their rounding
// would reach the SGN solves' right-hand side, and through the extrapolated starts cost
// iterations).
template <typename R>
__device__ __forceinline__ void flux32_body(const FluxTileArgs<float>& f, double* lh, double* lhu,
                                            double* lhv, int blk, int nblk) {
  double rh, rhu, rhv;
  int i, j;
  bool valid;
  flux_tile<float, R>(f, MODE_RATES, blk, nblk, rh, rhu, rhv, i, j, valid);
  if (!valid) return;
  const int k = j * f.g.w + i;
  store_nt(&lh[k], rh);
  store_nt(&lhu[k], rhu);
  store_nt(&lhv[k], rhv);
}

__global__ void __launch_bounds__(NT) k_flux_tile32(FluxTileArgs<float> f, double* lh,
                                                    double* lhu, double* lhv) {
  flux32_body<float>(f, lh, lhu, lhv, blockIdx.x, gridDim.x);
}

// The FP32 form's state reconstructed in FP64.
__global__ void __launch_bounds__(NT) k_flux_tile32r(FluxTileArgs<float> f, double* lh,
                                                     double* lhu, double* lhv) {
  flux32_body<double>(f, lh, lhu, lhv, blockIdx.x, gridDim.x);
}

template <typename R>
__device__ __forceinline__ void flux32_march(const FluxTileArgs<float>& f, double* lh,
                                             double* lhu, double* lhv, int rows, int blk) {
  flux_march<float, R>(f, rows, blk, [&](int, int, int k, double rh, double rhu, double rhv) {
    store_nt(&lh[k], rh);
    store_nt(&lhu[k], rhu);
    store_nt(&lhv[k], rhv);
  });
}

__global__ void __launch_bounds__(TX) k_flux_march32(FluxTileArgs<float> f, double* lh,
                                                     double* lhu, double* lhv, int rows) {
  flux32_march<float>(f, lh, lhu, lhv, rows, blockIdx.x);
}

__global__ void __launch_bounds__(TX, 3) k_flux_march32r(FluxTileArgs<float> f, double* lh,
                                                      double* lhu, double* lhv, int rows) {
  flux32_march<double>(f, lh, lhu, lhv, rows, blockIdx.x);
}

// Many instances in one launch (see kb_fill32).
template <typename R>
__device__ __forceinline__ void kb_flux32_body(const BFlux* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    const BFlux& a = args[q];
    flux32_body<R>(a.f, a.lh, a.lhu, a.lhv, vb - m.first[q], m.first[q + 1] - m.first[q]);
    __syncthreads();  // the next virtual block reuses the LDS
  }
}

__global__ void __launch_bounds__(NT) kb_flux32(const BFlux* __restrict__ args, BMap m) {
  kb_flux32_body<float>(args, m);
}

__global__ void __launch_bounds__(NT) kb_flux32r(const BFlux* __restrict__ args, BMap m) {
  kb_flux32_body<double>(args, m);
}

// The marching flux over many instances: instance q's virtual blocks are its strips times its
// bands of a.rows rows (the batch's tuned band height, at most its rows).
template <typename R>
__device__ __forceinline__ void kb_flux_march32_body(const BFlux* __restrict__ args, BMap m) {
  for (int vb = blockIdx.x; vb < m.total; vb += gridDim.x) {
    const int q = batch_instance(m, vb);
    const BFlux& a = args[q];
    flux32_march<R>(a.f, a.lh, a.lhu, a.lhv, a.rows, vb - m.first[q]);
    __syncthreads();  // the next virtual block reuses the LDS row
  }
}

__global__ void __launch_bounds__(TX) kb_flux_march32(const BFlux* __restrict__ args, BMap m) {
  kb_flux_march32_body<float>(args, m);
}

__global__ void __launch_bounds__(TX, 3) kb_flux_march32r(const BFlux* __restrict__ args, BMap m) {
  kb_flux_march32_body<double>(args, m);
}

}  // namespace wc
