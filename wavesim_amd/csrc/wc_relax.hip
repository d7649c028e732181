// Relaxation zones (forcing.rs:227-284) on the GPU: one thread per zone cell.
//
// relax_cell (forcing.rs:213-225) with the reference's operation order, then the dry-cell
// cleanup of Relaxation::apply (forcing.rs:276-280). The factor exp(-rate dt) comes from the
// host (glibc's exp, as the reference), one per distinct rate of the zones.

#include "wc_kernels.h"

namespace wc {

namespace {

__device__ __forceinline__ void relax(double f, double ht, double u, double v, double& h,
                                      double& hu, double& hv) {
  h = f * h + (1.0 - f) * ht;
  hu = f * hu + (1.0 - f) * ht * u;
  hv = f * hv + (1.0 - f) * ht * v;
  if (h < WC_H_DRY) {
    h = wc_max(h, 0.0);
    hu = 0.0;
    hv = 0.0;
  }
}

__device__ __forceinline__ void split_pair(double v, float& hi, float& lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}

}  // namespace

__global__ void __launch_bounds__(NT) k_relax64(RelaxArgs a, double* h, double* hu, double* hv) {
  const long long m = (long long)blockIdx.x * NT + threadIdx.x;
  if (m >= a.n) return;
  const int k = a.cell[m];
  const double f = a.factor[a.rate[m]];
  const double eta = a.target[3 * m], u = a.target[3 * m + 1], v = a.target[3 * m + 2];
  const double ht = wc_max(eta - a.b[k], 0.0);
  double nh = h[k], nhu = hu[k], nhv = hv[k];
  relax(f, ht, u, v, nh, nhu, nhv);
  h[k] = nh;
  hu[k] = nhu;
  hv[k] = nhv;
}

// The FP32 form: the state is (high, low) FP32 pairs; the relaxation runs in FP64 on their sum.
__global__ void __launch_bounds__(NT) k_relax32(RelaxArgs a, float* h, float* hl, float* hu,
                                                float* hul, float* hv, float* hvl) {
  const long long m = (long long)blockIdx.x * NT + threadIdx.x;
  if (m >= a.n) return;
  const int k = a.cell[m];
  const double f = a.factor[a.rate[m]];
  const double eta = a.target[3 * m], u = a.target[3 * m + 1], v = a.target[3 * m + 2];
  const double ht = wc_max(eta - a.b[k], 0.0);
  double nh = (double)h[k] + (double)hl[k];
  double nhu = (double)hu[k] + (double)hul[k];
  double nhv = (double)hv[k] + (double)hvl[k];
  relax(f, ht, u, v, nh, nhu, nhv);
  split_pair(nh, h[k], hl[k]);
  split_pair(nhu, hu[k], hul[k]);
  split_pair(nhv, hv[k], hvl[k]);
}

// wavesim's Sampler::sample (crates/wavesim/src/lib.rs:1592-1614) for zone cell m: the coarse
// water at the four corners idx[4m..4m+3] (padded coarse indices, ghosts read through the wall
// map of fill_boundaries(grid, false), state.rs:68-76) with bilinear weights w[4m..4m+3]; wet
// corners (h > 1e-3) only, renormalised; (-inf, 0, 0) where none is wet.
__global__ void __launch_bounds__(NT) k_sample(long long n, const int* idx, const double* w,
                                               Geo g, const double* h, const double* hu,
                                               const double* hv, const double* b, double* out) {
  const long long m = (long long)blockIdx.x * NT + threadIdx.x;
  if (m >= n) return;
  double eta = 0.0, u = 0.0, v = 0.0, weight = 0.0;
  for (int c = 0; c < 4; ++c) {
    const int k = idx[4 * m + c];
    const double wk = w[4 * m + c];
    const int i = k % g.w, j = k / g.w;
    const int mk = map_y(j, g.ny, 0) * g.w + map_x(i, g.nx);
    const double hk = h[mk];
    const double huk = is_ghost_x(i, g.nx) ? -hu[mk] : hu[mk];
    const double hvk = is_ghost_y(j, g.ny) ? -hv[mk] : hv[mk];
    if (hk > 1e-3 && wk > 0.0) {
      eta += wk * (hk + b[k]);
      u += wk * huk / hk;
      v += wk * hvk / hk;
      weight += wk;
    }
  }
  if (weight > 0.0) {
    out[3 * m] = eta / weight;
    out[3 * m + 1] = u / weight;
    out[3 * m + 2] = v / weight;
  } else {
    out[3 * m] = -INFINITY;
    out[3 * m + 1] = 0.0;
    out[3 * m + 2] = 0.0;
  }
}

// The nested run's targets between two coarse samples (lib.rs:1877-1890): linear in time where
// both are water, else the nearer one.
__global__ void __launch_bounds__(NT) k_lerp_targets(long long n, const double* from,
                                                     const double* to, double w, double* out) {
  const long long m = (long long)blockIdx.x * NT + threadIdx.x;
  if (m >= n) return;
  const double* a = from + 3 * m;
  const double* c = to + 3 * m;
  if (isfinite(a[0]) && isfinite(c[0])) {
    out[3 * m] = a[0] + w * (c[0] - a[0]);
    out[3 * m + 1] = a[1] + w * (c[1] - a[1]);
    out[3 * m + 2] = a[2] + w * (c[2] - a[2]);
  } else {
    const double* t = w < 0.5 ? a : c;
    out[3 * m] = t[0];
    out[3 * m + 1] = t[1];
    out[3 * m + 2] = t[2];
  }
}

}  // namespace wc
