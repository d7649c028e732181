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

}  // namespace wc
