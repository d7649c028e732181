// Breaking fronts followed from step to step and their eddy viscosity (crates/wavecore/src/
// eddy.rs, Kennedy et al. 2000), split from the step (solver.rs:512-532): FP64, one thread per
// padded cell, the reference's operation order, so the FP64 path stays equal to the oracle.

#include "wc_kernels.h"

namespace wc {

namespace {

constexpr double H_BREAKING_MIN = 0.01;  // eddy.rs:55
constexpr double BREAKING_ONSET = 0.8;   // eddy.rs:59

// EddyViscosity::threshold (eddy.rs:120-131)
__device__ __forceinline__ double threshold(const EddyP& e, double age, double h) {
  const double span = e.duration * sqrt(h / WC_G);
  if (age >= span) return e.settled;
  return e.initial + (e.settled - e.initial) * age / span;
}

__device__ __forceinline__ bool interior(const Geo& g, int i, int j) {
  return i >= GH && i < g.nx + GH && j >= GH && j < g.ny + GH;
}

}  // namespace

// The rate of rise -div(hu) (eddy.rs:169-178); 0 outside the interior.
__global__ void __launch_bounds__(NT) k_eddy_rise(Geo g, const double* hu, const double* hv,
                                                  double* rise) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.w * g.rows) return;
  const int i = t % g.w, j = t / g.w;
  double r = 0.0;
  if (interior(g, i, j)) {
    const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
    r = -((hu[t + 1] - hu[t - 1]) * ix + (hv[t + g.w] - hv[t - g.w]) * iy);
  }
  rise[t] = r;
}

// Where the fronts are now, and their viscosity (eddy.rs:179-219); 0 outside the interior.
__global__ void __launch_bounds__(NT) k_eddy_follow(Geo g, EddyP e, const double* h,
                                                    const double* old, const double* phi,
                                                    const double* rise, double dt,
                                                    double* memory, double* nu) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.w * g.rows) return;
  const int i = t % g.w, j = t / g.w;
  double m = 0.0, v = 0.0;
  if (interior(g, i, j) && h[t] >= H_BREAKING_MIN) {
    const double depth = h[t];
    const double c = sqrt(WC_G * depth);
    double front = 0.0;
    front = wc_max(front, old[t]);
    front = wc_max(front, old[t - 1]);
    front = wc_max(front, old[t + 1]);
    front = wc_max(front, old[t - g.w]);
    front = wc_max(front, old[t + g.w]);
    const double rk = rise[t];
    const bool starts = depth >= WC_H_DISPERSION_MIN && 1.0 - phi[t] >= BREAKING_ONSET && rk > 0.0;
    const bool goes_on = front > 0.0 && rk > threshold(e, front, depth) * c;
    const double age = (front > 0.0 && (starts || goes_on)) ? front + dt : starts ? dt : 0.0;
    if (age > 0.0) {
      double b = rk / (threshold(e, age, depth) * c) - 1.0;
      b = b < 0.0 ? 0.0 : (b > 1.0 ? 1.0 : b);
      m = age;
      v = b * e.mixing * e.mixing * depth * rk;
    } else if (old[t] > 0.0) {
      m = -e.linger * sqrt(depth / WC_G);
    } else if (old[t] < 0.0) {
      m = wc_min(old[t] + dt, 0.0);
    }
  }
  memory[t] = m;
  nu[t] = v;
}

// The largest viscosity over the interior (diffuse, eddy.rs:234-239): nu >= 0, so its bits
// order as the values do.
__global__ void __launch_bounds__(NT) k_eddy_top(Geo g, const double* nu,
                                                 unsigned long long* top) {
  const int t = blockIdx.x * NT + threadIdx.x;
  unsigned long long bits = 0ull;
  if (t < g.nx * g.ny) {
    const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
    bits = (unsigned long long)__double_as_longlong(nu[k]);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_xor(bits, off, 64);
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0 && bits != 0ull) atomicMax(top, bits);
}

// One explicit sub-step of the viscous terms (eddy.rs:254-295) into out_hu, out_hv (interior).
__global__ void __launch_bounds__(NT) k_eddy_diffuse(Geo g, const double* h, const double* hu,
                                                     const double* hv, const double* nu,
                                                     double sub, double* out_hu,
                                                     double* out_hv) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.nx * g.ny) return;
  const int w = g.w;
  const int k = (GH + t / g.nx) * w + GH + t % g.nx;
  if (h[k] < H_BREAKING_MIN) {
    out_hu[k] = hu[k];
    out_hv[k] = hv[k];
    return;
  }
  const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
  const double cross = 0.25 / (g.dx * g.dy);
  auto link = [&](int a, int b) {
    return (h[a] < H_BREAKING_MIN || h[b] < H_BREAKING_MIN) ? 0.0 : 0.5 * (nu[a] + nu[b]);
  };
  auto wet = [&](int a) { return h[a] < H_BREAKING_MIN ? 0.0 : nu[a]; };
  const int e = k + 1, west = k - 1, north = k + w, south = k - w;
  auto xx = [&](const double* q) {
    return (link(k, e) * (q[e] - q[k]) - link(k, west) * (q[k] - q[west])) * idx2;
  };
  auto yy = [&](const double* q) {
    return (link(k, north) * (q[north] - q[k]) - link(k, south) * (q[k] - q[south])) * idy2;
  };
  auto yx = [&](const double* q) {
    return (wet(north) * (q[north + 1] - q[north - 1]) -
            wet(south) * (q[south + 1] - q[south - 1])) * cross;
  };
  auto xy = [&](const double* q) {
    return (wet(e) * (q[e + w] - q[e - w]) - wet(west) * (q[west + w] - q[west - w])) * cross;
  };
  out_hu[k] = hu[k] + sub * (xx(hu) + 0.5 * (yy(hu) + yx(hv)));
  out_hv[k] = hv[k] + sub * (yy(hv) + 0.5 * (xx(hv) + xy(hu)));
}

// clean_dry (solver.rs:577-591) over the interior.
__global__ void __launch_bounds__(NT) k_clean_dry(Geo g, double* h, double* hu, double* hv) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= g.nx * g.ny) return;
  const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
  if (h[k] < WC_H_DRY) {
    h[k] = wc_max(h[k], 0.0);
    hu[k] = 0.0;
    hv[k] = 0.0;
  }
}

}  // namespace wc
