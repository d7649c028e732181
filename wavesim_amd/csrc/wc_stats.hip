// The run's per-step statistics and its frames on the GPU (crates/wavesim/src/lib.rs:849-937,
// 1099-1130, 1276-1300, 1348-1380): one kernel per step over the interior, nothing copied to
// the host until the statistics are read or a frame is taken. The indicator is breaking_into's
// (dispersion.rs:183-201), as k_breaking.

#include "wc_kernels.h"

namespace wc {

namespace {

// Stats::close_wave (lib.rs:918-932) for cell n.
__device__ __forceinline__ void close_wave_at(const RecordArgs& a, int n) {
  const double hi = a.wave_max[n], lo = a.wave_min[n];
  const double height = isfinite(hi) ? hi - lo : 0.0;
  a.wave_sum[n] += height;
  if (a.first_wave) a.wave_first[n] = height;
  a.wave_last[n] = height;
  a.wave_max[n] = -INFINITY;
  a.wave_min[n] = INFINITY;
}

}  // namespace

__global__ void __launch_bounds__(NT) k_record(RecordArgs a) {
  const Geo& g = a.g;
  const int n = blockIdx.x * NT + threadIdx.x;
  if (n >= g.nx * g.ny) return;
  const int i = GH + n % g.nx, j = GH + n / g.nx;
  const int k = j * g.w + i;
  if (a.close_wave) close_wave_at(a, n);
  const double now = breaking_at(a.h, a.b, g, a.br, i, j, k);
  const double eta = a.h[k] + (a.elevation ? (double)a.elevation[n] : a.b[k]);
  a.harm[3 * n] += a.w0 * eta;
  a.harm[3 * n + 1] += a.w1 * eta;
  a.harm[3 * n + 2] += a.w2 * eta;
  a.eta_max[n] = wc_max(a.eta_max[n], eta);
  a.eta_min[n] = wc_min(a.eta_min[n], eta);
  a.wave_max[n] = wc_max(a.wave_max[n], eta);
  a.wave_min[n] = wc_min(a.wave_min[n], eta);
  if (now >= 0.8) a.breaking_seconds[n] += a.inside;
}

__global__ void __launch_bounds__(NT) k_close_wave(RecordArgs a) {
  const int n = blockIdx.x * NT + threadIdx.x;
  if (n < a.g.nx * a.g.ny) close_wave_at(a, n);
}

__global__ void __launch_bounds__(NT) k_frame(FrameArgs a) {
  const Geo& g = a.g;
  const int n = blockIdx.x * NT + threadIdx.x;
  double rise = 0.0;
  bool bad = false;
  if (n < g.nx * g.ny) {
    const int i = GH + n % g.nx, j = GH + n / g.nx;
    const int k = j * g.w + i;
    const double hc = a.h[k];
    const float eta = (float)(hc + (double)a.elevation[n]);
    a.eta[n] = eta;
    a.breaking[n] = (float)breaking_at(a.h, a.b, g, a.br, i, j, k);
    // rise.max(eta - tide): a NaN difference leaves rise as it is, as f64::max does
    if (hc > 1e-3) {
      const double d = (double)eta - a.tide;
      if (d > rise) rise = d;
    }
    bad = !isfinite(eta);
  }
  // one atomic per wave: rise is >= 0 (or +inf), so its bits order as the values do
  unsigned long long bits = (unsigned long long)__double_as_longlong(rise);
  for (int m = 1; m < 64; m <<= 1) {
    const unsigned long long o = __shfl_xor(bits, m);
    bits = o > bits ? o : bits;
  }
  const unsigned long long any_bad = __ballot(bad);
  if ((threadIdx.x & 63) == 0) {
    if (bits) atomicMax(a.rise_bits, bits);
    if (any_bad) atomicOr(a.nonfinite, 1u);
  }
}

}  // namespace wc
