// The run's per-step statistics on the GPU (crates/wavesim/src/lib.rs:849-937, 1099-1130,
// 1276-1300): one kernel per step over the interior, nothing copied to the host until the
// statistics are read. The indicator is breaking_into's (dispersion.rs:183-201), as k_breaking.

#include "wc_kernels.h"

namespace wc {

__global__ void __launch_bounds__(NT) k_record(RecordArgs a) {
  const Geo& g = a.g;
  const int n = blockIdx.x * NT + threadIdx.x;
  if (n >= g.nx * g.ny) return;
  const int i = GH + n % g.nx, j = GH + n / g.nx;
  const int k = j * g.w + i;
  if (a.close_wave) {
    const double hi = a.wave_max[n], lo = a.wave_min[n];
    const double height = isfinite(hi) ? hi - lo : 0.0;
    a.wave_sum[n] += height;
    if (a.first_wave) a.wave_first[n] = height;
    a.wave_last[n] = height;
    a.wave_max[n] = -INFINITY;
    a.wave_min[n] = INFINITY;
  }
  // the indicator (breaking_into): 1 - phi with the default thresholds, 0 below 0.05 m
  const double hc = a.h[k];
  double now = 0.0;
  if (hc >= WC_H_DISPERSION_MIN) {
    const double ix = 0.5 / g.dx, iy = 0.5 / g.dy;
    const double ex =
        ((ghosted(a.h, g, i + 1, j) + a.b[k + 1]) - (ghosted(a.h, g, i - 1, j) + a.b[k - 1])) * ix;
    const double ey = ((ghosted(a.h, g, i, j + 1) + a.b[k + g.w]) -
                       (ghosted(a.h, g, i, j - 1) + a.b[k - g.w])) * iy;
    const double ratio = wc_max((hc + a.b[k]) / hc, 0.0);
    const double off = wc_max(wc_smoothstep(a.br.er_lo, a.br.er_hi, ratio),
                              wc_smoothstep(a.br.sl_lo, a.br.sl_hi, wc_hypot(ex, ey)));
    now = 1.0 - (1.0 - off);
  }
  const double eta = hc + (a.elevation ? (double)a.elevation[n] : a.b[k]);
  a.harm[3 * n] += a.w0 * eta;
  a.harm[3 * n + 1] += a.w1 * eta;
  a.harm[3 * n + 2] += a.w2 * eta;
  a.eta_max[n] = wc_max(a.eta_max[n], eta);
  a.eta_min[n] = wc_min(a.eta_min[n], eta);
  a.wave_max[n] = wc_max(a.wave_max[n], eta);
  a.wave_min[n] = wc_min(a.wave_min[n], eta);
  if (now >= 0.8) a.breaking_seconds[n] += a.inside;
}

}  // namespace wc
