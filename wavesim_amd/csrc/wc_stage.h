// Per-cell stage update shared by the flux kernels (wc_kernels.hip, wc_flux.hip).
#pragma once

#include "wc_kernels.h"

namespace wc {

// ---- the stage update ----------------------------------------------------------------------------
// clean_dry (solver.rs:577-591)
__device__ __forceinline__ void clean_dry(double& h, double& hu, double& hv) {
  if (h < WC_H_DRY) {
    h = wc_max(h, 0.0);
    hu = 0.0;
    hv = 0.0;
  }
}

// damp (solver.rs:535-574): the sponges in order, then Manning friction, then clean_dry.
__device__ __forceinline__ void damp(const StepArgs& a, double dt, int i, int j, double b,
                                     double& h, double& hu, double& hv) {
  const int ii = i - GH, jj = j - GH;
  for (int m = 0; m < a.n_sponges; ++m) {
    const SpongeDev& sp = a.sponges[m];
    long long d = 0x7fffffffffffffffLL;
    if (sp.west) d = d < ii ? d : ii;
    if (sp.east) d = d < (a.g.nx - 1 - ii) ? d : (a.g.nx - 1 - ii);
    if (sp.south) d = d < jj ? d : jj;
    if (sp.north) d = d < (a.g.ny - 1 - jj) ? d : (a.g.ny - 1 - jj);
    if (d < sp.len) {
      const double rate = a.sponge_rate[sp.offset + d];
      if (rate > 0.0) {
        const double f = a.prm[a.sponge_prm + sp.offset + d];
        const double ht = wc_max(sp.level - b, 0.0);
        h = f * h + (1.0 - f) * ht;
        hu = f * hu + (1.0 - f) * ht * 0.0;
        hv = f * hv + (1.0 - f) * ht * 0.0;
      }
    }
  }
  if (a.manning > 0.0 && h > WC_H_FRICTION_MIN) {
    const double speed = wc_hypot(hu, hv) / h;
    const double kk = WC_G * a.manning * a.manning / wc_pow43(h);
    const double fac = 1.0 / (1.0 + dt * kk * speed);
    hu *= fac;
    hv *= fac;
  }
  clean_dry(h, hu, hv);
}

// Where a stage's rates go (solver.rs:481-507, 594-616):
//   MODE_S1:    s1 = clean_dry(s + dt L)               (stage 1, and the Euler step)
//   MODE_FINAL: s = damp(clean_dry((s + clean_dry(s1 + dt L)) / 2))  (stage 2)
//   MODE_RATES: store L (the dispersive correction still has to be added)
__device__ __forceinline__ void stage_out(const StepArgs& a, int mode, int i, int j, int k,
                                          double rh, double rhu, double rhv) {
  if (mode == MODE_RATES) {  // the rates, read once by the SGN set-up: stored nontemporal
    store_nt(&a.lh[k], rh);
    store_nt(&a.lhu[k], rhu);
    store_nt(&a.lhv[k], rhv);
    return;
  }
  double h = a.in_h[k], hu = a.in_hu[k], hv = a.in_hv[k];
  h += a.dt * rh;
  hu += a.dt * rhu;
  hv += a.dt * rhv;
  clean_dry(h, hu, hv);
  if (mode == MODE_S1) {
    a.s1_h[k] = h;
    a.s1_hu[k] = hu;
    a.s1_hv[k] = hv;
    return;
  }
  double sh = 0.5 * (a.s_h[k] + h);
  double shu = 0.5 * (a.s_hu[k] + hu);
  double shv = 0.5 * (a.s_hv[k] + hv);
  clean_dry(sh, shu, shv);
  damp(a, a.dt, i, j, a.b[k], sh, shu, shv);
  a.s_h[k] = sh;
  a.s_hu[k] = shu;
  a.s_hv[k] = shv;
}

}  // namespace wc
