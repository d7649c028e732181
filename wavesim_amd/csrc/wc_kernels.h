// Kernel interfaces of the FP64 step: argument blocks and declarations (kernels in
// wc_kernels.hip, launches in wc_api.hip).
#pragma once

#include "wc_device.h"

namespace wc {

enum { MODE_S1 = 0, MODE_FINAL = 1, MODE_RATES = 2 };

struct Breaking {
  double er_lo, er_hi, sl_lo, sl_hi;
};

// One sponge as the kernels use it: its table of rates and per-step factors, indexed by the
// distance in cells to the nearest active edge (forcing.rs:181-206).
struct SpongeDev {
  long long offset, len;
  double level;
  int west, east, south, north;
};

constexpr int MAX_SPONGES = 16;

// Scalars of one PCG solve and the DispersionStats counters, in device memory.
struct CgScalars {
  double rz, target, alpha, beta;
  unsigned long long it;
  int converged, done;
  unsigned int ticket[4];
  unsigned long long solves, iterations, unconverged;
};

struct StepArgs {
  Geo g;
  int order;
  double dt;
  const double* b;
  const double* phi;
  // stage input (s for stage 1, s1 for stage 2)
  const double *in_h, *in_hu, *in_hv;
  double *s_h, *s_hu, *s_hv;     // the state s (read and written by MODE_FINAL)
  double *s1_h, *s1_hu, *s1_hv;  // written by MODE_S1
  double *lh, *lhu, *lhv;        // rates
  // sources: the wave-maker columns and the host-evaluated per-step values
  int has_drive;
  int src_i0, src_i1;
  const double* src_col;
  const double* prm;
  int ramp_prm, sin_prm, sponge_prm;
  int n_sponges;
  SpongeDev sponges[MAX_SPONGES];
  const double* sponge_rate;
  double manning;
};

struct SgnArgs {
  Geo g;
  const double *h, *hu, *hv, *phi;
  const double *lh, *lhu, *lhv;
  double *c, *dgx, *dgy, *bx, *by, *wx, *wy, *rx, *ry, *px, *py, *ax, *ay;
  double *zx, *zy;  // z = r / diag, kept between the two kernels of a fused iteration
  double* part;
  CgScalars* sc;
  int nonlinear;
  double tolerance;
  unsigned long long max_iterations;
};

__global__ void k_fill_state(Geo g, double* h, double* hu, double* hv, double* br, double* ah,
                             double* ahu, double* ahv);
__global__ void k_switch(Geo g, const double* h, const double* b, Breaking br,
                         const uint8_t* held, double* phi);
__global__ void k_breaking(Geo g, const double* h, const double* b, Breaking br, double* out);
__global__ void k_flux(StepArgs a, int mode);
__global__ void k_finish_euler(StepArgs a);
__global__ void k_sgn_setup(SgnArgs a);
__global__ void k_cg_init(SgnArgs a);
__global__ void k_cg_matvec(SgnArgs a);
__global__ void k_cg_update(SgnArgs a);
__global__ void k_cg_pupdate(SgnArgs a);
__global__ void k_cg_stats(CgScalars* sc);
__global__ void k_cg_matvec2(SgnArgs a, int first, const double* psx, const double* psy,
                             double* pdx, double* pdy);
__global__ void k_cg_update2(SgnArgs a, const double* pcx, const double* pcy);
__global__ void k_sgn_finish(StepArgs a, const double* wx, const double* wy, int mode);
__global__ void k_max_rate(Geo g, const double* h, const double* hu, const double* hv,
                           unsigned long long* out);
__global__ void k_limit_scan(Geo g, const double* h, const double* hu, const double* hv,
                             unsigned long long* maxbits, unsigned int* first_nan);
__global__ void k_limit_first(Geo g, const double* h, const double* hu, const double* hv,
                              const unsigned long long* maxbits, unsigned int* first);
__global__ void k_state_sums(Geo g, const double* h, const double* hu, const double* hv,
                             double* part, unsigned int* ticket, double* out,
                             unsigned long long* maxmom, unsigned long long* minh_key);

}  // namespace wc
