// stable_dt of a caller's device-resident state, allocating nothing (wc_stable_dt_noalloc).
//
// wc_stable_dt reduces the largest wave speed rate through a buffer the solver allocated. Some
// callers cannot allocate at all (code under HIP graph capture, tight serving loops, harnesses
// that forbid it): this variant takes the padded h, hu, hv from the caller's device memory and
// writes dt to the caller's device memory, and keeps its partial results in `scratch`, a
// file-scope device array of fixed size that each GPU holds from the library's first use on it
// until the process ends. Two launches, no atomics, nothing to clear: each block writes the
// largest rate bits of its 4096 cells into its own slot, and one block folds the slots and
// writes dt. The per-cell rate, the NaN-propagating maximum of the IEEE bit patterns and the
// final cfl / rate (solver.rs:95-100, 434-450) are those of wc_stable_dt, so the two agree bit
// for bit.

#include <mutex>
#include <string>

#include "../../include/wavecore_hip.h"
#include "wc_kernels.h"

namespace wc {

int api_fail(int code, const char* msg);  // wc_api.hip: sets wc_last_error

namespace {

// One slot per block of RATE_CELLS cells: 16384 slots (128 KiB) cover 67,108,864 cells.
constexpr int SCRATCH_SLOTS = 16384;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned long long scratch[SCRATCH_SLOTS];

// The scratch array's address on each device (its name is only a host-side stand-in).
unsigned long long* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_on(int dev, unsigned long long** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return api_fail(WC_EDEVICE, "wc_stable_dt_noalloc: hipGetSymbolAddress(scratch) failed");
    g_scratch[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch[dev];
  return WC_OK;
}

// The largest rate bits of block b's RATE_CELLS cells into part[b].
__global__ void __launch_bounds__(NT) k_rate_part(Geo g, const double* __restrict__ h,
                                                  const double* __restrict__ hu,
                                                  const double* __restrict__ hv,
                                                  unsigned long long* __restrict__ part) {
  __shared__ unsigned long long lds[NT / 64];
  const int n = g.nx * g.ny;
  unsigned long long bits = 0ull;
  for (int t = blockIdx.x * RATE_CELLS + threadIdx.x;
       t < min(n, (int)(blockIdx.x + 1) * RATE_CELLS); t += NT) {
    const int k = (GH + t / g.nx) * g.w + GH + t % g.nx;
    const double hc = h[k];
    const double c = sqrt(WC_G * hc);
    const double u = fabs(wc_velocity(hc, hu[k]));
    const double v = fabs(wc_velocity(hc, hv[k]));
    const double r = (u + c) / g.dx + (v + c) / g.dy;
    const unsigned long long b = (unsigned long long)__double_as_longlong(r);
    bits = b > bits ? b : bits;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_xor(bits, off, 64);
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; ++w) bits = lds[w] > bits ? lds[w] : bits;
    part[blockIdx.x] = bits;
  }
}

// The largest of part[0, nb) and dt = cfl / rate from it (NaN if NaN, +inf if not positive).
__global__ void __launch_bounds__(NT) k_rate_dt(const unsigned long long* __restrict__ part,
                                                int nb, double cfl, double* __restrict__ dt) {
  __shared__ unsigned long long lds[NT / 64];
  unsigned long long bits = 0ull;
  for (int b = threadIdx.x; b < nb; b += NT) bits = part[b] > bits ? part[b] : bits;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_xor(bits, off, 64);
    bits = o > bits ? o : bits;
  }
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; ++w) bits = lds[w] > bits ? lds[w] : bits;
    const double rate = __longlong_as_double((long long)bits);
    double out;
    if (isnan(rate))
      out = __longlong_as_double(0x7ff8000000000000ll);
    else if (rate > 0.0)
      out = cfl / rate;
    else
      out = __longlong_as_double(0x7ff0000000000000ll);
    *dt = out;
  }
}

}  // namespace
}  // namespace wc

extern "C" int wc_stable_dt_noalloc(const double* h, const double* hu, const double* hv,
                                    uint64_t nx, uint64_t ny, double dx, double dy, double cfl,
                                    double* dt_out, void* stream) {
  using namespace wc;
  if (!h || !hu || !hv || !dt_out)
    return api_fail(WC_EINVAL, "wc_stable_dt_noalloc: null array");
  if (nx < 3 || ny < 3)
    return api_fail(WC_EINVAL, "wc_stable_dt_noalloc: grids need at least 3 cells per axis");
  if (!(dx > 0.0) || !(dy > 0.0))
    return api_fail(WC_EINVAL, "wc_stable_dt_noalloc: cell sizes must be positive");
  const uint64_t cells = nx * ny;
  const uint64_t blocks = (cells + RATE_CELLS - 1) / RATE_CELLS;
  if (nx > (1u << 30) || ny > (1u << 30) || blocks > (uint64_t)SCRATCH_SLOTS)
    return api_fail(WC_EINVAL,
                    "wc_stable_dt_noalloc: at most 67,108,864 interior cells (16384 scratch "
                    "slots of 4096 cells)");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return api_fail(WC_EDEVICE, "wc_stable_dt_noalloc: no current device");
  unsigned long long* part = nullptr;
  const int rc = scratch_on(dev, &part);
  if (rc != WC_OK) return rc;
  Geo g{};
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.w = (int)nx + 2 * GH;
  g.rows = (int)ny + 2 * GH;
  g.dx = dx;
  g.dy = dy;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_rate_part, dim3((unsigned)blocks), dim3(NT), 0, st, g, h, hu, hv, part);
  hipLaunchKernelGGL(k_rate_dt, dim3(1), dim3(NT), 0, st, (const unsigned long long*)part,
                     (int)blocks, cfl, dt_out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return api_fail(WC_EDEVICE, (std::string("wc_stable_dt_noalloc: ") + hipGetErrorString(e)).c_str());
  return WC_OK;
}
