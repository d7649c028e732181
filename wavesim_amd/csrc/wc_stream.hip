// What plain streams reach on this GPU's HBM, the yardstick for the stencil kernels' bytes: a
// copy, reads alone, writes alone, and three reads with a write (the shape of a PCG update), each
// with plain and with nontemporal loads and stores. Grid-stride loops over arrays of double2
// (16 B a thread a step) in 256-thread blocks.

#include "wc_host.h"

namespace wc {

namespace {

template <bool NT>
__device__ __forceinline__ double2 get(const double2* p) {
  if constexpr (NT)
    return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
  else
    return *p;
}

template <bool NT>
__device__ __forceinline__ void put(double2* p, double2 v) {
  if constexpr (NT) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
  } else {
    *p = v;
  }
}

template <bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const double2* __restrict__ a,
                                                     double2* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    put<NT>(&b[i], get<NT>(&a[i]));
}

// The sum is stored only if it equals a value no input of zeros produces, so the loads stay.
template <bool NT>
__global__ void __launch_bounds__(256) k_stream_read(const double2* __restrict__ a, double* out,
                                                     long long n) {
  double s = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 v = get<NT>(&a[i]);
    s += v.x + v.y;
  }
  if (s == 12345.678) out[0] = s;
}

template <bool NT>
__global__ void __launch_bounds__(256) k_stream_write(double2* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    put<NT>(&b[i], make_double2(1.0, 2.0));
}

template <bool NT>
__global__ void __launch_bounds__(256) k_stream_triad(const double2* __restrict__ a,
                                                      const double2* __restrict__ c,
                                                      const double2* __restrict__ d,
                                                      double2* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double2 x = get<NT>(&a[i]), y = get<NT>(&c[i]), z = get<NT>(&d[i]);
    put<NT>(&b[i], make_double2(x.x + 0.5 * y.x + z.x, x.y + 0.5 * y.y + z.y));
  }
}

}  // namespace

}  // namespace wc

extern "C" int wc_stream_ceilings(uint64_t bytes, int32_t grid, int32_t reps, double* tbs) {
  if (!tbs) return fail(WC_EINVAL, "null output");
  if (bytes < 16 || bytes % 16 != 0 || bytes > (16ull << 30))
    return fail(WC_EINVAL, "bytes must be a multiple of 16 up to 16 GiB");
  if (grid < 1 || reps < 1) return fail(WC_EINVAL, "grid and reps must be positive");
  const long long n = (long long)(bytes / 16);
  double2* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  double* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto run = [&]() -> int {
    for (double2*& b : buf) {
      HIP_OK(hipMalloc(&b, bytes));
      HIP_OK(hipMemset(b, 0, bytes));
    }
    HIP_OK(hipMalloc(&out, sizeof(double)));
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const dim3 g((unsigned)grid), t(256);
    // each test: bytes moved per launch, and the launch
    auto time = [&](double moved, auto launch, double& result) -> int {
      launch();
      HIP_OK(hipDeviceSynchronize());
      HIP_OK(hipEventRecord(e0));
      for (int r = 0; r < reps; ++r) launch();
      HIP_OK(hipEventRecord(e1));
      HIP_OK(hipEventSynchronize(e1));
      HIP_OK(hipGetLastError());
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      result = moved * reps / (1e-3 * ms) / 1e12;
      return WC_OK;
    };
    const double B = (double)bytes;
    double2 *a = buf[0], *b = buf[1], *c = buf[2], *d = buf[3];
    TRY(time(2 * B, [&] { hipLaunchKernelGGL(k_stream_copy<false>, g, t, 0, 0, a, b, n); },
             tbs[0]));
    TRY(time(2 * B, [&] { hipLaunchKernelGGL(k_stream_copy<true>, g, t, 0, 0, a, b, n); },
             tbs[1]));
    TRY(time(B, [&] { hipLaunchKernelGGL(k_stream_read<false>, g, t, 0, 0, a, out, n); },
             tbs[2]));
    TRY(time(B, [&] { hipLaunchKernelGGL(k_stream_read<true>, g, t, 0, 0, a, out, n); },
             tbs[3]));
    TRY(time(B, [&] { hipLaunchKernelGGL(k_stream_write<false>, g, t, 0, 0, b, n); }, tbs[4]));
    TRY(time(B, [&] { hipLaunchKernelGGL(k_stream_write<true>, g, t, 0, 0, b, n); }, tbs[5]));
    TRY(time(4 * B, [&] { hipLaunchKernelGGL(k_stream_triad<false>, g, t, 0, 0, a, c, d, b, n); },
             tbs[6]));
    TRY(time(4 * B, [&] { hipLaunchKernelGGL(k_stream_triad<true>, g, t, 0, 0, a, c, d, b, n); },
             tbs[7]));
    return WC_OK;
  };
  const int rc = run();
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (double2* b : buf)
    if (b) (void)hipFree(b);
  if (out) (void)hipFree(out);
  return rc;
}
