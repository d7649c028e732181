// The learned preconditioner's network (include/wavecore_unet.h) on gfx950: BF16 MFMA
// convolutions with FP32 accumulation, forward and backward, and the training kernels.
//
// Layout. Activations and their gradients are NHWC BF16 with a one-pixel zero border, level l
// of size H0 / 2^l x W0 / 2^l (W0 = nx rounded up to 128, H0 = ny rounded up to 8, so every
// level's width is a multiple of 16). Pixel (y, x) of a C-channel tensor is at
// ((y + 1) (W + 2) + x + 1) C.
//
// Convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16, D (16 x 16) = A (16 x 32) B (32 x
// 16): A holds 16 output channels of the weights, B 16 pixels of a row, K runs over (tap, input
// channel) pairs, 32 per step: one tap and 32 channels when the input has 32 or more, two taps
// of 16 channels, or four taps of 8. Lane l holds row l % 16 of A and column l % 16 of B with
// the 8 consecutive k of group l / 16, so a lane's B operand is 8 channels of one pixel at one
// tap, a single 16-byte load, and its A operand one 16-byte load of the packed weights; D gives
// lane l output channels 4 (l / 16) .. +3 of pixel l % 16, stored as 8 bytes. A wave computes
// PB blocks of 16 pixels for all output channels. The transposed convolution of the backward
// pass (d input from d output) is the same kernel with the weights flipped and transposed at
// packing time.
//
// Weight gradients are GEMMs over the pixels (K = 32 pixels per MFMA step): a block stages 32
// pixels of the layer's input (shifted by its tap) and of the output gradient in LDS,
// channel-major, so each lane's operand is again one 16-byte read; blocks work on chunks of
// pixels and their partial sums are added in a fixed order by a second kernel.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/wavecore_hip.h"
#include "../../include/wavecore_unet.h"

namespace wc {
int api_fail(int code, const char* msg);  // wc_api.hip: sets wc_last_error
}

namespace wcu {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int CHUNK_STEPS = 64;  // weight-gradient chunk: 64 x 32 pixels

// ---- the architecture -------------------------------------------------------------------------
enum Layer { L_E0, L_E1, L_E2, L_E3, L_B, L_T3, L_D2, L_T2, L_D1, L_T1, L_D0, L_O, NLAYERS };
struct LayerDef {
  int taps, cin, cin_real, cout, level, relu;
};
constexpr LayerDef LAYERS[NLAYERS] = {
    {9, 8, 5, 16, 0, 1},      {9, 16, 16, 32, 1, 1}, {9, 32, 32, 64, 2, 1},
    {9, 64, 64, 128, 3, 1},   {9, 128, 128, 128, 3, 1}, {1, 128, 128, 64, 3, 0},
    {9, 64, 64, 64, 2, 1},    {1, 64, 64, 32, 2, 0},  {9, 32, 32, 32, 1, 1},
    {1, 32, 32, 16, 1, 0},    {9, 16, 16, 16, 0, 1},  {1, 16, 16, 2, 0, 0},
};

enum Tensor {
  X0, E0, P1, E1, P2, E2, P3, E3, Y3, T3, U2, Y2, T2, U1, Y1, T1, U0, Y0, NTENSORS
};
struct TensorDef {
  int level, c;
};
constexpr TensorDef TENSORS[NTENSORS] = {
    {0, 8},  {0, 16}, {1, 16}, {1, 32}, {2, 32}, {2, 64}, {3, 64}, {3, 128}, {3, 128},
    {3, 64}, {2, 64}, {2, 64}, {2, 32}, {1, 32}, {1, 32}, {1, 16}, {0, 16},  {0, 16},
};

uint64_t layer_params(int l) {
  const LayerDef& d = LAYERS[l];
  return (uint64_t)d.taps * d.cin_real * d.cout + d.cout;
}
uint64_t layer_offset(int l) {
  uint64_t o = 0;
  for (int m = 0; m < l; ++m) o += layer_params(m);
  return o;
}

// ---- K ordering of the implicit GEMM ----------------------------------------------------------
__host__ __device__ inline int ksteps(int cin, int taps) {
  return cin >= 32 ? taps * (cin / 32) : (taps + (32 / cin) - 1) / (32 / cin);
}
__host__ __device__ inline void kmap(int cin, int s, int g, int& tap, int& ch) {
  if (cin >= 32) {
    const int cpt = cin / 32;
    tap = s / cpt;
    ch = (s % cpt) * 32 + 8 * g;
  } else {
    const int gpt = cin / 8;
    tap = s * (32 / cin) + g / gpt;
    ch = 8 * (g % gpt);
  }
}

// ---- kernels ----------------------------------------------------------------------------------
struct ConvArgs {
  const bf16* in;
  int H, W;
  const bf16x8* wpk;
  const float* bias;  // may be null
  bf16* out;
  const bf16* mask;   // may be null: the output is kept only where mask > 0 (ReLU backward)
  const int* skip;    // may be null: the kernel does nothing while *skip != 0
};

template <int TAPS, int CIN, int COUT, int PB, bool RELU>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  constexpr int NB = COUT / 16;
  constexpr int STEPS = (CIN >= 32) ? TAPS * (CIN / 32) : (TAPS + (32 / CIN) - 1) / (32 / CIN);
  if (a.skip && *a.skip) return;
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int bpr = a.W / (16 * PB);
  const int y = wave / bpr, x0 = (wave % bpr) * 16 * PB;
  if (y >= a.H) return;
  const int pitch = a.W + 2;
  f32x4 acc[PB][NB];
#pragma unroll
  for (int pb = 0; pb < PB; ++pb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[pb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
  for (int s = 0; s < STEPS; ++s) {
    int tap, ch;
    kmap(CIN, s, g, tap, ch);
    bf16x8 b[PB];
    if (tap < TAPS) {
      const int dy = TAPS == 9 ? tap / 3 - 1 : 0, dx = TAPS == 9 ? tap % 3 - 1 : 0;
#pragma unroll
      for (int pb = 0; pb < PB; ++pb) {
        const size_t o =
            ((size_t)(y + 1 + dy) * pitch + (x0 + 16 * pb + c16 + 1 + dx)) * CIN + ch;
        b[pb] = *reinterpret_cast<const bf16x8*>(a.in + o);
      }
    } else {
#pragma unroll
      for (int pb = 0; pb < PB; ++pb)
#pragma unroll
        for (int e = 0; e < 8; ++e) b[pb][e] = (bf16)0.f;
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const bf16x8 w = a.wpk[(s * NB + nb) * 64 + lane];
#pragma unroll
      for (int pb = 0; pb < PB; ++pb)
        acc[pb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, b[pb], acc[pb][nb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int pb = 0; pb < PB; ++pb) {
    const size_t o = ((size_t)(y + 1) * pitch + (x0 + 16 * pb + c16 + 1)) * COUT;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = nb * 16 + 4 * g;
      bf16x4 m4;
      if (a.mask) m4 = *reinterpret_cast<const bf16x4*>(a.mask + o + co);
      bf16x4 v4;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = acc[pb][nb][i] + (a.bias ? a.bias[co + i] : 0.f);
        if (RELU) v = v > 0.f ? v : 0.f;
        if (a.mask && !((float)m4[i] > 0.f)) v = 0.f;
        v4[i] = (bf16)v;
      }
      *reinterpret_cast<bf16x4*>(a.out + o + co) = v4;
    }
  }
}

// Packs layer weights W[tap][ci][co] (ci < cin_real) into MFMA A fragments of the convolution
// (taps, cin, cout); transposed: the convolution from the layer's output channels to its input
// channels with the taps reversed.
__global__ void k_pack(const float* W, int taps, int cin, int cout, int lcin_real, int lcout,
                       int transposed, bf16x8* out) {
  const int nb_count = cout / 16, steps = ksteps(cin, taps);
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= steps * nb_count * 64) return;
  const int lane = idx % 64, nb = (idx / 64) % nb_count, s = idx / (64 * nb_count);
  const int g = lane >> 4, co = nb * 16 + (lane & 15);
  int tap, ch;
  kmap(cin, s, g, tap, ch);
  bf16x8 v;
  for (int e = 0; e < 8; ++e) {
    const int ci = ch + e;
    float w = 0.f;
    if (tap < taps) {
      if (!transposed) {
        if (ci < lcin_real && co < lcout) w = W[((size_t)tap * lcin_real + ci) * lcout + co];
      } else {
        // packed conv: input channels = the layer's outputs (ci), outputs = its inputs (co)
        if (co < lcin_real && ci < lcout)
          w = W[((size_t)(taps - 1 - tap) * lcin_real + co) * lcout + ci];
      }
    }
    v[e] = (bf16)w;
  }
  out[idx] = v;
}

__device__ inline size_t pix(int y, int x, int W, int C) {
  return ((size_t)(y + 1) * (W + 2) + x + 1) * C;
}

// 2x2 average: out (H/2 x W/2) from in (H x W).
__global__ void k_pool(const bf16* in, int H, int W, int C, bf16* out, const int* skip) {
  if (skip && *skip) return;
  const int c8 = C / 8;
  const long long n = (long long)(H / 2) * (W / 2) * c8;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int c = (int)(t % c8) * 8;
  const int x = (int)((t / c8) % (W / 2)), y = (int)(t / c8 / (W / 2));
  const bf16x8 a = *reinterpret_cast<const bf16x8*>(in + pix(2 * y, 2 * x, W, C) + c);
  const bf16x8 b = *reinterpret_cast<const bf16x8*>(in + pix(2 * y, 2 * x + 1, W, C) + c);
  const bf16x8 d = *reinterpret_cast<const bf16x8*>(in + pix(2 * y + 1, 2 * x, W, C) + c);
  const bf16x8 e = *reinterpret_cast<const bf16x8*>(in + pix(2 * y + 1, 2 * x + 1, W, C) + c);
  bf16x8 o;
  for (int i = 0; i < 8; ++i)
    o[i] = (bf16)((((float)a[i] + (float)b[i]) + ((float)d[i] + (float)e[i])) * 0.25f);
  *reinterpret_cast<bf16x8*>(out + pix(y, x, W / 2, C) + c) = o;
}

// out = up2(t) + e at level size H x W (t at H/2 x W/2).
__global__ void k_upadd(const bf16* t, const bf16* e, int H, int W, int C, bf16* out,
                        const int* skip) {
  if (skip && *skip) return;
  const int c8 = C / 8;
  const long long n = (long long)H * W * c8;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % c8) * 8;
  const int x = (int)((i / c8) % W), y = (int)(i / c8 / W);
  const bf16x8 a = *reinterpret_cast<const bf16x8*>(t + pix(y / 2, x / 2, W / 2, C) + c);
  const bf16x8 b = *reinterpret_cast<const bf16x8*>(e + pix(y, x, W, C) + c);
  bf16x8 o;
  for (int k = 0; k < 8; ++k) o[k] = (bf16)((float)a[k] + (float)b[k]);
  *reinterpret_cast<bf16x8*>(out + pix(y, x, W, C) + c) = o;
}

// dt (H/2 x W/2) = sum of du over each 2x2 block (the backward of up2).
__global__ void k_up_bwd(const bf16* du, int H, int W, int C, bf16* dt) {
  const int c8 = C / 8;
  const long long n = (long long)(H / 2) * (W / 2) * c8;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int c = (int)(t % c8) * 8;
  const int x = (int)((t / c8) % (W / 2)), y = (int)(t / c8 / (W / 2));
  const bf16x8 a = *reinterpret_cast<const bf16x8*>(du + pix(2 * y, 2 * x, W, C) + c);
  const bf16x8 b = *reinterpret_cast<const bf16x8*>(du + pix(2 * y, 2 * x + 1, W, C) + c);
  const bf16x8 d = *reinterpret_cast<const bf16x8*>(du + pix(2 * y + 1, 2 * x, W, C) + c);
  const bf16x8 e = *reinterpret_cast<const bf16x8*>(du + pix(2 * y + 1, 2 * x + 1, W, C) + c);
  bf16x8 o;
  for (int i = 0; i < 8; ++i)
    o[i] = (bf16)(((float)a[i] + (float)b[i]) + ((float)d[i] + (float)e[i]));
  *reinterpret_cast<bf16x8*>(dt + pix(y, x, W / 2, C) + c) = o;
}

// d pre-activation of an encoder layer: (skip + dp / 4 (the pooling's backward)) where act > 0.
__global__ void k_pool_bwd(const bf16* dp, const bf16* skip, const bf16* act, int H, int W,
                           int C, bf16* dx) {
  const int c8 = C / 8;
  const long long n = (long long)H * W * c8;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % c8) * 8;
  const int x = (int)((i / c8) % W), y = (int)(i / c8 / W);
  const bf16x8 p = *reinterpret_cast<const bf16x8*>(dp + pix(y / 2, x / 2, W / 2, C) + c);
  const bf16x8 s = *reinterpret_cast<const bf16x8*>(skip + pix(y, x, W, C) + c);
  const bf16x8 a = *reinterpret_cast<const bf16x8*>(act + pix(y, x, W, C) + c);
  bf16x8 o;
  for (int k = 0; k < 8; ++k)
    o[k] = (bf16)((float)a[k] > 0.f ? (float)s[k] + (float)p[k] * 0.25f : 0.f);
  *reinterpret_cast<bf16x8*>(dx + pix(y, x, W, C) + c) = o;
}

// The input channels from grid-layout arrays; zero outside the nx x ny cells.
__global__ void k_input(int H, int W, int nx, int ny, int pitch, int ghost, const float* h,
                        const float* bgx, const float* bgy, const float* rx, const float* ry,
                        float hs, float rs, bf16* x0) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)H * W) return;
  const int x = (int)(i % W), y = (int)(i / W);
  bf16x8 o;
  for (int k = 0; k < 8; ++k) o[k] = (bf16)0.f;
  if (x < nx && y < ny) {
    const size_t m = (size_t)(y + ghost) * pitch + x + ghost;
    o[0] = (bf16)(h[m] * hs);
    o[1] = (bf16)bgx[m];
    o[2] = (bf16)bgy[m];
    o[3] = (bf16)(rx[m] * rs);
    o[4] = (bf16)(ry[m] * rs);
  }
  *reinterpret_cast<bf16x8*>(x0 + pix(y, x, W, 8)) = o;
}

// The output layer O (1x1, 16 -> 2) on the vector units: out = b + W^T y0 at the nx x ny cells.
__global__ void k_out(const bf16* y0, int W, int nx, int ny, const float* wo, float* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nx * ny) return;
  const int x = (int)(i % nx), y = (int)(i / nx);
  const bf16x8 a = *reinterpret_cast<const bf16x8*>(y0 + pix(y, x, W, 16));
  const bf16x8 b = *reinterpret_cast<const bf16x8*>(y0 + pix(y, x, W, 16) + 8);
  float o0 = wo[32], o1 = wo[33];
  for (int c = 0; c < 8; ++c) {
    o0 += (float)a[c] * wo[2 * c];
    o1 += (float)a[c] * wo[2 * c + 1];
  }
  for (int c = 0; c < 8; ++c) {
    o0 += (float)b[c] * wo[2 * (c + 8)];
    o1 += (float)b[c] * wo[2 * (c + 8) + 1];
  }
  out[2 * i] = o0;
  out[2 * i + 1] = o1;
}

// Block sums in double of weight and weight * |out - target|^2 over the nx x ny cells.
__global__ void __launch_bounds__(256) k_loss_sums(const float* out, const float* target,
                                                   const float* weight, long long n,
                                                   double* part) {
  __shared__ double s[2][256];
  double sw = 0.0, se = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const double w = weight[i];
    const double d0 = (double)out[2 * i] - target[2 * i], d1 = (double)out[2 * i + 1] - target[2 * i + 1];
    sw += w;
    se += w * (d0 * d0 + d1 * d1);
  }
  s[0][threadIdx.x] = sw;
  s[1][threadIdx.x] = se;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] += s[0][threadIdx.x + h];
      s[1][threadIdx.x] += s[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s[0][0];
    part[2 * blockIdx.x + 1] = s[1][0];
  }
}

// d loss / d out = 2 weight (out - target) / sum(weight).
__global__ void k_loss_grad(const float* out, const float* target, const float* weight,
                            long long n, const double* part, int nblk, float* dout) {
  __shared__ double inv;
  if (threadIdx.x == 0) {
    double sw = 0.0;
    for (int b = 0; b < nblk; ++b) sw += part[2 * b];
    inv = sw > 0.0 ? 1.0 / sw : 0.0;
  }
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = (float)(2.0 * inv) * weight[i];
  dout[2 * i] = f * (out[2 * i] - target[2 * i]);
  dout[2 * i + 1] = f * (out[2 * i + 1] - target[2 * i + 1]);
}

// Backward of O: d pre-activation of D0 (g_y0 = W dout where y0 > 0) per cell, and per block the
// partial sums of y0[ci] dout[co] (32) and dout[co] (2) for O's weight and bias gradients.
__global__ void __launch_bounds__(256) k_out_bwd(const bf16* y0, int W, int nx, int ny,
                                                 const float* wo, const float* dout, bf16* gy0,
                                                 float* part) {
  __shared__ float s[34][256];
  float acc[34];
  for (int m = 0; m < 34; ++m) acc[m] = 0.f;
  const long long n = (long long)nx * ny;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const int x = (int)(i % nx), y = (int)(i / nx);
    const size_t o = pix(y, x, W, 16);
    const float d0 = dout[2 * i], d1 = dout[2 * i + 1];
    bf16x8 g[2];
    for (int h = 0; h < 2; ++h) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(y0 + o + 8 * h);
      for (int c = 0; c < 8; ++c) {
        const int ci = 8 * h + c;
        const float yv = (float)a[c];
        acc[2 * ci] += yv * d0;
        acc[2 * ci + 1] += yv * d1;
        const float gv = wo[2 * ci] * d0 + wo[2 * ci + 1] * d1;
        g[h][c] = (bf16)(yv > 0.f ? gv : 0.f);
      }
      *reinterpret_cast<bf16x8*>(gy0 + o + 8 * h) = g[h];
    }
    acc[32] += d0;
    acc[33] += d1;
  }
  for (int m = 0; m < 34; ++m) s[m][threadIdx.x] = acc[m];
  __syncthreads();
  if (threadIdx.x < 34) {
    float t = 0.f;
    for (int k = 0; k < 256; ++k) t += s[threadIdx.x][k];
    part[blockIdx.x * 34 + threadIdx.x] = t;
  }
}

// Adds the per-block partials (nblk x m) in block order into out[0..m).
__global__ void k_sum_parts(const float* part, int nblk, int m, float* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  float t = 0.f;
  for (int b = 0; b < nblk; ++b) t += part[(size_t)b * m + j];
  out[j] = t;
}

// Bias gradient partials: per block, per channel, the sum of dy over a chunk of pixels.
__global__ void __launch_bounds__(256) k_bias_part(const bf16* dy, int H, int W, int C,
                                                   int px_per_block, float* part) {
  __shared__ float s[256];
  const int lanes = 256 / C;  // pixel lanes
  const int c = threadIdx.x % C, l = threadIdx.x / C;
  const long long p0 = (long long)blockIdx.x * px_per_block;
  const long long p1 = std::min<long long>(p0 + px_per_block, (long long)H * W);
  float t = 0.f;
  for (long long p = p0 + l; p < p1; p += lanes) {
    const int x = (int)(p % W), y = (int)(p / W);
    t += (float)dy[pix(y, x, W, C) + c];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  if ((int)threadIdx.x < C) {
    float u = 0.f;
    for (int k = 0; k < lanes; ++k) u += s[k * C + threadIdx.x];
    part[(size_t)blockIdx.x * C + threadIdx.x] = u;
  }
}

// Weight-gradient partials of a convolution (TAPS, CIN, COUT): block (chunk, tap) adds
// x[p + d_tap][ci] dy[p][co] over its chunk's pixels into part[chunk][tap][co][ci16].
template <int TAPS, int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(const bf16* x, const bf16* dy, int H, int W,
                                               float* part) {
  constexpr int CI = CIN < 16 ? 16 : CIN;
  constexpr int NCO = COUT / 16, NCI = CI / 16, NTILE = NCO * NCI, TPW = (NTILE + 3) / 4;
  constexpr int RS = 40;  // LDS row: 32 pixels + padding, 80 bytes
  __shared__ __attribute__((aligned(16))) bf16 sdy[COUT][RS];
  __shared__ __attribute__((aligned(16))) bf16 sx[CI][RS];
  const int chunk = blockIdx.x, tap = blockIdx.y;
  const int dy_ = TAPS == 9 ? tap / 3 - 1 : 0, dx_ = TAPS == 9 ? tap % 3 - 1 : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
  const int npx = H * W;
  if (CIN < CI) {
    for (int t = threadIdx.x; t < (CI - CIN) * RS; t += 256) sx[CIN + t / RS][t % RS] = (bf16)0.f;
  }
  f32x4 acc[TPW];
  for (int k = 0; k < TPW; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < CHUNK_STEPS; ++ks) {
    const int px0 = (chunk * CHUNK_STEPS + ks) * 32;
    if (px0 >= npx) break;
    __syncthreads();
    // 32 pixels in row-major order from px0 (zero past the last pixel)
    for (int v = threadIdx.x; v < 32 * (COUT / 8); v += 256) {
      const int p = v / (COUT / 8), c8 = (v % (COUT / 8)) * 8, pa = px0 + p;
      bf16x8 val;
      if (pa < npx) {
        val = *reinterpret_cast<const bf16x8*>(dy + pix(pa / W, pa % W, W, COUT) + c8);
      } else {
        for (int e = 0; e < 8; ++e) val[e] = (bf16)0.f;
      }
      for (int e = 0; e < 8; ++e) sdy[c8 + e][p] = val[e];
    }
    for (int v = threadIdx.x; v < 32 * (CIN / 8); v += 256) {
      const int p = v / (CIN / 8), c8 = (v % (CIN / 8)) * 8, pa = px0 + p;
      bf16x8 val;
      if (pa < npx) {
        val = *reinterpret_cast<const bf16x8*>(
            x + pix(pa / W + dy_, pa % W + dx_, W, CIN) + c8);
      } else {
        for (int e = 0; e < 8; ++e) val[e] = (bf16)0.f;
      }
      for (int e = 0; e < 8; ++e) sx[c8 + e][p] = val[e];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
      const int tile = wave * TPW + k;
      if (tile < NTILE) {
        const int cb = tile / NCI, ib = tile % NCI;
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sdy[cb * 16 + c16][8 * g]);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sx[ib * 16 + c16][8 * g]);
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
      }
    }
  }
  float* out = part + ((size_t)chunk * TAPS + tap) * COUT * CI;
  for (int k = 0; k < TPW; ++k) {
    const int tile = wave * TPW + k;
    if (tile >= NTILE) continue;
    const int cb = tile / NCI, ib = tile % NCI;
    for (int i = 0; i < 4; ++i) out[(size_t)(cb * 16 + 4 * g + i) * CI + ib * 16 + c16] = acc[k][i];
  }
}

// dW[tap][ci][co] (ci < cin_real) = sum over chunks, in chunk order, of part[chunk][tap][co][ci].
__global__ void k_wgrad_sum(const float* part, int nchunks, int taps, int ci16, int cout,
                            int cin_real, float* dw) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (long long)taps * cin_real * cout) return;
  const int co = (int)(j % cout), ci = (int)((j / cout) % cin_real), tap = (int)(j / cout / cin_real);
  float t = 0.f;
  for (int c = 0; c < nchunks; ++c) t += part[(((size_t)c * taps + tap) * cout + co) * ci16 + ci];
  dw[j] = t;
}

__global__ void k_adam(long long n, float* p, const float* gr, float* m, float* v, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- host side --------------------------------------------------------------------------------
#define HIPC(x)                                                        \
  do {                                                                 \
    hipError_t e_ = (x);                                               \
    if (e_ != hipSuccess) return wc::api_fail(WC_EDEVICE, hipGetErrorString(e_)); \
  } while (0)

inline dim3 blocks_for(long long n, int per = 256) { return dim3((unsigned)((n + per - 1) / per)); }

int launch_conv(int taps, int cin, int cout, int relu, const ConvArgs& a, hipStream_t st) {
  const int pb = (a.W % 64 == 0 && a.H * a.W >= 65536) ? 4 : (a.W % 32 == 0 ? 2 : 1);
#define CONV(T, CI, CO, P, R)                                                                \
  if (taps == T && cin == CI && cout == CO && pb == P && relu == R) {                        \
    const long long waves = (long long)a.H * (a.W / (16 * P));                               \
    hipLaunchKernelGGL((k_conv<T, CI, CO, P, R>), blocks_for(waves, 4), dim3(256), 0, st, a); \
    HIPC(hipGetLastError());                                                                 \
    return WC_OK;                                                                            \
  }
#define CONV_ALL(T, CI, CO, R) CONV(T, CI, CO, 1, R) CONV(T, CI, CO, 2, R) CONV(T, CI, CO, 4, R)
  // forward
  CONV_ALL(9, 8, 16, 1) CONV_ALL(9, 16, 32, 1) CONV_ALL(9, 32, 64, 1) CONV_ALL(9, 64, 128, 1)
  CONV_ALL(9, 128, 128, 1) CONV_ALL(1, 128, 64, 0) CONV_ALL(9, 64, 64, 1) CONV_ALL(1, 64, 32, 0)
  CONV_ALL(9, 32, 32, 1) CONV_ALL(1, 32, 16, 0) CONV_ALL(9, 16, 16, 1)
  // transposed (backward data)
  CONV_ALL(9, 16, 16, 0) CONV_ALL(1, 16, 32, 0) CONV_ALL(9, 32, 32, 0) CONV_ALL(1, 32, 64, 0)
  CONV_ALL(9, 64, 64, 0) CONV_ALL(1, 64, 128, 0) CONV_ALL(9, 128, 128, 0) CONV_ALL(9, 128, 64, 0)
  CONV_ALL(9, 64, 32, 0) CONV_ALL(9, 32, 16, 0)
#undef CONV_ALL
#undef CONV
  return wc::api_fail(WC_EINVAL, "no convolution kernel for this shape");
}

int launch_wgrad(int taps, int cin, int cout, const bf16* x, const bf16* dy, int H, int W,
                 float* part, int nchunks, hipStream_t st) {
#define WG(T, CI, CO)                                                                    \
  if (taps == T && cin == CI && cout == CO) {                                            \
    hipLaunchKernelGGL((k_wgrad<T, CI, CO>), dim3(nchunks, T), dim3(256), 0, st, x, dy, H, W, \
                       part);                                                            \
    HIPC(hipGetLastError());                                                             \
    return WC_OK;                                                                        \
  }
  WG(9, 8, 16) WG(9, 16, 32) WG(9, 32, 64) WG(9, 64, 128) WG(9, 128, 128) WG(1, 128, 64)
  WG(9, 64, 64) WG(1, 64, 32) WG(9, 32, 32) WG(1, 32, 16) WG(9, 16, 16)
#undef WG
  return wc::api_fail(WC_EINVAL, "no weight-gradient kernel for this shape");
}

}  // namespace wcu

using namespace wcu;

struct wcu_net {
  int nx = 0, ny = 0;
  int H[4] = {}, W[4] = {};
  bf16* act[NTENSORS] = {};
  bf16* grad[NTENSORS] = {};
  bf16x8* wpk[NLAYERS] = {};
  bf16x8* wpkT[NLAYERS] = {};
  float* params = nullptr;  // own copy
  float* out = nullptr;     // 2 nx ny
  float* dout = nullptr;
  float* ws = nullptr;      // weight-gradient partials
  size_t ws_len = 0;
  float* part = nullptr;    // block partials
  double* lpart = nullptr;  // loss partials
  double* lpart_host = nullptr;
  int lblocks = 0;
  std::vector<void*> owned;

  size_t tensor_elems(int t) const {
    const TensorDef& d = TENSORS[t];
    return (size_t)(H[d.level] + 2) * (W[d.level] + 2) * d.c;
  }
};

namespace {

template <class T>
int nalloc(wcu_net* n, T** p, size_t count) {
  void* v = nullptr;
  HIPC(hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)));
  HIPC(hipMemset(v, 0, std::max<size_t>(count, 1) * sizeof(T)));
  n->owned.push_back(v);
  *p = static_cast<T*>(v);
  return WC_OK;
}

int conv_layer(wcu_net* n, int l, Tensor in, Tensor out, hipStream_t st,
               const int* skip = nullptr) {
  const LayerDef& d = LAYERS[l];
  ConvArgs a{n->act[in], n->H[d.level], n->W[d.level], n->wpk[l],
             n->params + layer_offset(l) + (size_t)d.taps * d.cin_real * d.cout, n->act[out],
             nullptr, skip};
  return launch_conv(d.taps, d.cin, d.cout, d.relu, a, st);
}

// d input = the transposed convolution of d output (d pre-activation of layer l), optionally
// kept where mask > 0.
int conv_back(wcu_net* n, int l, const bf16* dy, bf16* dx, const bf16* mask, hipStream_t st) {
  const LayerDef& d = LAYERS[l];
  ConvArgs a{dy, n->H[d.level], n->W[d.level], n->wpkT[l], nullptr, dx, mask, nullptr};
  return launch_conv(d.taps, d.cout, d.cin, 0, a, st);
}

int wgrad_layer(wcu_net* n, int l, const bf16* x, const bf16* dy, float* grads,
                hipStream_t st) {
  const LayerDef& d = LAYERS[l];
  const int H = n->H[d.level], W = n->W[d.level];
  const int npx = H * W, nchunks = (npx + 32 * CHUNK_STEPS - 1) / (32 * CHUNK_STEPS);
  const int ci16 = d.cin < 16 ? 16 : d.cin;
  int rc = launch_wgrad(d.taps, d.cin, d.cout, x, dy, H, W, n->ws, nchunks, st);
  if (rc) return rc;
  float* dw = grads + layer_offset(l);
  const long long m = (long long)d.taps * d.cin_real * d.cout;
  hipLaunchKernelGGL(k_wgrad_sum, blocks_for(m), dim3(256), 0, st, (const float*)n->ws, nchunks,
                     d.taps, ci16, d.cout, d.cin_real, dw);
  HIPC(hipGetLastError());
  // bias
  const int ppb = 1024, nblk = (npx + ppb - 1) / ppb;
  hipLaunchKernelGGL(k_bias_part, dim3(nblk), dim3(256), 0, st, dy, H, W, d.cout, ppb, n->part);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_sum_parts, blocks_for(d.cout), dim3(256), 0, st, (const float*)n->part,
                     nblk, d.cout, dw + m);
  HIPC(hipGetLastError());
  return WC_OK;
}

int elementwise_pool(wcu_net* n, Tensor in, Tensor out, hipStream_t st,
                     const int* skip = nullptr) {
  const TensorDef& d = TENSORS[in];
  const int H = n->H[d.level], W = n->W[d.level];
  hipLaunchKernelGGL(k_pool, blocks_for((long long)(H / 2) * (W / 2) * (d.c / 8)), dim3(256), 0,
                     st, (const bf16*)n->act[in], H, W, d.c, n->act[out], skip);
  HIPC(hipGetLastError());
  return WC_OK;
}

int elementwise_upadd(wcu_net* n, Tensor t, Tensor e, Tensor out, hipStream_t st,
                      const int* skip = nullptr) {
  const TensorDef& d = TENSORS[e];
  const int H = n->H[d.level], W = n->W[d.level];
  hipLaunchKernelGGL(k_upadd, blocks_for((long long)H * W * (d.c / 8)), dim3(256), 0, st,
                     (const bf16*)n->act[t], (const bf16*)n->act[e], H, W, d.c, n->act[out],
                     skip);
  HIPC(hipGetLastError());
  return WC_OK;
}

}  // namespace

namespace wc {
// The network's forward pass for the solver's preconditioner (wc_api.hip).
int unet_forward(wcu_net* n, float* out_xy, hipStream_t st);
}

extern "C" {

uint64_t wcu_param_count(void) { return layer_offset(NLAYERS); }

int wcu_layer(int32_t l, uint64_t* offset, uint64_t* weights, uint64_t* biases, int32_t* taps,
              int32_t* cin, int32_t* cout) {
  if (l < 0 || l >= NLAYERS) return wc::api_fail(WC_EINVAL, "layer index out of range");
  const LayerDef& d = LAYERS[l];
  if (offset) *offset = layer_offset(l);
  if (weights) *weights = (uint64_t)d.taps * d.cin_real * d.cout;
  if (biases) *biases = d.cout;
  if (taps) *taps = d.taps;
  if (cin) *cin = d.cin_real;
  if (cout) *cout = d.cout;
  return WC_OK;
}

int wcu_net_new(uint64_t nx, uint64_t ny, wcu_net** out) {
  if (!out) return wc::api_fail(WC_EINVAL, "null output");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nx > (1u << 16) || ny > (1u << 16))
    return wc::api_fail(WC_EINVAL, "network grid size out of range");
  auto n = std::make_unique<wcu_net>();
  n->nx = (int)nx;
  n->ny = (int)ny;
  n->W[0] = (int)((nx + 127) / 128 * 128);
  n->H[0] = (int)((ny + 7) / 8 * 8);
  for (int l = 1; l < 4; ++l) {
    n->W[l] = n->W[l - 1] / 2;
    n->H[l] = n->H[l - 1] / 2;
  }
  int rc;
  for (int t = 0; t < NTENSORS; ++t) {
    if ((rc = nalloc(n.get(), &n->act[t], n->tensor_elems(t)))) return rc;
    if (t != X0 && (rc = nalloc(n.get(), &n->grad[t], n->tensor_elems(t)))) return rc;
  }
  size_t ws = 0;
  for (int l = 0; l < NLAYERS; ++l) {
    const LayerDef& d = LAYERS[l];
    if (l != L_O) {
      if ((rc = nalloc(n.get(), &n->wpk[l], (size_t)ksteps(d.cin, d.taps) * (d.cout / 16) * 64)))
        return rc;
      if (l != L_E0 && (rc = nalloc(n.get(), &n->wpkT[l],
                                    (size_t)ksteps(d.cout, d.taps) * (d.cin / 16) * 64)))
        return rc;
    }
    const int npx = n->H[d.level] * n->W[d.level];
    const size_t nch = (npx + 32 * CHUNK_STEPS - 1) / (32 * CHUNK_STEPS);
    ws = std::max(ws, nch * d.taps * d.cout * (d.cin < 16 ? 16 : d.cin));
  }
  if ((rc = nalloc(n.get(), &n->ws, ws))) return rc;
  n->ws_len = ws;
  if ((rc = nalloc(n.get(), &n->params, wcu_param_count()))) return rc;
  if ((rc = nalloc(n.get(), &n->out, 2 * nx * ny))) return rc;
  if ((rc = nalloc(n.get(), &n->dout, 2 * nx * ny))) return rc;
  const size_t npx0 = (size_t)n->H[0] * n->W[0];
  if ((rc = nalloc(n.get(), &n->part, std::max<size_t>((npx0 / 1024 + 1) * 128, 4096 * 34))))
    return rc;
  n->lblocks = 1024;
  if ((rc = nalloc(n.get(), &n->lpart, 2 * (size_t)n->lblocks))) return rc;
  HIPC(hipHostMalloc(&n->lpart_host, 2 * (size_t)n->lblocks * sizeof(double), hipHostMallocDefault));
  // The zero fills above run on the null stream, which a caller's non-blocking stream does not
  // wait for: finish them before the network is used on any stream (the activations' zero
  // borders are read by the first forward pass).
  HIPC(hipDeviceSynchronize());
  *out = n.release();
  return WC_OK;
}

void wcu_net_free(wcu_net* n) {
  if (!n) return;
  (void)hipDeviceSynchronize();
  for (void* p : n->owned) (void)hipFree(p);
  if (n->lpart_host) (void)hipHostFree(n->lpart_host);
  delete n;
}

int wcu_net_set_params(wcu_net* n, const float* params, void* stream) {
  if (!n || !params) return wc::api_fail(WC_EINVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  HIPC(hipMemcpyAsync(n->params, params, wcu_param_count() * sizeof(float),
                      hipMemcpyDeviceToDevice, st));
  for (int l = 0; l < NLAYERS; ++l) {
    if (l == L_O) continue;
    const LayerDef& d = LAYERS[l];
    const float* W = n->params + layer_offset(l);
    const long long cnt = (long long)ksteps(d.cin, d.taps) * (d.cout / 16) * 64;
    hipLaunchKernelGGL(k_pack, blocks_for(cnt), dim3(256), 0, st, W, d.taps, d.cin, d.cout,
                       d.cin_real, d.cout, 0, n->wpk[l]);
    HIPC(hipGetLastError());
    if (l == L_E0) continue;
    const long long cntT = (long long)ksteps(d.cout, d.taps) * (d.cin / 16) * 64;
    hipLaunchKernelGGL(k_pack, blocks_for(cntT), dim3(256), 0, st, W, d.taps, d.cout, d.cin,
                       d.cin_real, d.cout, 1, n->wpkT[l]);
    HIPC(hipGetLastError());
  }
  return WC_OK;
}

int wcu_net_input(wcu_net* n, uint64_t nx, uint64_t ny, uint64_t pitch, uint64_t ghost,
                  const float* h, const float* bgx, const float* bgy, const float* rx,
                  const float* ry, float h_scale, float r_scale, void* stream) {
  if (!n || nx > (uint64_t)n->nx || ny > (uint64_t)n->ny)
    return wc::api_fail(WC_EINVAL, "input larger than the network instance");
  hipLaunchKernelGGL(k_input, blocks_for((long long)n->H[0] * n->W[0]), dim3(256), 0,
                     (hipStream_t)stream, n->H[0], n->W[0], (int)nx, (int)ny, (int)pitch,
                     (int)ghost, h, bgx, bgy, rx, ry, h_scale, r_scale, n->act[X0]);
  HIPC(hipGetLastError());
  return WC_OK;
}

int wcu_net_forward(wcu_net* n, float* out_xy, void* stream) {
  if (!n) return wc::api_fail(WC_EINVAL, "null network");
  return wc::unet_forward(n, out_xy, (hipStream_t)stream);
}

int wcu_net_loss(wcu_net* n, const float* target, const float* weight, double* loss_host,
                 void* stream) {
  if (!n || !target || !weight) return wc::api_fail(WC_EINVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const long long cells = (long long)n->nx * n->ny;
  hipLaunchKernelGGL(k_loss_sums, dim3(n->lblocks), dim3(256), 0, st, (const float*)n->out,
                     target, weight, cells, n->lpart);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_loss_grad, blocks_for(cells), dim3(256), 0, st, (const float*)n->out,
                     target, weight, cells, (const double*)n->lpart, n->lblocks, n->dout);
  HIPC(hipGetLastError());
  if (loss_host) {
    HIPC(hipMemcpyAsync(n->lpart_host, n->lpart, 2 * (size_t)n->lblocks * sizeof(double),
                        hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    double sw = 0.0, se = 0.0;
    for (int b = 0; b < n->lblocks; ++b) {
      sw += n->lpart_host[2 * b];
      se += n->lpart_host[2 * b + 1];
    }
    *loss_host = sw > 0.0 ? se / sw : 0.0;
  }
  return WC_OK;
}

int wcu_net_backward(wcu_net* n, float* grads, void* stream) {
  if (!n || !grads) return wc::api_fail(WC_EINVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  // O
  {
    const int nblk = 1024;
    const float* wo = n->params + layer_offset(L_O);
    hipLaunchKernelGGL(k_out_bwd, dim3(nblk), dim3(256), 0, st, (const bf16*)n->act[Y0], n->W[0],
                       n->nx, n->ny, wo, (const float*)n->dout, n->grad[Y0], n->part);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, (const float*)n->part, nblk, 34,
                       grads + layer_offset(L_O));
    HIPC(hipGetLastError());
  }
  // a decoder level: 3x3 layer D (input U, output Y, d pre in grad[Y]), its 1x1 T below it
  auto up_bwd = [&](Tensor du, Tensor dt) -> int {
    const TensorDef& d = TENSORS[du];
    const int H = n->H[d.level], W = n->W[d.level];
    hipLaunchKernelGGL(k_up_bwd, blocks_for((long long)(H / 2) * (W / 2) * (d.c / 8)), dim3(256),
                       0, st, (const bf16*)n->grad[du], H, W, d.c, n->grad[dt]);
    HIPC(hipGetLastError());
    return WC_OK;
  };
  auto pool_bwd = [&](Tensor dp, Tensor skip, Tensor act) -> int {
    const TensorDef& d = TENSORS[act];
    const int H = n->H[d.level], W = n->W[d.level];
    hipLaunchKernelGGL(k_pool_bwd, blocks_for((long long)H * W * (d.c / 8)), dim3(256), 0, st,
                       (const bf16*)n->grad[dp], (const bf16*)n->grad[skip],
                       (const bf16*)n->act[act], H, W, d.c, n->grad[act]);
    HIPC(hipGetLastError());
    return WC_OK;
  };
  // D0
  if ((rc = wgrad_layer(n, L_D0, n->act[U0], n->grad[Y0], grads, st))) return rc;
  if ((rc = conv_back(n, L_D0, n->grad[Y0], n->grad[U0], nullptr, st))) return rc;
  if ((rc = up_bwd(U0, T1))) return rc;
  // T1, D1
  if ((rc = wgrad_layer(n, L_T1, n->act[Y1], n->grad[T1], grads, st))) return rc;
  if ((rc = conv_back(n, L_T1, n->grad[T1], n->grad[Y1], n->act[Y1], st))) return rc;
  if ((rc = wgrad_layer(n, L_D1, n->act[U1], n->grad[Y1], grads, st))) return rc;
  if ((rc = conv_back(n, L_D1, n->grad[Y1], n->grad[U1], nullptr, st))) return rc;
  if ((rc = up_bwd(U1, T2))) return rc;
  // T2, D2
  if ((rc = wgrad_layer(n, L_T2, n->act[Y2], n->grad[T2], grads, st))) return rc;
  if ((rc = conv_back(n, L_T2, n->grad[T2], n->grad[Y2], n->act[Y2], st))) return rc;
  if ((rc = wgrad_layer(n, L_D2, n->act[U2], n->grad[Y2], grads, st))) return rc;
  if ((rc = conv_back(n, L_D2, n->grad[Y2], n->grad[U2], nullptr, st))) return rc;
  if ((rc = up_bwd(U2, T3))) return rc;
  // T3, B, E3
  if ((rc = wgrad_layer(n, L_T3, n->act[Y3], n->grad[T3], grads, st))) return rc;
  if ((rc = conv_back(n, L_T3, n->grad[T3], n->grad[Y3], n->act[Y3], st))) return rc;
  if ((rc = wgrad_layer(n, L_B, n->act[E3], n->grad[Y3], grads, st))) return rc;
  if ((rc = conv_back(n, L_B, n->grad[Y3], n->grad[E3], n->act[E3], st))) return rc;
  if ((rc = wgrad_layer(n, L_E3, n->act[P3], n->grad[E3], grads, st))) return rc;
  if ((rc = conv_back(n, L_E3, n->grad[E3], n->grad[P3], nullptr, st))) return rc;
  // E2, E1, E0
  if ((rc = pool_bwd(P3, U2, E2))) return rc;
  if ((rc = wgrad_layer(n, L_E2, n->act[P2], n->grad[E2], grads, st))) return rc;
  if ((rc = conv_back(n, L_E2, n->grad[E2], n->grad[P2], nullptr, st))) return rc;
  if ((rc = pool_bwd(P2, U1, E1))) return rc;
  if ((rc = wgrad_layer(n, L_E1, n->act[P1], n->grad[E1], grads, st))) return rc;
  if ((rc = conv_back(n, L_E1, n->grad[E1], n->grad[P1], nullptr, st))) return rc;
  if ((rc = pool_bwd(P1, U0, E0))) return rc;
  if ((rc = wgrad_layer(n, L_E0, n->act[X0], n->grad[E0], grads, st))) return rc;
  return WC_OK;
}

int wcu_adam(uint64_t count, float* params, const float* grads, float* m, float* v, float lr,
             float beta1, float beta2, float eps, int32_t t, void* stream) {
  if (t < 1) return wc::api_fail(WC_EINVAL, "Adam step must be >= 1");
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, blocks_for((long long)count), dim3(256), 0, (hipStream_t)stream,
                     (long long)count, params, grads, m, v, lr, beta1, beta2, eps, c1, c2);
  HIPC(hipGetLastError());
  return WC_OK;
}

int wcu_net_tensor(wcu_net* n, int32_t id, float* host, uint64_t* h, uint64_t* w, uint64_t* c) {
  if (!n) return wc::api_fail(WC_EINVAL, "null network");
  const bool is_grad = id >= 100;
  const int t = is_grad ? id - 100 : id;
  if (t < 0 || t >= NTENSORS || (is_grad && !n->grad[t]))
    return wc::api_fail(WC_EINVAL, "no such tensor");
  const TensorDef& d = TENSORS[t];
  const int H = n->H[d.level], W = n->W[d.level];
  if (h) *h = H;
  if (w) *w = W;
  if (c) *c = d.c;
  if (!host) return WC_OK;
  std::vector<uint16_t> raw(n->tensor_elems(t));
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(raw.data(), is_grad ? n->grad[t] : n->act[t], raw.size() * 2,
                 hipMemcpyDeviceToHost));
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      for (int k = 0; k < d.c; ++k) {
        const uint32_t bits = (uint32_t)raw[((size_t)(y + 1) * (W + 2) + x + 1) * d.c + k] << 16;
        float f;
        memcpy(&f, &bits, 4);
        host[((size_t)y * W + x) * d.c + k] = f;
      }
  return WC_OK;
}

}  // extern "C"

namespace wc {

// E0 .. D0 (everything but the output layer); every kernel does nothing while *skip != 0.
int unet_body(wcu_net* n, hipStream_t st, const int* skip) {
  int rc;
  if ((rc = conv_layer(n, L_E0, X0, E0, st, skip))) return rc;
  if ((rc = elementwise_pool(n, E0, P1, st, skip))) return rc;
  if ((rc = conv_layer(n, L_E1, P1, E1, st, skip))) return rc;
  if ((rc = elementwise_pool(n, E1, P2, st, skip))) return rc;
  if ((rc = conv_layer(n, L_E2, P2, E2, st, skip))) return rc;
  if ((rc = elementwise_pool(n, E2, P3, st, skip))) return rc;
  if ((rc = conv_layer(n, L_E3, P3, E3, st, skip))) return rc;
  if ((rc = conv_layer(n, L_B, E3, Y3, st, skip))) return rc;
  if ((rc = conv_layer(n, L_T3, Y3, T3, st, skip))) return rc;
  if ((rc = elementwise_upadd(n, T3, E2, U2, st, skip))) return rc;
  if ((rc = conv_layer(n, L_D2, U2, Y2, st, skip))) return rc;
  if ((rc = conv_layer(n, L_T2, Y2, T2, st, skip))) return rc;
  if ((rc = elementwise_upadd(n, T2, E1, U1, st, skip))) return rc;
  if ((rc = conv_layer(n, L_D1, U1, Y1, st, skip))) return rc;
  if ((rc = conv_layer(n, L_T1, Y1, T1, st, skip))) return rc;
  if ((rc = elementwise_upadd(n, T1, E0, U0, st, skip))) return rc;
  return conv_layer(n, L_D0, U0, Y0, st, skip);
}

// Where the solver's preconditioner writes the input and reads the last hidden layer.
void unet_views(wcu_net* n, void** x0, const void** y0, int* w0, int* h0, const float** wo) {
  *x0 = n->act[X0];
  *y0 = n->act[Y0];
  *w0 = n->W[0];
  *h0 = n->H[0];
  *wo = n->params + layer_offset(L_O);
}

int unet_forward(wcu_net* n, float* out_xy, hipStream_t st) {
  int rc;
  if ((rc = unet_body(n, st, nullptr))) return rc;
  hipLaunchKernelGGL(k_out, blocks_for((long long)n->nx * n->ny), dim3(256), 0, st,
                     (const bf16*)n->act[Y0], n->W[0], n->nx, n->ny,
                     (const float*)(n->params + layer_offset(L_O)), n->out);
  HIPC(hipGetLastError());
  if (out_xy && out_xy != n->out)
    HIPC(hipMemcpyAsync(out_xy, n->out, 2 * (size_t)n->nx * n->ny * sizeof(float),
                        hipMemcpyDeviceToDevice, st));
  return WC_OK;
}

}  // namespace wc
