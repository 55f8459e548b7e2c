// Per-atom surrogate energy: a small MLP on SOAP vectors,
//   e(q) = mu + sigma * (w3 . s(W2 s(W1 q + b1) + b2) + b3),  s(x) = x / (1 + e^-x),
// with its inference, gradient and Adam kernels (no reference counterpart:
// it screens Monte Carlo trials before the full GAP evaluation).
//
// Numerics: the descriptor and the two weight matrices enter as BF16 values
// (the FP32 master weights are rounded to BF16 copies after every update and
// q is rounded on load); products are accumulated in FP32.  Gradients are
// reduced over workgroups in a fixed order (no atomics), so training is
// deterministic.
//
// Parameter vector P (FP32 master), layout:
//   W1[H][D], b1[H], W2[H][H], b2[H], w3[H], b3
// BF16 copies: W1t[D][H] and W2t[H][H] (input-major, so that the threads of
// a workgroup, one per hidden unit, read consecutive words), and after W1t
// in the same buffer W1b[H][Dp], Dp = D rounded up to 32, zero-padded
// (unit-major: 16-byte k-fragments for the MFMA inference kernel).
#include <hip/hip_runtime.h>

#include "tgx_internal.h"

namespace tgx {

typedef unsigned short u16;

__device__ __forceinline__ float bf16_round(float f) {
  unsigned int u = __float_as_uint(f);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  return __uint_as_float(u);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned int)h << 16); }
__device__ __forceinline__ u16 f2bf16(float f) {
  unsigned int u = __float_as_uint(f);
  return (u16)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float sigm(float x) { return 1.f / (1.f + __expf(-x)); }

constexpr int MLP_H = 128;  // hidden width
constexpr int MLP_SB = 16;  // sites per workgroup
constexpr int MLP_DMAX = 448;  // static LDS budget of the gradient kernel

__host__ __device__ inline int mlp_dpad(int D) { return (D + 31) / 32 * 32; }

// Forward for n sites; rows of q (T) with leading dimension q_ld, optional
// site row list rows[] (else row = site).  Writes e[n]; with z1/z2 != null
// also the pre-activations (n x H each) for the gradient.
template <typename T>
__global__ __launch_bounds__(256) void mlp_fwd_kernel(int n, int D, const T* __restrict__ q,
                                                      int q_ld, const int* __restrict__ rows,
                                                      const float* __restrict__ P,
                                                      const u16* __restrict__ W1t,
                                                      const u16* __restrict__ W2t,
                                                      const float* __restrict__ norm,
                                                      double* __restrict__ e,
                                                      float* __restrict__ z1o,
                                                      float* __restrict__ z2o,
                                                      const int* __restrict__ guard) {
  if (guard && *guard == 0) return;
  __shared__ float sq[MLP_SB][MLP_DMAX + 1];
  __shared__ float sh1[MLP_SB][MLP_H + 1];
  __shared__ float sh2[MLP_SB][MLP_H + 1];
  const int H = MLP_H;
  const float mu = norm[0], sigma = norm[1];  // read here: captured graphs see updates
  const float* b1 = P + (size_t)H * D;
  const float* b2 = b1 + H + (size_t)H * H;
  const float* w3 = b2 + H;
  const float b3 = w3[H];
  const int t = threadIdx.x;
  const int s0 = blockIdx.x * MLP_SB;
  for (int k = t; k < MLP_SB * D; k += 256) {
    const int s = k / D, d = k - s * D;
    const int site = s0 + s;
    float v = 0.f;
    if (site < n) {
      const int r = rows ? rows[site] : site;
      v = bf16_round((float)q[(size_t)r * q_ld + d]);
    }
    sq[s][d] = v;
  }
  __syncthreads();
  const int u = t & (H - 1), half = t >> 7;  // unit, site half (8 sites each)
  {
    float acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = 0.f;
    for (int k = 0; k < D; k++) {
      const float w = bf2f(W1t[(size_t)k * H + u]);
#pragma unroll
      for (int j = 0; j < 8; j++) acc[j] += w * sq[half * 8 + j][k];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int s = half * 8 + j;
      const float z = acc[j] + b1[u];
      if (z1o && s0 + s < n) z1o[(size_t)(s0 + s) * H + u] = z;
      sh1[s][u] = z * sigm(z);
    }
  }
  __syncthreads();
  {
    float acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = 0.f;
    for (int k = 0; k < H; k++) {
      const float w = bf2f(W2t[(size_t)k * H + u]);
#pragma unroll
      for (int j = 0; j < 8; j++) acc[j] += w * sh1[half * 8 + j][k];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int s = half * 8 + j;
      const float z = acc[j] + b2[u];
      if (z2o && s0 + s < n) z2o[(size_t)(s0 + s) * H + u] = z;
      sh2[s][u] = z * sigm(z);
    }
  }
  __syncthreads();
  // output: wave w handles sites 4w..4w+3, lanes over units
  const int wave = t >> 6, lane = t & 63;
  for (int j = 0; j < 4; j++) {
    const int s = wave * 4 + j;
    float a = w3[lane] * sh2[s][lane] + w3[lane + 64] * sh2[s][lane + 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0 && s0 + s < n) e[s0 + s] = (double)mu + (double)sigma * (double)(a + b3);
  }
}

// Gradient of L = (1/n) sum_s (e_s - y_s)^2 / sigma^2 (the loss in the
// network's normalised units) for the sites of one workgroup, written to
// part[block][NP]; e, z1, z2 from mlp_fwd_kernel.
template <typename T>
__global__ __launch_bounds__(256) void mlp_grad_kernel(int n, int D, const T* __restrict__ q,
                                                       int q_ld, const int* __restrict__ rows,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ e,
                                                       const float* __restrict__ z1,
                                                       const float* __restrict__ z2,
                                                       const float* __restrict__ P,
                                                       const u16* __restrict__ W2t,
                                                       const float* __restrict__ norm,
                                                       float* __restrict__ part) {
  __shared__ float sq[MLP_SB][MLP_DMAX + 1];
  __shared__ float sh1[MLP_SB][MLP_H + 1];   // s(z1)
  __shared__ float sd1[MLP_SB][MLP_H + 1];   // dL/dz1
  __shared__ float sh2[MLP_SB][MLP_H + 1];   // s(z2)
  __shared__ float sd2[MLP_SB][MLP_H + 1];   // dL/dz2
  __shared__ float sg[MLP_SB];
  const int H = MLP_H;
  const float sigma = norm[1];
  const size_t NP = (size_t)H * D + H + (size_t)H * H + H + H + 1;
  const float* b1 = P + (size_t)H * D;
  const float* w3 = b1 + H + (size_t)H * H + H;
  float* g = part + (size_t)blockIdx.x * NP;
  float* gW1 = g;
  float* gb1 = gW1 + (size_t)H * D;
  float* gW2 = gb1 + H;
  float* gb2 = gW2 + (size_t)H * H;
  float* gw3 = gb2 + H;
  const int t = threadIdx.x;
  const int s0 = blockIdx.x * MLP_SB;
  for (int k = t; k < MLP_SB * D; k += 256) {
    const int s = k / D, d = k - s * D;
    const int site = s0 + s;
    float v = 0.f;
    if (site < n) {
      const int r = rows ? rows[site] : site;
      v = bf16_round((float)q[(size_t)r * q_ld + d]);
    }
    sq[s][d] = v;
  }
  if (t < MLP_SB) {
    const int site = s0 + t;
    sg[t] = site < n ? (float)(2.0 * (e[site] - y[site]) / ((double)sigma * sigma) / n) : 0.f;
  }
  for (int k = t; k < MLP_SB * H; k += 256) {
    const int s = k / H, u = k - s * H;
    const int site = s0 + s;
    float a1 = 0.f, a2 = 0.f, d2 = 0.f;
    if (site < n) {
      const float x1 = z1[(size_t)site * H + u], x2 = z2[(size_t)site * H + u];
      a1 = x1 * sigm(x1);
      a2 = x2 * sigm(x2);
      const float sg2 = sigm(x2);
      // de/dz2 = sigma w3 s'(z2); dL/de(normalised) = g * sigma
      d2 = w3[u] * sg2 * (1.f + x2 * (1.f - sg2));
    }
    sh1[s][u] = a1;
    sh2[s][u] = a2;
    sd2[s][u] = d2;
  }
  __syncthreads();
  for (int k = t; k < MLP_SB * H; k += 256) {
    const int s = k / H, u = k - s * H;
    sd2[s][u] *= sg[s] * sigma;
  }
  __syncthreads();
  // dL/dz1 = (W2^T dL/dz2) * s'(z1)
  {
    const int u = t & (H - 1), half = t >> 7;
    float acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = 0.f;
    for (int k = 0; k < H; k++) {
      // W2[k][u] = W2t[u][k]
      const float w = bf2f(W2t[(size_t)u * H + k]);
#pragma unroll
      for (int j = 0; j < 8; j++) acc[j] += w * sd2[half * 8 + j][k];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int s = half * 8 + j;
      const int site = s0 + s;
      float d1 = 0.f;
      if (site < n) {
        const float x1 = z1[(size_t)site * H + u];
        const float sg1 = sigm(x1);
        d1 = acc[j] * sg1 * (1.f + x1 * (1.f - sg1));
      }
      sd1[s][u] = d1;
    }
  }
  __syncthreads();
  // parameter gradients, summed over the workgroup's sites in order
  for (size_t k = t; k < (size_t)H * D; k += 256) {
    const int u = (int)(k / D), d = (int)(k - (size_t)u * D);
    float a = 0.f;
#pragma unroll
    for (int s = 0; s < MLP_SB; s++) a += sd1[s][u] * sq[s][d];
    gW1[k] = a;
  }
  for (int k = t; k < H * H; k += 256) {
    const int u = k / H, d = k - u * H;
    float a = 0.f;
#pragma unroll
    for (int s = 0; s < MLP_SB; s++) a += sd2[s][u] * sh1[s][d];
    gW2[k] = a;
  }
  if (t < H) {
    float a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
    for (int s = 0; s < MLP_SB; s++) {
      a1 += sd1[s][t];
      a2 += sd2[s][t];
      a3 += sg[s] * sigma * sh2[s][t];
    }
    gb1[t] = a1;
    gb2[t] = a2;
    gw3[t] = a3;
  }
  if (t == 0) {
    float a = 0.f;
    for (int s = 0; s < MLP_SB; s++) a += sg[s] * sigma;
    gw3[H] = a;  // b3
  }
}

// grad[k] = sum over blocks of part[b][k], fixed order.
__global__ void mlp_reduce_kernel(int nblocks, size_t NP, const float* part, float* grad) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NP) return;
  float a = 0.f;
  for (int b = 0; b < nblocks; b++) a += part[(size_t)b * NP + k];
  grad[k] = a;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights, then the BF16 copies.
__global__ void mlp_adam_kernel(int D, size_t NP, float* P, const float* grad, float* m,
                                float* v, float lr, float b1, float b2, float eps, int step,
                                u16* W1t, u16* W2t) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NP) return;
  const float g = grad[k];
  const float mk = b1 * m[k] + (1.f - b1) * g;
  const float vk = b2 * v[k] + (1.f - b2) * g * g;
  m[k] = mk;
  v[k] = vk;
  const float mh = mk / (1.f - powf(b1, (float)step));
  const float vh = vk / (1.f - powf(b2, (float)step));
  const float p = P[k] - lr * mh / (sqrtf(vh) + eps);
  P[k] = p;
  const int H = MLP_H;
  if (k < (size_t)H * D) {
    const int u = (int)(k / D), d = (int)(k - (size_t)u * D);
    W1t[(size_t)d * H + u] = f2bf16(p);
    W1t[(size_t)D * H + (size_t)u * mlp_dpad(D) + d] = f2bf16(p);  // W1b
  } else {
    const size_t o = k - (size_t)H * D - H;
    if (k >= (size_t)H * D + H && o < (size_t)H * H) {
      const int u = (int)(o / H), d = (int)(o - (size_t)u * H);
      W2t[(size_t)d * H + u] = f2bf16(p);
    }
  }
}

// BF16 copies from the master weights.
__global__ void mlp_cast_kernel(int D, const float* P, u16* W1t, u16* W2t) {
  const int H = MLP_H;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < (size_t)H * D) {
    const int u = (int)(k / D), d = (int)(k - (size_t)u * D);
    W1t[(size_t)d * H + u] = f2bf16(P[k]);
    W1t[(size_t)D * H + (size_t)u * mlp_dpad(D) + d] = f2bf16(P[k]);  // W1b
  }
  if (k < (size_t)H * H) {
    const int u = (int)(k / H), d = (int)(k - (size_t)u * H);
    W2t[(size_t)d * H + u] = f2bf16(P[(size_t)H * D + H + k]);
  }
}

// Inference on MFMA for 16 sites per workgroup (four waves, 32 hidden
// units each): layer 1 as BF16 products (q rounded to BF16 on load, W1b)
// on mfma_f32_16x16x32_bf16, layer 2 with the FP32 activations and the BF16
// W2 values on mfma_f32_16x16x4f32, so the products are those of the VALU
// kernel above (the sums run in another order).  No pre-activations are
// kept (the training path uses the VALU kernel).
typedef float f32x4m __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8m __attribute__((ext_vector_type(8)));

template <typename T>
__global__ __launch_bounds__(256) void mlp_fwd_mfma_kernel(
    int n, int D, const T* __restrict__ q, int q_ld, const int* __restrict__ rows,
    const float* __restrict__ P, const u16* __restrict__ W1t, const u16* __restrict__ W2t,
    const float* __restrict__ norm, double* __restrict__ e, const int* __restrict__ guard) {
  if (guard && *guard == 0) return;
  __shared__ __align__(16) u16 sq[MLP_SB][MLP_DMAX + 8];
  __shared__ float sh1[MLP_SB][MLP_H + 1];
  __shared__ float se[4][MLP_SB];
  const int H = MLP_H, Dp = mlp_dpad(D);
  const u16* W1b = W1t + (size_t)D * H;
  const float mu = norm[0], sigma = norm[1];
  const float* b1 = P + (size_t)H * D;
  const float* b2 = b1 + H + (size_t)H * H;
  const float* w3 = b2 + H;
  const float b3 = w3[H];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int lr = lane & 15, lg = lane >> 4;
  const int s0 = blockIdx.x * MLP_SB;
  for (int k = t; k < MLP_SB * Dp; k += 256) {
    const int s = k / Dp, d = k - s * Dp;
    const int site = s0 + s;
    float v = 0.f;
    if (site < n && d < D) {
      const int r = rows ? rows[site] : site;
      v = (float)q[(size_t)r * q_ld + d];
    }
    sq[s][d] = f2bf16(v);
  }
  __syncthreads();
  // layer 1: D[site][unit] over units 32 wave .. +31 (two 16-wide tiles)
  f32x4m a1[2] = {f32x4m{0, 0, 0, 0}, f32x4m{0, 0, 0, 0}};
  for (int k0 = 0; k0 < Dp; k0 += 32) {
    const bf16x8m af = *reinterpret_cast<const bf16x8m*>(&sq[lr][k0 + lg * 8]);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int u = wave * 32 + j * 16 + lr;
      const bf16x8m bw = *reinterpret_cast<const bf16x8m*>(W1b + (size_t)u * Dp + k0 + lg * 8);
      a1[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bw, a1[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int u = wave * 32 + j * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float z = a1[j][v] + b1[u];
      sh1[lg * 4 + v][u] = z * sigm(z);
    }
  }
  __syncthreads();
  // layer 2 (FP32 operands): D[site][unit], k over the 128 inputs
  f32x4m a2[2] = {f32x4m{0, 0, 0, 0}, f32x4m{0, 0, 0, 0}};
  for (int k0 = 0; k0 < H; k0 += 4) {
    const float ah = sh1[lr][k0 + lg];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int u = wave * 32 + j * 16 + lr;
      const float bw = bf2f(W2t[(size_t)(k0 + lg) * H + u]);
      a2[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bw, a2[j], 0, 0, 0);
    }
  }
  // output: w3 . s(z2) over this wave's units, then over the four waves
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int u = wave * 32 + j * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float z = a2[j][v] + b2[u];
      part[v] += w3[u] * (z * sigm(z));
    }
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    float x = part[v];
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    if (lr == 0) se[wave][lg * 4 + v] = x;
  }
  __syncthreads();
  if (t < MLP_SB && s0 + t < n) {
    const float a = ((se[0][t] + se[1][t]) + se[2][t]) + se[3][t];
    e[s0 + t] = (double)mu + (double)sigma * (double)(a + b3);
  }
}

size_t mlp_nparams(int D) { return (size_t)MLP_H * D + (size_t)MLP_H * MLP_H + 3 * MLP_H + 1; }
size_t mlp_w1_elems(int D) { return (size_t)MLP_H * D + (size_t)MLP_H * mlp_dpad(D); }
int mlp_hidden() { return MLP_H; }
int mlp_sites_per_block() { return MLP_SB; }
int mlp_max_dim() { return MLP_DMAX; }

template <typename T>
hipError_t launch_mlp_forward(int n, int D, const T* q, int q_ld, const int* rows,
                              const float* P, const u16* W1t, const u16* W2t, const float* norm,
                              double* e, float* z1, float* z2, const int* guard,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (D > MLP_DMAX) return hipErrorInvalidValue;
  if (!z1 && !z2)  // inference only: the MFMA kernel
    hipLaunchKernelGGL((mlp_fwd_mfma_kernel<T>), dim3((n + MLP_SB - 1) / MLP_SB), dim3(256), 0, st,
                       n, D, q, q_ld, rows, P, W1t, W2t, norm, e, guard);
  else
    hipLaunchKernelGGL((mlp_fwd_kernel<T>), dim3((n + MLP_SB - 1) / MLP_SB), dim3(256), 0, st, n,
                       D, q, q_ld, rows, P, W1t, W2t, norm, e, z1, z2, guard);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mlp_gradient(int n, int D, const T* q, int q_ld, const int* rows,
                               const double* y, const double* e, const float* z1,
                               const float* z2, const float* P, const u16* W2t,
                               const float* norm, float* part, float* grad, hipStream_t st) {
  if (n <= 0 || D > MLP_DMAX) return hipErrorInvalidValue;
  const int nb = (n + MLP_SB - 1) / MLP_SB;
  const size_t NP = mlp_nparams(D);
  hipLaunchKernelGGL((mlp_grad_kernel<T>), dim3(nb), dim3(256), 0, st, n, D, q, q_ld, rows, y, e,
                     z1, z2, P, W2t, norm, part);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, st, nb,
                     NP, part, grad);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(int D, float* P, const float* grad, float* m, float* v, float lr,
                           int step, u16* W1t, u16* W2t, hipStream_t st) {
  const size_t NP = mlp_nparams(D);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((NP + 255) / 256)), dim3(256), 0, st, D, NP,
                     P, grad, m, v, lr, 0.9f, 0.999f, 1e-8f, step, W1t, W2t);
  return hipGetLastError();
}

hipError_t launch_mlp_cast(int D, const float* P, u16* W1t, u16* W2t, hipStream_t st) {
  const size_t m = (size_t)MLP_H * (D > MLP_H ? D : MLP_H);
  hipLaunchKernelGGL(mlp_cast_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, D, P,
                     W1t, W2t);
  return hipGetLastError();
}

template hipError_t launch_mlp_forward<float>(int, int, const float*, int, const int*, const float*, const u16*, const u16*, const float*, double*, float*, float*, const int*, hipStream_t);
template hipError_t launch_mlp_forward<double>(int, int, const double*, int, const int*, const float*, const u16*, const u16*, const float*, double*, float*, float*, const int*, hipStream_t);
template hipError_t launch_mlp_gradient<float>(int, int, const float*, int, const int*, const double*, const double*, const float*, const float*, const float*, const u16*, const float*, float*, float*, hipStream_t);
template hipError_t launch_mlp_gradient<double>(int, int, const double*, int, const int*, const double*, const double*, const float*, const float*, const float*, const u16*, const float*, float*, float*, hipStream_t);

}  // namespace tgx
