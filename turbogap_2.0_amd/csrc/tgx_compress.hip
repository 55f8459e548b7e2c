// compress_soap: the descriptor of a block is q = P p / |P p|, p the power
// spectrum and P a sparse matrix (soap_turbo's P transformation, read by the
// reference at gap/read_gap.f90:485-549).  The SOAP forward writes the
// uncompressed q_f = p / |p| and |p|; these kernels form the compressed
// descriptor after it and take the kernel step's dE/dq back to the power
// spectrum before the adjoint:
//   forward   q = P q_f / |P q_f|,  |P p| = |p| |P q_f|;
//   adjoint   dE/dp = P^T (G - (G.q) q) / |P p|, handed to the unchanged
//             adjoint as G_f = |p| dE/dp, which is orthogonal to q_f (E
//             depends on p only through P p / |P p|), so the adjoint's own
//             projection (G_f - (G_f.q_f) q_f) / |p| returns dE/dp.
// One wave per site; the row it gathers from is staged in LDS; each output
// is one lane's sum over its row (forward, P by rows) or column (adjoint,
// P by columns) in a fixed order.
#include <hip/hip_runtime.h>

#include "tgx_internal.h"

namespace tgx {

namespace {

typedef unsigned short u16;

__device__ __forceinline__ u16 cmp_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  return (u16)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

template <typename T>
__device__ __forceinline__ T cmp_wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace

template <typename T>
__global__ __launch_bounds__(64) void compress_fwd_kernel(int dimf, int dimf_ld, const T* qf,
                                                          const T* pnf, int dim, int dim_ld,
                                                          const int* roff, const int* rcol,
                                                          const double* rval, T* q, u16* qb,
                                                          int qb_ld, T* pn) {
  extern __shared__ unsigned char cmp_smem[];
  T* sq = reinterpret_cast<T*>(cmp_smem);
  const int site = blockIdx.x, lane = threadIdx.x;
  const T* src = qf + (size_t)site * dimf_ld;
  for (int d = lane; d < dimf; d += 64) sq[d] = src[d];
  __syncthreads();
  T* dst = q + (size_t)site * dim_ld;
  T ss = 0;
  for (int i = lane; i < dim; i += 64) {
    T a = 0;
    for (int k = roff[i]; k < roff[i + 1]; k++) a += (T)rval[k] * sq[rcol[k]];
    dst[i] = a;
    ss += a * a;
  }
  ss = cmp_wave_sum(ss);
  const T inv = (T)1 / sqrt(ss);
  u16* qbs = qb ? qb + (size_t)site * qb_ld : nullptr;
  for (int i = lane; i < dim; i += 64) {
    const T v = dst[i] * inv;
    dst[i] = v;
    if (qbs) qbs[i] = cmp_bf16((float)v);
  }
  for (int i = dim + lane; i < dim_ld; i += 64) dst[i] = (T)0;
  if (qbs)
    for (int i = dim + lane; i < qb_ld; i += 64) qbs[i] = 0;
  if (lane == 0) pn[site] = pnf[site] * sqrt(ss);
}

template <typename T>
__global__ __launch_bounds__(64) void compress_bwd_kernel(int dim, int dim_ld, const T* q,
                                                          const T* pn, const T* G, int g_ld,
                                                          int dimf, int dimf_ld, const T* pnf,
                                                          const int* coff, const int* crow,
                                                          const double* cval, T* Gf) {
  extern __shared__ unsigned char cmp_smem[];
  T* sh = reinterpret_cast<T*>(cmp_smem);
  const int site = blockIdx.x, lane = threadIdx.x;
  const T* qs = q + (size_t)site * dim_ld;
  const T* gs = G + (size_t)site * g_ld;
  T gq = 0;
  for (int i = lane; i < dim; i += 64) gq += gs[i] * qs[i];
  gq = cmp_wave_sum(gq);
  const T inv = (T)1 / pn[site];
  for (int i = lane; i < dim; i += 64) sh[i] = (gs[i] - gq * qs[i]) * inv;
  __syncthreads();
  const T sc = pnf[site];
  T* out = Gf + (size_t)site * dimf_ld;
  for (int j = lane; j < dimf; j += 64) {
    T a = 0;
    for (int k = coff[j]; k < coff[j + 1]; k++) a += (T)cval[k] * sh[crow[k]];
    out[j] = sc * a;
  }
  for (int j = dimf + lane; j < dimf_ld; j += 64) out[j] = (T)0;
}

template <typename T>
hipError_t launch_compress_fwd(int n_sites, int dimf, int dimf_ld, const T* qf, const T* pnf,
                               int dim, int dim_ld, const int* roff, const int* rcol,
                               const double* rval, T* q, u16* qb, int qb_ld, T* pn,
                               hipStream_t st) {
  hipLaunchKernelGGL(compress_fwd_kernel<T>, dim3(n_sites), dim3(64), sizeof(T) * (size_t)dimf,
                     st, dimf, dimf_ld, qf, pnf, dim, dim_ld, roff, rcol, rval, q, qb, qb_ld, pn);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_compress_bwd(int n_sites, int dim, int dim_ld, const T* q, const T* pn,
                               const T* G, int g_ld, int dimf, int dimf_ld, const T* pnf,
                               const int* coff, const int* crow, const double* cval, T* Gf,
                               hipStream_t st) {
  hipLaunchKernelGGL(compress_bwd_kernel<T>, dim3(n_sites), dim3(64), sizeof(T) * (size_t)dim,
                     st, dim, dim_ld, q, pn, G, g_ld, dimf, dimf_ld, pnf, coff, crow, cval, Gf);
  return hipGetLastError();
}

template hipError_t launch_compress_fwd<float>(int, int, int, const float*, const float*, int,
                                               int, const int*, const int*, const double*,
                                               float*, u16*, int, float*, hipStream_t);
template hipError_t launch_compress_fwd<double>(int, int, int, const double*, const double*, int,
                                                int, const int*, const int*, const double*,
                                                double*, u16*, int, double*, hipStream_t);
template hipError_t launch_compress_bwd<float>(int, int, int, const float*, const float*,
                                               const float*, int, int, int, const float*,
                                               const int*, const int*, const double*, float*,
                                               hipStream_t);
template hipError_t launch_compress_bwd<double>(int, int, int, const double*, const double*,
                                                const double*, int, int, int, const double*,
                                                const int*, const int*, const double*, double*,
                                                hipStream_t);

}  // namespace tgx
