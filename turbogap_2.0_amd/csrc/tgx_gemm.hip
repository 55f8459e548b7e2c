// GAP kernel step on MFMA (reference gap/gap.f90:118-171):
//   K = q Qs^T                      (GEMM 1, epilogue: E partials, W)
//   E_i = delta^2 sum_s K_is^zeta alpha_s
//   W_is = zeta delta^2 alpha_s K_is^(zeta-1)
//   G = W Qs = dE/dq                (GEMM 2)
// Both are "NT" products C[M][N] = A[M][K] B[N][K]^T with K contiguous.
// Inputs FP64 (v_mfma_f64_16x16x4), FP32 (v_mfma_f32_16x16x4) or BF16
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "tgx_internal.h"

namespace tgx {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;

template <typename TI> struct MF;
template <> struct MF<double> {
  typedef double acc_t;
  typedef f64x4 vacc;
  static constexpr int BK = 8, KSTEP = 4, KPL = 1, PAD = 1;
  typedef double frag;
  __device__ static frag load(const double* s) { return *s; }
  __device__ static vacc mma(frag a, frag b, vacc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result row of accumulator element v in this lane (column = lane & 15)
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MF<float> {
  typedef float acc_t;
  typedef f32x4 vacc;
  static constexpr int BK = 16, KSTEP = 4, KPL = 1, PAD = 1;
  typedef float frag;
  __device__ static frag load(const float* s) { return *s; }
  __device__ static vacc mma(frag a, frag b, vacc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) * 4 + v; }
};
template <> struct MF<u16> {
  typedef float acc_t;
  typedef f32x4 vacc;
  static constexpr int BK = 32, KSTEP = 32, KPL = 8, PAD = 8;
  typedef bf16x8 frag;
  __device__ static frag load(const u16* s) {
    return *reinterpret_cast<const bf16x8*>(s);
  }
  __device__ static vacc mma(frag a, frag b, vacc c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) * 4 + v; }
};

// Round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ u16 to_bf16(float f) {
  return __builtin_bit_cast(u16, (__bf16)f);
}
template <typename TI> __device__ __forceinline__ TI cvt_out(float v);
template <> __device__ __forceinline__ float cvt_out<float>(float v) { return v; }
template <> __device__ __forceinline__ u16 cvt_out<u16>(float v) { return to_bf16(v); }
__device__ __forceinline__ double cvt_out_d(double v) { return v; }

struct EpiKernel {
  const void* alphas;  // acc_t, padded to N (zeros beyond n_sparse)
  int n_sparse;
  int zint;            // >0: integer zeta
  double zeta, zd2;    // zeta, zeta*delta^2
  void* W;             // TI [M][ldw]
  int ldw;
  void* epart;         // acc_t [gridDim.x][M]
  // optional: atom energies eatom[r] = e0[species[r] - 1] + esite[r] (every
  // atom a site, in order)
  double* eatom;
  const double* e0;
  const int* species;
};

__device__ __forceinline__ void put_site_energy(const EpiKernel& ek, double* esite, int r,
                                                double v) {
  esite[r] = v;
  if (ek.eatom) ek.eatom[r] = ek.e0[ek.species[r] - 1] + v;
}
struct EpiStore {
  void* C;  // acc_t [M][ldc]
  int ldc;
};

template <typename T>
__device__ __forceinline__ void kpow(T k, int zint, T zeta, T* kz, T* kz1) {
  if (zint > 0) {
    T p = 1;
    for (int i = 1; i < zint; i++) p *= k;
    *kz1 = p;
    *kz = p * k;
  } else {
    *kz = pow(k, zeta);
    *kz1 = pow(k, zeta - (T)1);
  }
}

// Workgroup tile: 2 x 2 waves, each TM x TN MFMA tiles of 16 x 16, k-slab BK
// (the built-in shapes; the autotuner may pick others per problem shape).
// The BF16 store GEMM (G = W Qs) uses 128 x 128 tiles (more MFMA work per
// LDS read); the kernel-epilogue GEMM keeps 64 x 64, where the per-element
// epilogue (powers, BF16 W stores, row sums) would otherwise dominate; both
// BF16 GEMMs take 64-deep slabs.  FP32 and FP64 use 64 x 64, FP64 with
// 16-deep slabs (measured on gfx950).
template <typename TI, int EPI> struct Tile {
  static constexpr int TM = 2, TN = 2, BK = MF<TI>::BK;
};
template <> struct Tile<u16, 1> { static constexpr int TM = 4, TN = 4, BK = 64; };
template <> struct Tile<u16, 0> { static constexpr int TM = 2, TN = 2, BK = 64; };
template <int EPI> struct Tile<double, EPI> { static constexpr int TM = 2, TN = 2, BK = 16; };

// EPI 0: kernel epilogue; EPI 1: plain store.  The next k-slab is read from
// HBM into registers while the MFMAs of the current one run.  TM_ x TN_
// MFMA tiles per wave (the autotuner's choice; Tile<> gives the default).
template <typename TI, int EPI, int TM_ = Tile<TI, EPI>::TM, int TN_ = Tile<TI, EPI>::TN>
__global__ __launch_bounds__(256) void gemm_nt_kernel(int M, int N, int K,
                                                     const TI* __restrict__ A, int lda,
                                                     const TI* __restrict__ B, int ldb,
                                                     EpiKernel ek, EpiStore es, int ntx,
                                                     int ntiles, const int* __restrict__ guard) {
  if (guard && *guard == 0) return;
  typedef MF<TI> F;
  typedef typename F::acc_t acc_t;
  typedef typename F::vacc vacc;
  constexpr int TM = TM_, TN = TN_, BK = Tile<TI, EPI>::BK;
  constexpr int BM = 32 * TM, BN = 32 * TN, LDS = BK + F::PAD;
  constexpr int EPR = 16 / sizeof(TI);     // elements per 16-byte load
  constexpr int TPR = BK / EPR;            // threads per tile row
  constexpr int RPP = 256 / TPR;           // rows per load pass
  constexpr int PA = BM / RPP, PB = BN / RPP;
  __shared__ __align__(16) TI sA[BM * LDS];
  __shared__ __align__(16) TI sB[BN * LDS];
  __shared__ acc_t sE[2][BM];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  // Workgroups are dealt round-robin to the 8 XCDs (id % 8).  Renumber so
  // that a contiguous run of tile ids, i.e. the column tiles of one row
  // strip, lands on one XCD and the strip of A is fetched into that XCD's L2
  // once instead of once per column tile.  gridDim.x is a multiple of 8.
  const int w = blockIdx.x;
  const int t = (w & 7) * (gridDim.x >> 3) + (w >> 3);
  if (t >= ntiles) return;
  const int tx = t % ntx, ty = t / ntx;
  const int row0 = ty * BM, col0 = tx * BN;
  const int lr = tid / TPR, lc = (tid % TPR) * EPR;

  vacc acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = vacc{0, 0, 0, 0};

  uint4 ra[PA], rb[PB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int p = 0; p < PA; p++) {
      const int r = row0 + p * RPP + lr;
      ra[p] = r < M ? *reinterpret_cast<const uint4*>(A + (size_t)r * lda + k0 + lc)
                    : uint4{0, 0, 0, 0};
    }
#pragma unroll
    for (int p = 0; p < PB; p++) {
      const int c = col0 + p * RPP + lr;
      rb[p] = c < N ? *reinterpret_cast<const uint4*>(B + (size_t)c * ldb + k0 + lc)
                    : uint4{0, 0, 0, 0};
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int p = 0; p < PA; p++) {
      const TI* v = reinterpret_cast<const TI*>(&ra[p]);
#pragma unroll
      for (int e = 0; e < EPR; e++) sA[(p * RPP + lr) * LDS + lc + e] = v[e];
    }
#pragma unroll
    for (int p = 0; p < PB; p++) {
      const TI* v = reinterpret_cast<const TI*>(&rb[p]);
#pragma unroll
      for (int e = 0; e < EPR; e++) sB[(p * RPP + lr) * LDS + lc + e] = v[e];
    }
    __syncthreads();
    if (k0 + BK < K) fetch(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += F::KSTEP) {
      const int kk = ks + (lane >> 4) * F::KPL;
      typename F::frag af[TM], bfr[TN];
#pragma unroll
      for (int mi = 0; mi < TM; mi++)
        af[mi] = F::load(sA + (wm * 16 * TM + mi * 16 + (lane & 15)) * LDS + kk);
#pragma unroll
      for (int ni = 0; ni < TN; ni++)
        bfr[ni] = F::load(sB + (wn * 16 * TN + ni * 16 + (lane & 15)) * LDS + kk);
#pragma unroll
      for (int mi = 0; mi < TM; mi++)
#pragma unroll
        for (int ni = 0; ni < TN; ni++) acc[mi][ni] = F::mma(af[mi], bfr[ni], acc[mi][ni]);
    }
    __syncthreads();
  }

  if (EPI == 1) {
    acc_t* C = reinterpret_cast<acc_t*>(es.C);
#pragma unroll
    for (int mi = 0; mi < TM; mi++)
#pragma unroll
      for (int ni = 0; ni < TN; ni++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = row0 + wm * 16 * TM + mi * 16 + F::row(lane, v);
          const int c = col0 + wn * 16 * TN + ni * 16 + (lane & 15);
          if (r < M && c < es.ldc) C[(size_t)r * es.ldc + c] = acc[mi][ni][v];
        }
    return;
  }
  // kernel epilogue
  const acc_t* al = reinterpret_cast<const acc_t*>(ek.alphas);
  const acc_t zeta = (acc_t)ek.zeta, zd2 = (acc_t)ek.zd2;
  TI* Wm = reinterpret_cast<TI*>(ek.W);
  acc_t rs[TM][4];
#pragma unroll
  for (int mi = 0; mi < TM; mi++)
#pragma unroll
    for (int v = 0; v < 4; v++) rs[mi][v] = 0;
#pragma unroll
  for (int ni = 0; ni < TN; ni++) {
    const int c = col0 + wn * 16 * TN + ni * 16 + (lane & 15);
    const bool valid = c < ek.n_sparse;
    const acc_t a = valid ? al[c] : (acc_t)0;
#pragma unroll
    for (int mi = 0; mi < TM; mi++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = row0 + wm * 16 * TM + mi * 16 + F::row(lane, v);
        const acc_t kv = acc[mi][ni][v];
        acc_t kz = 0, kz1 = 0;
        if (valid) kpow<acc_t>(kv, ek.zint, zeta, &kz, &kz1);
        rs[mi][v] += a * kz;
        const acc_t wv = valid ? zd2 * a * kz1 : (acc_t)0;
        if (r < M && c < ek.ldw) {
          if constexpr (sizeof(TI) == 8) Wm[(size_t)r * ek.ldw + c] = wv;
          else Wm[(size_t)r * ek.ldw + c] = cvt_out<TI>((float)wv);
        }
      }
  }
  // reduce row partials over the 16 lanes of a row group
#pragma unroll
  for (int mi = 0; mi < TM; mi++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      acc_t x = rs[mi][v];
      x += __shfl_xor(x, 1, 64);
      x += __shfl_xor(x, 2, 64);
      x += __shfl_xor(x, 4, 64);
      x += __shfl_xor(x, 8, 64);
      if ((lane & 15) == 0) sE[wn][wm * 16 * TM + mi * 16 + F::row(lane, v)] = x;
    }
  __syncthreads();
  if (tid < BM) {
    const int r = row0 + tid;
    if (r < M)
      reinterpret_cast<acc_t*>(ek.epart)[(size_t)tx * M + r] = sE[0][tid] + sE[1][tid];
  }
}

template <typename TA>
__global__ void esum_kernel(int M, int ntiles, const TA* epart, double d2,
                            double* esite, EpiKernel ek, const int* guard) {
  if (guard && *guard == 0) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  // partials added in tile order; their loads go out eight at a time
  double s = 0.0;
  int t = 0;
  for (; t + 8 <= ntiles; t += 8) {
    TA v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = epart[(size_t)(t + k) * M + r];
#pragma unroll
    for (int k = 0; k < 8; k++) s += (double)v[k];
  }
  for (; t < ntiles; t++) s += (double)epart[(size_t)t * M + r];
  put_site_energy(ek, esite, r, d2 * s);
}
// esum_kernel workgroup size: small, so that a few thousand sites still
// spread over many CUs (the kernel is bound by load latency)
#define TGX_ESUM_WG 64

// Both GEMMs of the BF16 kernel step in one workgroup per 128 sites, for
// models with at most 512 sparse points and 448 descriptor components:
//   S = q Qs^T (128 x N1, FP32 in registers), epilogue E_i and
//   W = zeta delta^2 alpha S^(zeta-1) as BF16 into LDS, then G = W Qs.
// W never goes to HBM and the energies need no separate row sum.  Eight
// waves: stage 1 splits the rows in two halves and the sparse points in
// four quarters, stage 2 the rows in two and the output columns in four.
// Every operand reaches LDS through coalesced 16-byte loads (whole rows of
// a 32-deep k-slab), the next slabs travelling in registers while the
// current one is multiplied: stage 1 double-buffers q and Qs slabs where W
// goes afterwards, stage 2 keeps one Qs slab next to W (rows padded by 16
// bytes); slab rows are 64 bytes with their 16-byte segments XOR-swizzled
// by row, so that the staging stores and the fragment reads are free of
// bank conflicts.  Both stages
// multiply with the Qs operand first (D = Qs-tile x q-tile), so that a lane
// owns four consecutive columns of one row: W goes to LDS as 8-byte stores
// and G to memory as float4.  Integer zeta up to 4 has its own epilogue
// (powers by multiplication).
#define TGX_FUSED_MAXT 8
#define TGX_FUSED_SLD 32  // stage-1 staged row: 32 bf16 (16-byte segments XOR-swizzled)
// Offset (in bf16) of 16-byte segment seg of staged row r: segments are
// XOR-swizzled by (r >> 2) & 3, so that both the coalesced stores (four rows
// of four segments per 16 lanes) and the fragment reads (sixteen rows of one
// segment) touch 64 distinct banks.
__device__ __forceinline__ int slab_off(int r, int seg) {
  return r * TGX_FUSED_SLD + ((seg ^ ((r >> 2) & 3)) << 3);
}
#define TGX_FUSED_PF 5    // stage-1 16-byte chunks per thread: (128 + 512) * 4 / 512

template <int ZI>
__device__ __forceinline__ void kpow_f(float k, int zint, float* kz, float* kz1) {
  float p = 1.f;
  if constexpr (ZI > 0) {
#pragma unroll
    for (int i = 1; i < ZI; i++) p *= k;
  } else {
#pragma unroll
    for (int i = 1; i < 8; i++)
      if (i < zint) p *= k;
  }
  *kz1 = p;
  *kz = p * k;
}

template <int V> struct ICz { static constexpr int v = V; };
template <int V> struct ICv { static constexpr int value = V; };

// LDS bytes of the W / stage-1 region and of the stage-2 slab region.
__host__ __device__ inline size_t fused_region_w(int N1) {
  const size_t w = (size_t)128 * (N1 + 8) * 2, s = (size_t)2 * (128 + N1) * TGX_FUSED_SLD * 2;
  return w > s ? w : s;
}
__host__ __device__ inline size_t fused_region_t(int K1) {
  const size_t t = (size_t)K1 * 64;
  return t > 4 * 128 * sizeof(float) ? t : 4 * 128 * sizeof(float);
}

template <int NT2, bool GRAD>
__global__ __launch_bounds__(512) void gap_fused_kernel(
    int M, int N1, int K1, const u16* __restrict__ q, int ldq, const u16* __restrict__ qs,
    int ldqs, const u16* __restrict__ qsT, int ldqt, EpiKernel ek, double d2,
    double* __restrict__ esite, float* __restrict__ G, int ldg, const int* __restrict__ guard) {
  if (guard && *guard == 0) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ldw = N1 + 8;
  u16* sW = reinterpret_cast<u16*>(smem_raw);  // [128][ldw]; stage-1 slabs before
  u16* sT = reinterpret_cast<u16*>(smem_raw + fused_region_w(N1));  // stage-2 slab [K1][32]
  float* sE = reinterpret_cast<float*>(sT);  // [4][128] row-sum partials, before stage 2
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int mh = wave >> 2, nq = wave & 3;
  const int row0 = blockIdx.x * 128, rw = row0 + mh * 64;
  const int lr = lane & 15, lk = (lane >> 4) * 8, lc = (lane >> 4) * 4;
  const int nt1 = N1 >> 6;

  // ---- stage 1: D[s][m] = S[m][s] over this wave's quarter of the sparse points
  f32x4 acc[4][TGX_FUSED_MAXT];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < TGX_FUSED_MAXT; b++) acc[a][b] = f32x4{0, 0, 0, 0};
  const int c1 = nq * (N1 >> 2);
  {
    const int stage_rows = 128 + N1, nchunk = stage_rows * 4;
    // two slabs in flight in registers (pa, pb) besides the one in LDS: the
    // q rows come from HBM and one slab of products does not cover that
    // latency; the loop runs in pairs of slabs so that the buffers keep
    // static names
    uint4 pa[TGX_FUSED_PF], pb[TGX_FUSED_PF];
    auto fetch = [&](uint4 (&pf)[TGX_FUSED_PF], int k0) {
#pragma unroll
      for (int p = 0; p < TGX_FUSED_PF; p++) {
        const int ch = tid + p * 512;
        const int r = ch >> 2, part = (ch & 3) * 8;
        uint4 v = uint4{0, 0, 0, 0};
        if (ch < nchunk) {
          if (r < 128) {
            if (row0 + r < M)
              v = *reinterpret_cast<const uint4*>(q + (size_t)(row0 + r) * ldq + k0 + part);
          } else {
            v = *reinterpret_cast<const uint4*>(qs + (size_t)(r - 128) * ldqs + k0 + part);
          }
        }
        pf[p] = v;
      }
    };
    auto stash = [&](const uint4 (&pf)[TGX_FUSED_PF], int buf) {
      u16* st = sW + (size_t)buf * stage_rows * TGX_FUSED_SLD;
#pragma unroll
      for (int p = 0; p < TGX_FUSED_PF; p++) {
        const int ch = tid + p * 512;
        if (ch < nchunk)
          *reinterpret_cast<uint4*>(st + slab_off(ch >> 2, ch & 3)) = pf[p];
      }
    };
    auto mma_slab = [&](int buf) {
      const u16* st = sW + (size_t)buf * stage_rows * TGX_FUSED_SLD;
      bf16x8 af[4];
#pragma unroll
      for (int mi = 0; mi < 4; mi++)
        af[mi] = *reinterpret_cast<const bf16x8*>(
            st + slab_off(mh * 64 + mi * 16 + lr, lane >> 4));
#pragma unroll
      for (int ni = 0; ni < TGX_FUSED_MAXT; ni++)
        if (ni < nt1) {
          const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(
              st + slab_off(128 + c1 + ni * 16 + lr, lane >> 4));
#pragma unroll
          for (int mi = 0; mi < 4; mi++)
            acc[mi][ni] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr, af[mi], acc[mi][ni], 0, 0, 0);
        }
    };
    const int nslab = K1 >> 5;
    fetch(pa, 0);
    stash(pa, 0);
    if (nslab > 1) fetch(pb, 32);
    __syncthreads();
    // slab sl lives in LDS buffer sl & 1
    for (int sl = 0; sl < nslab; sl += 2) {
      if (sl + 2 < nslab) fetch(pa, (sl + 2) * 32);
      mma_slab(0);
      if (sl + 1 < nslab) stash(pb, 1);
      __syncthreads();
      if (sl + 1 >= nslab) break;
      if (sl + 3 < nslab) fetch(pb, (sl + 3) * 32);
      mma_slab(1);
      if (sl + 2 < nslab) stash(pa, 0);
      __syncthreads();
    }
  }

  // ---- epilogue: E partials and W (BF16, four columns per 8-byte store)
  const float* al = reinterpret_cast<const float*>(ek.alphas);  // zero past n_sparse
  const float zd2 = (float)ek.zd2;
  float rs[4] = {0.f, 0.f, 0.f, 0.f};
  auto epilogue = [&](auto zi) {
    constexpr int ZI = decltype(zi)::v;
#pragma unroll
    for (int ni = 0; ni < TGX_FUSED_MAXT; ni++) {
      if (ni < nt1) {
        const int s0 = c1 + ni * 16 + lc;
        const float4 a4 = *reinterpret_cast<const float4*>(al + s0);
        const float av[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
        for (int mi = 0; mi < 4; mi++) {
          __bf16 w[4];
#pragma unroll
          for (int v = 0; v < 4; v++) {
            float kz, kz1;
            kpow_f<ZI>(acc[mi][ni][v], ek.zint, &kz, &kz1);
            rs[mi] += av[v] * kz;
            w[v] = (__bf16)(zd2 * av[v] * kz1);
          }
          if (GRAD)
            *reinterpret_cast<uint2*>(sW + (size_t)(mh * 64 + mi * 16 + lr) * ldw + s0) =
                __builtin_bit_cast(uint2, w);
        }
      }
    }
  };
  switch (ek.zint) {
    case 1: epilogue(ICz<1>{}); break;
    case 2: epilogue(ICz<2>{}); break;
    case 3: epilogue(ICz<3>{}); break;
    case 4: epilogue(ICz<4>{}); break;
    default: epilogue(ICz<0>{}); break;
  }
#pragma unroll
  for (int mi = 0; mi < 4; mi++) {
    float x = rs[mi];
    x += __shfl_xor(x, 16, 64);
    x += __shfl_xor(x, 32, 64);
    if (lane < 16) sE[nq * 128 + mh * 64 + mi * 16 + lr] = x;
  }
  __syncthreads();
  if (tid < 128 && row0 + tid < M)
    put_site_energy(ek, esite, row0 + tid,
                    d2 * ((double)sE[tid] + (double)sE[128 + tid] + (double)sE[256 + tid] +
                          (double)sE[384 + tid]));
  if (!GRAD) return;

  // ---- stage 2: D[c][m] = G[m][c] over this wave's quarter of the columns
  f32x4 acc2[4][NT2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < NT2; b++) acc2[a][b] = f32x4{0, 0, 0, 0};
  const int c2 = nq * (16 * NT2);
  {
    // K1 = 64 NT2 rows of four 16-byte chunks: NT2 / 2 chunks per thread
    constexpr int nchunk = 256 * NT2, PF2 = (NT2 + 1) / 2;
    if constexpr (NT2 <= 6) {
      // two slabs in flight in registers besides the one in LDS, as in stage 1
      uint4 pa[PF2], pb[PF2];
  #pragma unroll
      for (int p = 0; p < PF2; p++) pa[p] = pb[p] = uint4{0, 0, 0, 0};
      auto fetch = [&](uint4 (&pf)[PF2], int k0) {
  #pragma unroll
        for (int p = 0; p < PF2; p++) {
          const int ch = tid + p * 512;
          if (ch < nchunk)
            pf[p] = *reinterpret_cast<const uint4*>(qsT + (size_t)(ch >> 2) * ldqt + k0 +
                                                    (ch & 3) * 8);
        }
      };
      auto stash = [&](const uint4 (&pf)[PF2]) {
  #pragma unroll
        for (int p = 0; p < PF2; p++) {
          const int ch = tid + p * 512;
          const int r = ch >> 2;
          if (ch < nchunk)
            *reinterpret_cast<uint4*>(sT + (size_t)r * 32 + (((ch & 3) ^ ((r >> 2) & 3)) * 8)) =
                pf[p];
        }
      };
      auto mma_slab = [&](int k0) {
        bf16x8 aw[4];
  #pragma unroll
        for (int mi = 0; mi < 4; mi++)
          aw[mi] = *reinterpret_cast<const bf16x8*>(sW + (size_t)(mh * 64 + mi * 16 + lr) * ldw +
                                                    k0 + lk);
  #pragma unroll
        for (int nj = 0; nj < NT2; nj++) {
          const int r = c2 + nj * 16 + lr;
          const bf16x8 bq = *reinterpret_cast<const bf16x8*>(
              sT + (size_t)r * 32 + (((lane >> 4) ^ ((r >> 2) & 3)) * 8));
  #pragma unroll
          for (int mi = 0; mi < 4; mi++)
            acc2[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq, aw[mi], acc2[mi][nj], 0, 0, 0);
        }
      };
      fetch(pa, 0);
      if (32 < N1) fetch(pb, 32);
      for (int k0 = 0; k0 < N1; k0 += 64) {
        __syncthreads();  // the slab (and sE) is no longer read
        stash(pa);
        __syncthreads();
        if (k0 + 64 < N1) fetch(pa, k0 + 64);
        mma_slab(k0);
        if (k0 + 32 >= N1) break;
        __syncthreads();
        stash(pb);
        __syncthreads();
        if (k0 + 96 < N1) fetch(pb, k0 + 96);
        mma_slab(k0 + 32);
      }
    } else {  // wider G tiles: no registers left for a second slab
      uint4 pf[PF2];
  #pragma unroll
      for (int p = 0; p < PF2; p++) pf[p] = uint4{0, 0, 0, 0};
      auto fetch = [&](int k0) {
  #pragma unroll
        for (int p = 0; p < PF2; p++) {
          const int ch = tid + p * 512;
          if (ch < nchunk)
            pf[p] = *reinterpret_cast<const uint4*>(qsT + (size_t)(ch >> 2) * ldqt + k0 +
                                                    (ch & 3) * 8);
        }
      };
      auto stash = [&]() {
  #pragma unroll
        for (int p = 0; p < PF2; p++) {
          const int ch = tid + p * 512;
          const int r = ch >> 2;
          if (ch < nchunk)
            *reinterpret_cast<uint4*>(sT + (size_t)r * 32 + (((ch & 3) ^ ((r >> 2) & 3)) * 8)) =
                pf[p];
        }
      };
      fetch(0);
      for (int k0 = 0; k0 < N1; k0 += 32) {
        __syncthreads();  // the slab (and sE) is no longer read
        stash();
        __syncthreads();
        if (k0 + 32 < N1) fetch(k0 + 32);
        bf16x8 aw[4];
  #pragma unroll
        for (int mi = 0; mi < 4; mi++)
          aw[mi] = *reinterpret_cast<const bf16x8*>(sW + (size_t)(mh * 64 + mi * 16 + lr) * ldw +
                                                    k0 + lk);
  #pragma unroll
        for (int nj = 0; nj < NT2; nj++) {
          const int r = c2 + nj * 16 + lr;
          const bf16x8 bq = *reinterpret_cast<const bf16x8*>(
              sT + (size_t)r * 32 + (((lane >> 4) ^ ((r >> 2) & 3)) * 8));
  #pragma unroll
          for (int mi = 0; mi < 4; mi++)
            acc2[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq, aw[mi], acc2[mi][nj], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int mi = 0; mi < 4; mi++) {
    const int r = rw + mi * 16 + lr;
    if (r < M)
#pragma unroll
      for (int nj = 0; nj < NT2; nj++)
        *reinterpret_cast<f32x4*>(G + (size_t)r * ldg + c2 + nj * 16 + lc) = acc2[mi][nj];
  }
}

// The same for more than 512 sparse points: the sparse set is processed in
// chunks of 256 (stage 1, epilogue and stage 2 per chunk), G accumulating in
// registers across the chunks and the energies in the row sums.  The q rows
// are staged again for every chunk (from L2).
// Workgroups take 128 or 96 rows (MI = 4 or 3 row tiles per wave half); a
// row's arithmetic is the same either way.
#define TGX_FUSED_SC 256
#define TGX_FUSED_PFC 3  // stage-1 16-byte chunks per thread: (128 + 256) * 4 / 512
template <int NT2, bool GRAD, int MI>
__global__ __launch_bounds__(512) void gap_fused_chunked_kernel(
    int M, int NS, int K1, const u16* __restrict__ q, int ldq, const u16* __restrict__ qs,
    int ldqs, const u16* __restrict__ qsT, int ldqt, EpiKernel ek, double d2,
    double* __restrict__ esite, float* __restrict__ G, int ldg, const int* __restrict__ guard) {
  if (guard && *guard == 0) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int SC = TGX_FUSED_SC, T1 = SC / 64, ldw = SC + 8;
  // MI row tiles of 16 per wave half: HR rows per half, ROWS per workgroup
  constexpr int HR = 16 * MI, ROWS = 2 * HR;
  u16* sW = reinterpret_cast<u16*>(smem_raw);  // [ROWS][ldw]; stage-1 slabs before
  u16* sT = reinterpret_cast<u16*>(smem_raw + fused_region_w(SC));  // stage-2 slab [K1][32]
  float* sE = reinterpret_cast<float*>(sT);  // [4][ROWS] row-sum partials, at the end
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int mh = wave >> 2, nq = wave & 3;
  const int row0 = blockIdx.x * ROWS, rw = row0 + mh * HR;
  const int lr = lane & 15, lk = (lane >> 4) * 8, lc = (lane >> 4) * 4;
  const float* al = reinterpret_cast<const float*>(ek.alphas);  // zero past n_sparse
  const float zd2 = (float)ek.zd2;
  float rs[MI];
#pragma unroll
  for (int a = 0; a < MI; a++) rs[a] = 0.f;
  f32x4 acc2[MI][NT2];
#pragma unroll
  for (int a = 0; a < MI; a++)
#pragma unroll
    for (int b = 0; b < NT2; b++) acc2[a][b] = f32x4{0, 0, 0, 0};
  const int c2 = nq * (16 * NT2);
  for (int sc0 = 0; sc0 < NS; sc0 += SC) {
    const int N1 = min(SC, NS - sc0), nt1 = N1 >> 6;  // N1: a multiple of 64
    // ---- stage 1 over the chunk
    f32x4 acc[MI][T1];
#pragma unroll
    for (int a = 0; a < MI; a++)
#pragma unroll
      for (int b = 0; b < T1; b++) acc[a][b] = f32x4{0, 0, 0, 0};
    {
      const int stage_rows = ROWS + N1, nchunk = stage_rows * 4;
      uint4 pf[TGX_FUSED_PFC];
      auto fetch = [&](int k0) {
#pragma unroll
        for (int p = 0; p < TGX_FUSED_PFC; p++) {
          const int ch = tid + p * 512;
          const int r = ch >> 2, part = (ch & 3) * 8;
          uint4 v = uint4{0, 0, 0, 0};
          if (ch < nchunk) {
            if (r < ROWS) {
              if (row0 + r < M)
                v = *reinterpret_cast<const uint4*>(q + (size_t)(row0 + r) * ldq + k0 + part);
            } else {
              v = *reinterpret_cast<const uint4*>(qs + (size_t)(sc0 + r - ROWS) * ldqs + k0 +
                                                  part);
            }
          }
          pf[p] = v;
        }
      };
      auto stash = [&](int buf) {
        u16* st = sW + (size_t)buf * stage_rows * TGX_FUSED_SLD;
#pragma unroll
        for (int p = 0; p < TGX_FUSED_PFC; p++) {
          const int ch = tid + p * 512;
          if (ch < nchunk)
            *reinterpret_cast<uint4*>(st + slab_off(ch >> 2, ch & 3)) =
                pf[p];
        }
      };
      const int nslab = K1 >> 5;
      fetch(0);
      stash(0);
      __syncthreads();
      for (int sl = 0; sl < nslab; sl++) {
        if (sl + 1 < nslab) fetch((sl + 1) * 32);
        const u16* st = sW + (size_t)(sl & 1) * stage_rows * TGX_FUSED_SLD;
        bf16x8 af[MI];
#pragma unroll
        for (int mi = 0; mi < MI; mi++)
          af[mi] = *reinterpret_cast<const bf16x8*>(
              st + slab_off(mh * HR + mi * 16 + lr, lane >> 4));
#pragma unroll
        for (int ni = 0; ni < T1; ni++)
          if (ni < nt1) {
            const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(
                st + slab_off(ROWS + (N1 >> 2) * nq + ni * 16 + lr, lane >> 4));
#pragma unroll
            for (int mi = 0; mi < MI; mi++)
              acc[mi][ni] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr, af[mi], acc[mi][ni], 0, 0, 0);
          }
        if (sl + 1 < nslab) stash((sl + 1) & 1);
        __syncthreads();
      }
    }
    // ---- epilogue of the chunk: energy partials and W into LDS
    const int ldwc = N1 + 8;
    auto epilogue = [&](auto zi) {
      constexpr int ZI = decltype(zi)::v;
#pragma unroll
      for (int ni = 0; ni < T1; ni++) {
        if (ni < nt1) {
          const int sl0 = (N1 >> 2) * nq + ni * 16 + lc;  // column within the chunk
          const float4 a4 = *reinterpret_cast<const float4*>(al + sc0 + sl0);
          const float av[4] = {a4.x, a4.y, a4.z, a4.w};
#pragma unroll
          for (int mi = 0; mi < MI; mi++) {
            __bf16 w[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
              float kz, kz1;
              kpow_f<ZI>(acc[mi][ni][v], ek.zint, &kz, &kz1);
              rs[mi] += av[v] * kz;
              w[v] = (__bf16)(zd2 * av[v] * kz1);
            }
            if (GRAD)
              *reinterpret_cast<uint2*>(sW + (size_t)(mh * HR + mi * 16 + lr) * ldwc + sl0) =
                  __builtin_bit_cast(uint2, w);
          }
        }
      }
    };
    switch (ek.zint) {
      case 1: epilogue(ICz<1>{}); break;
      case 2: epilogue(ICz<2>{}); break;
      case 3: epilogue(ICz<3>{}); break;
      case 4: epilogue(ICz<4>{}); break;
      default: epilogue(ICz<0>{}); break;
    }
    if (GRAD) {
      // ---- stage 2 of the chunk: G += W_chunk Qs_chunk
      constexpr int nchunk2 = 256 * NT2, PF2 = (NT2 + 1) / 2;
      uint4 pf[PF2];
#pragma unroll
      for (int p = 0; p < PF2; p++) pf[p] = uint4{0, 0, 0, 0};
      auto fetch = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PF2; p++) {
          const int ch = tid + p * 512;
          if (ch < nchunk2)
            pf[p] = *reinterpret_cast<const uint4*>(qsT + (size_t)(ch >> 2) * ldqt + sc0 + k0 +
                                                    (ch & 3) * 8);
        }
      };
      auto stash = [&]() {
#pragma unroll
        for (int p = 0; p < PF2; p++) {
          const int ch = tid + p * 512;
          const int r = ch >> 2;
          if (ch < nchunk2)
            *reinterpret_cast<uint4*>(sT + (size_t)r * 32 + (((ch & 3) ^ ((r >> 2) & 3)) * 8)) =
                pf[p];
        }
      };
      fetch(0);
      for (int k0 = 0; k0 < N1; k0 += 32) {
        __syncthreads();  // W complete; the slab is no longer read
        stash();
        __syncthreads();
        if (k0 + 32 < N1) fetch(k0 + 32);
        bf16x8 aw[MI];
#pragma unroll
        for (int mi = 0; mi < MI; mi++)
          aw[mi] = *reinterpret_cast<const bf16x8*>(sW + (size_t)(mh * HR + mi * 16 + lr) * ldwc +
                                                    k0 + lk);
#pragma unroll
        for (int nj = 0; nj < NT2; nj++) {
          const int r = c2 + nj * 16 + lr;
          const bf16x8 bq = *reinterpret_cast<const bf16x8*>(
              sT + (size_t)r * 32 + (((lane >> 4) ^ ((r >> 2) & 3)) * 8));
#pragma unroll
          for (int mi = 0; mi < MI; mi++)
            acc2[mi][nj] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq, aw[mi], acc2[mi][nj], 0, 0, 0);
        }
      }
    }
    __syncthreads();  // the next chunk's staging overwrites W and the slab
  }
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    float x = rs[mi];
    x += __shfl_xor(x, 16, 64);
    x += __shfl_xor(x, 32, 64);
    if (lane < 16) sE[nq * ROWS + mh * HR + mi * 16 + lr] = x;
  }
  __syncthreads();
  if (tid < ROWS && row0 + tid < M)
    put_site_energy(ek, esite, row0 + tid,
                    d2 * ((double)sE[tid] + (double)sE[ROWS + tid] + (double)sE[2 * ROWS + tid] +
                          (double)sE[3 * ROWS + tid]));
  if (!GRAD) return;
#pragma unroll
  for (int mi = 0; mi < MI; mi++) {
    const int r = rw + mi * 16 + lr;
    if (r < M)
#pragma unroll
      for (int nj = 0; nj < NT2; nj++)
        *reinterpret_cast<f32x4*>(G + (size_t)r * ldg + c2 + nj * 16 + lc) = acc2[mi][nj];
  }
}

template <int NT2, bool GRAD, int MI>
static hipError_t launch_fused_chunked(int M, int sp_pad, int dim_pad, const void* q,
                                       const void* qs, const void* qsT, const EpiKernel& ek,
                                       double delta, double* esite, void* G, const int* guard,
                                       hipStream_t st) {
  static size_t sm_set = 0;
  // sized for 128-row tiles, which also holds 96
  const size_t sm = fused_region_w(TGX_FUSED_SC) + fused_region_t(dim_pad);
  if (sm > sm_set) {
    hipError_t e = hipFuncSetAttribute((const void*)gap_fused_chunked_kernel<NT2, GRAD, MI>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
    sm_set = sm;
  }
  constexpr int ROWS = 32 * MI;
  hipLaunchKernelGGL((gap_fused_chunked_kernel<NT2, GRAD, MI>), dim3((M + ROWS - 1) / ROWS),
                     dim3(512), sm, st, M, sp_pad, dim_pad, (const u16*)q, dim_pad,
                     (const u16*)qs, dim_pad, (const u16*)qsT, sp_pad, ek, delta * delta, esite,
                     (float*)G, dim_pad, guard);
  return hipGetLastError();
}

// Rows per workgroup, 128 or 96: one workgroup fits a CU (LDS), so the
// launch takes ceil(workgroups / CUs) rounds of a workgroup's time, which
// scales with its rows; 96-row tiles win where 128-row ones leave the last
// round mostly empty (71,255 rows on 256 CUs: 3 rounds of 96 rows against 3
// of 128).
static int chunked_tile_rows(int M) {
  static int ncu = 0;
  if (ncu <= 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
      ncu = 256;
  }
  auto cost = [&](long rows) { return ((M + rows - 1) / rows + ncu - 1) / ncu * rows; };
  return cost(96) < cost(128) ? 96 : 128;
}

template <int NT2, bool GRAD>
static hipError_t run_fused_chunked(int M, int sp_pad, int dim_pad, const void* q, const void* qs,
                                    const void* qsT, const EpiKernel& ek, double delta,
                                    double* esite, void* G, const int* guard, hipStream_t st) {
  if (chunked_tile_rows(M) == 96)
    return launch_fused_chunked<NT2, GRAD, 3>(M, sp_pad, dim_pad, q, qs, qsT, ek, delta, esite,
                                              G, guard, st);
  return launch_fused_chunked<NT2, GRAD, 4>(M, sp_pad, dim_pad, q, qs, qsT, ek, delta, esite, G,
                                            guard, st);
}

template <int NT2, bool GRAD>
static hipError_t run_fused(int M, int sp_pad, int dim_pad, const void* q, const void* qs,
                            const void* qsT, const EpiKernel& ek, double delta, double* esite,
                            void* G, const int* guard, hipStream_t st) {
  static size_t sm_set = 0;
  const size_t sm = fused_region_w(sp_pad) + fused_region_t(dim_pad);
  if (sm > sm_set) {
    hipError_t e = hipFuncSetAttribute((const void*)gap_fused_kernel<NT2, GRAD>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
    sm_set = sm;
  }
  hipLaunchKernelGGL((gap_fused_kernel<NT2, GRAD>), dim3((M + 127) / 128), dim3(512), sm, st, M, sp_pad,
                     dim_pad, (const u16*)q, dim_pad, (const u16*)qs, dim_pad, (const u16*)qsT,
                     sp_pad, ek, delta * delta, esite, (float*)G, dim_pad, guard);
  return hipGetLastError();
}

// Energies only, for at most 128 sites (the affected sites of a Monte Carlo
// trial), where one workgroup per 128 rows would leave the chip idle: one
// workgroup per 32 sparse points, its four waves owning 16 * MT rows each
// and reading their MFMA fragments straight from memory (the few q rows and
// the Qs slab are L2-resident).  Each workgroup writes the row sums of
// alpha_s K_is^zeta over its sparse points to epart[workgroup][row]; the
// row-sum kernel adds them in workgroup order.  (Summing them in the last
// workgroup to finish instead needs device-scope fences, which write back
// and invalidate L2 and cost more than the extra launch.)
#define TGX_FEW_NC 2  // 16-column tiles per workgroup
template <int MT>
__global__ __launch_bounds__(256) void gap_energy_few_kernel(
    int M, int dim_pad, const u16* __restrict__ q, const u16* __restrict__ qs, EpiKernel ek,
    float* __restrict__ part, const int* __restrict__ guard) {
  if (guard && *guard == 0) return;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col0 = blockIdx.x * 16 * TGX_FEW_NC;
  const int kq = (lane >> 4) * 8;
  f32x4 acc[MT][TGX_FEW_NC];
#pragma unroll
  for (int mi = 0; mi < MT; mi++)
#pragma unroll
    for (int nj = 0; nj < TGX_FEW_NC; nj++) acc[mi][nj] = f32x4{0.f, 0.f, 0.f, 0.f};
  // rows past M read row M - 1 (a row of D depends only on its row of A;
  // those rows are never stored)
  const u16* ap[MT];
#pragma unroll
  for (int mi = 0; mi < MT; mi++)
    ap[mi] = q + (size_t)min((wave * MT + mi) * 16 + (lane & 15), M - 1) * dim_pad + kq;
  const u16* bp[TGX_FEW_NC];
#pragma unroll
  for (int nj = 0; nj < TGX_FEW_NC; nj++)
    bp[nj] = qs + (size_t)(col0 + nj * 16 + (lane & 15)) * dim_pad + kq;
  // KC k-steps of fragments are loaded before their products, so that a
  // wave waits for memory once per KC steps (q was just written by another
  // kernel and comes from the memory-side cache); steps past dim_pad re-read
  // the last slab and multiply a zero Qs fragment
  constexpr int KC = 6;
  for (int k0 = 0; k0 < dim_pad; k0 += 32 * KC) {
    bf16x8 a[KC][MT], b[KC][TGX_FEW_NC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
      const int kk = min(k0 + 32 * kc, dim_pad - 32);
#pragma unroll
      for (int nj = 0; nj < TGX_FEW_NC; nj++)
        b[kc][nj] = *reinterpret_cast<const bf16x8*>(bp[nj] + kk);
#pragma unroll
      for (int mi = 0; mi < MT; mi++) a[kc][mi] = *reinterpret_cast<const bf16x8*>(ap[mi] + kk);
    }
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
      if (k0 + 32 * kc >= dim_pad) {
#pragma unroll
        for (int nj = 0; nj < TGX_FEW_NC; nj++) b[kc][nj] = bf16x8{};
      }
#pragma unroll
      for (int mi = 0; mi < MT; mi++)
#pragma unroll
        for (int nj = 0; nj < TGX_FEW_NC; nj++)
          acc[mi][nj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kc][mi], b[kc][nj], acc[mi][nj], 0, 0, 0);
    }
  }
  // D[row][col]: lane holds rows (lane >> 4) * 4 + v of column lane & 15
  const float* al = reinterpret_cast<const float*>(ek.alphas);
  float rs[MT][4];
#pragma unroll
  for (int mi = 0; mi < MT; mi++)
#pragma unroll
    for (int v = 0; v < 4; v++) rs[mi][v] = 0.f;
#pragma unroll
  for (int nj = 0; nj < TGX_FEW_NC; nj++) {
    const int c = col0 + nj * 16 + (lane & 15);
    const float w = c < ek.n_sparse ? al[c] : 0.f;
#pragma unroll
    for (int mi = 0; mi < MT; mi++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        float kz, kz1;
        kpow_f<0>(acc[mi][nj][v], ek.zint, &kz, &kz1);
        rs[mi][v] += w * kz;
      }
  }
#pragma unroll
  for (int mi = 0; mi < MT; mi++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      float x = rs[mi][v];
      x += __shfl_xor(x, 1, 64);
      x += __shfl_xor(x, 2, 64);
      x += __shfl_xor(x, 4, 64);
      x += __shfl_xor(x, 8, 64);
      const int r = (wave * MT + mi) * 16 + (lane >> 4) * 4 + v;
      if ((lane & 15) == 0 && r < M) part[(size_t)blockIdx.x * M + r] = x;
    }
}

template <typename TI>
static hipError_t run(int n_sites, int dim, int dim_pad, int n_sparse, int sp_pad,
                      const void* q, const void* qs, const void* qsT, const void* alphas,
                      double delta, double zeta, void* wmat, void* epart, double* esite,
                      void* G, int want_grad, const int* guard, double* eatom,
                      const double* e0, const int* species, int* few_parts,
                      const GemmPlan* plan, hipStream_t st) {
  typedef typename MF<TI>::acc_t acc_t;
  const int M = n_sites;
  if (few_parts) *few_parts = 0;
  if (M == 0) return hipSuccess;
  const int zint = fabs(zeta - (double)(int)zeta) < 1e-5 ? (int)zeta : 0;
  const double zz = zint > 0 ? (double)zint : zeta;
  EpiKernel ek{alphas, n_sparse, zint, zz, zz * delta * delta, wmat, sp_pad, epart,
               eatom, e0, species};
  EpiStore es{G, dim_pad};
  if constexpr (sizeof(TI) == 2) {
    // TGX_GEMM_UNFUSED=1 selects the two separate GEMMs (for comparisons)
    // TGX_GEMM_UNFUSED=1 or plan path 2: the separate GEMMs
    const char* ufe = getenv("TGX_GEMM_UNFUSED");
    const bool uf_sep = (ufe && ufe[0] == '1') || (plan && plan->path == 2);
    const char* uf = uf_sep ? "1" : nullptr;
    if (!want_grad && zint >= 1 && zint <= 8 && M <= 128 && sp_pad % 32 == 0 &&
        dim_pad % 32 == 0 && !(uf && uf[0] == '1')) {
      const int nwg = sp_pad / (16 * TGX_FEW_NC);
      if (M <= 64)
        hipLaunchKernelGGL(gap_energy_few_kernel<1>, dim3(nwg), dim3(256), 0, st, M, dim_pad,
                           (const u16*)q, (const u16*)qs, ek, (float*)epart, guard);
      else
        hipLaunchKernelGGL(gap_energy_few_kernel<2>, dim3(nwg), dim3(256), 0, st, M, dim_pad,
                           (const u16*)q, (const u16*)qs, ek, (float*)epart, guard);
      if (few_parts)  // the caller sums the partials itself
        *few_parts = nwg;
      else
        hipLaunchKernelGGL(esum_kernel<float>, dim3((M + TGX_ESUM_WG - 1) / TGX_ESUM_WG), dim3(TGX_ESUM_WG), 0, st, M, nwg,
                           (const float*)epart, delta * delta, esite, ek, guard);
      return hipGetLastError();
    }
    if (zint >= 1 && zint <= 8 && sp_pad > 64 * TGX_FUSED_MAXT &&
        dim_pad <= 64 * TGX_FUSED_MAXT && !(uf && uf[0] == '1')) {
      if (!want_grad)
        return run_fused_chunked<1, false>(M, sp_pad, dim_pad, q, qs, qsT, ek, delta, esite, G,
                                           guard, st);
      switch (dim_pad / 64) {
#define TGX_FUSEDC_CASE(t) \
  case t:                  \
    return run_fused_chunked<t, true>(M, sp_pad, dim_pad, q, qs, qsT, ek, delta, esite, G, guard, st);
        TGX_FUSEDC_CASE(1) TGX_FUSEDC_CASE(2) TGX_FUSEDC_CASE(3) TGX_FUSEDC_CASE(4)
        TGX_FUSEDC_CASE(5) TGX_FUSEDC_CASE(6) TGX_FUSEDC_CASE(7)
#undef TGX_FUSEDC_CASE
      }
    }
    if (zint >= 1 && zint <= 8 && sp_pad <= 64 * TGX_FUSED_MAXT &&
        dim_pad <= 64 * TGX_FUSED_MAXT &&
        fused_region_w(sp_pad) + fused_region_t(dim_pad) <= 160 * 1024 &&
        !(uf && uf[0] == '1')) {
      if (!want_grad)  // energies only: stage 1 and the row sums
        return run_fused<1, false>(M, sp_pad, dim_pad, q, qs, qsT, ek, delta, esite, G, guard,
                                   st);
      switch (dim_pad / 64) {
#define TGX_FUSED_CASE(t) \
  case t:                 \
    return run_fused<t, true>(M, sp_pad, dim_pad, q, qs, qsT, ek, delta, esite, G, guard, st);
        TGX_FUSED_CASE(1) TGX_FUSED_CASE(2) TGX_FUSED_CASE(3) TGX_FUSED_CASE(4)
        TGX_FUSED_CASE(5) TGX_FUSED_CASE(6) TGX_FUSED_CASE(7) TGX_FUSED_CASE(8)
#undef TGX_FUSED_CASE
      }
    }
  }
  // separate GEMMs, tiles per the plan (epilogue GEMM t0: 64x64, 128x64,
  // 64x128; store GEMM t1: the default of the type or its alternative)
  const int t0 = plan ? plan->t0 : 0, t1 = plan ? plan->t1 : 0;
  int nt1 = 0;
  auto gemm0 = [&](auto TMc, auto TNc) {
    constexpr int BM = 32 * decltype(TMc)::value, BN = 32 * decltype(TNc)::value;
    const int mt = (M + BM - 1) / BM;
    nt1 = (sp_pad + BN - 1) / BN;
    hipLaunchKernelGGL((gemm_nt_kernel<TI, 0, decltype(TMc)::value, decltype(TNc)::value>),
                       dim3((nt1 * mt + 7) / 8 * 8), dim3(256), 0, st, M, sp_pad, dim_pad,
                       (const TI*)q, dim_pad, (const TI*)qs, dim_pad, ek, es, nt1, nt1 * mt,
                       guard);
  };
  if (t0 == 1) gemm0(ICv<4>{}, ICv<2>{});
  else if (t0 == 2) gemm0(ICv<2>{}, ICv<4>{});
  else gemm0(ICv<2>{}, ICv<2>{});
  hipLaunchKernelGGL(esum_kernel<acc_t>, dim3((M + TGX_ESUM_WG - 1) / TGX_ESUM_WG), dim3(TGX_ESUM_WG), 0, st, M, nt1,
                     (const acc_t*)epart, delta * delta, esite, ek, guard);
  if (want_grad) {
    auto gemm1 = [&](auto TMc, auto TNc) {
      constexpr int BM = 32 * decltype(TMc)::value, BN = 32 * decltype(TNc)::value;
      const int mt = (M + BM - 1) / BM;
      const int nt2 = (dim_pad + BN - 1) / BN;
      hipLaunchKernelGGL((gemm_nt_kernel<TI, 1, decltype(TMc)::value, decltype(TNc)::value>),
                         dim3((nt2 * mt + 7) / 8 * 8), dim3(256), 0, st, M, dim_pad, sp_pad,
                         (const TI*)wmat, sp_pad, (const TI*)qsT, sp_pad, ek, es, nt2, nt2 * mt,
                         guard);
    };
    constexpr int DM = Tile<TI, 1>::TM, DN = Tile<TI, 1>::TN;
    if (t1 == 1 && sizeof(TI) < 8) {
      if constexpr (DM == 4) gemm1(ICv<2>{}, ICv<2>{});
      else gemm1(ICv<4>{}, ICv<4>{});
    } else {
      gemm1(ICv<DM>{}, ICv<DN>{});
    }
  }
  return hipGetLastError();
}

hipError_t launch_gap_kernel(int mode, int n_sites, int dim, int dim_pad, int n_sparse,
                             int sp_pad, const void* q, const void* qs, const void* qsT,
                             const void* alphas, double delta, double zeta, void* wmat,
                             void* epart, double* esite, void* G, int want_grad,
                             const int* guard, hipStream_t st, double* eatom, const double* e0,
                             const int* species, int* few_parts, const GemmPlan* plan) {
  if (mode == TGX_FP64)
    return run<double>(n_sites, dim, dim_pad, n_sparse, sp_pad, q, qs, qsT, alphas, delta, zeta,
                       wmat, epart, esite, G, want_grad, guard, eatom, e0, species, few_parts, plan, st);
  if (mode == TGX_FP32)
    return run<float>(n_sites, dim, dim_pad, n_sparse, sp_pad, q, qs, qsT, alphas, delta, zeta,
                      wmat, epart, esite, G, want_grad, guard, eatom, e0, species, few_parts, plan, st);
  return run<u16>(n_sites, dim, dim_pad, n_sparse, sp_pad, q, qs, qsT, alphas, delta, zeta, wmat,
                  epart, esite, G, want_grad, guard, eatom, e0, species, few_parts, plan, st);
}

}  // namespace tgx
