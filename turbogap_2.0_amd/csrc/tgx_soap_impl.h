// SOAP expansion (forward) and its adjoint (forces, virial) on the device:
// kernel templates, shared by the single-cell and the batched translation
// units (each instantiates one pair geometry).
//
// Layout: a workgroup runs up to four independent 64-lane waves (fewer when
// a large model needs more LDS per site); a wave serves two sites with 32
// lanes each (one site with 64 lanes for models with more than 16 radial
// channels), each site with its own LDS region, so nothing waits across
// sites.  Forward, per site:
//   A  lanes = neighbour slots: minimum-image vector, radial values R_n(r)
//      and A_l(r) Y_lm(u) for the valid pairs, compacted into LDS in chunks
//      (FP32 with two sites per wave: 28 pairs a chunk, the harmonics one
//      block of degrees at a time, each block followed by its B step);
//   B  lanes = lm: c[n][lm] += R_n(j) A_l Y_lm(j) accumulated in registers;
//   C  power spectrum p[(n<=n'),l], its norm and q = p/|p|.
// Adjoint, per site: dE/dp from G = dE/dq, dE/dc (lanes = lm), then lanes =
// pairs recompute the basis with derivatives and contract it with dE/dc into
// g_ij = dE_i/dr_ij; the site virial is reduced over the wave.
// l_max and n_max are compile-time bounds (loops fully unrolled, guarded by
// the runtime values) so that the recurrences live in registers.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

typedef unsigned short u16;

template <typename T> struct Radial;
template <> struct Radial<double> {
  // FP64: closed form with the FP64 W for every pair.
  template <int NMX>
  __device__ __forceinline__ static void eval(const SoapDev* sd, const double* W, const float2*, int s,
                              double r, double* R, double* dR) {
    double Bv[NMX], dB[NMX], e, de;
    radial_basis_fp64<NMX>(sd, W, s, r, Bv, dB);
    radial_envelope<double>(sd, s, r, &e, &de);
#pragma unroll
    for (int n = 0; n < NMX; n++) {
      R[n] = e * Bv[n];
      dR[n] = de * Bv[n] + e * dB[n];
    }
  }
};
template <> struct Radial<float> {
  // FP32: table of the smooth part (built in FP64), analytic envelope.
  template <int NMX>
  __device__ __forceinline__ static void eval(const SoapDev* sd, const double*, const float2* tab, int s,
                              float r, float* R, float* dR) {
    float Bv[NMX], dB[NMX], e, de;
    radial_basis_tab<NMX>(sd, tab, s, r, Bv, dB, sd->nmax[s]);
    radial_envelope<float>(sd, s, r, &e, &de);
#pragma unroll
    for (int n = 0; n < NMX; n++) {
      R[n] = e * Bv[n];
      dR[n] = de * Bv[n] + e * dB[n];
    }
  }
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ u16 f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  return (u16)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Minimum-image pair vector by rounding (the neighbour list already fixed the
// pair set with the reference's arithmetic; here only the vector is needed).
__device__ __forceinline__ double pair_vec_fast(const double* __restrict__ pos, int i,
                                                int packed, const Cell& c, double* dv) {
  const int j = packed & TGX_ATOM_MASK, img = packed >> TGX_IMG_SHIFT;
  double pj[3];
  if (img == 0) {
    pj[0] = pos[3 * j];
    pj[1] = pos[3 * j + 1];
    pj[2] = pos[3 * j + 2];
  } else {
    image_pos(pos, j, img, c, pj);
  }
  double dx = pj[0] - pos[3 * i], dy = pj[1] - pos[3 * i + 1], dz = pj[2] - pos[3 * i + 2];
  dx -= c.La * rint(dx * c.iLa);
  dy -= c.Lb * rint(dy * c.iLb);
  dz -= c.Lc * rint(dz * c.iLc);
  dv[0] = dx;
  dv[1] = dy;
  dv[2] = dz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Sites per 64-lane wave: two (32 lanes each) when the model has at most 16
// radial channels, else one.  Pairs per LDS chunk and the per-site LDS
// footprint (elements of T).
template <int NRP> struct Spw { static constexpr int v = NRP <= 16 ? 2 : 1; };
// FP32 with two sites per wave: the forward evaluates the harmonics of a chunk
// in blocks of degrees (hblk_*), so that LDS holds one block of a chunk at a
// time and a chunk can take 28 pairs instead of 12.
template <typename T, int NRP> struct FwdBlocked {
  static constexpr bool v = sizeof(T) == 4 && Spw<NRP>::v == 2;
};
template <typename T, int NRP> struct FwdChunk {
  static constexpr int v = FwdBlocked<T, NRP>::v ? (NRP <= 8 ? 28 : 24)
                                                 : (sizeof(T) == 8 ? 16 : Spw<NRP>::v == 2 ? 12 : 32);
};
// Harmonic blocks: consecutive degrees with at most 32 (l, m) values each
// (l_max 8: l = 0-4, 5-6, 7-8; l_max 10 adds 9 and 10).  A pair's block row
// in LDS has the odd stride HSTR, so lanes storing different pairs' values
// hit different banks.
__host__ __device__ constexpr int hblk_of(int l) {
  int b = 0, l0 = 0;
  for (int k = 1; k <= l; k++)
    if ((k + 1) * (k + 1) - l0 * l0 > 32) {
      b++;
      l0 = k;
    }
  return b;
}
__host__ __device__ constexpr int hblk_first(int b) {
  int l = 0;
  while (hblk_of(l) < b) l++;
  return l;
}
constexpr int HSTR = 33;

// Compacted valid-pair slots per site (two scan rounds of the site's lanes).
template <int NRP> struct Capv { static constexpr int v = 2 * (64 / Spw<NRP>::v); };

// Per-site LDS footprint in bytes (16-byte multiple): forward = harmonics
// and radial values of a chunk of pairs (the first part reused for c) plus
// the compacted list entries of the valid pairs (32 bits); adjoint = dE/dc [L2][NRP] (rows
// of species sub-rows at 4-aligned offsets), the upper triangles of the
// coupling matrices M_l and the list entries and slots of the valid pairs.
// Forward region A holds the harmonics of a chunk (CH x L2, or one block of
// them, CH x HSTR), later c with rows padded to a multiple of four (NRP x L2P).
template <typename T, int NRP>
__host__ __device__ constexpr size_t fwd_region_a(int L2) {
  constexpr int CH = FwdChunk<T, NRP>::v;
  const size_t rows = (size_t)CH * (FwdBlocked<T, NRP>::v ? HSTR : L2);
  return rows > (size_t)NRP * ((L2 + 3) & ~3) ? rows : (size_t)NRP * ((L2 + 3) & ~3);
}
template <typename T, int NRP>
__host__ __device__ constexpr size_t fwd_site_bytes(int L2) {
  constexpr int CH = FwdChunk<T, NRP>::v;
  return ((sizeof(T) * (fwd_region_a<T, NRP>(L2) + (size_t)CH * NRP) + 4 * Capv<NRP>::v) + 15) /
         16 * 16;
}
// The symmetric coupling matrices M_l keep their upper triangle (n <= n2) in
// LDS, tri(n, n2) = n (2 NRP - n + 1) / 2 + n2 - n; the stride per l is odd so
// that lanes reading the same entry of different l hit different banks.
template <int NRP> struct MStride {
  static constexpr int t = NRP * (NRP + 1) / 2;
  static constexpr int v = t % 2 ? t : t + 1;
};
template <int NRP>
__host__ __device__ constexpr int tri(int a, int b) {
  return a <= b ? a * (2 * NRP - a + 1) / 2 + (b - a) : b * (2 * NRP - b + 1) / 2 + (a - b);
}
template <typename T, int NRP>
__host__ __device__ constexpr size_t bwd_site_bytes(int L2, int Lp) {
  return ((sizeof(T) * ((size_t)NRP * L2 + (size_t)Lp * MStride<NRP>::v) + 6 * Capv<NRP>::v) +
          15) / 16 * 16 + 64;  // + the site's running virial and 2-body energy (8 doubles)
}

// Appends, in slot order, the pairs within the neighbour species' cutoff --
// their slots to slots[] and their list entries to ents[], either of which
// may be null -- scanning the list from *next in steps of LW slots while at
// least LW entries are free.  Returns the count appended; the same on every
// lane of the site.  on_skip(slot) runs for slots that are scanned but not
// appended.
template <int LW, int SPW, typename GEO, typename F>
__device__ __forceinline__ int scan_pairs(const SoapDev* sd, const double* pos,
                                          const int* species, const Cell& cell, const int* row,
                                          int i, int cnt, int* next, u16* slots, int* ents,
                                          int capv, int sub, int lane, F on_skip) {
  int nv = 0;
  while (*next < cnt && nv <= capv - LW) {
    const int slot = *next + lane;
    bool valid = false;
    int e = 0;
    if (slot < cnt) {
      e = row[slot];
      double dv[3];
      const double rd = pair_vec_fast(pos, i, e, cell, dv);
      valid = rd < sd->rh[species[e & TGX_ATOM_MASK] - 1];
      if (!valid) on_skip(slot);
    }
    unsigned long long mask = __ballot(valid);
    if (SPW == 2) mask = (mask >> (sub * 32)) & 0xffffffffull;
    const int jc = __popcll(mask & ((1ull << lane) - 1ull));
    if (valid && slots) slots[nv + jc] = (u16)slot;
    if (valid && ents) ents[nv + jc] = e;
    nv += __popcll(mask);
    *next += LW;
  }
  return nv;
}

// N values from LDS at p, which is 16-byte aligned, as 16-byte loads.
template <typename T, int N>
__device__ __forceinline__ void lds_row16(const T* p, T* out) {
  typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V;
  constexpr int PER = 16 / sizeof(T);
  static_assert(N % PER == 0, "row of whole 16-byte units");
  const V* v = reinterpret_cast<const V*>(p);
#pragma unroll
  for (int k = 0; k < N / PER; k++) {
    const V x = v[k];
    __builtin_memcpy(out + k * PER, &x, 16);
  }
}

// Sum over the LW lanes of a sub-wave.
template <int LW, typename T>
__device__ __forceinline__ T sub_sum(T v) {
#pragma unroll
  for (int o = LW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// A_l = t^(l(l+1)), t = exp(-sig_t^2/(2 r^2)), for l = 0..LMX.
template <typename T, int LMX>
__device__ __forceinline__ void ang_factors(T log2t, T* Al) {
  const T t2 = exp2_t((T)2 * log2t);
  T Tl = 1;
  Al[0] = 1;
#pragma unroll
  for (int l = 1; l <= LMX; l++) {
    Tl *= t2;
    Al[l] = Al[l - 1] * Tl;
  }
}

// Basis of one pair for the forward expansion: R_n(r) into rr[0..NRP) (at
// the neighbour species' radial offset, zeros elsewhere) and A_l(r) Y_lm(u)
// into ay[lm], the real harmonics by the m-outer recurrence in l.
template <typename T, int NRP, int NMX, int LMX>
__device__ __forceinline__ void fwd_pair_basis(const SoapDev* sd, const double* W,
                                               const float2* rtab, const double* pos,
                                               const int* species, const Cell& cell, int i,
                                               int e, int L, T* ay, T* rr) {
  double dv[3];
  const double rd = pair_vec_fast(pos, i, e, cell, dv);
  const int s = species[e & TGX_ATOM_MASK] - 1;
  const T r = (T)rd;
  const T ir = (T)1 / r;
  const T ux = (T)dv[0] * ir, uy = (T)dv[1] * ir, uz = (T)dv[2] * ir;
  T R[NMX], dR[NMX];
  Radial<T>::template eval<NMX>(sd, W, rtab, s, r, R, dR);
  const int o = sd->off[s];
  // R[n] = 0 for n >= n_max(s) and o + NMX <= NRP: no guard needed
#pragma unroll
  for (int n = 0; n < NRP; n++) rr[n] = (T)0;
#pragma unroll
  for (int n = 0; n < NMX; n++) rr[o + n] = R[n];
  T log2t, D;
  ang_params<T>(sd, s, r, &log2t, &D);
  T Al[LMX + 1];
  ang_factors<T, LMX>(log2t, Al);
  T Cm = 1, Sm = 0;
#pragma unroll
  for (int m = 0; m <= LMX; m++) {
    if (m <= L) {
      T pm2 = 0, pm1 = SphK<T>::P(m);
#pragma unroll
      for (int l = m; l <= LMX; l++) {
        if (l <= L) {
          T P = pm1;
          if (l > m) {
            const int ix = l * (TGX_MAX_LMAX + 1) + m;
            P = SphK<T>::A(ix) * uz * pm1 - SphK<T>::B(ix) * pm2;
            pm2 = pm1;
            pm1 = P;
          }
          const T AP = Al[l] * P;
          ay[l * l + l + m] = AP * Cm;
          if (m > 0) ay[l * l + l - m] = AP * Sm;
        }
      }
      const T Cn = ux * Cm - uy * Sm;
      Sm = ux * Sm + uy * Cm;
      Cm = Cn;
    }
  }
}

// The same basis evaluated block by block (blocked forward): the pair state
// that carries from one harmonic block to the next -- u_z, A_l, the trig
// factors C_m + i S_m = (u_x + i u_y)^m and the two last Legendre values of
// every m-chain -- stays in registers.  The values and the arithmetic are
// those of fwd_pair_basis.
template <typename T, int LMX>
struct PairHarm {
  T uz;
  T Al[LMX + 1], C[LMX + 1], S[LMX + 1], p1[LMX + 1], p2[LMX + 1];
};

// R_n(r) into rr[0..NRP) and the pair state of the harmonic recurrences.
template <typename T, int NRP, int NMX, int LMX>
__device__ __forceinline__ void fwd_pair_setup(const SoapDev* sd, const double* W,
                                               const float2* rtab, const double* pos,
                                               const int* species, const Cell& cell, int i,
                                               int e, T* rr, PairHarm<T, LMX>& h) {
  double dv[3];
  const double rd = pair_vec_fast(pos, i, e, cell, dv);
  const int s = species[e & TGX_ATOM_MASK] - 1;
  const T r = (T)rd;
  const T ir = (T)1 / r;
  const T ux = (T)dv[0] * ir, uy = (T)dv[1] * ir;
  h.uz = (T)dv[2] * ir;
  T R[NMX], dR[NMX];
  Radial<T>::template eval<NMX>(sd, W, rtab, s, r, R, dR);
  const int o = sd->off[s];
#pragma unroll
  for (int n = 0; n < NRP; n++) rr[n] = (T)0;
#pragma unroll
  for (int n = 0; n < NMX; n++) rr[o + n] = R[n];
  T log2t, D;
  ang_params<T>(sd, s, r, &log2t, &D);
  ang_factors<T, LMX>(log2t, h.Al);
  h.C[0] = 1;
  h.S[0] = 0;
#pragma unroll
  for (int m = 0; m < LMX; m++) {
    h.C[m + 1] = ux * h.C[m] - uy * h.S[m];
    h.S[m + 1] = ux * h.S[m] + uy * h.C[m];
  }
#pragma unroll
  for (int m = 0; m <= LMX; m++) {
    h.p1[m] = SphK<T>::P(m);
    h.p2[m] = 0;
  }
}

// State of a chunk lane without a pair: zero radial row and zero harmonics.
template <typename T, int NRP, int LMX>
__device__ __forceinline__ void fwd_pair_zero(T* rr, PairHarm<T, LMX>& h) {
#pragma unroll
  for (int n = 0; n < NRP; n++) rr[n] = (T)0;
  h.uz = 0;
#pragma unroll
  for (int m = 0; m <= LMX; m++) h.Al[m] = h.C[m] = h.S[m] = h.p1[m] = h.p2[m] = 0;
}

// Harmonic block B (degrees hblk_first(B) .. hblk_first(B+1)-1, at most
// l_max): A_l Y_lm into ay[lm - l0^2], advancing each m-chain's recurrence.
template <typename T, int LMX, int B>
__device__ __forceinline__ void fwd_harm_block(PairHarm<T, LMX>& h, int L, T* ay) {
  constexpr int l0 = hblk_first(B), l1 = hblk_first(B + 1) - 1 < LMX ? hblk_first(B + 1) - 1 : LMX;
  constexpr int base = l0 * l0;
#pragma unroll
  for (int m = 0; m <= l1; m++) {
    if (m <= L) {
#pragma unroll
      for (int l = m > l0 ? m : l0; l <= l1; l++) {
        if (l <= L) {
          T P = h.p1[m];
          if (l > m) {
            const int ix = l * (TGX_MAX_LMAX + 1) + m;
            P = SphK<T>::A(ix) * h.uz * h.p1[m] - SphK<T>::B(ix) * h.p2[m];
            h.p2[m] = h.p1[m];
            h.p1[m] = P;
          }
          const T AP = h.Al[l] * P;
          ay[l * l + l + m - base] = AP * h.C[m];
          if (m > 0) ay[l * l + l - m - base] = AP * h.S[m];
        }
      }
    }
  }
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Minimum waves per SIMD the FP32 forward kernel is register-limited to
// (measured on gfx950 at 32k sites: 98.5 us at four waves with 28-pair
// chunks, 151 us at five with 20-pair chunks, which spills); FP64 is left
// unconstrained.
template <typename T> struct FwdWaves { static constexpr int v = 4; };
template <> struct FwdWaves<double> { static constexpr int v = 1; };
// The same for the FP32 adjoint (4: 148 us at 32k atoms vs 163 unconstrained
// at 3 waves and 197 at 5 with spills).
template <typename T> struct BwdWaves { static constexpr int v = 4; };
template <> struct BwdWaves<double> { static constexpr int v = 1; };

template <typename T, int NRP, int NMX, int LMX, bool EXL, typename GEO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FwdWaves<T>::v))) void soap_fwd_kernel(
    const SoapDev* __restrict__ sd, const double* __restrict__ W,
    const float2* __restrict__ rtab, const double* __restrict__ pos,
    const int* __restrict__ species, GEO geo, int maxnb, const int* __restrict__ count,
    const int* __restrict__ list, const int* __restrict__ site_atom, int n_sites,
    const int* __restrict__ dmap, T* __restrict__ q, int q_ld, u16* __restrict__ qb,
    int qb_ld, T* __restrict__ cbuf, T* __restrict__ pnorm, int wpb) {
  constexpr int SPW = Spw<NRP>::v, LW = 64 / SPW;
  constexpr bool BLK = FwdBlocked<T, NRP>::v;
  constexpr int CH = FwdChunk<T, NRP>::v;
  // lm slots per lane: lm = lane + k LW, or (blocked) the lane-th value of
  // harmonic block k
  constexpr int NS = BLK ? hblk_of(LMX) + 1 : ((LMX + 1) * (LMX + 1) + LW - 1) / LW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // EXL: l_max equals the bound LMX, so L and L2 are compile-time constants
  const int L = EXL ? LMX : sd->l_max, L2 = EXL ? (LMX + 1) * (LMX + 1) : sd->L2;
  const int wave = threadIdx.x >> 6, lane64 = threadIdx.x & 63;
  const int sub = lane64 / LW, lane = lane64 % LW;
  constexpr int CAPV = Capv<NRP>::v;
  T* sAY = reinterpret_cast<T*>(smem_raw + (size_t)(wave * SPW + sub) * fwd_site_bytes<T, NRP>(L2));
  T* sR = sAY + fwd_region_a<T, NRP>(L2);
  // list entries of the valid pairs: the chunk lanes take them from LDS
  // instead of reloading the list row
  int* sEnt = reinterpret_cast<int*>(sR + CH * NRP);
  T* sC = sAY;  // reused once the accumulation is done
  const int site = (blockIdx.x * wpb + wave) * SPW + sub;
  if (site >= n_sites) return;
  const int i = site_atom ? site_atom[site] : site;  // null: site s is atom s
  const Cell cell = geo.of(i);
  const int si = species[i] - 1;
  const int cnt = min(count[i], maxnb);
  const int* row = list + (size_t)i * maxnb;

  T acc[NS][NRP];
#pragma unroll
  for (int k = 0; k < NS; k++)
#pragma unroll
    for (int n = 0; n < NRP; n++) acc[k][n] = (T)0;

  int next = 0;
  while (next < cnt) {
    // valid pairs first, so that the expensive basis evaluation runs on
    // packed lanes (list slots beyond the cutoff are skipped)
    const int nv = scan_pairs<LW, SPW, GEO>(sd, pos, species, cell, row, i, cnt, &next, nullptr,
                                            sEnt, CAPV, sub, lane, [](int) {});
    wave_sync();
    for (int c0 = 0; c0 < nv; c0 += CH) {
      const bool act = lane < CH && c0 + lane < nv;
      const int nj = min(CH, nv - c0);
      if constexpr (BLK) {
        // one harmonic block of the chunk in LDS at a time: lanes = pairs
        // advance their recurrences by one block, then lanes = the block's
        // lm accumulate it
        // lanes of the chunk without a pair write zero rows, so that the
        // accumulation runs over whole groups of four rows (CH is a multiple
        // of four) with its loads issued together
        PairHarm<T, LMX> h;
        if (act)
          fwd_pair_setup<T, NRP, NMX, LMX>(sd, W, rtab, pos, species, cell, i,
                                           sEnt[c0 + lane], sR + lane * NRP, h);
        else if (lane < CH)
          fwd_pair_zero<T, NRP, LMX>(sR + lane * NRP, h);
        const int nj4 = (nj + 3) & ~3;
        static_for<0, NS>([&](auto bc) {
          constexpr int B = decltype(bc)::value;
          constexpr int base = hblk_first(B) * hblk_first(B);
          constexpr int end = hblk_first(B + 1) * hblk_first(B + 1);
          const int nval = min(end, L2) - base;
          if (nval > 0) {
            if (lane < CH) fwd_harm_block<T, LMX, B>(h, L, sAY + lane * HSTR);
            wave_sync();
            for (int j = 0; j < nj4; j += 4) {
              // unconditional loads: a lane past the block reads a valid
              // entry and its accumulator is never stored
              T y[4];
#pragma unroll
              for (int u = 0; u < 4; u++) y[u] = sAY[(j + u) * HSTR + min(lane, nval - 1)];
#pragma unroll
              for (int u = 0; u < 4; u++)
#pragma unroll
                for (int n = 0; n < NRP; n++) acc[B][n] += sR[(j + u) * NRP + n] * y[u];
            }
            wave_sync();
          }
        });
      } else {
        if (act)
          fwd_pair_basis<T, NRP, NMX, LMX>(sd, W, rtab, pos, species, cell, i,
                                           sEnt[c0 + lane], L, sAY + lane * L2,
                                           sR + lane * NRP);
        wave_sync();
        for (int j = 0; j < nj; j++) {
          // unconditional loads: a slot past L2 reads a valid entry and its
          // accumulator is never stored
          T y[NS];
#pragma unroll
          for (int k = 0; k < NS; k++) y[k] = sAY[j * L2 + min(lane + k * LW, L2 - 1)];
#pragma unroll
          for (int n = 0; n < NRP; n++) {
            const T rv = sR[j * NRP + n];
#pragma unroll
            for (int k = 0; k < NS; k++) acc[k][n] += rv * y[k];
          }
        }
        wave_sync();
      }
    }
  }
  // central atom: l = 0 only
  if (lane == 0) {
    const int o = sd->off[si];
    for (int n = 0; n < sd->nmax[si]; n++) {
#pragma unroll
      for (int k = 0; k < NRP; k++)
        if (k == o + n) acc[0][k] += (T)sd->rself[si * TGX_MAX_NR + n];
    }
  }
  // c to HBM for the adjoint ([n][lm]) and to LDS with rows padded to a
  // multiple of four for the power spectrum
  const int L2P = (L2 + 3) & ~3;
  T* cs = cbuf + (size_t)site * NRP * L2;
  if constexpr (BLK) {
    static_for<0, NS>([&](auto bc) {
      constexpr int B = decltype(bc)::value;
      constexpr int base = hblk_first(B) * hblk_first(B);
      constexpr int end = hblk_first(B + 1) * hblk_first(B + 1);
      const int lm = base + lane;
      if (lm < min(end, L2)) {
#pragma unroll
        for (int n = 0; n < NRP; n++) {
          sC[n * L2P + lm] = acc[B][n];
          cs[n * L2 + lm] = acc[B][n];
        }
      }
    });
  } else {
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int lm = lane + k * LW;
      if (lm < L2) {
#pragma unroll
        for (int n = 0; n < NRP; n++) {
          sC[n * L2P + lm] = acc[k][n];
          cs[n * L2 + lm] = acc[k][n];
        }
      }
    }
  }
  wave_sync();
  // power spectrum, lanes = radial pairs (n1 <= n2): p_l = w sum_m
  // c[n1][l,m] c[n2][l,m] for every l, w = 1 or sqrt 2 (no divergence across
  // lanes), kept in registers; then the norm and q = p / |p| written as T
  // and BF16
  const int dim = sd->dim, Lp = L + 1, npair = sd->npair;
  constexpr int KPI = (NRP * (NRP + 1) / 2 + LW - 1) / LW;  // radial pairs per lane
  T* qs = q + (size_t)site * q_ld;
  u16* qbs = qb ? qb + (size_t)site * qb_ld : nullptr;
  T pv[KPI][LMX + 1];
  T ss = 0;
#pragma unroll
  for (int t = 0; t < KPI; t++) {
    const int kp = lane + t * LW;
    if (kp < npair) {
      const int dm = dmap[kp * Lp];
      const T* ca = sC + (dm & 255) * L2P;
      const T* cb = sC + ((dm >> 8) & 255) * L2P;
      const T w = (dm & 255) == ((dm >> 8) & 255) ? (T)1 : (T)1.41421356237309504880;
#pragma unroll
      for (int l = 0; l <= LMX; l++) {
        if (l <= L) {
          T a = 0;
#pragma unroll
          for (int m = 0; m <= 2 * l; m++) a += ca[l * l + m] * cb[l * l + m];
          pv[t][l] = w * a;
          ss += pv[t][l] * pv[t][l];
        }
      }
    }
  }
  ss = sub_sum<LW>(ss);
  const T inv = (T)1 / sqrt(ss);
  if ((size_t)q_ld <= fwd_region_a<T, NRP>(L2) && (!qbs || qb_ld <= q_ld)) {
    // q (with its zero padding) staged in LDS, where c was, and written in
    // whole 16-byte units: lanes writing their own Lp consecutive components
    // would leave every store uncoalesced
    T* sQ = sAY;
    wave_sync();
#pragma unroll
    for (int t = 0; t < KPI; t++) {
      const int kp = lane + t * LW;
      if (kp < npair) {
#pragma unroll
        for (int l = 0; l <= LMX; l++)
          if (l <= L) sQ[kp * Lp + l] = pv[t][l] * inv;
      }
    }
    for (int d = dim + lane; d < q_ld; d += LW) sQ[d] = (T)0;
    wave_sync();
    typedef typename std::conditional<sizeof(T) == 4, float4, double2>::type V;
    constexpr int PER = 16 / sizeof(T);
    for (int d = lane * PER; d < q_ld; d += LW * PER)
      *reinterpret_cast<V*>(qs + d) = *reinterpret_cast<const V*>(sQ + d);
    if (qbs) {
      for (int d = lane * 8; d < qb_ld; d += LW * 8) {
        unsigned int w[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
          w[k] = (unsigned int)f2bf((float)sQ[d + 2 * k]) |
                 ((unsigned int)f2bf((float)sQ[d + 2 * k + 1]) << 16);
        *reinterpret_cast<uint4*>(qbs + d) = uint4{w[0], w[1], w[2], w[3]};
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < KPI; t++) {
      const int kp = lane + t * LW;
      if (kp < npair) {
#pragma unroll
        for (int l = 0; l <= LMX; l++) {
          if (l <= L) {
            const T v = pv[t][l] * inv;
            qs[kp * Lp + l] = v;
            if (qbs) qbs[kp * Lp + l] = f2bf((float)v);
          }
        }
      }
    }
    for (int d = dim + lane; d < q_ld; d += LW) qs[d] = (T)0;
    if (qbs)
      for (int d = dim + lane; d < qb_ld; d += LW) qbs[d] = 0;
  }
  if (lane == 0) pnorm[site] = sqrt(ss);
}

// Forward expansion of few sites (up to TGX_FWD_SPLIT_MAX: the affected
// sites of a Monte Carlo trial, on request), where the one-wave-per-site
// kernel is a long serial chain on an idle chip: one workgroup per site, its four
// waves taking a contiguous quarter of the neighbour list each (scan, basis
// into LDS in chunks of CH pairs, c partials in registers with lanes = lm),
// the partials added in wave order, then the power spectrum with threads
// over the descriptor components and a workgroup reduction of the norm.
// Output layout as soap_fwd_kernel's.  FP32 only.
#define TGX_FWD_SPLIT_MAX 256
template <int NRP, int LMX>
__host__ __device__ constexpr size_t fwd_split_bytes(int L2) {
  return sizeof(float) * ((size_t)4 * 16 * (L2 + NRP) + (size_t)5 * NRP * ((L2 + 3) & ~3)) +
         2 * 4 * 64;
}
template <typename T, int NRP, int NMX, int LMX, bool EXL, typename GEO>
__global__ __launch_bounds__(256) void soap_fwd_split_kernel(
    const SoapDev* __restrict__ sd, const double* __restrict__ W,
    const float2* __restrict__ rtab, const double* __restrict__ pos,
    const int* __restrict__ species, GEO geo, int maxnb, const int* __restrict__ count,
    const int* __restrict__ list, const int* __restrict__ site_atom, int n_sites,
    const int* __restrict__ dmap, T* __restrict__ q, int q_ld, u16* __restrict__ qb,
    int qb_ld, T* __restrict__ cbuf, T* __restrict__ pnorm) {
  constexpr int CH = 16;
  constexpr int NS = ((LMX + 1) * (LMX + 1) + 63) / 64;  // lm slots per lane
  constexpr int IPT = (NRP * (NRP + 1) / 2 * (LMX + 1) + 255) / 256;  // components per thread
  const int L = EXL ? LMX : sd->l_max, L2 = EXL ? (LMX + 1) * (LMX + 1) : sd->L2;
  const int L2P = (L2 + 3) & ~3;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sAY = reinterpret_cast<T*>(smem_raw) + (size_t)wave * CH * (L2 + NRP);  // [CH][L2]
  T* sR = sAY + CH * L2;                                                     // [CH][NRP]
  T* sP = reinterpret_cast<T*>(smem_raw) + (size_t)4 * CH * (L2 + NRP);     // [4][NRP][L2P]
  T* sC = sP + (size_t)4 * NRP * L2P;                                        // [NRP][L2P]
  u16* sIdx = reinterpret_cast<u16*>(sC + (size_t)NRP * L2P) + wave * 64;
  __shared__ T sred[4];
  const int site = blockIdx.x;
  if (site >= n_sites) return;
  const int i = site_atom ? site_atom[site] : site;  // null: site s is atom s
  const Cell cell = geo.of(i);
  const int si = species[i] - 1;
  const int cnt = min(count[i], maxnb);
  const int* row = list + (size_t)i * maxnb;
  T acc[NS][NRP];
#pragma unroll
  for (int k = 0; k < NS; k++)
#pragma unroll
    for (int n = 0; n < NRP; n++) acc[k][n] = (T)0;
  const int quarter = (cnt + 3) / 4;
  const int lo = min(cnt, wave * quarter), hi = min(cnt, lo + quarter);
  for (int base = lo; base < hi; base += 64) {
    // valid pairs of the next 64 slots, compacted in slot order
    const int slot = base + lane;
    bool valid = false;
    if (slot < hi) {
      const int e = row[slot];
      double dv[3];
      const double rd = pair_vec_fast(pos, i, e, cell, dv);
      valid = rd < sd->rh[species[e & TGX_ATOM_MASK] - 1];
    }
    const unsigned long long mask = __ballot(valid);
    if (valid) sIdx[__popcll(mask & ((1ull << lane) - 1ull))] = (u16)slot;
    const int nv = __popcll(mask);
    wave_sync();
    for (int c0 = 0; c0 < nv; c0 += CH) {
      if (lane < CH && c0 + lane < nv)
        fwd_pair_basis<T, NRP, NMX, LMX>(sd, W, rtab, pos, species, cell, i, row[sIdx[c0 + lane]],
                                         L, sAY + lane * L2, sR + lane * NRP);
      wave_sync();
      const int nj = min(CH, nv - c0);
      for (int j = 0; j < nj; j++) {
        T y[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) y[k] = sAY[j * L2 + min(lane + k * 64, L2 - 1)];
#pragma unroll
        for (int n = 0; n < NRP; n++) {
          const T rv = sR[j * NRP + n];
#pragma unroll
          for (int k = 0; k < NS; k++) acc[k][n] += rv * y[k];
        }
      }
      wave_sync();
    }
  }
  // central atom: l = 0 only
  if (wave == 0 && lane == 0) {
    const int o = sd->off[si];
    for (int n = 0; n < sd->nmax[si]; n++) {
#pragma unroll
      for (int k = 0; k < NRP; k++)
        if (k == o + n) acc[0][k] += (T)sd->rself[si * TGX_MAX_NR + n];
    }
  }
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const int lm = lane + k * 64;
    if (lm < L2) {
#pragma unroll
      for (int n = 0; n < NRP; n++) sP[((size_t)wave * NRP + n) * L2P + lm] = acc[k][n];
    }
  }
  __syncthreads();
  T* cs = cbuf + (size_t)site * NRP * L2;
  for (int x = tid; x < NRP * L2; x += 256) {
    const int n = x / L2, lm = x - n * L2;
    const T v = ((sP[(size_t)n * L2P + lm] + sP[((size_t)NRP + n) * L2P + lm]) +
                 sP[((size_t)2 * NRP + n) * L2P + lm]) +
                sP[((size_t)3 * NRP + n) * L2P + lm];
    sC[n * L2P + lm] = v;
    cs[x] = v;
  }
  __syncthreads();
  // power spectrum p_d, d = (n1 <= n2, l): w sum_m c[n1][l,m] c[n2][l,m]
  const int dim = sd->dim;
  T pv[IPT];
  T ss = 0;
#pragma unroll
  for (int t = 0; t < IPT; t++) {
    const int d = tid + t * 256;
    pv[t] = 0;
    if (d < dim) {
      const int dm = dmap[d];
      const int n1 = dm & 255, n2 = (dm >> 8) & 255, l = dm >> 16;
      const T* ca = sC + n1 * L2P + l * l;
      const T* cb = sC + n2 * L2P + l * l;
      T a = 0;
      for (int m = 0; m <= 2 * l; m++) a += ca[m] * cb[m];
      pv[t] = (n1 == n2 ? (T)1 : (T)1.41421356237309504880) * a;
      ss += pv[t] * pv[t];
    }
  }
  ss = wave_sum(ss);
  if (lane == 0) sred[wave] = ss;
  __syncthreads();
  ss = ((sred[0] + sred[1]) + sred[2]) + sred[3];
  const T inv = (T)1 / sqrt(ss);
  T* qs = q + (size_t)site * q_ld;
  u16* qbs = qb ? qb + (size_t)site * qb_ld : nullptr;
#pragma unroll
  for (int t = 0; t < IPT; t++) {
    const int d = tid + t * 256;
    if (d < dim) {
      const T v = pv[t] * inv;
      qs[d] = v;
      if (qbs) qbs[d] = f2bf((float)v);
    }
  }
  for (int d = dim + tid; d < q_ld; d += 256) qs[d] = (T)0;
  if (qbs)
    for (int d = dim + tid; d < qb_ld; d += 256) qbs[d] = 0;
  if (tid == 0) pnorm[site] = sqrt(ss);
}

template <typename T, int NRP, int NMX, int LMX, bool EXL, typename GEO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BwdWaves<T>::v))) void soap_bwd_kernel(
    const SoapDev* __restrict__ sd, const double* __restrict__ W,
    const float2* __restrict__ rtab, const double* __restrict__ pos,
    const int* __restrict__ species, GEO geo, int maxnb, const int* __restrict__ count,
    const int* __restrict__ list, const int* __restrict__ site_atom, int n_sites,
    const int* __restrict__ dmap, const T* __restrict__ q, int q_ld,
    const T* __restrict__ cbuf, const T* __restrict__ pnorm, const T* __restrict__ G, int g_ld,
    T* __restrict__ gpair, double* __restrict__ vsite, const TwoBodyDev* __restrict__ twob,
    int n2b, double* __restrict__ eatom, int wpb) {
  constexpr int SPW = Spw<NRP>::v, LW = 64 / SPW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = EXL ? LMX : sd->l_max, L2 = EXL ? (LMX + 1) * (LMX + 1) : sd->L2;
  const int dim = sd->dim, Lp = L + 1;
  const int wave = threadIdx.x >> 6, lane64 = threadIdx.x & 63;
  const int sub = lane64 / LW, lane = lane64 % LW;
  constexpr int CAPV = Capv<NRP>::v;
  T* sGC = reinterpret_cast<T*>(smem_raw + (size_t)(wave * SPW + sub) * bwd_site_bytes<T, NRP>(L2, Lp));
  // sGC [L2][NRP] = dE/dc, species s at offa[s]; sM [Lp][NRP][NRP]: dE/dc = M_l c
  T* sM = sGC + NRP * L2;
  constexpr int MS = MStride<NRP>::v;
  // compacted valid pairs: list entries (so that the pair lanes need not
  // reload the list row) and slots (where their gradients go)
  int* sEnt = reinterpret_cast<int*>(sM + Lp * MS);
  u16* sIdx = reinterpret_cast<u16*>(sEnt + CAPV);
  const int site = (blockIdx.x * wpb + wave) * SPW + sub;
  if (site >= n_sites) return;
  const int i = site_atom ? site_atom[site] : site;  // null: site s is atom s
  const Cell cell = geo.of(i);
  const int cnt = min(count[i], maxnb);
  const int* row = list + (size_t)i * maxnb;
  const T* cs = cbuf + (size_t)site * NRP * L2;
  for (int k = lane; k < Lp * MS; k += LW) sM[k] = (T)0;
  // row slots that no channel fills are read with R_n = 0: they must be finite
  if (sd->lds_gap)
    for (int k = lane; k < L2 * NRP; k += LW) sGC[k] = (T)0;
  const T* qs = q + (size_t)site * q_ld;
  const T* gs = G + (size_t)site * g_ld;
  T gq = 0;
  for (int d = lane; d < dim; d += LW) gq += gs[d] * qs[d];
  gq = sub_sum<LW>(gq);
  const T inv = (T)1 / pnorm[site];
  wave_sync();
  // dE/dp -> M_l[n][n2] = w dE/dp[(n,n2),l] (symmetric; w = 2 on the diagonal,
  // sqrt(2) off it, the derivative of the multiplicity-weighted product)
  for (int d = lane; d < dim; d += LW) {
    const T gp = (gs[d] - gq * qs[d]) * inv;
    const int dm = dmap[d];
    const int n1 = dm & 255, n2 = (dm >> 8) & 255, l = dm >> 16;
    T* Ml = sM + l * MS;
    const T w = (n1 == n2 ? (T)2 : (T)1.41421356237309504880) * gp;
    Ml[tri<NRP>(n1, n2)] = w;
  }
  wave_sync();
  for (int lm = lane; lm < L2; lm += LW) {
    const int l = kSph.lm2l[lm];
    const T* Ml = sM + l * MS;
    T cl[NRP];
#pragma unroll
    for (int n2 = 0; n2 < NRP; n2++) cl[n2] = cs[n2 * L2 + lm];
#pragma unroll
    for (int n = 0; n < NRP; n++) {
      T a = 0;
#pragma unroll
      for (int n2 = 0; n2 < NRP; n2++) a += Ml[tri<NRP>(n, n2)] * cl[n2];
      const int dst = sd->chana[n];
      if (dst >= 0) sGC[lm * NRP + dst] = a;
    }
  }
  wave_sync();
  // the site's virial and 2-body energy, summed per round of pairs into LDS
  // (not carried in registers through the pair loop)
  double* sV = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(sGC) +
                                         bwd_site_bytes<T, NRP>(L2, Lp) - 64);
  if (lane < 7) sV[lane] = 0.0;
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type T4;
  T4* gp_out = reinterpret_cast<T4*>(gpair) + (size_t)i * maxnb;  // (g, 0) per slot
  int next = 0;
  while (next < cnt) {
    // valid pairs are packed onto lanes; skipped slots get a zero gradient
    const int nv = scan_pairs<LW, SPW, GEO>(sd, pos, species, cell, row, i, cnt, &next, sIdx,
                                            sEnt, CAPV, sub, lane, [&](int slot) {
                                              gp_out[slot] = T4{0, 0, 0, 0};
                                            });
    wave_sync();
    for (int c0 = 0; c0 < nv; c0 += LW) {
      double v[6] = {0, 0, 0, 0, 0, 0};
      double e2b = 0.0;
      if (c0 + lane < nv) {
      const int slot = sIdx[c0 + lane];
      const int e = sEnt[c0 + lane];
      double dv[3];
      const double rd = pair_vec_fast(pos, i, e, cell, dv);
      const int s = species[e & TGX_ATOM_MASK] - 1;
      T g0 = 0, g1 = 0, g2 = 0;
      {
      const T r = (T)rd;
      const T ir = (T)1 / r;
      const T ux = (T)dv[0] * ir, uy = (T)dv[1] * ir, uz = (T)dv[2] * ir;
      T R[NMX], dR[NMX];
      Radial<T>::template eval<NMX>(sd, W, rtab, s, r, R, dR);
      const int o = sd->offa[s];
      T log2t, D;
      ang_params<T>(sd, s, r, &log2t, &D);
      T Al[LMX + 1];
      ang_factors<T, LMX>(log2t, Al);
      T s1 = 0, Vx = 0, Vy = 0, Vz = 0;
      T Cm = 1, Sm = 0, Cp = 0, Sp = 0;  // C_m, S_m and C_{m-1}, S_{m-1}
      T kap[LMX + 1];  // (dA_l/dr) / A_l = l(l+1) D
#pragma unroll
      for (int l = 0; l <= LMX; l++) kap[l] = (T)(l * (l + 1)) * D;
      // per (l, +-m): the radial contractions h = sum_n dE/dc R_n and
      // hd = sum_n dE/dc dR_n of the cosine and sine rows, then the radial
      // (s1) and angular (V) parts of the gradient with the two rows combined
      // and A_l, dA_l/dr = A_l kap_l factored out of both
#pragma unroll
      for (int m = 0; m <= LMX; m++) {
        if (m <= L) {
          T pm2 = 0, pm1 = SphK<T>::P(m), dm2 = 0, dm1 = 0;
#pragma unroll
          for (int l = m; l <= LMX; l++) {
            if (l <= L) {
              T P = pm1, dP = dm1;
              if (l > m) {
                const int ix = l * (TGX_MAX_LMAX + 1) + m;
                const T a = SphK<T>::A(ix), b = SphK<T>::B(ix);
                P = a * uz * pm1 - b * pm2;
                dP = a * (pm1 + uz * dm1) - b * dm2;
                pm2 = pm1;
                pm1 = P;
                dm2 = dm1;
                dm1 = dP;
              }
              const T A = Al[l];
              T gcc[NMX];
              lds_row16<T, NMX>(sGC + (l * l + l + m) * NRP + o, gcc);
              T hc = 0, hdc = 0;
#pragma unroll
              for (int n = 0; n < NMX; n++) {
                hc += gcc[n] * R[n];
                hdc += gcc[n] * dR[n];
              }
              const T PA = P * A, dPA = dP * A;
              if (m == 0) {
                s1 += PA * (hdc + kap[l] * hc);
                Vz += dPA * hc;
              } else {
                T gcs[NMX];
                lds_row16<T, NMX>(sGC + (l * l + l - m) * NRP + o, gcs);
                T hs = 0, hds = 0;
#pragma unroll
                for (int n = 0; n < NMX; n++) {
                  hs += gcs[n] * R[n];
                  hds += gcs[n] * dR[n];
                }
                const T t1 = hc * Cm + hs * Sm, t2 = hdc * Cm + hds * Sm;
                const T mPA = PA * (T)m;
                s1 += PA * (t2 + kap[l] * t1);
                Vx += mPA * (hc * Cp + hs * Sp);
                Vy += mPA * (hs * Cp - hc * Sp);
                Vz += dPA * t1;
              }
            }
          }
          const T Cn = ux * Cm - uy * Sm;
          const T Sn = ux * Sm + uy * Cm;
          Cp = Cm;
          Sp = Sm;
          Cm = Cn;
          Sm = Sn;
        }
      }
      const T uV = ux * Vx + uy * Vy + uz * Vz;
      g0 = s1 * ux + (Vx - uV * ux) * ir;
      g1 = s1 * uy + (Vy - uV * uy) * ir;
      g2 = s1 * uz + (Vz - uV * uz) * ir;
      double gd[3] = {(double)g0, (double)g1, (double)g2};
      if (n2b) {
        // distance_2b blocks whose cutoff lies inside the SOAP cutoff: the
        // pair term phi(r) enters E_i, its gradient phi'(r) r/|r| the slot
        const int si1 = species[i], sj1 = s + 1;
        double gsc = 0.0;
        for (int b = 0; b < n2b; b++) {
          const TwoBodyDev& B = twob[b];
          if (!((si1 == B.sp1 && sj1 == B.sp2) || (si1 == B.sp2 && sj1 == B.sp1))) continue;
          double phi, dphi;
          two_body_pair<sizeof(T) == 8>(B, rd, &phi, &dphi);
          e2b += phi;
          gsc += dphi;
        }
        const double ird = 1.0 / rd;
        gd[0] += gsc * dv[0] * ird;
        gd[1] += gsc * dv[1] * ird;
        gd[2] += gsc * dv[2] * ird;
        g0 = (T)gd[0];
        g1 = (T)gd[1];
        g2 = (T)gd[2];
      }
      // virial += 1/2 (F x + x F), F = -g the force on the neighbour
      v[0] -= gd[0] * dv[0];
      v[1] -= gd[1] * dv[1];
      v[2] -= gd[2] * dv[2];
      v[3] -= 0.5 * (gd[0] * dv[1] + gd[1] * dv[0]);
      v[4] -= 0.5 * (gd[0] * dv[2] + gd[2] * dv[0]);
      v[5] -= 0.5 * (gd[1] * dv[2] + gd[2] * dv[1]);
    }
      gp_out[slot] = T4{g0, g1, g2, (T)0};
      }
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const double t = sub_sum<LW>(v[k]);
        if (lane == 0) sV[k] += t;
      }
      if (n2b) {
        const double t = sub_sum<LW>(e2b);
        if (lane == 0) sV[6] += t;
      }
    }
    wave_sync();  // the slot list is refilled by the next scan round
  }
  if (lane < 6) vsite[(size_t)site * 6 + lane] = sV[lane];
  if (n2b && lane == 0) eatom[i] += sV[6];
}

// Sites per workgroup: as many waves as fit in 64 KB of LDS (at least one).
static int waves_per_block(size_t per_wave_bytes) {
  int w = 4;
  while (w > 1 && (size_t)w * per_wave_bytes > 65536) w >>= 1;
  return w;
}
// The autotuner's request (1, 2 or 4 waves) when it fits, else the built-in
// choice: one wave per workgroup (measured fastest for the FP32 kernels and
// the FP64 forward at 32k sites, 8 % on the MD step) unless dflt_max asks
// for as many as fit (the FP64 adjoint).
static int pick_wpb(size_t per_wave_bytes, int req, bool dflt_max) {
  const int w = waves_per_block(per_wave_bytes);
  if ((req == 1 || req == 2 || req == 4) && req <= w) return req;
  return dflt_max ? w : 1;
}

template <typename K>
static hipError_t ensure_lds(K kernel, size_t sm, size_t* set) {
  if (sm <= *set) return hipSuccess;
  if (sm > 160 * 1024) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // do not leave the failure as the sticky error
    return e;
  }
  *set = sm;
  return hipSuccess;
}

template <typename T, int NRP, int NMX, int LMX, bool EXL, typename GEO>
static hipError_t fwd_inst(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const GEO& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, T* q, int q_ld,
                           u16* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st, int split,
                           int wpb_req) {
  constexpr int SPW = Spw<NRP>::v;
  if (maxnb > 65535) return hipErrorInvalidValue;  // 16-bit slot lists
  if constexpr (sizeof(T) == 4 && NRP <= 16) {
    if (split && n_sites <= TGX_FWD_SPLIT_MAX) {
      const size_t sm = fwd_split_bytes<NRP, LMX>(sdh.L2);
      static size_t sm_split = 65536;
      hipError_t e = ensure_lds(soap_fwd_split_kernel<T, NRP, NMX, LMX, EXL, GEO>, sm, &sm_split);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((soap_fwd_split_kernel<T, NRP, NMX, LMX, EXL, GEO>), dim3(n_sites),
                         dim3(256), sm, st, sd, W, reinterpret_cast<const float2*>(rtab), pos,
                         species, cell, maxnb, count, list, site_atom, n_sites, dmap, q, q_ld, qb,
                         qb_ld, cbuf, pnorm);
      return hipGetLastError();
    }
  }
  const size_t per_wave = SPW * fwd_site_bytes<T, NRP>(sdh.L2);
  const int wpb = pick_wpb(per_wave, wpb_req, false);
  const size_t sm = wpb * per_wave;
  static size_t sm_set = 65536;
  hipError_t e = ensure_lds(soap_fwd_kernel<T, NRP, NMX, LMX, EXL, GEO>, sm, &sm_set);
  if (e != hipSuccess) return e;
  const int per_block = wpb * SPW;
  hipLaunchKernelGGL((soap_fwd_kernel<T, NRP, NMX, LMX, EXL, GEO>), dim3((n_sites + per_block - 1) / per_block),
                     dim3(64 * wpb), sm, st, sd, W, reinterpret_cast<const float2*>(rtab), pos,
                     species, cell, maxnb, count, list, site_atom, n_sites, dmap, q, q_ld, qb,
                     qb_ld, cbuf, pnorm, wpb);
  return hipGetLastError();
}

template <typename T, int NRP, int NMX, int LMX, bool EXL, typename GEO>
static hipError_t bwd_inst(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const GEO& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, const T* q,
                           int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                           T* gpair, double* vsite, const TwoBodyDev* twob, int n2b,
                           double* eatom, hipStream_t st, int wpb_req) {
  constexpr int SPW = Spw<NRP>::v;
  if (maxnb > 65535) return hipErrorInvalidValue;  // 16-bit slot lists
  const size_t per_wave = SPW * bwd_site_bytes<T, NRP>(sdh.L2, sdh.l_max + 1);
  const int wpb = pick_wpb(per_wave, wpb_req, sizeof(T) == 8);
  const size_t sm = wpb * per_wave;
  static size_t sm_set = 65536;
  hipError_t e = ensure_lds(soap_bwd_kernel<T, NRP, NMX, LMX, EXL, GEO>, sm, &sm_set);
  if (e != hipSuccess) return e;
  const int per_block = wpb * SPW;
  hipLaunchKernelGGL((soap_bwd_kernel<T, NRP, NMX, LMX, EXL, GEO>), dim3((n_sites + per_block - 1) / per_block),
                     dim3(64 * wpb), sm, st, sd, W, reinterpret_cast<const float2*>(rtab), pos,
                     species, cell, maxnb, count, list, site_atom, n_sites, dmap, q, q_ld, cbuf,
                     pnorm, G, g_ld, gpair, vsite, twob, n2b, eatom, wpb);
  return hipGetLastError();
}

// Compile-time shape of a model: NRP = padded radial channels (SoapDev::nrp),
// NMX = per-species n_max bound (8 or 16), LMX = l_max bound (8 or 10), EXL =
// l_max equals LMX.  f is called with the four as integral constants.
template <int V> using IC = std::integral_constant<int, V>;
template <typename F>
static hipError_t soap_dispatch(const SoapDev& sdh, F&& f) {
  const int nrp = sdh.nrp, nmx = sdh.nmx, lmax = sdh.l_max;
  auto by_l = [&](auto NR, auto NM) -> hipError_t {
    if (lmax == 8) return f(NR, NM, IC<8>{}, std::true_type{});
    if (lmax == 10) return f(NR, NM, IC<10>{}, std::true_type{});
    if (lmax < 8) return f(NR, NM, IC<8>{}, std::false_type{});
    return f(NR, NM, IC<10>{}, std::false_type{});
  };
  if (nrp == 8) return by_l(IC<8>{}, IC<8>{});
  if (nrp == 16 && nmx == 8) return by_l(IC<16>{}, IC<8>{});
  if (nrp == 16) return by_l(IC<16>{}, IC<16>{});
  if (nrp == 24 && nmx == 8) return by_l(IC<24>{}, IC<8>{});
  if (nrp == 24) return by_l(IC<24>{}, IC<16>{});
  if (nrp == 32 && nmx == 8) return by_l(IC<32>{}, IC<8>{});
  if (nrp == 32) return by_l(IC<32>{}, IC<16>{});
  return hipErrorInvalidValue;
}

template <typename T, typename GEO>
hipError_t launch_soap_fwd_g(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const GEO& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, T* q, int q_ld,
                           u16* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st,
                           int split, int wpb) {
  if (n_sites == 0) return hipSuccess;
  return soap_dispatch(sdh, [&](auto NR, auto NM, auto LM, auto EX) {
    return fwd_inst<T, decltype(NR)::value, decltype(NM)::value, decltype(LM)::value,
                    decltype(EX)::value, GEO>(sdh, sd, W, rtab, pos, species, cell, maxnb, count,
                                              list, site_atom, n_sites, dmap, q, q_ld, qb, qb_ld,
                                              cbuf, pnorm, st, split, wpb);
  });
}

template <typename T, typename GEO>
hipError_t launch_soap_bwd_g(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const GEO& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, const T* q,
                           int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                           T* gpair, double* vsite, const TwoBodyDev* twob, int n2b,
                           double* eatom, hipStream_t st, int wpb) {
  if (n_sites == 0) return hipSuccess;
  return soap_dispatch(sdh, [&](auto NR, auto NM, auto LM, auto EX) {
    return bwd_inst<T, decltype(NR)::value, decltype(NM)::value, decltype(LM)::value,
                    decltype(EX)::value, GEO>(sdh, sd, W, rtab, pos, species, cell, maxnb, count,
                                              list, site_atom, n_sites, dmap, q, q_ld, cbuf, pnorm,
                                              G, g_ld, gpair, vsite, twob, n2b, eatom, st,
                                              wpb);
  });
}

}  // namespace tgx
