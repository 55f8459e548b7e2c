// Device-side math shared by the tgx kernels: minimum-image distance with the
// reference's FP64 operation order, poly3 radial closed form, radial table
// lookup, and real spherical-harmonic recurrences.
#pragma once

#include <hip/hip_runtime.h>

#include "tgx_internal.h"

namespace tgx {

// Normalised associated-Legendre recurrence constants (real harmonics without
// Condon-Shortley phase): P(l,m) = A[l][m] z P(l-1,m) - B[l][m] P(l-2,m),
// P(m,m) = Pmm[m]; the sqrt(2) of m > 0 is folded into Pmm.  With
// K(l,m) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!):
//   A = sqrt((2l+1)(2l-1)/((l+m)(l-m))),
//   B = sqrt((2l+1)(l+m-1)(l-m-1)/((2l-3)(l+m)(l-m))),
//   Pmm = K(m,m) (2m-1)!! (times sqrt 2 for m > 0).
// The table is a compile-time constant: in the fully unrolled recurrences the
// entries fold into instruction literals instead of occupying scalar registers.
constexpr double cx_sqrt(double a) {
  if (a <= 0.0) return 0.0;
  double x = a > 1.0 ? a : 1.0, prev = 0.0;
  for (int it = 0; it < 200 && x != prev; it++) {
    prev = x;
    x = 0.5 * (x + a / x);
  }
  return x;
}
struct SphTab {
  float A[TGX_L2MAX], B[TGX_L2MAX], Pmm[TGX_MAX_LMAX + 1];
  double Ad[TGX_L2MAX], Bd[TGX_L2MAX], Pmmd[TGX_MAX_LMAX + 1];
  int lm2l[TGX_L2MAX];  // l of the combined index lm = l*l + l + m
};
constexpr SphTab make_sph_tab() {
  SphTab t{};
  const double pi = 3.14159265358979323846;
  double df = 1.0;  // (2m-1)!!
  for (int m = 0; m <= TGX_MAX_LMAX; m++) {
    if (m > 0) df *= (2.0 * m - 1.0);
    double f = 1.0;  // (l-m)!/(l+m)! at l = m
    for (int k = 1; k <= 2 * m; k++) f /= (double)k;
    double p = cx_sqrt((2.0 * m + 1.0) / (4.0 * pi) * f) * df;
    if (m > 0) p *= cx_sqrt(2.0);
    t.Pmmd[m] = p;
    t.Pmm[m] = (float)p;
    for (int l = m + 1; l <= TGX_MAX_LMAX; l++) {
      const int ix = l * (TGX_MAX_LMAX + 1) + m;
      const double a = cx_sqrt((2.0 * l + 1.0) * (2.0 * l - 1.0) / ((double)(l + m) * (l - m)));
      const double b = l - 2 >= m ? cx_sqrt((2.0 * l + 1.0) * (l + m - 1.0) * (l - m - 1.0) /
                                            ((2.0 * l - 3.0) * (l + m) * (l - m)))
                                  : 0.0;
      t.Ad[ix] = a;
      t.Bd[ix] = b;
      t.A[ix] = (float)a;
      t.B[ix] = (float)b;
    }
  }
  for (int l = 0; l <= TGX_MAX_LMAX; l++)
    for (int k = l * l; k < (l + 1) * (l + 1); k++) t.lm2l[k] = l;
  return t;
}
constexpr SphTab kSph = make_sph_tab();

template <typename T> struct SphK;
template <> struct SphK<float> {
  __device__ static constexpr float A(int i) { return kSph.A[i]; }
  __device__ static constexpr float B(int i) { return kSph.B[i]; }
  __device__ static constexpr float P(int m) { return kSph.Pmm[m]; }
};
template <> struct SphK<double> {
  __device__ static constexpr double A(int i) { return kSph.Ad[i]; }
  __device__ static constexpr double B(int i) { return kSph.Bd[i]; }
  __device__ static constexpr double P(int m) { return kSph.Pmmd[m]; }
};

// Pair geometry as seen by a kernel: of(i) is the periodic cell of atom i.
struct CellOne {
  Cell c;
  __device__ __forceinline__ const Cell& of(int) const { return c; }
};
struct CellMany {
  const Cell* cells;
  const int* sid;
  __device__ __forceinline__ Cell of(int i) const { return cells[sid[i]]; }
};

// Workgroups of the deterministic energy/virial reduction (each writes one
// row of 8 partials).
#define TGX_RB 128

// Partial row b of the deterministic energy/virial reduction: a fixed
// partition of the sites (6-component virials vsite) and atoms (virials
// vatom, energies eatom; each input may be null) over TGX_RB workgroups of
// 256 threads, each summed by a fixed tree into part[b*8 + 0..6].  sh holds
// 256 x 7 doubles.  Every thread of the workgroup must call it.
__device__ inline void reduce_part_rows(int b, int n_sites, const double* vsite, int n,
                                        const double* vatom, const double* eatom,
                                        double* part, double (*sh)[7]) {
  const int t = threadIdx.x;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  if (vsite) {
    const int per = (n_sites + TGX_RB - 1) / TGX_RB, lo = b * per, hi = min(n_sites, lo + per);
    for (int s = lo + t; s < hi; s += 256)
      for (int k = 0; k < 6; k++) a[k] += vsite[(size_t)s * 6 + k];
  }
  {
    const int per = (n + TGX_RB - 1) / TGX_RB, lo = b * per, hi = min(n, lo + per);
    for (int s = lo + t; s < hi; s += 256) {
      if (vatom)
        for (int k = 0; k < 6; k++) a[k] += vatom[(size_t)s * 6 + k];
      if (eatom) a[6] += eatom[s];
    }
  }
  for (int k = 0; k < 7; k++) sh[t][k] = a[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
      for (int k = 0; k < 7; k++) sh[t][k] += sh[t + off][k];
    __syncthreads();
  }
  if (t < 7) part[b * 8 + t] = sh[0][t];
}

// Fixed-order sum of the TGX_RB partial rows into out[0..8] (virial) and
// out[9] (energy) by the first TGX_RB threads of a workgroup; sh holds
// TGX_RB x 7 doubles.  Every thread of the workgroup must call it.
__device__ inline void reduce_final_rows(const double* part, double* out, double (*sh)[7]) {
  const int t = threadIdx.x;
  if (t < TGX_RB)
    for (int k = 0; k < 7; k++) sh[t][k] = part[t * 8 + k];
  __syncthreads();
  for (int off = TGX_RB / 2; off > 0; off >>= 1) {
    if (t < off)
      for (int k = 0; k < 7; k++) sh[t][k] += sh[t + off][k];
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sh[0][0];
    out[4] = sh[0][1];
    out[8] = sh[0][2];
    out[1] = out[3] = sh[0][3];
    out[2] = out[6] = sh[0][4];
    out[5] = out[7] = sh[0][5];
    out[9] = sh[0][6];
  }
}

// LDS written by other lanes of this wave is visible after this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// fmod(a, p) for p > 0, bit-identical to the C library's for |a/p| < 2^40:
// a itself when |a| < p; otherwise the remainder a - n p of the true quotient n = trunc(a/p) is exactly
// representable, so one fused multiply-add gives it exactly; a quotient
// estimate off by one (a/p within rounding of an integer) puts the result
// outside [0, p) (or (-p, 0] for negative a) and is corrected once.
__host__ __device__ __forceinline__ double rem_exact(double a, double p) {
  if (fabs(a) < p) return a;  // the common case of the minimum image: no division
  const double n = trunc(a / p);
  double r = fma(-n, p, a);
  if (a >= 0.0) {
    if (r < 0.0) r = fma(-(n - 1.0), p, a);
    else if (r >= p) r = fma(-(n + 1.0), p, a);
  } else {
    if (r > 0.0) r = fma(-(n + 1.0), p, a);
    else if (r <= -p) r = fma(-(n - 1.0), p, a);
  }
  return r == 0.0 ? copysign(0.0, a) : r;  // fmod's zero carries the sign of a
}

// Fortran MODULO(a,p) for reals and p > 0 in gfortran's formulation (fmod
// plus sign fix), then the L/2 fold and the x,y,z sum of squares in that
// order (get_distance, reference neighbors_interface.f90:565-581).
// Contraction is off so the FP64 result is bit-identical to the reference
// arithmetic.
__host__ __device__ __forceinline__ double f_modulo(double a, double p) {
#pragma clang fp contract(off)
  double r = rem_exact(a, p);
  if (r != 0.0) {
    if (a < 0.0) r = r + p;
  } else {
    r = copysign(0.0, p);
  }
  return r;
}

__device__ __forceinline__ double mic(const double* pi, const double* pj,
                                      double La, double Lb, double Lc,
                                      double* dv) {
#pragma clang fp contract(off)
  double L[3] = {La, Lb, Lc};
  double d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double t = pj[k] - pi[k];
    if (L[k] > 0.0) {  // an edge of 0 (open cell): the vector as it is
      t = f_modulo(t, L[k]);
      if (t > L[k] / 2.0) t = t - L[k];
    }
    dv[k] = t;
    d2 = d2 + t * t;
  }
  return sqrt(d2);
}

// Position of supercell image `img` of atom j, summed as set_supercell does
// (xyz/read_xyz.f90:336-339): ((p + i2 a) + j2 b) + k2 c.
__device__ __forceinline__ void image_pos(const double* pos, int j, int img,
                                          const Cell& c, double* out) {
#pragma clang fp contract(off)
  const double* p = pos + 3 * j;
  if (img == 0) {
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
    return;
  }
  int k2 = img % c.ic;
  int j2 = (img / c.ic) % c.ib;
  int i2 = img / (c.ic * c.ib);
  // orthorhombic: a = (a,0,0), b = (0,b,0), c = (0,0,c)
  out[0] = ((p[0] + (double)i2 * c.a) + (double)j2 * 0.0) + (double)k2 * 0.0;
  out[1] = ((p[1] + (double)i2 * 0.0) + (double)j2 * c.b) + (double)k2 * 0.0;
  out[2] = ((p[2] + (double)i2 * 0.0) + (double)j2 * 0.0) + (double)k2 * c.c;
}

// Pair vector r_j(img) - r_i with the reference's minimum image.
__device__ __forceinline__ double pair_vec(const double* pos, int i, int packed,
                                           const Cell& c, double* dv) {
  int j = packed & TGX_ATOM_MASK;
  int img = packed >> TGX_IMG_SHIFT;
  double pj[3];
  image_pos(pos, j, img, c, pj);
  return mic(pos + 3 * i, pj, c.La, c.Lb, c.Lc, dv);
}

// ---------------------------------------------------------------------------
// Radial basis
// ---------------------------------------------------------------------------
#define TGX_JMAX (TGX_MAX_NR + 8)

// Unnormalised poly3 integrals combined with W for species s:
// B_n(r) = sum_a W[n][a] N_a int_0^rh rho^p (rh-rho)^(a+2) G(rho; r, sig(r))
// (G the unit-normalised 1-D Gaussian) and dB_n/dr.  FP64 throughout.
template <int NMX>
__device__ inline void radial_basis_fp64(const SoapDev* sd, const double* W,
                                         int s, double r, double* Bv,
                                         double* dB) {
  const int nmax = sd->nmax[s];
  const int p = sd->radial_enhancement;
  const double rh = sd->rh[s];
  const double sig0 = sd->sig_r[s], sigs = sd->sig_rs[s];
  const double sg = sig0 + sigs * r;
  const int kmax = nmax + 6;
  double J[NMX + 8];
  const double c = rh - r, s2 = sg * sg;
  const double er = exp(-r * r / (2.0 * s2)), ec = exp(-c * c / (2.0 * s2));
  J[0] = sg * 1.2533141373155002512 *
         (erf(r / (sg * 1.4142135623730950488)) +
          erf(c / (sg * 1.4142135623730950488)));
  J[1] = c * J[0] + s2 * (ec - er);
  double rhp = 1.0;
#pragma unroll
  for (int k = 2; k < NMX + 8; k++) {
    rhp *= rh;
    J[k] = k <= kmax ? c * J[k - 1] + (k - 1) * s2 * J[k - 2] - s2 * rhp * er : 0.0;
  }
  const double pre = 1.0 / (2.5066282746310005024 * sg);
  double I[NMX], dI[NMX];
  // poly3gauss: slot 0 is the Gaussian, slot a the poly3 function a
  const int g = sd->basis == 1 ? 1 : 0;
#pragma unroll
  for (int a = 1; a <= NMX; a++) {
    if (a > nmax) {
      I[a - 1] = dI[a - 1] = 0.0;
      continue;
    }
    const int d = a + 2 - g;  // degree: (rh - r)^(d) of poly3 function a - g
    const int ak = a - g;
    double Na = sqrt((2.0 * ak + 5.0) / pow(rh, 2.0 * ak + 5.0));
    double T, Tc, Ts;
    double j0 = J[d], j1 = J[d + 1], j2 = J[d + 2], j3 = J[d + 3], j4 = J[d + 4];
    // dJ_k/dc = (J_{k+1} - c J_k)/s2 ; dJ_k/dsig = (J_{k+2} - 2c J_{k+1} + c^2 J_k)/s2/sg
    double c0 = (j1 - c * j0) / s2, c1 = (j2 - c * j1) / s2, c2 = (j3 - c * j2) / s2;
    double s0 = (j2 - 2.0 * c * j1 + c * c * j0) / (s2 * sg);
    double s1 = (j3 - 2.0 * c * j2 + c * c * j1) / (s2 * sg);
    double s2_ = (j4 - 2.0 * c * j3 + c * c * j2) / (s2 * sg);
    if (p == 0) {
      T = j0; Tc = c0; Ts = s0;
    } else if (p == 1) {
      T = rh * j0 - j1; Tc = rh * c0 - c1; Ts = rh * s0 - s1;
    } else {
      T = rh * rh * j0 - 2.0 * rh * j1 + j2;
      Tc = rh * rh * c0 - 2.0 * rh * c1 + c2;
      Ts = rh * rh * s0 - 2.0 * rh * s1 + s2_;
    }
    I[a - 1] = Na * pre * T;
    dI[a - 1] = Na * (pre * (-Tc + Ts * sigs) - pre / sg * sigs * T);
  }
  if (g) {
    // slot 0: int_0^rh exp(-u^2/(2 sb^2)) u^p G(u; r, sg) du, sb = sig_r.
    // The product of the two Gaussians is exp(-r^2/(2 Sg)), Sg = sb^2 +
    // sg^2, times a Gaussian of centre mu = r sb^2/Sg and width t = sb sg /
    // sqrt(Sg); M_k = int_0^rh u^k exp(-(u-mu)^2/(2 t^2)) du by parts:
    // M_k = mu M_(k-1) + (k-1) t^2 M_(k-2) - t^2 [u^(k-1) e]_0^rh.  The r
    // derivative goes through mu, t (sg = sig_r + sig_rs r) and the factor.
    const double sb = sd->sig_r[s], sb2 = sb * sb, Sg = sb2 + s2;
    const double A = exp(-r * r / (2.0 * Sg));
    const double mu = r * sb2 / Sg, t2 = sb2 * s2 / Sg, t = sqrt(t2);
    const double e0 = exp(-mu * mu / (2.0 * t2));
    const double eh = exp(-(rh - mu) * (rh - mu) / (2.0 * t2));
    double M[5];
    M[0] = t * 1.2533141373155002512 *
           (erf((rh - mu) / (t * 1.4142135623730950488)) + erf(mu / (t * 1.4142135623730950488)));
    M[1] = mu * M[0] - t2 * (eh - e0);
    double rp = 1.0;
#pragma unroll
    for (int k = 2; k < 5; k++) {
      rp *= rh;
      M[k] = mu * M[k - 1] + (k - 1) * t2 * M[k - 2] - t2 * rp * eh;
    }
    const double Mp = M[p], Mp1 = M[p + 1], Mp2 = M[p + 2];
    const double dMmu = (Mp1 - mu * Mp) / t2;
    const double dMt = (Mp2 - 2.0 * mu * Mp1 + mu * mu * Mp) / (t2 * t);
    const double dmu = sb2 / Sg - r * sb2 * 2.0 * sg * sigs / (Sg * Sg);
    const double dt = sigs * sb * sb2 / (Sg * sqrt(Sg));
    const double dA = A * (-r / Sg + r * r * sg * sigs / (Sg * Sg));
    const double ng = sd->gnorm[s];
    I[0] = ng * pre * A * Mp;
    dI[0] = ng * (-pre * sigs / sg * A * Mp + pre * dA * Mp + pre * A * (dMmu * dmu + dMt * dt));
  }
  const double* Ws = W + s * TGX_MAX_NR * TGX_MAX_NR;
#pragma unroll
  for (int n = 0; n < NMX; n++) {
    double v = 0.0, dv = 0.0;
#pragma unroll
    for (int a = 0; a < NMX; a++) {
      if (a < nmax) {
        v += Ws[n * TGX_MAX_NR + a] * I[a];
        dv += Ws[n * TGX_MAX_NR + a] * dI[a];
      }
    }
    Bv[n] = n < nmax ? v : 0.0;
    dB[n] = n < nmax ? dv : 0.0;
  }
}

// Amplitude (polynomial scaling of the reference's scaling_mode
// "polynomial", times global_scaling) and the soft cutoff between rcut_soft
// and rcut_hard.  Evaluated in T.
template <typename T>
__device__ __forceinline__ void radial_envelope(const SoapDev* sd, int s, T r,
                                                T* e, T* de) {
  const T rh = (T)sd->rh[s], rs = (T)sd->rs[s];
  const T x = r / rh;
  const T poly = (T)1 + (T)2 * x * x * x - (T)3 * x * x;
  const T dpoly = ((T)6 * x * x - (T)6 * x) / rh;
  const T as = (T)sd->amp_scal[s];
  T amp, damp;
  if (as == (T)0) {
    amp = 1;
    damp = 0;
  } else if (as == (T)1) {
    amp = poly;
    damp = dpoly;
  } else {
    amp = pow(poly, as);
    damp = poly > (T)0 ? as * pow(poly, as - (T)1) * dpoly : (T)0;
  }
  T fc = 1, dfc = 0;
  if (r >= rh) {
    fc = 0;
  } else if (rs < rh && r > rs) {
    const T y = (r - rs) / (rh - rs);
    fc = (T)1 - (T)3 * y * y + (T)2 * y * y * y;
    dfc = ((T)-6 * y + (T)6 * y * y) / (rh - rs);
  }
  const T g = (T)sd->glob[s];
  *e = g * amp * fc;
  *de = g * (damp * fc + amp * dfc);
}

// FP32 lookup of B_n(r), dB_n/dr from the cubic-Hermite table
// tab[s][k][n] = (B_n(r_k), h dB_n/dr(r_k)), r_k = k h, h = rh/TGX_RTAB_N;
// entries with n >= n_max(s) are zero.
template <int NMX>
__device__ __forceinline__ void radial_basis_tab(const SoapDev* sd,
                                                 const float2* tab, int s,
                                                 float r, float* Bv,
                                                 float* dB, int nmax) {
  const float rh = (float)sd->rh[s];
  const float hinv = (float)TGX_RTAB_N / rh;
  float x = r * hinv;
  int k = (int)x;
  k = k < 0 ? 0 : (k > TGX_RTAB_N - 1 ? TGX_RTAB_N - 1 : k);
  const float t = x - (float)k;
  const float t2 = t * t, t3 = t2 * t;
  const float h00 = 2.f * t3 - 3.f * t2 + 1.f, h10 = t3 - 2.f * t2 + t;
  const float h01 = -2.f * t3 + 3.f * t2, h11 = t3 - t2;
  const float g00 = 6.f * t2 - 6.f * t, g10 = 3.f * t2 - 4.f * t + 1.f;
  const float g01 = -6.f * t2 + 6.f * t, g11 = 3.f * t2 - 2.f * t;
  const float2* r0 = tab + ((size_t)s * (TGX_RTAB_N + 1) + k) * TGX_MAX_NR;
  const float2* r1 = r0 + TGX_MAX_NR;
#pragma unroll
  for (int n = 0; n < NMX; n++) {
    // rows are zero beyond the species' n_max, so no guard is needed
    const float2 a = r0[n], b = r1[n];
    Bv[n] = h00 * a.x + h10 * a.y + h01 * b.x + h11 * b.y;
    dB[n] = (g00 * a.x + g10 * a.y + g01 * b.x + g11 * b.y) * hinv;
  }
}

// Angular smearing A_l = exp(-l(l+1) sig_t^2 / (2 r^2)), sig_t = sig_t0 +
// sig_ts r.  Returns log2 of t = exp(-sig_t^2/(2 r^2)) and the factor D with
// dA_l/dr = A_l l(l+1) D.
template <typename T>
__device__ __forceinline__ void ang_params(const SoapDev* sd, int s, T r,
                                           T* log2t, T* D) {
  const T st = (T)sd->sig_t[s] + (T)sd->sig_ts[s] * r;
  const T sts = (T)sd->sig_ts[s];
  const T ir = (T)1 / r;
  *log2t = -(T)0.72134752044448170368 * st * st * ir * ir;  // -st^2/(2r^2)/ln2
  *D = st * st * ir * ir * ir - st * sts * ir * ir;
}

__device__ __forceinline__ float exp2_t(float x) { return exp2f(x); }
__device__ __forceinline__ double exp2_t(double x) { return exp2(x); }

// distance_2b pair term (gap.f90:383-435): phi(r) = delta^2 fcut(r)
// sum_s alpha_s c_s exp(-(r-Q_s)^2 / (2 sigma^2)) for one side of the pair and
// dphi/dr.  EXACT: the reference sum over sparse points in FP64; otherwise the
// smooth part from the FP32 cubic-Hermite table with the cosine cutoff applied
// analytically.  Zero beyond rcut.
template <int EXACT>
__device__ __forceinline__ void two_body_pair(const TwoBodyDev& b, double r, double* phi,
                                              double* dphi) {
  const double pi = 3.14159265358979323846;
  if (!(r < b.rcut)) {
    *phi = *dphi = 0.0;
    return;
  }
  if (EXACT) {
    double fcut, dfcut;
    if (r < b.rcut - b.buffer) {
      fcut = 1.0;
      dfcut = 0.0;
    } else {
      const double arg = pi * (r - b.rcut + b.buffer) / b.buffer;
      fcut = (cos(arg) + 1.0) / 2.0;
      dfcut = pi / 2.0 / b.buffer * sin(arg);
    }
    const double d2 = b.delta * b.delta, is2 = 1.0 / (b.sigma * b.sigma);
    double es = 0, fs = 0;
    for (int k = 0; k < b.n_sparse; k++) {
      const double dr = r - b.Qs[k];
      const double ex = exp(-0.5 * dr * dr * is2);
      es += b.alphas[k] * ex;
      fs += b.alphas[k] * ex * (dr * is2 * fcut + dfcut);
    }
    *phi = d2 * fcut * es;
    *dphi = -d2 * fs;
  } else {
    const float rf = (float)r;
    const float hinv = (float)b.nt / (float)b.rcut;
    const float xg = rf * hinv;
    int k = (int)xg;
    k = k > b.nt - 1 ? b.nt - 1 : k;
    const float t = xg - (float)k, t2 = t * t, t3 = t2 * t;
    const float2 p0 = b.tab[k], p1 = b.tab[k + 1];
    const float g = (2.f * t3 - 3.f * t2 + 1.f) * p0.x + (t3 - 2.f * t2 + t) * p0.y +
                    (-2.f * t3 + 3.f * t2) * p1.x + (t3 - t2) * p1.y;
    const float dg = ((6.f * t2 - 6.f * t) * p0.x + (3.f * t2 - 4.f * t + 1.f) * p0.y +
                      (-6.f * t2 + 6.f * t) * p1.x + (3.f * t2 - 2.f * t) * p1.y) * hinv;
    float fc = 1.f, dfc = 0.f;
    const float r1 = (float)(b.rcut - b.buffer);
    if (rf >= r1) {
      const float arg = (float)pi * (rf - r1) / (float)b.buffer;
      fc = 0.5f * (__cosf(arg) + 1.f);
      dfc = -0.5f * (float)pi / (float)b.buffer * __sinf(arg);
    }
    *phi = (double)(fc * g);
    *dphi = (double)(dfc * g + fc * dg);
  }
}

}  // namespace tgx
