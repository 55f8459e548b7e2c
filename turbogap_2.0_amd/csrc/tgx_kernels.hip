// Kernels of the GAP/SOAP path except the kernel-step GEMMs: neighbour list,
// radial table, SOAP expansion forward and adjoint, force gather, 2-body
// term, energy assembly and deterministic reductions.
#include <hip/hip_runtime.h>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

__constant__ SphConst c_sph;

hipError_t upload_sph_constants(int /*lmax*/, hipStream_t st) {
  static SphConst h;
  static bool ready = false;
  if (!ready) {
    sph_constants(TGX_MAX_LMAX, h.A, h.B, h.Pmm, h.Ad, h.Bd, h.Pmmd);
    ready = true;
  }
  return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sph), &h, sizeof(h), 0,
                                hipMemcpyHostToDevice, st);
}

// ---------------------------------------------------------------------------
// Radial table: tab[s][k][n] = (B_n(r_k), h dB_n/dr(r_k)) in FP32, evaluated
// in FP64 from the closed form with the FP64 orthonormalisation matrix.
// ---------------------------------------------------------------------------
__global__ void build_rtab_kernel(const SoapDev* sd, const double* W,
                                  float2* tab) {
  int s = blockIdx.y;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sd->n_species || k > TGX_RTAB_N) return;
  double h = sd->rh[s] / TGX_RTAB_N;
  double r = k * h;
  double Bv[TGX_MAX_NR], dB[TGX_MAX_NR];
  radial_basis_fp64<TGX_MAX_NR>(sd, W, s, r, Bv, dB);
  float2* row = tab + ((size_t)s * (TGX_RTAB_N + 1) + k) * TGX_MAX_NR;
  for (int n = 0; n < sd->nmax[s]; n++) row[n] = make_float2((float)Bv[n], (float)(h * dB[n]));
}

hipError_t launch_build_rtab(const SoapDev* sd_dev, int n_species,
                             const double* W, float* tab, hipStream_t st) {
  dim3 grid((TGX_RTAB_N + 1 + 127) / 128, n_species);
  hipLaunchKernelGGL(build_rtab_kernel, grid, dim3(128), 0, st, sd_dev, W,
                     reinterpret_cast<float2*>(tab));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Neighbour list
// ---------------------------------------------------------------------------

// Linked-cell index of atom i with the reference's arithmetic
// (neighbors_interface.f90:306-321): fold into the box around its centre,
// shift back, then int(x/(L+tol)*m).
__device__ __forceinline__ void cell_coords(const double* p, const Cell& c,
                                            int* cx, int* cy, int* cz) {
#pragma clang fp contract(off)
  const double tol = 1e-10;
  double ctr[3] = {c.La / 2.0, c.Lb / 2.0, c.Lc / 2.0};
  double dv[3];
  mic(ctr, p, c.La, c.Lb, c.Lc, dv);
  for (int k = 0; k < 3; k++) dv[k] = dv[k] + ctr[k];
  *cx = (int)(dv[0] / (c.La + tol) * c.mx);
  *cy = (int)(dv[1] / (c.Lb + tol) * c.my);
  *cz = (int)(dv[2] / (c.Lc + tol) * c.mz);
}

__device__ __forceinline__ int imodd(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

#define GUARD_RETURN() \
  if (guard && *guard == 0) return

__global__ void zero_ints_kernel(int* p, int m, const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] = 0;
}

__global__ void bin_atoms_kernel(const double* pos, int n, Cell c, int* cell_of,
                                 int* cell_count, const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx, cy, cz;
  cell_coords(pos + 3 * i, c, &cx, &cy, &cz);
  int cell = imodd(cx, c.mx) + imodd(cy, c.my) * c.mx + imodd(cz, c.mz) * c.mx * c.my;
  cell_of[i] = cell;
  atomicAdd(cell_count + cell, 1);
}

// Exclusive scan of ncells counts by one workgroup (ncells is small).
__global__ void scan_cells_kernel(const int* count, int ncells, int* start,
                                  int* fill, const int* guard) {
  GUARD_RETURN();
  __shared__ int part[1024];
  int t = threadIdx.x;
  int per = (ncells + 1023) / 1024;
  int b = t * per, e = min(ncells, b + per);
  int s = 0;
  for (int k = b; k < e; k++) s += count[k];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int k = b; k < e; k++) {
    start[k] = run;
    fill[k] = 0;
    run += count[k];
  }
  if (t == 1023) start[ncells] = part[1023];
}

__global__ void scatter_cells_kernel(int n, const int* cell_of, const int* start,
                                     int* fill, int* cell_atoms, const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of[i];
  int slot = atomicAdd(fill + c, 1);
  cell_atoms[start[c] + slot] = i;
}

// Candidates of atom i within rcut, unsorted, packed (img<<26 | j).
__global__ void find_neighbors_kernel(const double* pos, int n, Cell c,
                                      double rcut, const int* cell_start,
                                      const int* cell_atoms, int* count,
                                      int maxnb, int* list, int* flag,
                                      const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cnt = 0;
  int* out = list + (size_t)i * maxnb;
  const double* pi = pos + 3 * i;
  if (!c.small) {
    int cx, cy, cz;
    cell_coords(pi, c, &cx, &cy, &cz);
    int seen[27];
    int ns = 0;
    for (int k3 = cz - 1; k3 <= cz + 1; k3++)
      for (int j3 = cy - 1; j3 <= cy + 1; j3++)
        for (int i3 = cx - 1; i3 <= cx + 1; i3++) {
          int cell = imodd(i3, c.mx) + imodd(j3, c.my) * c.mx + imodd(k3, c.mz) * c.mx * c.my;
          bool dup = false;
          for (int q = 0; q < ns; q++) dup |= (seen[q] == cell);
          if (dup) continue;
          seen[ns++] = cell;
          for (int a = cell_start[cell]; a < cell_start[cell + 1]; a++) {
            int k = cell_atoms[a];
            if (k == i) continue;
            double dv[3];
            double d = mic(pi, pos + 3 * k, c.La, c.Lb, c.Lc, dv);
            if (d < rcut) {
              if (cnt < maxnb) out[cnt] = k;
              cnt++;
            }
          }
        }
  } else {
    for (int img = 0; img < c.nrep; img++)
      for (int k = 0; k < n; k++) {
        if (img == 0 && k == i) continue;
        double pk[3], dv[3];
        image_pos(pos, k, img, c, pk);
        double d = mic(pi, pk, c.La, c.Lb, c.Lc, dv);
        if (d < rcut) {
          if (cnt < maxnb) out[cnt] = (img << TGX_IMG_SHIFT) | k;
          cnt++;
        }
      }
  }
  count[i] = cnt;
  if (cnt > maxnb) atomicMax(flag, cnt);
}

// Sort each atom's list ascending (== ascending reference supercell id).
__global__ void sort_lists_kernel(int n, int maxnb, const int* count, int* list,
                                  const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cnt = min(count[i], maxnb);
  int* a = list + (size_t)i * maxnb;
  for (int p = 1; p < cnt; p++) {
    int v = a[p];
    int q = p - 1;
    while (q >= 0 && a[q] > v) {
      a[q + 1] = a[q];
      q--;
    }
    a[q + 1] = v;
  }
}

// rev[i][slot] = flat index of the entry (neg img, i) in j's list.
__global__ void reverse_kernel(int n, int maxnb, Cell c, const int* count,
                               const int* list, int* rev, int* flag,
                               const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x;
  int cnt = min(count[i], maxnb);
  for (int s = threadIdx.x; s < cnt; s += blockDim.x) {
    int e = list[(size_t)i * maxnb + s];
    int j = e & TGX_ATOM_MASK, img = e >> TGX_IMG_SHIFT;
    int nimg = 0;
    if (img) {
      int k2 = img % c.ic, j2 = (img / c.ic) % c.ib, i2 = img / (c.ic * c.ib);
      int ni2 = (c.ia - i2) % c.ia, nj2 = (c.ib - j2) % c.ib, nk2 = (c.ic - k2) % c.ic;
      nimg = (ni2 * c.ib + nj2) * c.ic + nk2;
    }
    int key = (nimg << TGX_IMG_SHIFT) | i;
    const int* a = list + (size_t)j * maxnb;
    int lo = 0, hi = min(count[j], maxnb) - 1, found = -1;
    while (lo <= hi) {
      int mid = (lo + hi) >> 1;
      int v = a[mid];
      if (v == key) {
        found = mid;
        break;
      }
      if (v < key) lo = mid + 1;
      else hi = mid - 1;
    }
    if (found < 0) atomicOr(flag + 1, 1);
    rev[(size_t)i * maxnb + s] = found < 0 ? (int)((size_t)i * maxnb + s) : (int)((size_t)j * maxnb + found);
  }
}

// CSR output in the reference's layout: self first, then the sorted list.
__global__ void csr_out_kernel(const double* pos, int n, Cell c, int maxnb,
                               const int* count, const int* list,
                               const long long* off, int* out_list,
                               double* rjs, double* xyz, double* th,
                               double* ph) {
#pragma clang fp contract(off)
  int i = blockIdx.x;
  long long base = off[i];
  int cnt = count[i];
  for (int s = threadIdx.x; s <= cnt; s += blockDim.x) {
    long long k = base + s;
    if (s == 0) {
      out_list[k] = i + 1;
      rjs[k] = 0.0;
      xyz[3 * k] = xyz[3 * k + 1] = xyz[3 * k + 2] = 0.0;
      th[k] = 0.0;
      ph[k] = 0.0;
      continue;
    }
    int e = list[(size_t)i * maxnb + s - 1];
    int j = e & TGX_ATOM_MASK, img = e >> TGX_IMG_SHIFT;
    double dv[3];
    double d = pair_vec(pos, i, e, c, dv);
    out_list[k] = img * n + j + 1;
    rjs[k] = d;
    xyz[3 * k] = dv[0];
    xyz[3 * k + 1] = dv[1];
    xyz[3 * k + 2] = dv[2];
    if (fabs(dv[2]) >= d)
      th[k] = dv[2] > 0.0 ? 0.0 : acos(-1.0);
    else
      th[k] = acos(dv[2] / d);
    ph[k] = atan2(dv[1], dv[0]);
  }
}

// The whole list build; with guard != nullptr every kernel returns at once
// unless *guard is non-zero (rebuild requested on the device), so the
// sequence can sit in a captured graph.
hipError_t launch_neighbors(const double* pos, int n, const Cell& c, double rcut,
                            int* cell_of, int* cell_count, int* cell_start,
                            int* cell_fill, int* cell_atoms, int* count,
                            int maxnb, int* list, int* rev, int* flag,
                            hipStream_t st, const int* guard) {
  const int B = 256;
  int nb = (n + B - 1) / B;
  if (!c.small) {
    int ncell = c.mx * c.my * c.mz;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((ncell + B - 1) / B), dim3(B), 0, st, cell_count,
                       ncell, guard);
    hipLaunchKernelGGL(bin_atoms_kernel, dim3(nb), dim3(B), 0, st, pos, n, c, cell_of, cell_count,
                       guard);
    hipLaunchKernelGGL(scan_cells_kernel, dim3(1), dim3(1024), 0, st, cell_count, ncell, cell_start,
                       cell_fill, guard);
    hipLaunchKernelGGL(scatter_cells_kernel, dim3(nb), dim3(B), 0, st, n, cell_of, cell_start,
                       cell_fill, cell_atoms, guard);
  }
  hipLaunchKernelGGL(find_neighbors_kernel, dim3(nb), dim3(B), 0, st, pos, n, c, rcut, cell_start,
                     cell_atoms, count, maxnb, list, flag, guard);
  hipLaunchKernelGGL(sort_lists_kernel, dim3(nb), dim3(B), 0, st, n, maxnb, count, list, guard);
  if (rev)
    hipLaunchKernelGGL(reverse_kernel, dim3(n), dim3(64), 0, st, n, maxnb, c, count, list, rev,
                       flag, guard);
  return hipGetLastError();
}

hipError_t launch_csr_out(const double* pos, int n, const Cell& c, int maxnb,
                          const int* count, const int* list,
                          const long long* off, int* out_list, double* rjs,
                          double* xyz, double* th, double* ph, hipStream_t st) {
  hipLaunchKernelGGL(csr_out_kernel, dim3(n), dim3(64), 0, st, pos, n, c, maxnb, count, list, off,
                     out_list, rjs, xyz, th, ph);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SOAP expansion
// ---------------------------------------------------------------------------

template <typename T> struct Radial;
template <> struct Radial<double> {
  // FP64: closed form with the FP64 W for every pair.
  template <int NMX>
  __device__ static void eval(const SoapDev* sd, const double* W, const float2*,
                              int s, double r, double* R, double* dR) {
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
  __device__ static void eval(const SoapDev* sd, const double*, const float2* tab,
                              int s, float r, float* R, float* dR) {
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

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  unsigned int r = 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)((u + r) >> 16);
}
__device__ __forceinline__ unsigned short f2bf(double f) { return f2bf((float)f); }

// One 64-lane workgroup per site.  Phase A (lane = list slot): pair geometry,
// radial values R_n and A_l Y_lm into LDS.  Phase B (lane = lm): c_nlm =
// sum_j R_n(j) A_l Y_lm(j) in registers.  Phase C: power spectrum, norm.
template <typename T, int NRP, int NMX>
__global__ __launch_bounds__(64) void soap_fwd_kernel(
    const SoapDev* __restrict__ sd, const double* __restrict__ W,
    const float2* __restrict__ rtab, const double* __restrict__ pos,
    const int* __restrict__ species, Cell cell, int maxnb,
    const int* __restrict__ count, const int* __restrict__ list,
    const int* __restrict__ site_atom, int n_sites,
    const int2* __restrict__ pairn, T* __restrict__ q, int q_ld,
    unsigned short* __restrict__ qb, int qb_ld, T* __restrict__ cbuf,
    T* __restrict__ pnorm) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = sd->l_max, L2 = sd->L2;
  T* sAY = reinterpret_cast<T*>(smem_raw);  // [64][L2]
  T* sR = sAY + 64 * L2;                    // [64][NRP]
  T* sC = sR + 64 * NRP;                    // [NRP][L2]
  const int site = blockIdx.x;
  const int lane = threadIdx.x;
  const int i = site_atom[site];
  const int si = species[i] - 1;
  const int cnt = min(count[i], maxnb);
  const double rcm = sd->rcut_max;
  const int lm0 = lane, lm1 = lane + 64;

  T acc0[NRP], acc1[NRP];
#pragma unroll
  for (int n = 0; n < NRP; n++) acc0[n] = acc1[n] = (T)0;

  for (int base = 0; base < cnt; base += 64) {
    const int slot = base + lane;
    double dv[3] = {0, 0, 0};
    double rd = 1e30;
    int s = 0;
    if (slot < cnt) {
      int e = list[(size_t)i * maxnb + slot];
      rd = pair_vec(pos, i, e, cell, dv);
      s = species[e & TGX_ATOM_MASK] - 1;
    }
    bool valid = slot < cnt && rd < rcm && rd < sd->rh[s];
    unsigned long long mask = __ballot(valid);
    int jc = __popcll(mask & ((1ull << lane) - 1ull));
    int npv = __popcll(mask);
    if (valid) {
      const T r = (T)rd;
      const T ir = (T)1 / r;
      const T ux = (T)dv[0] * ir, uy = (T)dv[1] * ir, uz = (T)dv[2] * ir;
      T R[NMX], dR[NMX];
      Radial<T>::template eval<NMX>(sd, W, rtab, s, r, R, dR);
      const int o = sd->off[s], nm = sd->nmax[s];
#pragma unroll
      for (int n = 0; n < NRP; n++) sR[jc * NRP + n] = (T)0;
#pragma unroll
      for (int n = 0; n < NMX; n++)
        if (n < nm) sR[jc * NRP + o + n] = R[n];
      T log2t, D;
      ang_params<T>(sd, s, r, &log2t, &D);
      T Cm = 1, Sm = 0;
      T* ay = sAY + jc * L2;
      for (int m = 0; m <= L; m++) {
        T A = exp2_t((T)(m * (m + 1)) * log2t);
        T Tl = exp2_t((T)(2 * m) * log2t);
        const T t2 = exp2_t((T)2 * log2t);
        T pm2 = 0, pm1 = SphK<T>::P(m);
        for (int l = m; l <= L; l++) {
          T P;
          if (l == m) {
            P = pm1;
          } else {
            Tl *= t2;
            A *= Tl;
            const int ix = l * (TGX_MAX_LMAX + 1) + m;
            P = SphK<T>::A(ix) * uz * pm1 - SphK<T>::B(ix) * pm2;
            pm2 = pm1;
            pm1 = P;
          }
          const int lmc = l * l + l + m;
          ay[lmc] = A * P * Cm;
          if (m > 0) ay[l * l + l - m] = A * P * Sm;
        }
        const T Cn = ux * Cm - uy * Sm;
        const T Sn = ux * Sm + uy * Cm;
        Cm = Cn;
        Sm = Sn;
      }
    }
    __syncthreads();
    for (int j = 0; j < npv; j++) {
      const T y0 = sAY[j * L2 + lm0];
      const T y1 = lm1 < L2 ? sAY[j * L2 + lm1] : (T)0;
#pragma unroll
      for (int n = 0; n < NRP; n++) {
        const T rv = sR[j * NRP + n];
        acc0[n] += rv * y0;
        acc1[n] += rv * y1;
      }
    }
    __syncthreads();
  }
  // central atom: l = 0 only
  if (lane == 0) {
    const int o = sd->off[si];
    for (int n = 0; n < sd->nmax[si]; n++) {
#pragma unroll
      for (int k = 0; k < NRP; k++)
        if (k == o + n) acc0[k] += (T)sd->rself[si * TGX_MAX_NR + n];
    }
  }
  T* cs = cbuf + (size_t)site * NRP * L2;
#pragma unroll
  for (int n = 0; n < NRP; n++) {
    if (lm0 < L2) {
      sC[n * L2 + lm0] = acc0[n];
      cs[n * L2 + lm0] = acc0[n];
    }
    if (lm1 < L2) {
      sC[n * L2 + lm1] = acc1[n];
      cs[n * L2 + lm1] = acc1[n];
    }
  }
  __syncthreads();
  // power spectrum
  const int dim = sd->dim, Lp = L + 1;
  T* qs = q + (size_t)site * q_ld;
  T ss = 0;
  for (int d = lane; d < dim; d += 64) {
    const int kp = d / Lp, l = d - kp * Lp;
    const int2 nn = pairn[kp];
    T a = 0;
    for (int m = -l; m <= l; m++) {
      const int lm = l * l + l + m;
      a += sC[nn.x * L2 + lm] * sC[nn.y * L2 + lm];
    }
    const T p = (nn.x == nn.y ? (T)1 : (T)1.41421356237309504880) * a;
    qs[d] = p;
    ss += p * p;
  }
  ss = wave_sum(ss);
  const T nrm = sqrt(ss);
  const T inv = (T)1 / nrm;
  for (int d = lane; d < q_ld; d += 64) {
    T v = d < dim ? qs[d] * inv : (T)0;
    qs[d] = v;
    if (qb) qb[(size_t)site * qb_ld + d] = f2bf(v);
  }
  if (qb)
    for (int d = q_ld + lane; d < qb_ld; d += 64) qb[(size_t)site * qb_ld + d] = 0;
  if (lane == 0) pnorm[site] = nrm;
}

// Adjoint: from G = dE/dq of a site, the force contributions
// g_ij = dE_i/dr_ij of every pair and the site's virial.
template <typename T, int NRP, int NMX>
__global__ __launch_bounds__(64) void soap_bwd_kernel(
    const SoapDev* __restrict__ sd, const double* __restrict__ W,
    const float2* __restrict__ rtab, const double* __restrict__ pos,
    const int* __restrict__ species, Cell cell, int maxnb,
    const int* __restrict__ count, const int* __restrict__ list,
    const int* __restrict__ site_atom, int n_sites,
    const int2* __restrict__ pairn, const T* __restrict__ q, int q_ld,
    const T* __restrict__ cbuf, const T* __restrict__ pnorm,
    const T* __restrict__ G, int g_ld, T* __restrict__ gpair,
    double* __restrict__ vsite) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = sd->l_max, L2 = sd->L2, NR = sd->nr, dim = sd->dim, Lp = L + 1;
  T* sC = reinterpret_cast<T*>(smem_raw);  // [NRP][L2]
  T* sGC = sC + NRP * L2;                  // [L2][NRP]
  T* sGP = sGC + NRP * L2;                 // [dim]
  const int site = blockIdx.x;
  const int lane = threadIdx.x;
  const int i = site_atom[site];
  const int cnt = min(count[i], maxnb);
  const T* cs = cbuf + (size_t)site * NRP * L2;
  for (int k = lane; k < NRP * L2; k += 64) sC[k] = cs[k];
  const T* qs = q + (size_t)site * q_ld;
  const T* gs = G + (size_t)site * g_ld;
  T gq = 0;
  for (int d = lane; d < dim; d += 64) gq += gs[d] * qs[d];
  gq = wave_sum(gq);
  const T inv = (T)1 / pnorm[site];
  for (int d = lane; d < dim; d += 64) sGP[d] = (gs[d] - gq * qs[d]) * inv;
  __syncthreads();
  // g_c[n][lm] = sum_n2 M_l[n][n2] c[n2][lm]
  for (int lm = lane; lm < L2; lm += 64) {
    int l = 0;
    while ((l + 1) * (l + 1) <= lm) l++;
    for (int n = 0; n < NRP; n++) {
      T a = 0;
      if (n < NR) {
        for (int n2 = 0; n2 < NR; n2++) {
          const int lo = n < n2 ? n : n2, hi = n < n2 ? n2 : n;
          const int kp = lo * NR - lo * (lo - 1) / 2 + (hi - lo);
          const T w = (n == n2) ? (T)2 : (T)1.41421356237309504880;
          a += w * sGP[kp * Lp + l] * sC[n2 * L2 + lm];
        }
      }
      sGC[lm * NRP + n] = a;
    }
  }
  __syncthreads();
  double v[6] = {0, 0, 0, 0, 0, 0};
  T* gp_out = gpair + (size_t)i * maxnb * 3;
  for (int base = 0; base < cnt; base += 64) {
    const int slot = base + lane;
    if (slot >= cnt) continue;
    const int e = list[(size_t)i * maxnb + slot];
    double dv[3];
    const double rd = pair_vec(pos, i, e, cell, dv);
    const int s = species[e & TGX_ATOM_MASK] - 1;
    T g0 = 0, g1 = 0, g2 = 0;
    if (rd < sd->rcut_max && rd < sd->rh[s]) {
      const T r = (T)rd;
      const T ir = (T)1 / r;
      const T ux = (T)dv[0] * ir, uy = (T)dv[1] * ir, uz = (T)dv[2] * ir;
      T R[NMX], dR[NMX];
      Radial<T>::template eval<NMX>(sd, W, rtab, s, r, R, dR);
      const int o = sd->off[s], nm = sd->nmax[s];
      T log2t, D;
      ang_params<T>(sd, s, r, &log2t, &D);
      T s1 = 0, Vx = 0, Vy = 0, Vz = 0;
      T Cm = 1, Sm = 0, Cp = 0, Sp = 0;  // C_m, S_m and C_{m-1}, S_{m-1}
      for (int m = 0; m <= L; m++) {
        T A = exp2_t((T)(m * (m + 1)) * log2t);
        T Tl = exp2_t((T)(2 * m) * log2t);
        const T t2 = exp2_t((T)2 * log2t);
        T pm2 = 0, pm1 = SphK<T>::P(m), dm2 = 0, dm1 = 0;
        for (int l = m; l <= L; l++) {
          T P, dP;
          if (l == m) {
            P = pm1;
            dP = dm1;
          } else {
            Tl *= t2;
            A *= Tl;
            const int ix = l * (TGX_MAX_LMAX + 1) + m;
            const T a = SphK<T>::A(ix), b = SphK<T>::B(ix);
            P = a * uz * pm1 - b * pm2;
            dP = a * (pm1 + uz * dm1) - b * dm2;
            pm2 = pm1;
            pm1 = P;
            dm2 = dm1;
            dm1 = dP;
          }
          const T dA = A * (T)(l * (l + 1)) * D;
          // cos (or m = 0) component
          {
            const T* gc = sGC + (l * l + l + m) * NRP + o;
            T h = 0, hd = 0;
#pragma unroll
            for (int n = 0; n < NMX; n++) {
              if (n < nm) {
                h += gc[n] * R[n];
                hd += gc[n] * dR[n];
              }
            }
            const T Y = P * Cm;
            s1 += (hd * A + h * dA) * Y;
            const T hA = h * A;
            if (m > 0) {
              Vx += hA * P * (T)m * Cp;
              Vy -= hA * P * (T)m * Sp;
            }
            Vz += hA * dP * Cm;
          }
          if (m > 0) {
            const T* gc = sGC + (l * l + l - m) * NRP + o;
            T h = 0, hd = 0;
#pragma unroll
            for (int n = 0; n < NMX; n++) {
              if (n < nm) {
                h += gc[n] * R[n];
                hd += gc[n] * dR[n];
              }
            }
            const T Y = P * Sm;
            s1 += (hd * A + h * dA) * Y;
            const T hA = h * A;
            Vx += hA * P * (T)m * Sp;
            Vy += hA * P * (T)m * Cp;
            Vz += hA * dP * Sm;
          }
        }
        const T Cn = ux * Cm - uy * Sm;
        const T Sn = ux * Sm + uy * Cm;
        Cp = Cm;
        Sp = Sm;
        Cm = Cn;
        Sm = Sn;
      }
      const T uV = ux * Vx + uy * Vy + uz * Vz;
      g0 = s1 * ux + (Vx - uV * ux) * ir;
      g1 = s1 * uy + (Vy - uV * uy) * ir;
      g2 = s1 * uz + (Vz - uV * uz) * ir;
      const double gd[3] = {(double)g0, (double)g1, (double)g2};
      // virial += 1/2 (F x + x F), F = -g the force on the neighbour
      v[0] -= gd[0] * dv[0];
      v[1] -= gd[1] * dv[1];
      v[2] -= gd[2] * dv[2];
      v[3] -= 0.5 * (gd[0] * dv[1] + gd[1] * dv[0]);
      v[4] -= 0.5 * (gd[0] * dv[2] + gd[2] * dv[0]);
      v[5] -= 0.5 * (gd[1] * dv[2] + gd[2] * dv[1]);
    }
    gp_out[slot * 3 + 0] = g0;
    gp_out[slot * 3 + 1] = g1;
    gp_out[slot * 3 + 2] = g2;
  }
  for (int k = 0; k < 6; k++) {
    double t = wave_sum(v[k]);
    if (lane == 0) vsite[(size_t)site * 6 + k] = t;
  }
}

template <typename T, int NRP, int NMX>
static hipError_t fwd_inst(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Cell& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int2* pairn, T* q, int q_ld,
                           unsigned short* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st) {
  size_t sm = sizeof(T) * ((size_t)64 * sdh.L2 + 64 * NRP + (size_t)NRP * sdh.L2);
  static size_t sm_set = 65536;
  if (sm > sm_set) {
    hipError_t e = hipFuncSetAttribute((const void*)soap_fwd_kernel<T, NRP, NMX>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
    sm_set = sm;
  }
  hipLaunchKernelGGL((soap_fwd_kernel<T, NRP, NMX>), dim3(n_sites), dim3(64), sm, st, sd, W,
                     reinterpret_cast<const float2*>(rtab), pos, species, cell, maxnb, count,
                     list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_soap_fwd(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Cell& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int2* pairn, T* q, int q_ld,
                           unsigned short* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st) {
  if (n_sites == 0) return hipSuccess;
  int nrp = (sdh.nr + 7) / 8 * 8;
  int nmx = 0;
  for (int s = 0; s < sdh.n_species; s++) nmx = sdh.nmax[s] > nmx ? sdh.nmax[s] : nmx;
  nmx = nmx <= 8 ? 8 : 16;
  if (nrp == 8) return fwd_inst<T, 8, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  if (nrp == 16 && nmx == 8) return fwd_inst<T, 16, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  if (nrp == 16) return fwd_inst<T, 16, 16>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  if (nrp == 24 && nmx == 8) return fwd_inst<T, 24, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  if (nrp == 24) return fwd_inst<T, 24, 16>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  if (nrp == 32 && nmx == 8) return fwd_inst<T, 32, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  if (nrp == 32) return fwd_inst<T, 32, 16>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, qb, qb_ld, cbuf, pnorm, st);
  return hipErrorInvalidValue;
}

template <typename T, int NRP, int NMX>
static hipError_t bwd_inst(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Cell& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int2* pairn, const T* q,
                           int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                           T* gpair, double* vsite, hipStream_t st) {
  size_t sm = sizeof(T) * ((size_t)2 * NRP * sdh.L2 + sdh.dim);
  static size_t sm_set = 65536;
  if (sm > sm_set) {
    hipError_t e = hipFuncSetAttribute((const void*)soap_bwd_kernel<T, NRP, NMX>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
    sm_set = sm;
  }
  hipLaunchKernelGGL((soap_bwd_kernel<T, NRP, NMX>), dim3(n_sites), dim3(64), sm, st, sd, W,
                     reinterpret_cast<const float2*>(rtab), pos, species, cell, maxnb, count,
                     list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair,
                     vsite);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_soap_bwd(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Cell& cell, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int2* pairn, const T* q,
                           int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                           T* gpair, double* vsite, hipStream_t st) {
  if (n_sites == 0) return hipSuccess;
  int nrp = (sdh.nr + 7) / 8 * 8;
  int nmx = 0;
  for (int s = 0; s < sdh.n_species; s++) nmx = sdh.nmax[s] > nmx ? sdh.nmax[s] : nmx;
  nmx = nmx <= 8 ? 8 : 16;
  if (nrp == 8) return bwd_inst<T, 8, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  if (nrp == 16 && nmx == 8) return bwd_inst<T, 16, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  if (nrp == 16) return bwd_inst<T, 16, 16>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  if (nrp == 24 && nmx == 8) return bwd_inst<T, 24, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  if (nrp == 24) return bwd_inst<T, 24, 16>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  if (nrp == 32 && nmx == 8) return bwd_inst<T, 32, 8>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  if (nrp == 32) return bwd_inst<T, 32, 16>(sdh, sd, W, rtab, pos, species, cell, maxnb, count, list, site_atom, n_sites, pairn, q, q_ld, cbuf, pnorm, G, g_ld, gpair, vsite, st);
  return hipErrorInvalidValue;
}

template hipError_t launch_soap_fwd<float>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Cell&, int, const int*, const int*, const int*, int, const int2*, float*, int, unsigned short*, int, float*, float*, hipStream_t);
template hipError_t launch_soap_fwd<double>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Cell&, int, const int*, const int*, const int*, int, const int2*, double*, int, unsigned short*, int, double*, double*, hipStream_t);
template hipError_t launch_soap_bwd<float>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Cell&, int, const int*, const int*, const int*, int, const int2*, const float*, int, const float*, const float*, const float*, int, float*, double*, hipStream_t);
template hipError_t launch_soap_bwd<double>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Cell&, int, const int*, const int*, const int*, int, const int2*, const double*, int, const double*, const double*, const double*, int, double*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// Forces: F_k = sum_slots (g[k][slot] - g[rev[k][slot]]), fixed order, FP64.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void force_gather_kernel(int n, int maxnb, const int* count,
                                    const int* list, const int* rev,
                                    const int* site_of_atom, const T* gpair,
                                    double* forces) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int cnt = min(count[k], maxnb);
  const bool ksite = site_of_atom[k] >= 0;
  double f0 = 0, f1 = 0, f2 = 0;
  for (int s = 0; s < cnt; s++) {
    const size_t p = (size_t)k * maxnb + s;
    if (ksite) {
      f0 += (double)gpair[p * 3];
      f1 += (double)gpair[p * 3 + 1];
      f2 += (double)gpair[p * 3 + 2];
    }
    const int j = list[p] & TGX_ATOM_MASK;
    if (site_of_atom[j] >= 0) {
      const size_t rp = (size_t)rev[p];
      f0 -= (double)gpair[rp * 3];
      f1 -= (double)gpair[rp * 3 + 1];
      f2 -= (double)gpair[rp * 3 + 2];
    }
  }
  forces[3 * k] += f0;
  forces[3 * k + 1] += f1;
  forces[3 * k + 2] += f2;
}

template <typename T>
hipError_t launch_force_gather(int n, int maxnb, const int* count, const int* list,
                               const int* rev, const int* site_of_atom, const T* gpair,
                               double* forces, hipStream_t st) {
  hipLaunchKernelGGL(force_gather_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, maxnb,
                     count, list, rev, site_of_atom, gpair, forces);
  return hipGetLastError();
}
template hipError_t launch_force_gather<float>(int, int, const int*, const int*, const int*, const int*, const float*, double*, hipStream_t);
template hipError_t launch_force_gather<double>(int, int, const int*, const int*, const int*, const int*, const double*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// distance_2b term (gap.f90:329-460): each pair counted from both atoms,
// forces on the central atom, virial -1/2 sym(F x) then x 1/2.
// ---------------------------------------------------------------------------
__global__ void two_body_kernel(TwoBodyDev b, const double* pos, const int* species,
                                int n, Cell c, int maxnb, const int* count,
                                const int* list, double* eatom, double* forces,
                                double* vatom, int do_forces) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double pi = 3.14159265358979323846;
  const int si = species[i];
  double e = 0, f[3] = {0, 0, 0}, v[6] = {0, 0, 0, 0, 0, 0};
  if (si == b.sp1 || si == b.sp2) {
    const int cnt = min(count[i], maxnb);
    for (int s = 0; s < cnt; s++) {
      const int ent = list[(size_t)i * maxnb + s];
      const int sj = species[ent & TGX_ATOM_MASK];
      if (!((si == b.sp1 && sj == b.sp2) || (si == b.sp2 && sj == b.sp1))) continue;
      double x[3];
      const double r = pair_vec(pos, i, ent, c, x);
      if (!(r < b.rcut)) continue;
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
      e += d2 * fcut * es;
      if (do_forces) {
        const double pre = -2.0 * d2 * fs / r;
        const double ff[3] = {pre * x[0], pre * x[1], pre * x[2]};
        f[0] += ff[0];
        f[1] += ff[1];
        f[2] += ff[2];
        v[0] -= ff[0] * x[0];
        v[1] -= ff[1] * x[1];
        v[2] -= ff[2] * x[2];
        v[3] -= 0.5 * (ff[0] * x[1] + ff[1] * x[0]);
        v[4] -= 0.5 * (ff[0] * x[2] + ff[2] * x[0]);
        v[5] -= 0.5 * (ff[1] * x[2] + ff[2] * x[1]);
      }
    }
  }
  eatom[i] += e;
  if (do_forces) {
    forces[3 * i] += f[0];
    forces[3 * i + 1] += f[1];
    forces[3 * i + 2] += f[2];
    for (int k = 0; k < 6; k++) vatom[(size_t)i * 6 + k] += 0.5 * v[k];
  }
}

hipError_t launch_two_body(const TwoBodyDev& b, const double* pos, const int* species, int n,
                           const Cell& c, int maxnb, const int* count, const int* list,
                           double* eatom, double* forces, double* vatom, int do_forces,
                           hipStream_t st) {
  hipLaunchKernelGGL(two_body_kernel, dim3((n + 127) / 128), dim3(128), 0, st, b, pos, species, n,
                     c, maxnb, count, list, eatom, forces, vatom, do_forces);
  return hipGetLastError();
}

// e_atom = e0(species) + SOAP site energy; forces and per-atom virial zeroed.
__global__ void energy_init_kernel(int n, const int* species, const double* e0,
                                   const int* site_of_atom, const double* esite,
                                   double* eatom, double* forces, double* vatom) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = site_of_atom[i];
  eatom[i] = e0[species[i] - 1] + (s >= 0 && esite ? esite[s] : 0.0);
  if (forces) forces[3 * i] = forces[3 * i + 1] = forces[3 * i + 2] = 0.0;
  if (vatom)
    for (int k = 0; k < 6; k++) vatom[(size_t)i * 6 + k] = 0.0;
}

hipError_t launch_energy_init(int n, const int* species, const double* e0,
                              const int* site_of_atom, const double* esite, double* eatom,
                              double* forces, double* vatom, hipStream_t st) {
  hipLaunchKernelGGL(energy_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, species, e0,
                     site_of_atom, esite, eatom, forces, vatom);
  return hipGetLastError();
}

// Deterministic reduction of per-site and per-atom 6-component virials and of
// the atom energies: one workgroup, fixed thread/element assignment.
__global__ void reduce_kernel(int n_sites, const double* vsite, int n,
                              const double* vatom, const double* eatom,
                              double* out /* 9 virial + 1 energy */) {
  __shared__ double sh[256][7];
  const int t = threadIdx.x;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  if (vsite)
    for (int s = t; s < n_sites; s += 256)
      for (int k = 0; k < 6; k++) a[k] += vsite[(size_t)s * 6 + k];
  if (vatom)
    for (int s = t; s < n; s += 256)
      for (int k = 0; k < 6; k++) a[k] += vatom[(size_t)s * 6 + k];
  if (eatom)
    for (int s = t; s < n; s += 256) a[6] += eatom[s];
  for (int k = 0; k < 7; k++) sh[t][k] = a[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
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

hipError_t launch_reduce(int n_sites, const double* vsite, int n, const double* vatom,
                         const double* eatom, double* out, hipStream_t st) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, st, n_sites, vsite, n, vatom, eatom,
                     out);
  return hipGetLastError();
}

}  // namespace tgx
