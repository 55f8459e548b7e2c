// Metropolis Monte Carlo with single-atom moves on the device (reference:
// mc/mc_interface.f90:183-447, mc/mc_utils.f90:172-180, 215-277, 696-698).
//
// A trial is a fixed sequence of launches, so a block of trials can be
// captured once in a hipGraph and replayed:
//   propose   draws the atom and displacement from the counter-based stream,
//             moves the atom in place and raises the list-rebuild flag when
//             it left the skin sphere of the last build;
//   (guarded neighbour rebuild and reference copy)
//   affected  lists the sites whose environment changed: the moved atom and
//             its list neighbours (the list holds every atom within
//             rcut + skin of the build positions, which covers rcut around
//             both the old and the new position);
//   SOAP forward and the energy GEMM on those sites (library kernels);
//   decide    dE = sum of site-energy changes + the exact 2-body change,
//             p = min(1, exp(-dE/(kB T))), accept when p > u; on rejection
//             the atom goes back, on acceptance the site energies are kept.
// Sequential semantics are exact: every trial sees the state left by the
// previous one.
#include <hip/hip_runtime.h>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

// u(seed, t, k) in [0, 1): the top 53 bits of the splitmix64 finaliser of
// seed + 0x9E3779B97F4A7C15 * (5t + k + 1).  k = 0 atom, 1..3 displacement,
// 4 acceptance.
__device__ __forceinline__ double mc_uniform(unsigned long long seed, long long t, int k) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(5 * t + k + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void mc_propose_kernel(int n, double* pos, const double* pref,
                                  unsigned long long seed, double move_max, double lim2,
                                  int* ctl, double* dctl) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const int t = ctl[MC_STEP];
  const int live = t < ctl[MC_END];
  ctl[MC_LIVE] = live;
  ctl[MC_REBUILD] = 0;
  if (!live) return;
  int m = (int)floor(mc_uniform(seed, t, 0) * (double)n);
  if (m > n - 1) m = n - 1;
  ctl[MC_ATOM] = m;
  double d2 = 0.0;
  for (int k = 0; k < 3; k++) {
    const double disp = move_max * (mc_uniform(seed, t, 1 + k) - 0.5);
    const double old = pos[3 * m + k];
    dctl[MC_OLD + k] = old;
    pos[3 * m + k] = old + disp;
    const double dx = pos[3 * m + k] - pref[3 * m + k];
    d2 += dx * dx;
  }
  if (d2 > lim2) {
    ctl[MC_REBUILD] = 1;
    ctl[MC_NREB] += 1;
  }
}

// aff[0] = moved atom, aff[1..cnt] = its list neighbours; the remaining
// capacity repeats the moved atom (computed, never summed).  With a central
// species only atoms of that species are sites.
__global__ void mc_affected_kernel(const int* ctl, const int* count, const int* list,
                                   int maxnb, const int* species, int central, int* aff,
                                   int cap, int* n_aff) {
  if (!ctl[MC_LIVE]) return;
  const int m = ctl[MC_ATOM];
  const int cnt = min(count[m], maxnb);
  __shared__ int keep[1024];
  __shared__ int nk;
  const int t = threadIdx.x;
  if (t == 0) {
    // fixed order: the moved atom, then the list in slot order
    int k = 0;
    if (central == 0 || species[m] == central) keep[k++] = m;
    for (int s = 0; s < cnt; s++) {
      const int j = list[(size_t)m * maxnb + s] & TGX_ATOM_MASK;
      if (central == 0 || species[j] == central) keep[k++] = j;
    }
    nk = k;
    *n_aff = k;
  }
  __syncthreads();
  for (int k = t; k < cap; k += blockDim.x) aff[k] = k < nk ? keep[k] : m;
}

// Exact FP64 pair energy of one distance_2b block (gap.f90:383-399).
__device__ double mc_phi(const TwoBodyDev& b, double r) {
  double phi, dphi;
  two_body_pair<1>(b, r, &phi, &dphi);
  return phi;
}

__global__ __launch_bounds__(64) void mc_decide_kernel(
    int n, double* pos, const int* species, Cell cell, int maxnb, const int* count,
    const int* list, const int* aff, const int* n_aff, const double* esite, double* e_cur,
    const TwoBodyDev* twob, int n2b, unsigned long long seed, double kT, int* ctl, double* dctl,
    double* trace) {
  if (!ctl[MC_LIVE]) return;
  const int lane = threadIdx.x;
  const int t = ctl[MC_STEP];
  const int m = ctl[MC_ATOM];
  const int na = *n_aff;
  double de = 0.0;
  for (int k = lane; k < na; k += 64) de += esite[k] - e_cur[aff[k]];
  // 2-body: every pair (m, j) enters E_m and E_j, so dE = 2 sum_j dphi
  const int cnt = min(count[m], maxnb);
  const double* pm = pos + 3 * m;
  const double po[3] = {dctl[MC_OLD], dctl[MC_OLD + 1], dctl[MC_OLD + 2]};
  for (int b = 0; b < n2b; b++) {
    const TwoBodyDev& B = twob[b];
    const int sm = species[m];
    if (sm != B.sp1 && sm != B.sp2) continue;
    for (int s = lane; s < cnt; s += 64) {
      const int j = list[(size_t)m * maxnb + s] & TGX_ATOM_MASK;
      const int sj = species[j];
      if (!((sm == B.sp1 && sj == B.sp2) || (sm == B.sp2 && sj == B.sp1))) continue;
      double dv[3];
      const double rn = mic(pm, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
      const double ro = mic(po, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
      de += 2.0 * (mc_phi(B, rn) - mc_phi(B, ro));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) de += __shfl_xor(de, o, 64);
  __shared__ int acc;
  if (lane == 0) {
    double p = exp(-de / kT);
    p = p < 1.0 ? p : 1.0;
    acc = p > mc_uniform(seed, t, 4);
    if (acc) {
      dctl[MC_ETOT] += de;
      ctl[MC_NACC] += 1;
    } else {
      for (int k = 0; k < 3; k++) pos[3 * m + k] = dctl[MC_OLD + k];
    }
    if (trace) trace[t + 1] = dctl[MC_ETOT];
    ctl[MC_STEP] = t + 1;
  }
  __syncthreads();
  if (acc)
    for (int k = lane; k < na; k += 64) e_cur[aff[k]] = esite[k];
}

// e_cur[atom] = SOAP site energy of the atom (0 for atoms that are no site).
__global__ void mc_init_kernel(int n, const int* site_of_atom, const double* esite,
                               double* e_cur) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = site_of_atom[i];
  e_cur[i] = (s >= 0 && esite) ? esite[s] : 0.0;
}

hipError_t launch_mc_propose(int n, double* pos, const double* pref, unsigned long long seed,
                             double move_max, double lim2, int* ctl, double* dctl,
                             hipStream_t st) {
  hipLaunchKernelGGL(mc_propose_kernel, dim3(1), dim3(64), 0, st, n, pos, pref, seed, move_max,
                     lim2, ctl, dctl);
  return hipGetLastError();
}

hipError_t launch_mc_affected(const int* ctl, const int* count, const int* list, int maxnb,
                              const int* species, int central, int* aff, int cap, int* n_aff,
                              hipStream_t st) {
  if (cap > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mc_affected_kernel, dim3(1), dim3(256), 0, st, ctl, count, list, maxnb,
                     species, central, aff, cap, n_aff);
  return hipGetLastError();
}

hipError_t launch_mc_decide(int n, double* pos, const int* species, const Cell& cell, int maxnb,
                            const int* count, const int* list, const int* aff, const int* n_aff,
                            const double* esite, double* e_cur, const TwoBodyDev* twob, int n2b,
                            unsigned long long seed, double kT, int* ctl, double* dctl,
                            double* trace, hipStream_t st) {
  hipLaunchKernelGGL(mc_decide_kernel, dim3(1), dim3(64), 0, st, n, pos, species, cell, maxnb,
                     count, list, aff, n_aff, esite, e_cur, twob, n2b, seed, kT, ctl, dctl, trace);
  return hipGetLastError();
}

hipError_t launch_mc_init(int n, const int* site_of_atom, const double* esite, double* e_cur,
                          hipStream_t st) {
  hipLaunchKernelGGL(mc_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, site_of_atom,
                     esite, e_cur);
  return hipGetLastError();
}

}  // namespace tgx
