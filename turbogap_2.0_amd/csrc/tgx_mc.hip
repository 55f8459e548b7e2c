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
//   SOAP forward on those sites (library kernel);
//   stage 1   (delayed acceptance only) the surrogate MLP energy change
//             dE_s of the same sites plus the exact 2-body change; accepted
//             with min(1, exp(-dE_s/(kB T))) > u_4, else the trial ends here;
//   energy GEMM on the sites (skipped by a device flag after a stage-1
//             rejection) and, while training data are collected, a copy of
//             (q, GAP site energy) into the sample ring;
//   stage 2   dE = sum of site-energy changes + the exact 2-body change;
//             plain Metropolis accepts with min(1, exp(-dE/(kB T))) > u_4,
//             delayed acceptance (Christen & Fox, J. Comput. Graph. Stat. 14,
//             795 (2005)) with min(1, exp(-(dE - dE_s)/(kB T))) > u_5.  On
//             rejection the atom goes back; on acceptance the site energies
//             are kept.
// Sequential semantics are exact: every trial sees the state left by the
// previous one, and the chain samples exp(-E/(kB T)) in both modes.
#include <hip/hip_runtime.h>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

// u(seed, t, k) in [0, 1): the top 53 bits of the splitmix64 finaliser of
// seed + 0x9E3779B97F4A7C15 * (6t + k + 1).  k = 0 atom, 1..3 displacement,
// 4 acceptance.
__device__ __forceinline__ double mc_uniform(unsigned long long seed, long long t, int k) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(6 * t + k + 1);
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

// aff[0] = moved atom, then the list neighbours within rsoap of its old or
// new position (the only sites whose descriptor changes); the remaining
// capacity repeats the moved atom (computed, never summed).  With a central
// species only atoms of that species are sites.
__device__ double mc_de2b(const double* pos, const int* species, const Cell& cell, int maxnb,
                          const int* count, const int* list, const TwoBodyDev* twob, int n2b,
                          int m, const double* pm, const double* po, int lane, int w, double* red,
                          bool exact);

// With WITH_2B the same workgroup then computes the 2-body change of the move
// into dctl[MC_DE2B] (and opens the GAP gate when open_gate: plain
// Metropolis, whose stage 1 is nothing else).
template <bool WITH_2B>
__global__ __launch_bounds__(256) void mc_affected_kernel(
    int* ctl, double* dctl, const double* pos, Cell cell, double rsoap, const int* count,
    const int* list, int maxnb, const int* species, int central, int* aff, int cap, int* n_aff,
    const TwoBodyDev* twob, int n2b, int exact, int open_gate) {
  if (!ctl[MC_LIVE]) return;
  const int m = ctl[MC_ATOM];
  const int cnt = min(count[m], maxnb);
  __shared__ int keep[1024];
  __shared__ int nk;
  const int t = threadIdx.x;
  if (t < 64) {
    // fixed order: the moved atom, then the list in slot order (the first
    // wave tests 64 slots at a time and compacts them with a ballot)
    const double po[3] = {dctl[MC_OLD], dctl[MC_OLD + 1], dctl[MC_OLD + 2]};
    int k = (central == 0 || species[m] == central) ? 1 : 0;
    if (t == 0 && k) keep[0] = m;
    for (int base = 0; base < cnt; base += 64) {
      const int s = base + t;
      bool ok = false;
      int j = 0;
      if (s < cnt) {
        j = list[(size_t)m * maxnb + s] & TGX_ATOM_MASK;
        if (central == 0 || species[j] == central) {
          double dv[3];
          const double rn = mic(pos + 3 * m, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
          const double ro = mic(po, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
          ok = rn < rsoap || ro < rsoap;
        }
      }
      const unsigned long long b = __ballot(ok);
      if (ok) keep[k + __popcll(b & ((1ull << t) - 1ull))] = j;
      k += __popcll(b);
    }
    if (t == 0) {
      nk = k;
      *n_aff = k;
    }
  }
  __syncthreads();
  for (int k = t; k < cap; k += blockDim.x) aff[k] = k < nk ? keep[k] : m;
  if constexpr (WITH_2B) {
    __shared__ double red[4];
    const double de2 = mc_de2b(pos, species, cell, maxnb, count, list, twob, n2b, m, pos + 3 * m,
                               dctl + MC_OLD, t & 63, t >> 6, red, exact != 0);
    if (t == 0) {
      dctl[MC_DE2B] = de2;
      if (open_gate) ctl[MC_GATE] = 1;
    }
  }
}

// Exact FP64 pair energy of one distance_2b block (gap.f90:383-399),
// restricted to the sparse distances k0 <= k < k1 (the cutoff factor
// multiplies the whole sum, so partial sums add up to phi).
__device__ double mc_phi_part(const TwoBodyDev& b, double r, int k0, int k1) {
  const double pi = 3.14159265358979323846;
  if (!(r < b.rcut)) return 0.0;
  double fcut = 1.0;
  if (!(r < b.rcut - b.buffer)) fcut = (cos(pi * (r - b.rcut + b.buffer) / b.buffer) + 1.0) / 2.0;
  const double is2 = 1.0 / (b.sigma * b.sigma);
  double es = 0;
  for (int k = k0; k < k1; k++) {
    const double dr = r - b.Qs[k];
    es += b.alphas[k] * exp(-0.5 * dr * dr * is2);
  }
  return b.delta * b.delta * fcut * es;
}

// Exact 2-body energy change of moving atom m from po to pm:
// every pair (m, j) enters E_m and E_j, so dE = 2 sum_j dphi.  The four
// waves of the workgroup take a quarter of the sparse distances each, the
// lanes the neighbour slots; fixed-order sum over lanes, then waves.
__device__ double mc_de2b(const double* pos, const int* species, const Cell& cell, int maxnb,
                          const int* count, const int* list, const TwoBodyDev* twob, int n2b,
                          int m, const double* pm, const double* po, int lane, int w, double* red,
                          bool exact) {
  double de = 0.0;
  const int cnt = min(count[m], maxnb);
  for (int b = 0; b < n2b; b++) {
    const TwoBodyDev& B = twob[b];
    const int sm = species[m];
    if (sm != B.sp1 && sm != B.sp2) continue;
    // exact: every wave takes a quarter of the sparse distances for all
    // pairs; otherwise (the FP32/mixed modes, as their full evaluation) the
    // tabulated phi, the waves taking every fourth group of 64 pairs
    const int qn = (B.n_sparse + 3) / 4, k0 = min(B.n_sparse, w * qn),
              k1 = min(B.n_sparse, k0 + qn);
    for (int s = exact ? lane : lane + 64 * w; s < cnt; s += exact ? 64 : 256) {
      const int j = list[(size_t)m * maxnb + s] & TGX_ATOM_MASK;
      const int sj = species[j];
      if (!((sm == B.sp1 && sj == B.sp2) || (sm == B.sp2 && sj == B.sp1))) continue;
      double dv[3];
      const double rn = mic(pm, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
      const double ro = mic(po, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
      if (exact) {
        de += 2.0 * (mc_phi_part(B, rn, k0, k1) - mc_phi_part(B, ro, k0, k1));
      } else {
        double pn, po_, d;
        two_body_pair<false>(B, rn, &pn, &d);
        two_body_pair<false>(B, ro, &po_, &d);
        de += 2.0 * (pn - po_);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) de += __shfl_xor(de, o, 64);
  if (lane == 0) red[w] = de;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

__device__ void mc_finish_trial(int t, double* trace, int* ctl, const double* dctl) {
  if (trace) trace[t + 1] = dctl[MC_ETOT];
  ctl[MC_STEP] = t + 1;
}

// Stage 1: the exact 2-body change always; with delayed acceptance active the
// surrogate change of the affected sites and the first accept/reject.
__global__ __launch_bounds__(256) void mc_stage1_kernel(
    double* pos, const int* species, Cell cell, int maxnb, const int* count, const int* list,
    const int* aff, const int* n_aff, const double* esur, const double* esur_cur,
    const TwoBodyDev* twob, int n2b, int exact, int have_2b, unsigned long long seed, double kT,
    int* ctl, double* dctl, double* trace) {
  if (!ctl[MC_LIVE]) return;
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = ctl[MC_STEP];
  const int m = ctl[MC_ATOM];
  // have_2b: the affected-site kernel already computed the 2-body change
  const double de2 =
      have_2b ? dctl[MC_DE2B]
              : mc_de2b(pos, species, cell, maxnb, count, list, twob, n2b, m, pos + 3 * m,
                        dctl + MC_OLD, lane, w, red, exact);
  if (w != 0) return;
  const bool da = ctl[MC_DA] && !ctl[MC_FALLBACK];
  double ds = 0.0;
  if (da) {
    const int na = *n_aff;
    for (int k = lane; k < na; k += 64) ds += esur[k] - esur_cur[aff[k]];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ds += __shfl_xor(ds, o, 64);
  }
  if (lane == 0) {
    dctl[MC_DE2B] = de2;
    ctl[MC_GATE] = 1;
    if (da) {
      const double dsur = ds + de2;
      dctl[MC_DESUR] = dsur;
      double p = exp(-dsur / kT);
      p = p < 1.0 ? p : 1.0;
      if (!(p > mc_uniform(seed, t, 4))) {
        ctl[MC_GATE] = 0;
        ctl[MC_NREJ1] += 1;
        for (int k = 0; k < 3; k++) pos[3 * m + k] = dctl[MC_OLD + k];
        mc_finish_trial(t, trace, ctl, dctl);
      }
    }
  }
}

// Stage 2 (GAP): plain Metropolis, or the delayed-acceptance correction.
// New site energy k of the trial: esite[k], or, from the few-site kernel's
// FP32 partials, d2 times their sum in part order (the row-sum kernel's
// arithmetic).
__device__ __forceinline__ double mc_site_e(const double* esite, const float* parts, int nparts,
                                            int part_ld, double d2, int k) {
  if (!parts) return esite[k];
  double s = 0.0;
  for (int t = 0; t < nparts; t++) s += (double)parts[(size_t)t * part_ld + k];
  return d2 * s;
}

__global__ __launch_bounds__(64) void mc_stage2_kernel(
    double* pos, const int* aff, const int* n_aff, const double* esite, const float* parts,
    int nparts, int part_ld, double d2, double* e_cur, const double* esur, double* esur_cur,
    unsigned long long seed, double kT, int* ctl, double* dctl, double* trace) {
  if (!ctl[MC_LIVE] || !ctl[MC_GATE]) return;
  const int lane = threadIdx.x;
  const int t = ctl[MC_STEP];
  const int m = ctl[MC_ATOM];
  const int na = *n_aff;
  double de = 0.0;
  if (esite)
    for (int k = lane; k < na; k += 64)
      de += mc_site_e(esite, parts, nparts, part_ld, d2, k) - e_cur[aff[k]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) de += __shfl_xor(de, o, 64);
  const bool da = ctl[MC_DA] && !ctl[MC_FALLBACK];
  __shared__ int acc;
  if (lane == 0) {
    de += dctl[MC_DE2B];
    ctl[MC_NGAP] += 1;
    double p;
    if (da) {
      const double d = de - dctl[MC_DESUR];
      p = exp(-d / kT);
      dctl[MC_EMA] = 0.98 * dctl[MC_EMA] + 0.02 * d * d;
      if (dctl[MC_EMA] > dctl[MC_MAXMS]) ctl[MC_FALLBACK] = 1;
    } else {
      p = exp(-de / kT);
    }
    p = p < 1.0 ? p : 1.0;
    acc = p > mc_uniform(seed, t, da ? 5 : 4);
    if (acc) {
      dctl[MC_ETOT] += de;
      ctl[MC_NACC] += 1;
    } else {
      for (int k = 0; k < 3; k++) pos[3 * m + k] = dctl[MC_OLD + k];
    }
    mc_finish_trial(t, trace, ctl, dctl);
  }
  __syncthreads();
  if (acc) {
    if (esite)
      for (int k = lane; k < na; k += 64)
        e_cur[aff[k]] = mc_site_e(esite, parts, nparts, part_ld, d2, k);
    if (da)
      for (int k = lane; k < na; k += 64) esur_cur[aff[k]] = esur[k];
  }
}

// Training samples: the descriptors and GAP energies of the trial's sites
// (rows 0..n_aff-1 of q and esite) appended to a ring of cap_s samples.
template <typename T>
__global__ void mc_collect_kernel(const int* ctl, const int* n_aff, const T* q, int q_ld, int D,
                                  const double* esite, float* sq, double* sy, int cap_s,
                                  int* cursor) {
  if (!ctl[MC_LIVE] || !ctl[MC_GATE] || !ctl[MC_COLLECT]) return;
  const int na = *n_aff;
  const int base = *cursor;
  for (int k = 0; k < na; k++) {
    const int slot = (base + k) % cap_s;
    for (int d = threadIdx.x; d < D; d += blockDim.x)
      sq[(size_t)slot * D + d] = (float)q[(size_t)k * q_ld + d];
    if (threadIdx.x == 0) sy[slot] = esite[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) *cursor = base + na;
}

// e_cur[atom] = SOAP site energy of the atom (0 for atoms that are no site).
__global__ void mc_init_kernel(int n, const int* site_of_atom, const double* esite,
                               double* e_cur) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = site_of_atom[i];
  e_cur[i] = (s >= 0 && esite) ? esite[s] : 0.0;
}

// ---------------------------------------------------------------------------
// Speculative batches: nb trials proposed and evaluated at once against the
// current state, then decided in trial order by one wave.  Trial t of a
// batch sees the state its predecessors left only if none of the accepted
// ones moved an atom within rdep of its moved atom (the range its energy
// change depends on); the walk stops at the first trial for which that
// fails, or that needs a list rebuild, and the next batch starts there,
// with the same random numbers.  The decisions are those of the one-trial
// sequence.
//
// Each trial's new descriptors come from virtual atoms: ghost n + t holds
// the moved atom at its new position; virtual atom vb + t cap + k holds the
// trial's k-th affected site (the moved atom at its new position, the others
// where they are) with its list row, in which the moved atom is replaced by
// the ghost.  The SOAP and kernel-step launches then run unchanged on the
// virtual sites.
// ---------------------------------------------------------------------------
enum { MB_ATOM = 0, MB_NAFF, MB_REB, MB_LIVE, MB_NI = 4 };
enum { MB_OLD = 0, MB_NEW = 3, MB_DE2B = 6, MB_ND = 8 };

__global__ void mcb_propose_kernel(int n, int nb, double* pos, int* species, const double* pref,
                                   unsigned long long seed, double move_max, double lim2,
                                   int* ctl, int* mi, double* md) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int T = ctl[MC_STEP] + t;
  __syncthreads();  // every lane has read the step before the flag is cleared
  if (t == 0) ctl[MC_REBUILD] = 0;  // the batch's rebuild (if any) has run
  if (t >= nb) return;
  const int live = T < ctl[MC_END];
  mi[t * MB_NI + MB_LIVE] = live;
  mi[t * MB_NI + MB_REB] = 0;
  if (!live) {
    mi[t * MB_NI + MB_ATOM] = 0;
    return;
  }
  int m = (int)floor(mc_uniform(seed, T, 0) * (double)n);
  if (m > n - 1) m = n - 1;
  mi[t * MB_NI + MB_ATOM] = m;
  double d2 = 0.0;
  for (int k = 0; k < 3; k++) {
    const double disp = move_max * (mc_uniform(seed, T, 1 + k) - 0.5);
    const double old = pos[3 * m + k];
    const double nw = old + disp;
    md[t * MB_ND + MB_OLD + k] = old;
    md[t * MB_ND + MB_NEW + k] = nw;
    pos[3 * (n + t) + k] = nw;  // ghost
    const double dx = nw - pref[3 * m + k];
    d2 += dx * dx;
  }
  species[n + t] = species[m];
  if (d2 > lim2) mi[t * MB_NI + MB_REB] = 1;
}

// One workgroup per trial: affected sites (the moved atom, then its list
// neighbours within rsoap of its old or new position, in slot order, as the
// one-trial scan), the virtual atoms and rows, and the exact 2-body change.
// Trials that do not run get isolated virtual sites (no neighbours).
__global__ __launch_bounds__(256) void mcb_affected_kernel(
    int n, int nb, int cap, int vb, double* pos, int* species, int* count, int* list, int maxnb,
    Cell cell, double rsoap, int central, int* site_atom, int* vatom, int* mi, double* md,
    const TwoBodyDev* twob, int n2b, int exact) {
  const int t = blockIdx.x;
  const int th = threadIdx.x;
  __shared__ int keep[1024];
  __shared__ int nk;
  __shared__ double red[4];
  const int live = mi[t * MB_NI + MB_LIVE];
  const int m = mi[t * MB_NI + MB_ATOM];
  const int ghost = n + t;
  if (!live) {
    if (th == 0) {
      nk = 0;
      mi[t * MB_NI + MB_NAFF] = 0;
    }
  } else if (th < 64) {
    const int cnt = min(count[m], maxnb);
    const double* po = md + t * MB_ND + MB_OLD;
    const double* pn = md + t * MB_ND + MB_NEW;
    int k = (central == 0 || species[m] == central) ? 1 : 0;
    if (th == 0 && k) keep[0] = m;
    for (int base = 0; base < cnt; base += 64) {
      const int s = base + th;
      bool ok = false;
      int j = 0;
      if (s < cnt) {
        j = list[(size_t)m * maxnb + s] & TGX_ATOM_MASK;
        if (central == 0 || species[j] == central) {
          double dv[3];
          const double rn = mic(pn, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
          const double ro = mic(po, pos + 3 * j, cell.La, cell.Lb, cell.Lc, dv);
          ok = rn < rsoap || ro < rsoap;
        }
      }
      const unsigned long long b = __ballot(ok);
      if (ok) keep[k + __popcll(b & ((1ull << th) - 1ull))] = j;
      k += __popcll(b);
    }
    if (th == 0) {
      nk = k;
      mi[t * MB_NI + MB_NAFF] = k;
    }
  }
  __syncthreads();
  const int na = nk;
  // virtual atoms: position, species, count; site list and real atom
  for (int k = th; k < cap; k += blockDim.x) {
    const int V = vb + t * cap + k;
    int a, c;
    const double* p;
    if (!live || k >= na) {
      // padding sites (computed, never summed): isolated copies
      a = live ? m : 0;
      c = 0;
      p = pos + 3 * a;
    } else {
      a = keep[k];
      c = min(count[a], maxnb);
      p = a == m ? md + t * MB_ND + MB_NEW : pos + 3 * a;
    }
    pos[3 * V] = p[0];
    pos[3 * V + 1] = p[1];
    pos[3 * V + 2] = p[2];
    species[V] = species[a];
    count[V] = c;
    site_atom[t * cap + k] = V;
    vatom[t * cap + k] = a;
  }
  // rows of the affected sites: the real row with the moved atom replaced
  // by the ghost; a wave per site, lanes over its slots
  if (live) {
    for (int k = th >> 6; k < na; k += blockDim.x >> 6) {
      const int a = keep[k];
      const int c = min(count[a], maxnb);
      const int* src = list + (size_t)a * maxnb;
      int* dst = list + (size_t)(vb + t * cap + k) * maxnb;
      for (int s = th & 63; s < c; s += 64) {
        const int ent = src[s];
        dst[s] = (ent & TGX_ATOM_MASK) == m ? ghost : ent;
      }
    }
  }
  if (n2b && live) {
    const double de2 = mc_de2b(pos, species, cell, maxnb, count, list, twob, n2b, m,
                               md + t * MB_ND + MB_NEW, md + t * MB_ND + MB_OLD, th & 63, th >> 6,
                               red, exact != 0);
    if (th == 0) md[t * MB_ND + MB_DE2B] = de2;
  } else if (th == 0) {
    md[t * MB_ND + MB_DE2B] = 0.0;
  }
}

// The walk: trials in order, each decided exactly as the one-trial kernels
// decide it (same sums, same random numbers), until one is stale or needs a
// rebuild.  One wave.
__global__ __launch_bounds__(64) void mcb_walk_kernel(
    int nb, int cap, double* pos, const Cell cell, double rdep, const int* vatom,
    const double* esite, double* e_cur, const double* esur, double* esur_cur,
    unsigned long long seed, double kT, int* ctl, double* dctl, const int* mi, const double* md,
    double* trace) {
  const int lane = threadIdx.x;
  __shared__ double acc_old[64 * 3], acc_new[64 * 3];
  __shared__ int acc_atom[64];
  __shared__ int n_acc, take;
  if (lane == 0) n_acc = 0;
  __syncthreads();
  int t = 0;
  for (; t < nb; t++) {
    if (!mi[t * MB_NI + MB_LIVE]) break;
    const int m = mi[t * MB_NI + MB_ATOM];
    const double* po = md + t * MB_ND + MB_OLD;
    if (mi[t * MB_NI + MB_REB]) {
      if (lane == 0) {
        ctl[MC_REBUILD] = 1;
        ctl[MC_NREB] += 1;
      }
      break;
    }
    // stale: an accepted trial of this batch moved an atom within rdep +
    // |d_t| (the trial's sites lie within rcut + |d_t| of its old position,
    // their environments within rcut of those)
    const double* pn0 = md + t * MB_ND + MB_NEW;
    const double dd = sqrt((pn0[0] - po[0]) * (pn0[0] - po[0]) + (pn0[1] - po[1]) * (pn0[1] - po[1]) +
                           (pn0[2] - po[2]) * (pn0[2] - po[2]));
    const double rt = rdep + dd, rdep2 = rt * rt;
    bool st = false;
    for (int a = lane; a < n_acc && !st; a += 64) {
      double dv[3];
      const double ro = mic(acc_old + 3 * a, po, cell.La, cell.Lb, cell.Lc, dv);
      const double rn = mic(acc_new + 3 * a, po, cell.La, cell.Lb, cell.Lc, dv);
      st = acc_atom[a] == m || ro * ro < rdep2 || rn * rn < rdep2;
    }
    if (__any(st)) break;
    const int T = ctl[MC_STEP];
    const int na = mi[t * MB_NI + MB_NAFF];
    const double de2 = md[t * MB_ND + MB_DE2B];
    const bool da = ctl[MC_DA] && !ctl[MC_FALLBACK];
    // stage 1 (delayed acceptance): surrogate change of the sites
    double dsur = 0.0;
    bool gate = true;
    if (da) {
      double ds = 0.0;
      for (int k = lane; k < na; k += 64) ds += esur[t * cap + k] - esur_cur[vatom[t * cap + k]];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ds += __shfl_xor(ds, o, 64);
      dsur = ds + de2;
      double p = exp(-dsur / kT);
      p = p < 1.0 ? p : 1.0;
      gate = p > mc_uniform(seed, T, 4);
    }
    bool accept = false;
    double de = 0.0;
    if (gate) {
      for (int k = lane; k < na; k += 64) de += esite[t * cap + k] - e_cur[vatom[t * cap + k]];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) de += __shfl_xor(de, o, 64);
      de += de2;
      double p;
      if (da) {
        const double d = de - dsur;
        p = exp(-d / kT);
      } else {
        p = exp(-de / kT);
      }
      p = p < 1.0 ? p : 1.0;
      accept = p > mc_uniform(seed, T, da ? 5 : 4);
    }
    if (lane == 0) {
      if (!gate) {
        ctl[MC_NREJ1] += 1;
      } else {
        ctl[MC_NGAP] += 1;
        if (da) {
          const double d = de - dsur;
          dctl[MC_EMA] = 0.98 * dctl[MC_EMA] + 0.02 * d * d;
          if (dctl[MC_EMA] > dctl[MC_MAXMS]) ctl[MC_FALLBACK] = 1;
        }
      }
      if (accept) {
        dctl[MC_ETOT] += de;
        ctl[MC_NACC] += 1;
        const double* pn = md + t * MB_ND + MB_NEW;
        const int a = n_acc;
        acc_atom[a] = m;
        for (int k = 0; k < 3; k++) {
          acc_old[3 * a + k] = po[k];
          acc_new[3 * a + k] = pn[k];
          pos[3 * m + k] = pn[k];
        }
        n_acc = a + 1;
      }
      if (trace) trace[T + 1] = dctl[MC_ETOT];
      ctl[MC_STEP] = T + 1;
      take = accept ? 1 : 0;
    }
    __syncthreads();
    if (take) {
      for (int k = lane; k < na; k += 64) e_cur[vatom[t * cap + k]] = esite[t * cap + k];
      if (da)
        for (int k = lane; k < na; k += 64) esur_cur[vatom[t * cap + k]] = esur[t * cap + k];
    }
    __syncthreads();
  }
}

hipError_t launch_mcb_propose(int n, int nb, double* pos, int* species, const double* pref,
                              unsigned long long seed, double move_max, double lim2,
                              int* ctl, int* mi, double* md, hipStream_t st) {
  if (nb > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mcb_propose_kernel, dim3(1), dim3(64), 0, st, n, nb, pos, species, pref, seed,
                     move_max, lim2, ctl, mi, md);
  return hipGetLastError();
}

hipError_t launch_mcb_affected(int n, int nb, int cap, int vb, double* pos, int* species,
                               int* count, int* list, int maxnb, const Cell& cell, double rsoap,
                               int central, int* site_atom, int* vatom, int* mi, double* md,
                               const TwoBodyDev* twob, int n2b, int exact, hipStream_t st) {
  if (cap > 1024 || nb > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mcb_affected_kernel, dim3(nb), dim3(256), 0, st, n, nb, cap, vb, pos, species,
                     count, list, maxnb, cell, rsoap, central, site_atom, vatom, mi, md, twob, n2b,
                     exact);
  return hipGetLastError();
}

hipError_t launch_mcb_walk(int nb, int cap, double* pos, const Cell& cell, double rdep,
                           const int* vatom, const double* esite, double* e_cur,
                           const double* esur, double* esur_cur, unsigned long long seed,
                           double kT, int* ctl, double* dctl, const int* mi, const double* md,
                           double* trace, hipStream_t st) {
  if (nb > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mcb_walk_kernel, dim3(1), dim3(64), 0, st, nb, cap, pos, cell, rdep, vatom,
                     esite, e_cur, esur, esur_cur, seed, kT, ctl, dctl, mi, md, trace);
  return hipGetLastError();
}

hipError_t launch_mc_propose(int n, double* pos, const double* pref, unsigned long long seed,
                             double move_max, double lim2, int* ctl, double* dctl,
                             hipStream_t st) {
  hipLaunchKernelGGL(mc_propose_kernel, dim3(1), dim3(64), 0, st, n, pos, pref, seed, move_max,
                     lim2, ctl, dctl);
  return hipGetLastError();
}

hipError_t launch_mc_affected(int* ctl, double* dctl, const double* pos, const Cell& cell,
                              double rsoap, const int* count, const int* list, int maxnb,
                              const int* species, int central, int* aff, int cap, int* n_aff,
                              const TwoBodyDev* twob, int n2b, int exact, int with_2b,
                              int open_gate, hipStream_t st) {
  if (cap > 1024) return hipErrorInvalidValue;
  if (with_2b)
    hipLaunchKernelGGL(mc_affected_kernel<true>, dim3(1), dim3(256), 0, st, ctl, dctl, pos, cell,
                       rsoap, count, list, maxnb, species, central, aff, cap, n_aff, twob, n2b,
                       exact, open_gate);
  else
    hipLaunchKernelGGL(mc_affected_kernel<false>, dim3(1), dim3(256), 0, st, ctl, dctl, pos, cell,
                       rsoap, count, list, maxnb, species, central, aff, cap, n_aff, twob, n2b,
                       exact, open_gate);
  return hipGetLastError();
}

hipError_t launch_mc_stage1(double* pos, const int* species, const Cell& cell, int maxnb,
                            const int* count, const int* list, const int* aff, const int* n_aff,
                            const double* esur, const double* esur_cur, const TwoBodyDev* twob,
                            int n2b, int exact, int have_2b, unsigned long long seed, double kT,
                            int* ctl, double* dctl, double* trace, hipStream_t st) {
  hipLaunchKernelGGL(mc_stage1_kernel, dim3(1), dim3(256), 0, st, pos, species, cell, maxnb, count,
                     list, aff, n_aff, esur, esur_cur, twob, n2b, exact, have_2b, seed, kT, ctl,
                     dctl, trace);
  return hipGetLastError();
}

hipError_t launch_mc_stage2(double* pos, const int* aff, const int* n_aff, const double* esite,
                            const float* parts, int nparts, int part_ld, double d2, double* e_cur,
                            const double* esur, double* esur_cur, unsigned long long seed,
                            double kT, int* ctl, double* dctl, double* trace, hipStream_t st) {
  hipLaunchKernelGGL(mc_stage2_kernel, dim3(1), dim3(64), 0, st, pos, aff, n_aff, esite, parts,
                     nparts, part_ld, d2, e_cur, esur, esur_cur, seed, kT, ctl, dctl, trace);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mc_collect(const int* ctl, const int* n_aff, const T* q, int q_ld, int D,
                             const double* esite, float* sq, double* sy, int cap_s, int* cursor,
                             hipStream_t st) {
  hipLaunchKernelGGL((mc_collect_kernel<T>), dim3(1), dim3(256), 0, st, ctl, n_aff, q, q_ld, D,
                     esite, sq, sy, cap_s, cursor);
  return hipGetLastError();
}
template hipError_t launch_mc_collect<float>(const int*, const int*, const float*, int, int, const double*, float*, double*, int, int*, hipStream_t);
template hipError_t launch_mc_collect<double>(const int*, const int*, const double*, int, int, const double*, float*, double*, int, int*, hipStream_t);

hipError_t launch_mc_init(int n, const int* site_of_atom, const double* esite, double* e_cur,
                          hipStream_t st) {
  hipLaunchKernelGGL(mc_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, site_of_atom,
                     esite, e_cur);
  return hipGetLastError();
}

}  // namespace tgx
