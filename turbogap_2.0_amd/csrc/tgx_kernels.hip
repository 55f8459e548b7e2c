// Kernels of the GAP/SOAP path except the kernel-step GEMMs: neighbour list,
// radial table, SOAP expansion forward and adjoint, force gather, 2-body
// term, energy assembly and deterministic reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {


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

// Wrapped FP32 position of atom i (x - L floor(x/L) in FP64, then rounded)
// with its id in w.
__device__ __forceinline__ float4 wrapped_pos4(const double* pos, int i, const Cell& c) {
  const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  return make_float4((float)(x - c.La * floor(x * c.iLa)), (float)(y - c.Lb * floor(y * c.iLb)),
                     (float)(z - c.Lc * floor(z * c.iLc)), __int_as_float(i));
}

// Also clears the counts the scan consumed: every binning leaves cell_count
// all zero, ready for the next one.  With cpos != nullptr the wrapped FP32
// positions are stored in cell order too.
__global__ void scatter_cells_kernel(int n, const int* cell_of, const int* start,
                                     int* fill, int* cell_atoms, int* cell_count,
                                     const double* pos, Cell cl, float4* cpos,
                                     const int* guard) {
  GUARD_RETURN();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of[i];
  int slot = atomicAdd(fill + c, 1);
  cell_atoms[start[c] + slot] = i;
  if (cpos) cpos[start[c] + slot] = wrapped_pos4(pos, i, cl);
  cell_count[c] = 0;
}

// Candidates of atom i within rcut, unsorted, packed (img<<26 | j).  Four
// lanes per atom split the 27 neighbouring cells; a rounding minimum image
// screens candidates and only those within rcut (plus a margin far above
// rounding) are decided with the reference's arithmetic (mic), so the pair
// set is exactly the reference's.  count[] must be zero on entry.
#define TGX_NB_LANES 4
__global__ void find_neighbors_kernel(const double* pos, int n, Cell c,
                                      double rcut, const int* cell_start,
                                      const int* cell_atoms, int* count,
                                      int maxnb, int* list, int* flag,
                                      const int* guard) {
  GUARD_RETURN();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / TGX_NB_LANES, part = t % TGX_NB_LANES;
  if (i >= n) return;
  int* out = list + (size_t)i * maxnb;
  const double* pi = pos + 3 * i;
  const double screen2 = (rcut + 1e-6) * (rcut + 1e-6);
  if (!c.small) {
    int cx, cy, cz;
    cell_coords(pi, c, &cx, &cy, &cz);
    for (int q = part; q < 27; q += TGX_NB_LANES) {
      const int i3 = cx - 1 + q % 3, j3 = cy - 1 + (q / 3) % 3, k3 = cz - 1 + q / 9;
      const int cell = imodd(i3, c.mx) + imodd(j3, c.my) * c.mx + imodd(k3, c.mz) * c.mx * c.my;
      // skip a cell already visited by a smaller q (grids of 1 or 2 cells)
      bool dup = false;
      for (int q2 = 0; q2 < q; q2++) {
        const int a3 = cx - 1 + q2 % 3, b3 = cy - 1 + (q2 / 3) % 3, c3 = cz - 1 + q2 / 9;
        dup |= (imodd(a3, c.mx) + imodd(b3, c.my) * c.mx + imodd(c3, c.mz) * c.mx * c.my) == cell;
      }
      if (dup) continue;
      for (int a = cell_start[cell]; a < cell_start[cell + 1]; a++) {
        const int k = cell_atoms[a];
        if (k == i) continue;
        const double* pk = pos + 3 * k;
        double dx = pk[0] - pi[0], dy = pk[1] - pi[1], dz = pk[2] - pi[2];
        dx -= c.La * rint(dx * c.iLa);
        dy -= c.Lb * rint(dy * c.iLb);
        dz -= c.Lc * rint(dz * c.iLc);
        if (dx * dx + dy * dy + dz * dz > screen2) continue;
        double dv[3];
        const double d = mic(pi, pk, c.La, c.Lb, c.Lc, dv);
        if (d < rcut) {
          const int slot = atomicAdd(count + i, 1);
          if (slot < maxnb) out[slot] = k;
          else atomicMax(flag, slot + 1);
        }
      }
    }
  } else {
    for (int img = 0; img < c.nrep; img++)
      for (int k = part; k < n; k += TGX_NB_LANES) {
        if (img == 0 && k == i) continue;
        double pk[3], dv[3];
        image_pos(pos, k, img, c, pk);
        const double d = mic(pi, pk, c.La, c.Lb, c.Lc, dv);
        if (d < rcut) {
          const int slot = atomicAdd(count + i, 1);
          if (slot < maxnb) out[slot] = (img << TGX_IMG_SHIFT) | k;
          else atomicMax(flag, slot + 1);
        }
      }
  }
}

// Ascending bitonic sort of 64 values held one per lane.
__device__ __forceinline__ int wave_sort64(int v, int lane) {
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int o = __shfl_xor(v, j, 64);
      const bool up = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      v = (lower == up) ? min(v, o) : max(v, o);
    }
  }
  return v;
}

// Ascending bitonic sort of 128 values, v0 at index lane, v1 at lane + 64.
__device__ __forceinline__ void wave_sort128(int& v0, int& v1, int lane) {
#pragma unroll
  for (int k = 2; k <= 128; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j == 64) {
        // partners sit in the same lane; k == 128 here, so ascending
        const int lo = min(v0, v1), hi = max(v0, v1);
        v0 = lo;
        v1 = hi;
      } else {
        const int o0 = __shfl_xor(v0, j, 64), o1 = __shfl_xor(v1, j, 64);
        const bool lower = (lane & j) == 0;
        const bool up0 = (lane & k) == 0, up1 = ((lane + 64) & k) == 0;
        v0 = (lower == up0) ? min(v0, o0) : max(v0, o0);
        v1 = (lower == up1) ? min(v1, o1) : max(v1, o1);
      }
    }
  }
}

// The linked-cell list build for grids of at least 3 cells per direction
// in one pass, one wave per atom: the atoms of the 27 neighbouring cells
// are numbered across the wave and screened 64 at a time with a rounding
// minimum image on their cell-ordered FP32 copies (coalesced reads); the
// survivors (within rcut plus a margin far above rounding) queue in LDS and are decided 64 at a time with the reference's
// arithmetic (mic), so the pair set is exactly the reference's.  Hits are
// compacted with a ballot into LDS, sorted ascending (the reference
// supercell id order) and written with their count; count needs no clearing
// and no atomics are used.  A list longer than maxnb raises flag[0] to its
// length.  With pref != nullptr the atom's position is also copied there
// (the reference point of the skin test).  Every wave of the block reaches
// both barriers.
#define TGX_NB_CAPL 128
__device__ __forceinline__ void nb_decide(const double* pos, const double* pi, const Cell& c,
                                          double rcut, const int* q, int m, int lane,
                                          int* buf, int* out, int maxnb, int* found) {
  bool hit = false;
  int k = 0;
  if (lane < m) {
    k = q[lane];
    double dv[3];
    hit = mic(pi, pos + 3 * k, c.La, c.Lb, c.Lc, dv) < rcut;
  }
  const unsigned long long b = __ballot(hit);
  if (hit) {
    const int slot = *found + __popcll(b & ((1ull << lane) - 1ull));
    if (slot < TGX_NB_CAPL) buf[slot] = k;
    else if (slot < maxnb) out[slot] = k;
  }
  *found += __popcll(b);
}

// Per-wave LDS of the one-pass build.
struct NbWave {
  int pre[28], base[28];  // first candidate index and first slot of the 27 cells
  int buf[TGX_NB_CAPL];   // hits
  int q[128];             // screened candidates awaiting the exact decision
};

__device__ void nb_finish(NbWave& L, int i, int found, int maxnb, int* count, int* flag,
                          int* out, int lane);

// One atom's list by one wave (all 64 lanes; LDS of this wave only).
// own = the atom's linked cell as binned, cell_start the cell offsets.
__device__ void nb_find_one(const double* pos, int i, int own, const Cell& c, double rcut,
                            const int* cell_start, const float4* cpos, int* count, int maxnb,
                            int* list, int* flag, double* pref, NbWave& L, int lane) {
  const double* pi = pos + 3 * i;
  const int cx = own % c.mx, cy = (own / c.mx) % c.my, cz = own / (c.mx * c.my);
  // lanes 0..26: cell q = lane, its atom count and first slot
  int cnt_q = 0, base_q = 0;
  if (lane < 27) {
    const int q = lane;
    const int i3 = cx - 1 + q % 3, j3 = cy - 1 + (q / 3) % 3, k3 = cz - 1 + q / 9;
    const int cell = imodd(i3, c.mx) + imodd(j3, c.my) * c.mx + imodd(k3, c.mz) * c.mx * c.my;
    base_q = cell_start[cell];
    cnt_q = cell_start[cell + 1] - base_q;
  }
  int incl = cnt_q;
#pragma unroll
  for (int d = 1; d < 32; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  const int total = __shfl(incl, 26, 64);
  if (lane < 27) {
    L.pre[lane] = incl - cnt_q;
    L.base[lane] = base_q;
  }
  if (pref && lane < 3) pref[3 * i + lane] = pi[lane];
  wave_sync();
  // FP32 screen on wrapped coordinates: its error (a few 1e-5 A for boxes
  // of 100 A) is far inside a 1e-3 A margin, and images near L/2 are far
  // beyond rcut since L >= 3 rcut.  Candidates within rcut - 1e-3 are hits
  // for certain; those within 1e-3 of rcut take the exact FP64 test.
  const float rs = (float)rcut + 1e-3f, screen2 = rs * rs;
  const float ri = (float)rcut - 1e-3f, inner2 = ri > 0.f ? ri * ri : 0.f;
  const float4 pw = wrapped_pos4(pos, i, c);
  const float La = (float)c.La, Lb = (float)c.Lb, Lc = (float)c.Lc;
  const float iLa = (float)c.iLa, iLb = (float)c.iLb, iLc = (float)c.iLc;
  int found = 0, nq = 0;
  int* out = list + (size_t)i * maxnb;
  for (int idx0 = 0; idx0 < total; idx0 += 64) {
    const int idx = idx0 + lane;
    bool near = false, sure = false;
    int k = 0;
    if (idx < total) {
      // largest q with pre[q] <= idx
      int lo = 0, hi = 26;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (L.pre[mid] <= idx) lo = mid;
        else hi = mid - 1;
      }
      const float4 pk = cpos[L.base[lo] + idx - L.pre[lo]];
      k = __float_as_int(pk.w);
      if (k != i) {
        float dx = pk.x - pw.x, dy = pk.y - pw.y, dz = pk.z - pw.z;
        dx -= La * rintf(dx * iLa);
        dy -= Lb * rintf(dy * iLb);
        dz -= Lc * rintf(dz * iLc);
        const float d2 = dx * dx + dy * dy + dz * dz;
        sure = d2 <= inner2;
        near = !sure && d2 <= screen2;
      }
    }
    // certain hits straight to the list (its final sort fixes the order);
    // the band of +-1e-3 A around rcut to the exact FP64 test
    const unsigned long long ms = __ballot(sure);
    if (sure) {
      const int slot = found + __popcll(ms & ((1ull << lane) - 1ull));
      if (slot < TGX_NB_CAPL) L.buf[slot] = k;
      else if (slot < maxnb) out[slot] = k;
    }
    found += __popcll(ms);
    const unsigned long long m = __ballot(near);
    if (near) L.q[nq + __popcll(m & ((1ull << lane) - 1ull))] = k;  // nq < 64 here
    nq += __popcll(m);
    if (nq >= 64) {
      wave_sync();
      nb_decide(pos, pi, c, rcut, L.q, 64, lane, L.buf, out, maxnb, &found);
      const int rest = nq - 64 > lane ? L.q[64 + lane] : 0;
      wave_sync();
      if (lane < nq - 64) L.q[lane] = rest;
      nq -= 64;
    }
  }
  wave_sync();
  if (nq > 0) nb_decide(pos, pi, c, rcut, L.q, nq, lane, L.buf, out, maxnb, &found);
  nb_finish(L, i, found, maxnb, count, flag, out, lane);
}

// The list's count, the overflow flag and the sorted entries (the first
// TGX_NB_CAPL hits in L.buf, any further ones already in out).
__device__ void nb_finish(NbWave& L, int i, int found, int maxnb, int* count, int* flag,
                          int* out, int lane) {
  wave_sync();
  if (lane == 0) {
    count[i] = found;
    if (found > maxnb) atomicMax(flag, found);
  }
  const int keep = min(found, maxnb);
  if (found <= 64) {
    int v = lane < found ? L.buf[lane] : 0x7fffffff;
    v = wave_sort64(v, lane);
    if (lane < keep) out[lane] = v;
  } else if (found <= TGX_NB_CAPL) {
    int v0 = L.buf[lane];
    int v1 = lane + 64 < found ? L.buf[lane + 64] : 0x7fffffff;
    wave_sort128(v0, v1, lane);
    if (lane < keep) out[lane] = v0;
    if (lane + 64 < keep) out[lane + 64] = v1;
  } else {
    for (int s = lane; s < TGX_NB_CAPL && s < keep; s += 64) out[s] = L.buf[s];
    __threadfence_block();
    if (lane == 0) {
      for (int p = 1; p < keep; p++) {
        const int v = out[p];
        int q = p - 1;
        while (q >= 0 && out[q] > v) {
          out[q + 1] = out[q];
          q--;
        }
        out[q + 1] = v;
      }
    }
  }
}

__global__ __launch_bounds__(256) void find_sorted_kernel(
    const double* pos, int n, Cell c, double rcut, const int* cell_of, const int* cell_start,
    const float4* cpos, int* count, int maxnb, int* list, int* flag, double* pref,
    const int* guard) {
  GUARD_RETURN();
  __shared__ NbWave L[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 4 + w;
  if (i >= n) return;
  nb_find_one(pos, i, cell_of[i], c, rcut, cell_start, cpos, count, maxnb, list, flag, pref,
              L[w], lane);
}

// Batched structures, one wave per atom: the candidates are the atoms of
// the (deduplicated) 27 neighbouring bins of the atom's structure, or for a
// structure smaller than 2 rcut every atom of every supercell image; the
// exact FP64 test (mic) decides, hits are compacted with a ballot and
// sorted as in the single-structure build.
__global__ __launch_bounds__(256) void find_sorted_many_kernel(
    const double* pos, int n, const Cell* cells, const int* sid, const int* boff,
    const int* aoff, double rcut, const int* cell_of, const int* cell_start,
    const int* cell_atoms, int* count, int maxnb, int* list, int* flag) {
  __shared__ NbWave Ls[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 4 + w;
  if (i >= n) return;
  NbWave& L = Ls[w];
  const int s = sid[i];
  const Cell c = cells[s];
  const double* pi = pos + 3 * i;
  int* out = list + (size_t)i * maxnb;
  int found = 0;
  auto append = [&](bool hit, int e) {
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const int slot = found + __popcll(m & ((1ull << lane) - 1ull));
      if (slot < TGX_NB_CAPL) L.buf[slot] = e;
      else if (slot < maxnb) out[slot] = e;
    }
    found += __popcll(m);
  };
  if (c.small) {
    const int k0 = aoff[s], ns = aoff[s + 1] - k0, total = c.nrep * ns;
    for (int idx0 = 0; idx0 < total; idx0 += 64) {
      const int idx = idx0 + lane;
      bool hit = false;
      int e = 0;
      if (idx < total) {
        const int img = idx / ns, k = k0 + (idx - img * ns);
        if (!(img == 0 && k == i)) {
          double pk[3], dv[3];
          image_pos(pos, k, img, c, pk);
          hit = mic(pi, pk, c.La, c.Lb, c.Lc, dv) < rcut;
          e = (img << TGX_IMG_SHIFT) | k;
        }
      }
      append(hit, e);
    }
  } else {
    const int own = cell_of[i] - boff[s];
    const int cx = own % c.mx, cy = (own / c.mx) % c.my, cz = own / (c.mx * c.my);
    int cell = -1, cnt_q = 0, base_q = 0;
    if (lane < 27) {
      const int q = lane;
      const int i3 = cx - 1 + q % 3, j3 = cy - 1 + (q / 3) % 3, k3 = cz - 1 + q / 9;
      cell = imodd(i3, c.mx) + imodd(j3, c.my) * c.mx + imodd(k3, c.mz) * c.mx * c.my;
    }
    // grids of 1 or 2 bins per direction repeat bins: keep the first
    bool dup = false;
    for (int q2 = 0; q2 < 26; q2++) {
      const int other = __shfl(cell, q2, 64);
      dup |= (lane < 27 && q2 < lane && other == cell);
    }
    if (lane < 27 && !dup) {
      base_q = cell_start[boff[s] + cell];
      cnt_q = cell_start[boff[s] + cell + 1] - base_q;
    }
    int incl = cnt_q;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
      const int o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, 26, 64);
    if (lane < 27) {
      L.pre[lane] = incl - cnt_q;
      L.base[lane] = base_q;
    }
    wave_sync();
    const double screen2 = (rcut + 1e-6) * (rcut + 1e-6);
    for (int idx0 = 0; idx0 < total; idx0 += 64) {
      const int idx = idx0 + lane;
      bool hit = false;
      int k = 0;
      if (idx < total) {
        int lo = 0, hi = 26;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (L.pre[mid] <= idx) lo = mid;
          else hi = mid - 1;
        }
        k = cell_atoms[L.base[lo] + idx - L.pre[lo]];
        if (k != i) {
          const double* pk = pos + 3 * k;
          double dx = pk[0] - pi[0], dy = pk[1] - pi[1], dz = pk[2] - pi[2];
          dx -= c.La * rint(dx * c.iLa);
          dy -= c.Lb * rint(dy * c.iLb);
          dz -= c.Lc * rint(dz * c.iLc);
          if (dx * dx + dy * dy + dz * dz <= screen2) {
            double dv[3];
            hit = mic(pi, pk, c.La, c.Lb, c.Lc, dv) < rcut;
          }
        }
      }
      append(hit, k);
    }
  }
  nb_finish(L, i, found, maxnb, count, flag, out, lane);
}

// The whole build for small systems in one workgroup of 1024 threads:
// binning, cell offsets and the cell-ordered copies in LDS and memory,
// then the one-pass list of every atom, one wave per atom in turn.  The
// same lists as the multi-kernel build; one launch instead of four (used
// for builds captured in the Monte Carlo trial graph).
#define TGX_NB_SMALL_N 4096
#define TGX_NB_SMALL_CELLS 4096
__global__ __launch_bounds__(1024) void nb_build_small_kernel(
    const double* pos, int n, Cell c, double rcut, int* cell_of, int* cell_atoms, float4* cpos,
    int* count, int maxnb, int* list, int* flag, double* pref, const int* guard) {
  GUARD_RETURN();
  __shared__ int s_cnt[TGX_NB_SMALL_CELLS];
  __shared__ int s_start[TGX_NB_SMALL_CELLS + 1];
  __shared__ int s_part[1024];
  __shared__ NbWave L[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int ncell = c.mx * c.my * c.mz;
  for (int k = t; k < ncell; k += 1024) s_cnt[k] = 0;
  __syncthreads();
  for (int i = t; i < n; i += 1024) {
    int cx, cy, cz;
    cell_coords(pos + 3 * i, c, &cx, &cy, &cz);
    const int cell = imodd(cx, c.mx) + imodd(cy, c.my) * c.mx + imodd(cz, c.mz) * c.mx * c.my;
    cell_of[i] = cell;
    atomicAdd(s_cnt + cell, 1);
  }
  __syncthreads();
  // exclusive scan: chunks of per cells per thread, then the chunk sums
  const int per = (ncell + 1023) / 1024, b0 = t * per, b1 = min(ncell, b0 + per);
  int sum = 0;
  for (int k = b0; k < b1; k++) sum += s_cnt[k];
  s_part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? s_part[t - off] : 0;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  int run = s_part[t] - sum;
  for (int k = b0; k < b1; k++) {
    s_start[k] = run;
    run += s_cnt[k];
    s_cnt[k] = 0;  // becomes the fill counter
  }
  if (t == 1023) s_start[ncell] = s_part[1023];
  __syncthreads();
  for (int i = t; i < n; i += 1024) {
    const int cell = cell_of[i];
    const int slot = s_start[cell] + atomicAdd(s_cnt + cell, 1);
    cell_atoms[slot] = i;
    cpos[slot] = wrapped_pos4(pos, i, c);
  }
  __syncthreads();  // also makes cell_of and cpos visible across the workgroup
  for (int i = w; i < n; i += 16)
    nb_find_one(pos, i, cell_of[i], c, rcut, s_start, cpos, count, maxnb, list, flag, pref, L[w],
                lane);
}

// Sort each atom's list ascending (== ascending reference supercell id):
// one wave per atom, bitonic network over 64 lanes; lists longer than 64
// entries fall back to an insertion sort by lane 0.
__global__ __launch_bounds__(256) void sort_lists_kernel(int n, int maxnb, const int* count,
                                                         int* list, const int* guard) {
  GUARD_RETURN();
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int cnt = min(count[i], maxnb);
  int* a = list + (size_t)i * maxnb;
  if (cnt <= 64) {
    int v = lane < cnt ? a[lane] : 0x7fffffff;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int o = __shfl_xor(v, j, 64);
        const bool up = (lane & k) == 0;
        const bool lower = (lane & j) == 0;
        v = (lower == up) ? min(v, o) : max(v, o);
      }
    }
    if (lane < cnt) a[lane] = v;
  } else if (lane == 0) {
    for (int p = 1; p < cnt; p++) {
      const int v = a[p];
      int q = p - 1;
      while (q >= 0 && a[q] > v) {
        a[q + 1] = a[q];
        q--;
      }
      a[q + 1] = v;
    }
  }
}

// rev[i][slot] = flat index of the entry (neg img, i) in j's list; one
// wave per atom, four atoms per workgroup.
template <typename G>
__global__ __launch_bounds__(256) void reverse_kernel(int n, int maxnb, G geo, const int* count,
                                                      const int* list, int* rev, int* flag,
                                                      const int* guard) {
  GUARD_RETURN();
  // 32 lanes per atom, eight atoms per workgroup
  const int i = blockIdx.x * 8 + (threadIdx.x >> 5);
  if (i >= n) return;
  const Cell c = geo.of(i);
  int cnt = min(count[i], maxnb);
  for (int s = threadIdx.x & 31; s < cnt; s += 32) {
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

// ---- Cells of any shape: the reference's brute-force search -------------

// Supercell image img = (i2 ib + j2) ic + k2 of atom j, summed as
// set_supercell does with the unit-cell vectors: ((p + i2 a) + j2 b) + k2 c.
__device__ __forceinline__ void gen_image_pos(const double* pos, int j, int img,
                                              const GenCell& g, double* out) {
#pragma clang fp contract(off)
  const double* p = pos + 3 * j;
  const int k2 = img % g.ic, j2 = (img / g.ic) % g.ib, i2 = img / (g.ic * g.ib);
#pragma unroll
  for (int k = 0; k < 3; k++)
    out[k] = ((p[k] + (double)i2 * g.u[k]) + (double)j2 * g.u[3 + k]) + (double)k2 * g.u[6 + k];
}

__device__ __forceinline__ double dot3d(const double* x, const double* y) {
#pragma clang fp contract(off)
  return x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
}

// get_distance's general branch (neighbors_interface.f90:582-629) against
// the supercell vectors: the real shift from the normal equations, its
// nearest integer (half away from zero) and the best of the 27 integer
// shifts around it, the first minimum in i, j, k order.
__device__ __forceinline__ double gen_mic(const double* pi, const double* pj, const GenCell& g,
                                          double* dv) {
#pragma clang fp contract(off)
  double dd[3], v[3];
  for (int k = 0; k < 3; k++) dd[k] = pj[k] - pi[k];
  v[0] = -1.0 * dot3d(dd, g.s);
  v[1] = -1.0 * dot3d(dd, g.s + 3);
  v[2] = -1.0 * dot3d(dd, g.s + 6);
  int nn[3];
  for (int i = 0; i < 3; i++) {
    const double x = (g.mi[3 * i] * v[0] + g.mi[3 * i + 1] * v[1]) + g.mi[3 * i + 2] * v[2];
    nn[i] = (int)round(x);
  }
  double res_opt = 1e10, b0 = dd[0], b1 = dd[1], b2 = dd[2];
  for (int i = nn[0] - 1; i <= nn[0] + 1; i++)
    for (int j = nn[1] - 1; j <= nn[1] + 1; j++)
      for (int k = nn[2] - 1; k <= nn[2] + 1; k++) {
        double t[3];
        for (int q = 0; q < 3; q++)
          t[q] = ((dd[q] + (double)i * g.s[q]) + (double)j * g.s[3 + q]) + (double)k * g.s[6 + q];
        const double res = dot3d(t, t);
        if (res < res_opt) {
          res_opt = res;
          b0 = t[0];
          b1 = t[1];
          b2 = t[2];
        }
      }
  dv[0] = b0;
  dv[1] = b1;
  dv[2] = b2;
  return sqrt(dot3d(dv, dv));
}

// One wave per atom i of the unit cell: every supercell atom, 64 at a time
// in ascending supercell id, tested with the general minimum image; hits
// are compacted in that order (so rows are sorted, packed as atom | image
// << TGX_IMG_SHIFT).  A row longer than maxnb raises flag[0] to its length.
__global__ __launch_bounds__(64) void nb_general_kernel(const double* pos, int n, GenCell g,
                                                        double rcut, int maxnb, int* count,
                                                        int* list, int* flag) {
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= n) return;
  const int nsc = n * g.nrep;
  double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  int found = 0;
  int* row = list + (size_t)i * maxnb;
  for (int base = 0; base < nsc; base += 64) {
    const int J = base + lane;
    bool hit = false;
    if (J < nsc && J != i) {
      const int img = J / n, j = J - img * n;
      double pj[3], dv[3];
      gen_image_pos(pos, j, img, g, pj);
      hit = gen_mic(pi, pj, g, dv) < rcut;
    }
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const int slot = found + __popcll(m & ((1ull << lane) - 1ull));
      const int img = J / n;
      if (slot < maxnb) row[slot] = (J - img * n) | (img << TGX_IMG_SHIFT);
    }
    found += __popcll(m);
  }
  if (lane == 0) {
    count[i] = found;
    if (found > maxnb) atomicMax(flag, found);
  }
}

// CSR export of the general-cell rows: self first, then the entries with
// their supercell ids, rjs, xyz (the second get_distance pass,
// neighbors_interface.f90:470-500) and the angles with the pole guard.
__global__ void csr_out_general_kernel(const double* pos, int n, GenCell g, int maxnb,
                                       const int* count, const int* list, const long long* off,
                                       int* out_list, double* rjs, double* xyz, double* th,
                                       double* ph) {
#pragma clang fp contract(off)
  const int i = blockIdx.x;
  const long long base = off[i];
  const int cnt = count[i];
  for (int s = threadIdx.x; s <= cnt; s += blockDim.x) {
    const long long k = base + s;
    if (s == 0) {
      out_list[k] = i + 1;
      rjs[k] = 0.0;
      xyz[3 * k] = xyz[3 * k + 1] = xyz[3 * k + 2] = 0.0;
      th[k] = 0.0;
      ph[k] = 0.0;
      continue;
    }
    const int e = list[(size_t)i * maxnb + s - 1];
    const int j = e & TGX_ATOM_MASK, img = e >> TGX_IMG_SHIFT;
    double pj[3], dv[3];
    gen_image_pos(pos, j, img, g, pj);
    const double d = gen_mic(pos + 3 * i, pj, g, dv);
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

hipError_t launch_neighbors_general(const double* pos, int n, const GenCell& g, double rcut,
                                    int maxnb, int* count, int* list, int* flag,
                                    hipStream_t st) {
  hipLaunchKernelGGL(nb_general_kernel, dim3(n), dim3(64), 0, st, pos, n, g, rcut, maxnb, count,
                     list, flag);
  return hipGetLastError();
}

hipError_t launch_csr_out_general(const double* pos, int n, const GenCell& g, int maxnb,
                                  const int* count, const int* list, const long long* off,
                                  int* out_list, double* rjs, double* xyz, double* th,
                                  double* ph, hipStream_t st) {
  hipLaunchKernelGGL(csr_out_general_kernel, dim3(n), dim3(64), 0, st, pos, n, g, maxnb, count,
                     list, off, out_list, rjs, xyz, th, ph);
  return hipGetLastError();
}

// ---- Cells of any shape for energies and forces: image atoms -------------

// Lattice translation of image img (img 0 = t 0): t1 a + t2 b + t3 c.  The
// translation of -t is the exact negation (every operation commutes with it).
__device__ __forceinline__ void img_translation(const ImgCell& g, int img, double* t) {
#pragma clang fp contract(off)
  const int n2 = 2 * g.m[1] + 1, n3 = 2 * g.m[2] + 1;
  int idx = img + (g.nt - 1) / 2;
  if (idx >= g.nt) idx -= g.nt;
  const double t3 = (double)(idx % n3 - g.m[2]);
  const double t2 = (double)((idx / n3) % n2 - g.m[1]);
  const double t1 = (double)(idx / (n2 * n3) - g.m[0]);
#pragma unroll
  for (int k = 0; k < 3; k++) t[k] = (t1 * g.u[k] + t2 * g.u[3 + k]) + t3 * g.u[6 + k];
}

// Image atoms J = img * n + j: atom j wrapped into the cell (fractional
// coordinates floored), plus translation img; species copied.
__global__ void image_atoms_kernel(const double* pos0, const int* sp0, int n, ImgCell g,
                                   double* pos, int* sp) {
#pragma clang fp contract(off)
  const long long total = (long long)n * g.nt;
  for (long long J = blockIdx.x * (long long)blockDim.x + threadIdx.x; J < total;
       J += (long long)gridDim.x * blockDim.x) {
    const int img = (int)(J / n), j = (int)(J - (long long)img * n);
    const double* p = pos0 + 3 * j;
    double f[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      f[k] = (p[0] * g.ui[k] + p[1] * g.ui[3 + k]) + p[2] * g.ui[6 + k];
      f[k] = f[k] - floor(f[k]);
    }
    double t[3] = {0.0, 0.0, 0.0};
    if (img) img_translation(g, img, t);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double w = (f[0] * g.u[k] + f[1] * g.u[3 + k]) + f[2] * g.u[6 + k];
      pos[3 * J + k] = img ? w + t[k] : w;
    }
    sp[J] = sp0[j];
  }
}

// Rows of the image-atom list: one wave per atom i, images in ascending J
// (translation-major), pair vector (w_j - w_i) + t so that the pair (j, -t)
// of atom j sees exactly the negated vector and the list is symmetric.
// A row longer than maxnb raises flag[0] to its length.
__global__ __launch_bounds__(64) void nb_image_kernel(const double* pos, int n, ImgCell g,
                                                      double rcut2, int maxnb, int* count,
                                                      int* list, int* flag) {
#pragma clang fp contract(off)
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= n) return;
  const double pi0 = pos[3 * i], pi1 = pos[3 * i + 1], pi2 = pos[3 * i + 2];
  int found = 0;
  int* row = list + (size_t)i * maxnb;
  for (int img = 0; img < g.nt; img++) {
    double t[3] = {0.0, 0.0, 0.0};
    if (img) img_translation(g, img, t);
    for (int jb = 0; jb < n; jb += 64) {
      const int j = jb + lane;
      const int J = img * n + j;
      bool hit = false;
      if (j < n && J != i) {
        const double d0 = (pos[3 * j] - pi0) + t[0];
        const double d1 = (pos[3 * j + 1] - pi1) + t[1];
        const double d2 = (pos[3 * j + 2] - pi2) + t[2];
        hit = (d0 * d0 + d1 * d1) + d2 * d2 < rcut2;
      }
      const unsigned long long m = __ballot(hit);
      if (hit) {
        const int slot = found + __popcll(m & ((1ull << lane) - 1ull));
        if (slot < maxnb) row[slot] = J;
      }
      found += __popcll(m);
    }
  }
  if (lane == 0) {
    count[i] = found;
    if (found > maxnb) atomicMax(flag, found);
  }
}

// rev[i][slot] = flat index of the entry (image of -t, atom i) in row j;
// one wave per atom, four atoms per workgroup.  A missing counterpart sets
// flag[1] and points the slot at itself.
__global__ __launch_bounds__(256) void reverse_image_kernel(int n, int nt, int maxnb,
                                                            const int* count, const int* list,
                                                            int* rev, int* flag) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int cnt = min(count[i], maxnb);
  for (int s = threadIdx.x & 63; s < cnt; s += 64) {
    const int J = list[(size_t)i * maxnb + s];
    const int img = J / n, j = J - img * n;
    const int key = (img ? nt - img : 0) * n + i;
    const int* a = list + (size_t)j * maxnb;
    int lo = 0, hi = min(count[j], maxnb) - 1, found = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = a[mid];
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

hipError_t launch_image_atoms(const double* pos0, const int* sp0, int n, const ImgCell& g,
                              double* pos, int* sp, hipStream_t st) {
  const long long total = (long long)n * g.nt;
  const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(image_atoms_kernel, dim3(blocks), dim3(256), 0, st, pos0, sp0, n, g, pos, sp);
  return hipGetLastError();
}

hipError_t launch_neighbors_image(const double* pos, int n, const ImgCell& g, double rcut,
                                  int maxnb, int* count, int* list, int* rev, int* flag,
                                  hipStream_t st) {
  hipLaunchKernelGGL(nb_image_kernel, dim3(n), dim3(64), 0, st, pos, n, g, rcut * rcut, maxnb,
                     count, list, flag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !rev) return e;
  hipLaunchKernelGGL(reverse_image_kernel, dim3((n + 3) / 4), dim3(256), 0, st, n, g.nt, maxnb,
                     count, list, rev, flag);
  return hipGetLastError();
}

__global__ void copy_pos_kernel(int n, const double* pos, double* pref, const int* guard) {
  GUARD_RETURN();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 3 * n) pref[k] = pos[k];
}

// The whole list build; with guard != nullptr every kernel returns at once
// unless *guard is non-zero (rebuild requested on the device), so the
// sequence can sit in a captured graph.  pref (optional) receives a copy of
// the positions the list was built from.

hipError_t launch_neighbors(const double* pos, int n, const Cell& c, double rcut,
                            int* cell_of, int* cell_count, int* cell_start,
                            int* cell_fill, int* cell_atoms, int* count,
                            int maxnb, int* list, int* rev, int* flag,
                            hipStream_t st, const int* guard, double* pref) {
  const int B = 256;
  int nb = (n + B - 1) / B;
  // the one-pass build screens on FP32 wrapped positions with a 1e-3 A
  // margin either side of the cutoff: boxes up to 1,024 A keep the FP32
  // error below a third of it; larger ones take the FP64 build
  const bool fp32_ok = c.La < 1024.0 && c.Lb < 1024.0 && c.Lc < 1024.0;
  const bool fused = !c.small && c.mx >= 3 && c.my >= 3 && c.mz >= 3 && fp32_ok;
  // wrapped FP32 positions in cell order, after the ids (cell_atoms_bytes)
  float4* cpos = reinterpret_cast<float4*>(cell_atoms + ((n + 3) & ~3));
  // captured (guarded) builds of small systems: one launch instead of four,
  // since such builds are rare and their launches mostly exit at once; a
  // single workgroup is slower than the grid-wide build when it does run
  if (guard && fused && n <= TGX_NB_SMALL_N && c.mx * c.my * c.mz <= TGX_NB_SMALL_CELLS) {
    hipLaunchKernelGGL(nb_build_small_kernel, dim3(1), dim3(1024), 0, st, pos, n, c, rcut,
                       cell_of, cell_atoms, cpos, count, maxnb, list, flag, pref, guard);
    if (rev)
      hipLaunchKernelGGL(reverse_kernel<CellOne>, dim3((n + 7) / 8), dim3(256), 0, st, n, maxnb,
                         CellOne{c}, count, list, rev, flag, guard);
    return hipGetLastError();
  }
  if (!c.small) {
    int ncell = c.mx * c.my * c.mz;
    // a guarded (captured) build relies on the previous build of the same
    // grid having left cell_count zero
    if (!guard)
      hipLaunchKernelGGL(zero_ints_kernel, dim3((ncell + B - 1) / B), dim3(B), 0, st, cell_count,
                         ncell, nullptr);
    hipLaunchKernelGGL(bin_atoms_kernel, dim3(nb), dim3(B), 0, st, pos, n, c, cell_of, cell_count,
                       guard);
    hipLaunchKernelGGL(scan_cells_kernel, dim3(1), dim3(1024), 0, st, cell_count, ncell, cell_start,
                       cell_fill, guard);
    hipLaunchKernelGGL(scatter_cells_kernel, dim3(nb), dim3(B), 0, st, n, cell_of, cell_start,
                       cell_fill, cell_atoms, cell_count, pos, c, fused ? cpos : nullptr, guard);
  }
  if (fused) {
    hipLaunchKernelGGL(find_sorted_kernel, dim3((n + 3) / 4), dim3(256), 0, st, pos, n, c, rcut,
                       cell_of, cell_start, cpos, count, maxnb, list, flag, pref, guard);
  } else {
    hipLaunchKernelGGL(zero_ints_kernel, dim3(nb), dim3(B), 0, st, count, n, guard);
    hipLaunchKernelGGL(find_neighbors_kernel, dim3((n * TGX_NB_LANES + B - 1) / B), dim3(B), 0,
                       st, pos, n, c, rcut, cell_start, cell_atoms, count, maxnb, list, flag,
                       guard);
    hipLaunchKernelGGL(sort_lists_kernel, dim3((n + 3) / 4), dim3(256), 0, st, n, maxnb, count,
                       list, guard);
    if (pref)
      hipLaunchKernelGGL(copy_pos_kernel, dim3((3 * n + B - 1) / B), dim3(B), 0, st, n, pos,
                         pref, guard);
  }
  if (rev)
    hipLaunchKernelGGL(reverse_kernel<CellOne>, dim3((n + 7) / 8), dim3(256), 0, st, n, maxnb,
                       CellOne{c},
                       count, list, rev, flag, guard);
  return hipGetLastError();
}

// Batched structures.  Every structure s has its own cell cells[s], atoms
// aoff[s]..aoff[s+1]-1 and, for a box of at least 2 rcut, its own linked-cell
// grid at bins boff[s]..; a smaller box gets one bin and brute force over
// its atoms and supercell images, as for a single structure.
__global__ void bin_atoms_many_kernel(const double* pos, int n, const Cell* cells,
                                      const int* sid, const int* boff, int* cell_of,
                                      int* cell_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = sid[i];
  const Cell c = cells[s];
  int cell = boff[s];
  if (!c.small) {
    int cx, cy, cz;
    cell_coords(pos + 3 * i, c, &cx, &cy, &cz);
    cell += imodd(cx, c.mx) + imodd(cy, c.my) * c.mx + imodd(cz, c.mz) * c.mx * c.my;
  }
  cell_of[i] = cell;
  atomicAdd(cell_count + cell, 1);
}

hipError_t launch_neighbors_many(const double* pos, int n, const Cell* cells, const int* sid,
                                 const int* boff, const int* aoff, int n_bins, double rcut,
                                 int* cell_of, int* cell_count, int* cell_start, int* cell_fill,
                                 int* cell_atoms, int* count, int maxnb, int* list, int* rev,
                                 int* flag, hipStream_t st) {
  const int B = 256;
  const int nb = (n + B - 1) / B;
  hipLaunchKernelGGL(zero_ints_kernel, dim3((n_bins + B - 1) / B), dim3(B), 0, st, cell_count,
                     n_bins, nullptr);
  hipLaunchKernelGGL(bin_atoms_many_kernel, dim3(nb), dim3(B), 0, st, pos, n, cells, sid, boff,
                     cell_of, cell_count);
  hipLaunchKernelGGL(scan_cells_kernel, dim3(1), dim3(1024), 0, st, cell_count, n_bins,
                     cell_start, cell_fill, nullptr);
  hipLaunchKernelGGL(scatter_cells_kernel, dim3(nb), dim3(B), 0, st, n, cell_of, cell_start,
                     cell_fill, cell_atoms, cell_count, pos, Cell{}, nullptr, nullptr);
  hipLaunchKernelGGL(find_sorted_many_kernel, dim3((n + 3) / 4), dim3(256), 0, st, pos, n, cells,
                     sid, boff, aoff, rcut, cell_of, cell_start, cell_atoms, count, maxnb, list,
                     flag);
  if (rev)
    hipLaunchKernelGGL(reverse_kernel<CellMany>, dim3((n + 7) / 8), dim3(256), 0, st, n, maxnb,
                       CellMany{cells, sid}, count, list, rev, flag, nullptr);
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

// ---------------------------------------------------------------------------
// Forces: F_k = sum_slots (g[k][slot] - g[rev[k][slot]]), fixed order, FP64.
// g holds one 4-vector (gx, gy, gz, 0) per slot: one 16- or 32-byte load per
// gathered slot.  all_sites: every atom is a SOAP site (no site checks).
// ---------------------------------------------------------------------------
// With red.part, the first TGX_RB workgroups compute the partial rows of
// the energy/virial reduction instead (the same rows as reduce_part_kernel).
template <typename T>
__global__ __launch_bounds__(256) void force_gather_kernel(int n, int maxnb, const int* count,
                                                           const int* list, const int* rev,
                                                           const int* site_of_atom, int all_sites,
                                                           const T* gpair, double* forces,
                                                           int overwrite, ReduceArgs red) {
  int blk = blockIdx.x;
  if (red.part) {
    if (blk < TGX_RB) {
      __shared__ double sh[256][7];
      reduce_part_rows(blk, red.n_sites, red.vsite, red.n, red.vatom, red.eatom, red.part, sh);
      return;
    }
    blk -= TGX_RB;
  }
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type T4;
  const T4* g4 = reinterpret_cast<const T4*>(gpair);
  // 32 lanes per atom (a typical list has about 30 slots), eight atoms per
  // workgroup
  const int lane = threadIdx.x & 31;
  const int k = blk * 8 + (threadIdx.x >> 5);
  if (k >= n) return;
  const int cnt = min(count[k], maxnb);
  const bool ksite = all_sites || site_of_atom[k] >= 0;
  double f0 = 0, f1 = 0, f2 = 0;
  for (int s = lane; s < cnt; s += 32) {
    const size_t p = (size_t)k * maxnb + s;
    if (ksite) {
      const T4 g = g4[p];
      f0 += (double)g.x;
      f1 += (double)g.y;
      f2 += (double)g.z;
    }
    if (all_sites || site_of_atom[list[p] & TGX_ATOM_MASK] >= 0) {
      const T4 g = g4[rev[p]];
      f0 -= (double)g.x;
      f1 -= (double)g.y;
      f2 -= (double)g.z;
    }
  }
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) {
    f0 += __shfl_xor(f0, o, 64);
    f1 += __shfl_xor(f1, o, 64);
    f2 += __shfl_xor(f2, o, 64);
  }
  if (lane == 0) {
    if (overwrite) {
      forces[3 * k] = f0;
      forces[3 * k + 1] = f1;
      forces[3 * k + 2] = f2;
    } else {
      forces[3 * k] += f0;
      forces[3 * k + 1] += f1;
      forces[3 * k + 2] += f2;
    }
  }
}

template <typename T>
hipError_t launch_force_gather(int n, int maxnb, const int* count, const int* list,
                               const int* rev, const int* site_of_atom, int all_sites,
                               const T* gpair, double* forces, hipStream_t st, int overwrite,
                               const ReduceArgs* red) {
  const ReduceArgs r = red ? *red : ReduceArgs{0, 0, nullptr, nullptr, nullptr, nullptr};
  hipLaunchKernelGGL(force_gather_kernel<T>, dim3((n + 7) / 8 + (r.part ? TGX_RB : 0)), dim3(256),
                     0, st, n, maxnb, count, list, rev, site_of_atom, all_sites, gpair, forces,
                     overwrite, r);
  return hipGetLastError();
}
template hipError_t launch_force_gather<float>(int, int, const int*, const int*, const int*, const int*, int, const float*, double*, hipStream_t, int, const ReduceArgs*);
template hipError_t launch_force_gather<double>(int, int, const int*, const int*, const int*, const int*, int, const double*, double*, hipStream_t, int, const ReduceArgs*);

// ---------------------------------------------------------------------------
// distance_2b term (gap.f90:329-460): each pair counted from both atoms,
// force on the central atom, virial -1/2 sym(F x) then x 1/2.  One wave per
// atom, lanes over its neighbour slots, wave reductions at the end.
// FP64 mode evaluates the reference sum over sparse points per pair; the
// other modes read the smooth part g(r) = delta^2 sum_s alpha_s c_s
// exp(-(r-Q_s)^2/2 sigma^2) from an FP32 cubic-Hermite table (built in FP64)
// and apply the cosine cutoff analytically.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int EXACT, typename G>
__global__ __launch_bounds__(256) void two_body_kernel(TwoBodyDev b, const double* pos,
                                                       const int* species, int n, G geo,
                                                       int maxnb, const int* count,
                                                       const int* list, double* eatom,
                                                       double* forces, double* vatom,
                                                       int do_forces) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int si = species[i];
  if (si != b.sp1 && si != b.sp2) return;
  const Cell c = geo.of(i);
  const int cnt = min(count[i], maxnb);
  double e = 0, f[3] = {0, 0, 0}, v[6] = {0, 0, 0, 0, 0, 0};
  for (int s = lane; s < cnt; s += 64) {
    const int ent = list[(size_t)i * maxnb + s];
    const int sj = species[ent & TGX_ATOM_MASK];
    if (!((si == b.sp1 && sj == b.sp2) || (si == b.sp2 && sj == b.sp1))) continue;
    double x[3];
    const double r = pair_vec(pos, i, ent, c, x);
    if (!(r < b.rcut)) continue;
    double phi, dphi;  // pair energy (one side) and d phi/dr
    two_body_pair<EXACT>(b, r, &phi, &dphi);
    e += phi;
    if (do_forces) {
      // reference: F_i += -2 delta^2 alpha c exp(.) x/r (...) = 2 phi'(r) x/r
      const double pre = 2.0 * dphi / r;
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
  e = wsum(e);
  if (lane == 0) eatom[i] += e;
  if (do_forces) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double t = wsum(f[k]);
      if (lane == 0) forces[3 * i + k] += t;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double t = wsum(v[k]);
      if (lane == 0) vatom[(size_t)i * 6 + k] += 0.5 * t;
    }
  }
}

template <typename G>
static void two_body_go(const TwoBodyDev& b, const double* pos, const int* species, int n,
                        const G& g, int maxnb, const int* count, const int* list, double* eatom,
                        double* forces, double* vatom, int do_forces, int exact,
                        hipStream_t st) {
  const dim3 grid((n + 3) / 4), block(256);
  if (exact || !b.tab)
    hipLaunchKernelGGL((two_body_kernel<1, G>), grid, block, 0, st, b, pos, species, n, g, maxnb,
                       count, list, eatom, forces, vatom, do_forces);
  else
    hipLaunchKernelGGL((two_body_kernel<0, G>), grid, block, 0, st, b, pos, species, n, g, maxnb,
                       count, list, eatom, forces, vatom, do_forces);
}

hipError_t launch_two_body(const TwoBodyDev& b, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           double* eatom, double* forces, double* vatom, int do_forces,
                           int exact, hipStream_t st) {
  if (g.many)
    two_body_go(b, pos, species, n, CellMany{g.many, g.sid}, maxnb, count, list, eatom, forces,
                vatom, do_forces, exact, st);
  else
    two_body_go(b, pos, species, n, CellOne{g.one}, maxnb, count, list, eatom, forces, vatom,
                do_forces, exact, st);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// core_pot term (gap.f90:552-686): cubic spline of a tabulated pair
// potential, each pair counted from both atoms (E_i += V/2), force on the
// central atom V'(r) x/r, virial -1/2 sym(F x) then x 1/2.  One wave per
// atom, as the distance_2b kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double core_spline(const CorePotDev& b, double r, double* dV) {
  const double* x = b.x;
  const int n = b.n;
  if (r < x[0]) {
    *dV = b.yp1;
    return b.V[0] + (r - x[0]) * b.yp1;
  }
  if (r > x[n - 1]) {
    *dV = b.ypn;
    return b.V[n - 1] + (r - x[n - 1]) * b.ypn;
  }
  // first interval j with r < x[j+1] (the last one when r = x[n-1])
  int lo = 0, hi = n - 2;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (r < x[mid + 1]) hi = mid;
    else lo = mid + 1;
  }
  const double h = x[lo + 1] - x[lo];
  const double A = (x[lo + 1] - r) / h, B = 1.0 - A;
  const double y0 = b.V[lo], y1 = b.V[lo + 1], m0 = b.d2[lo], m1 = b.d2[lo + 1];
  *dV = (y1 - y0) / h + ((1.0 - 3.0 * A * A) * m0 + (3.0 * B * B - 1.0) * m1) * (h / 6.0);
  return A * y0 + B * y1 + ((A * A * A - A) * m0 + (B * B * B - B) * m1) * (h * h / 6.0);
}

template <typename G>
__global__ __launch_bounds__(256) void core_pot_kernel(CorePotDev b, const double* pos,
                                                       const int* species, int n, G geo,
                                                       int maxnb, const int* count,
                                                       const int* list, double* eatom,
                                                       double* forces, double* vatom,
                                                       int do_forces) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int si = species[i];
  if (si != b.sp1 && si != b.sp2) return;
  const Cell c = geo.of(i);
  const int cnt = min(count[i], maxnb);
  double e = 0, f[3] = {0, 0, 0}, v[6] = {0, 0, 0, 0, 0, 0};
  for (int s = lane; s < cnt; s += 64) {
    const int ent = list[(size_t)i * maxnb + s];
    const int sj = species[ent & TGX_ATOM_MASK];
    if (!((si == b.sp1 && sj == b.sp2) || (si == b.sp2 && sj == b.sp1))) continue;
    double x[3];
    const double r = pair_vec(pos, i, ent, c, x);
    if (!(r < b.rcut)) continue;
    double dV;
    e += 0.5 * core_spline(b, r, &dV);
    if (do_forces) {
      const double pre = dV / r;
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
  e = wsum(e);
  if (lane == 0) eatom[i] += e;
  if (do_forces) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double t = wsum(f[k]);
      if (lane == 0) forces[3 * i + k] += t;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double t = wsum(v[k]);
      if (lane == 0) vatom[(size_t)i * 6 + k] += 0.5 * t;
    }
  }
}

// ---------------------------------------------------------------------------
// angle_3b term (gap.f90:756-1015): for a centre of species sp0, every
// unordered pair of neighbours within rcut of species (sp1, sp2) in either
// order forms q = (r12 + r13, (r12 - r13)^2, r23); energy fcut12 fcut13
// sum_s pref_s k(|q - Q_s|_sigma).  One wave per centre; lane a owns
// neighbour slot a and sums, over every partner b, the force on its atom,
// so each slot's force is written once, in a fixed order: g[i][a] = -(force
// on slot a's atom from centre i), which the force gather turns into the
// forces on all three atoms (the centre's is minus the sum of its slots').
// Per-atom virial: sum_a sym(F_a x_a)/2, the reference's sym(F2 x12 + F3
// x13)/2 per triplet.  The energy counts each pair once (b > a).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cut3b(double r, double rcut, double buffer, double* fc,
                                      double* dfc) {
  if (r < rcut - buffer) {
    *fc = 1.0;
    *dfc = 0.0;
  } else {
    const double t = M_PI * (r - rcut + buffer) / buffer;
    *fc = (cos(t) + 1.0) / 2.0;
    *dfc = sin(t) / 2.0 * M_PI / buffer;  // -d fc / dr
  }
}

template <typename G>
__global__ __launch_bounds__(256) void angle3b_kernel(Angle3bDev b, const double* pos,
                                                      const int* species, int n, G geo,
                                                      int maxnb, const int* count,
                                                      const int* list, double* eatom,
                                                      double* vatom, double4* g3, int do_forces) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const Cell c = geo.of(i);
  const int cnt = min(count[i], maxnb);
  const bool centre = species[i] == b.sp0;
  const int* row = list + (size_t)i * maxnb;
  double e = 0, v[6] = {0, 0, 0, 0, 0, 0};
  for (int a = lane; a < cnt; a += 64) {
    double F[3] = {0, 0, 0}, xa[3] = {0, 0, 0};
    const int ea = row[a];
    const int sa = species[ea & TGX_ATOM_MASK];
    const double ra = centre ? pair_vec(pos, i, ea, c, xa) : 0.0;
    if (centre && (sa == b.sp1 || sa == b.sp2) && ra < b.rcut) {
      double fa, sfa;
      cut3b(ra, b.rcut, b.buffer, &fa, &sfa);
      const double ua[3] = {xa[0] / ra, xa[1] / ra, xa[2] / ra};
      for (int bb = 0; bb < cnt; bb++) {
        if (bb == a) continue;
        const int eb = row[bb];
        const int sb = species[eb & TGX_ATOM_MASK];
        if (!((sa == b.sp1 && sb == b.sp2) || (sa == b.sp2 && sb == b.sp1))) continue;
        double xb[3];
        const double rb = pair_vec(pos, i, eb, c, xb);
        if (!(rb < b.rcut)) continue;
        double fb, sfb;
        cut3b(rb, b.rcut, b.buffer, &fb, &sfb);
        const double x23[3] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]};
        const double r23 = sqrt(x23[0] * x23[0] + x23[1] * x23[1] + x23[2] * x23[2]);
        const double q1 = ra + rb, q2 = (ra - rb) * (ra - rb), q3 = r23;
        double ks = 0.0, C1 = 0.0, C2 = 0.0, C3 = 0.0;
        for (int s = 0; s < b.n_sparse; s++) {
          const double4 Q = reinterpret_cast<const double4*>(b.Qp)[s];
          const double d1 = (q1 - Q.x) * b.is2[0], d2 = (q2 - Q.y) * b.is2[1],
                       d3 = (q3 - Q.z) * b.is2[2];
          const double rr2 = d1 * (q1 - Q.x) + d2 * (q2 - Q.y) + d3 * (q3 - Q.z);
          double k, kd;  // kernel and the factor of its q-gradient
          if (b.kernel == 0) {
            k = Q.w * exp(-0.5 * rr2);
            kd = k;
          } else {
            const double rs = sqrt(rr2);
            if (rs >= 1.0) {
              k = 0.0;
              kd = 0.0;
            } else {
              const double om = 1.0 - rs;
              k = rs <= 0.0 ? Q.w : Q.w * om * om * om * om * (4.0 * rs + 1.0);
              kd = rs <= 0.0 ? 0.0 : 20.0 * Q.w * om * om * om;  // -cov_pp_der
            }
          }
          ks += k;
          C1 += kd * d1;
          C2 += kd * d2;
          C3 += kd * d3;
        }
        const double fcut = fa * fb;
        if (bb > a) e += fcut * ks;
        if (do_forces) {
          // force on slot a's atom, -d(fa fb ks)/dx_a: d fa/dx_a = -sfa ua and
          // d ks/dx_a = -(C . dq/dx_a), dq/dx_a = (ua, 2 (ra - rb) ua, -u23)
          const double t = C1 + 2.0 * (ra - rb) * C2;
#pragma unroll
          for (int l = 0; l < 3; l++) {
            const double u23 = x23[l] / r23;
            F[l] += fb * ks * sfa * ua[l] + fcut * (t * ua[l] - C3 * u23);
          }
        }
      }
    }
    if (do_forces) {
      g3[(size_t)i * maxnb + a] = make_double4(-F[0], -F[1], -F[2], 0.0);
      v[0] += F[0] * xa[0];
      v[1] += F[1] * xa[1];
      v[2] += F[2] * xa[2];
      v[3] += 0.5 * (F[0] * xa[1] + F[1] * xa[0]);
      v[4] += 0.5 * (F[0] * xa[2] + F[2] * xa[0]);
      v[5] += 0.5 * (F[1] * xa[2] + F[2] * xa[1]);
    }
  }
  e = wsum(e);
  if (lane == 0 && centre) eatom[i] += e;
  if (do_forces) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double t = wsum(v[k]);
      if (lane == 0) vatom[(size_t)i * 6 + k] += t;
    }
  }
}

hipError_t launch_core_pot(const CorePotDev& b, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           double* eatom, double* forces, double* vatom, int do_forces,
                           hipStream_t st) {
  const dim3 grid((n + 3) / 4), block(256);
  if (g.many)
    hipLaunchKernelGGL(core_pot_kernel<CellMany>, grid, block, 0, st, b, pos, species, n,
                       CellMany{g.many, g.sid}, maxnb, count, list, eatom, forces, vatom,
                       do_forces);
  else
    hipLaunchKernelGGL(core_pot_kernel<CellOne>, grid, block, 0, st, b, pos, species, n,
                       CellOne{g.one}, maxnb, count, list, eatom, forces, vatom, do_forces);
  return hipGetLastError();
}

hipError_t launch_angle_3b(const Angle3bDev& b, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           double* eatom, double* vatom, double* g3, int do_forces,
                           hipStream_t st) {
  const dim3 grid((n + 3) / 4), block(256);
  double4* g4 = reinterpret_cast<double4*>(g3);
  if (g.many)
    hipLaunchKernelGGL(angle3b_kernel<CellMany>, grid, block, 0, st, b, pos, species, n,
                       CellMany{g.many, g.sid}, maxnb, count, list, eatom, vatom, g4, do_forces);
  else
    hipLaunchKernelGGL(angle3b_kernel<CellOne>, grid, block, 0, st, b, pos, species, n,
                       CellOne{g.one}, maxnb, count, list, eatom, vatom, g4, do_forces);
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

// Adds one SOAP block's site energies and site virials to their atoms (the
// sites of one block are distinct atoms).
__global__ void site_add_kernel(int n_sites, const int* site_atom, const double* esite,
                                const double* vsite, double* eatom, double* vatom) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_sites) return;
  const int i = site_atom[k];
  if (esite) eatom[i] += esite[k];
  if (vsite)
    for (int t = 0; t < 6; t++) vatom[(size_t)i * 6 + t] += vsite[(size_t)k * 6 + t];
}

hipError_t launch_site_add(int n_sites, const int* site_atom, const double* esite,
                           const double* vsite, double* eatom, double* vatom, hipStream_t st) {
  if (n_sites == 0) return hipSuccess;
  hipLaunchKernelGGL(site_add_kernel, dim3((n_sites + 255) / 256), dim3(256), 0, st, n_sites,
                     site_atom, esite, vsite, eatom, vatom);
  return hipGetLastError();
}

// Deterministic reduction of the per-site and per-atom 6-component virials
// and of the atom energies: a fixed partition over RB workgroups (each a
// fixed tree), then one workgroup sums the RB partials in order.
__global__ void reduce_part_kernel(int n_sites, const double* vsite, int n, const double* vatom,
                                   const double* eatom, double* part /* [RB][8] */) {
  __shared__ double sh[256][7];
  reduce_part_rows(blockIdx.x, n_sites, vsite, n, vatom, eatom, part, sh);
}

__global__ void reduce_final_kernel(const double* part, double* out /* 9 virial + energy */) {
  __shared__ double sh[TGX_RB][7];
  reduce_final_rows(part, out, sh);
}

// The final step alone, the partial rows at out + 16 already written.
hipError_t launch_reduce_final(double* out, hipStream_t st) {
  hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(TGX_RB), 0, st, out + 16, out);
  return hipGetLastError();
}

// out must hold 10 + 8*TGX_RB doubles (the partials live after the result).
// with_final = false leaves the last step to the caller's next kernel
// (reduce_final_rows on out + 16).
hipError_t launch_reduce(int n_sites, const double* vsite, int n, const double* vatom,
                         const double* eatom, double* out, hipStream_t st, bool with_final) {
  hipLaunchKernelGGL(reduce_part_kernel, dim3(TGX_RB), dim3(256), 0, st, n_sites, vsite, n, vatom,
                     eatom, out + 16);
  if (with_final)
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(TGX_RB), 0, st, out + 16, out);
  return hipGetLastError();
}

// Batched structures: per-structure energy and virial, one workgroup per
// structure summing its atoms in a fixed order (out[10*s + 0..8] virial,
// out[10*s + 9] energy).
__global__ void reduce_many_kernel(const int* aoff, const int* site_of_atom,
                                   const double* vsite, const double* vatom,
                                   const double* eatom, double* out) {
  __shared__ double sh[256][7];
  const int t = threadIdx.x, s = blockIdx.x;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = aoff[s] + t; i < aoff[s + 1]; i += 256) {
    if (vsite) {
      const int st = site_of_atom[i];
      if (st >= 0)
        for (int k = 0; k < 6; k++) a[k] += vsite[(size_t)st * 6 + k];
    }
    if (vatom)
      for (int k = 0; k < 6; k++) a[k] += vatom[(size_t)i * 6 + k];
    a[6] += eatom[i];
  }
  for (int k = 0; k < 7; k++) sh[t][k] = a[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
      for (int k = 0; k < 7; k++) sh[t][k] += sh[t + off][k];
    __syncthreads();
  }
  if (t == 0) {
    double* o = out + 10 * (size_t)s;
    o[0] = sh[0][0];
    o[4] = sh[0][1];
    o[8] = sh[0][2];
    o[1] = o[3] = sh[0][3];
    o[2] = o[6] = sh[0][4];
    o[5] = o[7] = sh[0][5];
    o[9] = sh[0][6];
  }
}

hipError_t launch_reduce_many(int n_struct, const int* aoff, const int* site_of_atom,
                              const double* vsite, const double* vatom, const double* eatom,
                              double* out, hipStream_t st) {
  hipLaunchKernelGGL(reduce_many_kernel, dim3(n_struct), dim3(256), 0, st, aoff, site_of_atom,
                     vsite, vatom, eatom, out);
  return hipGetLastError();
}

}  // namespace tgx
