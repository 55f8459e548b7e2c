// Tkatchenko-Scheffler dispersion on Hirshfeld volumes (get_ts_energy_and_
// forces, reference vdw/vdw_interface.f90:127-385).  The volumes V_i come from
// a local-property model on a SOAP block (its kernel step's site values plus
// V0, clipped at 0); dV_i/dr_j per list slot from that block's adjoint.
//
// Three launches per call, one list at max(rcut, descriptor cutoff):
//   ts_volume_kernel   V per atom;
//   ts_pair_kernel     one wave per atom i over its slots: the pair energy,
//                      the pair force on i, the pair virial and the
//                      prefactor dE/dV_i (all FP64, wave sums in lane order);
//   ts_gather_kernel   one wave per atom a over its slots: the Hirshfeld force
//                      dV_i/dr_a * pref_i of every neighbour i, read through
//                      the reverse index at i's own slot (its distance and
//                      vector as i's row has them), so the force is gathered
//                      without atomics.  The self entry (r = 0) is left out,
//                      as the reference's condition rcut_inner < r leaves it.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

namespace {

__device__ __forceinline__ double ts_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// r^-6 with the cubic switch of the inner and outer buffers, and the
// reference's derivative: the unswitched slope plus r^-6 times the switch's
// slope (vdw_interface.f90:199-226).
__device__ __forceinline__ void ts_r6(const TsDev& t, double r, double* r6o, double* r6d) {
  const double r2 = r * r;
  double r6 = 1.0 / (r2 * r2 * r2);
  double d = -6.0 / r * r6;
  const double rc = t.rcut, b = t.buffer, rci = t.rcut_inner, bi = t.buffer_inner;
  if ((b > 0.0 || bi > 0.0) && ((r > rc - b && r < rc) || (r < rci + bi && r > rci))) {
    double x;
    if (r <= rci + bi) {
      x = (rci + bi - r) / bi;
      d += 6.0 * r6 / (-bi) * (-x + x * x);
    } else {
      x = (r - rc + b) / b;
      d += 6.0 * r6 / b * (-x + x * x);
    }
    r6 *= 1.0 - 3.0 * x * x + 2.0 * x * x * x;
  }
  *r6o = r6;
  *r6d = d;
}

}  // namespace

__global__ __launch_bounds__(256) void ts_volume_kernel(int n, const int* site_of_atom,
                                                        const double* esite, double v0,
                                                        double* V) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int s = site_of_atom[a];
  double v = s >= 0 ? esite[s] + v0 : 0.0;
  if (v < 0.0) v = 0.0;
  V[a] = v;
}

template <typename G>
__global__ __launch_bounds__(256) void ts_pair_kernel(TsDev t, const double* pos,
                                                      const int* species, int n, G geo,
                                                      int maxnb, const int* count,
                                                      const int* list, const double* V,
                                                      double* eatom, double* forces,
                                                      double* vatom, double* pref,
                                                      int do_forces) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const Cell c = geo.of(i);
  const int cnt = min(count[i], maxnb);
  const int si = species[i] - 1;
  const double vi = V[i];
  const double c6i = t.c6[si] * vi * vi;
  const double r0i = t.r0[si] * pow(vi, 1.0 / 3.0);
  const double ai = t.a0[si] * vi;
  double e = 0, p1 = 0, p2 = 0, f[3] = {0, 0, 0}, v[6] = {0, 0, 0, 0, 0, 0};
  for (int s = lane; s < cnt; s += 64) {
    const int ent = list[(size_t)i * maxnb + s];
    double dv[3];
    const double r = pair_vec(pos, i, ent, c, dv);
    if (!(r > t.rcut_inner && r < t.rcut)) continue;
    const int j = (ent & TGX_ATOM_MASK) % n;
    const int sj = species[j] - 1;
    const double vj = V[j];
    const double c6j = t.c6[sj] * vj * vj;
    if (c6i == 0.0 || c6j == 0.0) continue;  // C6_ij = 0: no energy, force or prefactor
    const double aj = t.a0[sj] * vj;
    const double c6ij = 2.0 * c6i * c6j / (aj / ai * c6i + ai / aj * c6j);
    const double r0ij = r0i + t.r0[sj] * pow(vj, 1.0 / 3.0);
    double r6, r6d;
    ts_r6(t, r, &r6, &r6d);
    const double ed = exp(-t.d * (r / (t.sR * r0ij) - 1.0));
    const double fd = 1.0 / (1.0 + ed);
    e += c6ij * r6 * fd;
    if (!do_forces) continue;
    p1 += c6ij * fd * r6;
    p2 -= c6ij * fd * fd * r * r6 * ed * t.d / t.sR / (r0ij * r0ij);
    const double sc = c6ij / r * fd * (-r6d - r6 * fd * ed * t.d / t.sR / r0ij);
    const double fx = sc * dv[0], fy = sc * dv[1], fz = sc * dv[2];
    if (j != i) {  // the periodic images of i itself put no force on it
      f[0] += fx;
      f[1] += fy;
      f[2] += fz;
    }
    // virial - 1/4 sym(F x)
    v[0] -= 0.5 * fx * dv[0];
    v[1] -= 0.5 * fy * dv[1];
    v[2] -= 0.5 * fz * dv[2];
    v[3] -= 0.25 * (fx * dv[1] + fy * dv[0]);
    v[4] -= 0.25 * (fx * dv[2] + fz * dv[0]);
    v[5] -= 0.25 * (fy * dv[2] + fz * dv[1]);
  }
  e = ts_wave_sum(e);
  if (lane == 0) eatom[i] = -0.5 * e;
  if (!do_forces) return;
  p1 = ts_wave_sum(p1);
  p2 = ts_wave_sum(p2);
#pragma unroll
  for (int k = 0; k < 3; k++) f[k] = ts_wave_sum(f[k]);
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = ts_wave_sum(v[k]);
  if (lane == 0) {
    pref[i] = vi == 0.0 ? 0.0 : p1 / vi + p2 * t.r0[si] / 3.0 / pow(vi, 2.0 / 3.0);
    forces[3 * i] = f[0];
    forces[3 * i + 1] = f[1];
    forces[3 * i + 2] = f[2];
#pragma unroll
    for (int k = 0; k < 6; k++) vatom[6 * (size_t)i + k] = v[k];
  }
}

// Row a's slot p holds neighbour i; rev[p] is the slot of the same pair in
// row i, where gpair holds dV_i/dr_a (zero beyond the descriptor cutoff).
template <typename T, typename G>
__global__ __launch_bounds__(256) void ts_gather_kernel(TsDev t, const double* pos, int n,
                                                        G geo, int maxnb, const int* count,
                                                        const int* list, const int* rev,
                                                        const T* gpair, const double* pref,
                                                        double* forces, double* vatom) {
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type T4;
  const T4* g4 = reinterpret_cast<const T4*>(gpair);
  const int lane = threadIdx.x & 63;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a >= n) return;
  const int cnt = min(count[a], maxnb);
  double f[3] = {0, 0, 0}, v[6] = {0, 0, 0, 0, 0, 0};
  for (int s = lane; s < cnt; s += 64) {
    const int q = rev[(size_t)a * maxnb + s];
    const int i = q / maxnb;
    const double pi = pref[i];
    if (pi == 0.0) continue;
    const T4 g = g4[q];
    if (g.x == (T)0 && g.y == (T)0 && g.z == (T)0) continue;
    double dv[3];
    const double r = pair_vec(pos, i, list[q], geo.of(i), dv);
    if (!(r > t.rcut_inner && r < t.rcut)) continue;
    const double fx = (double)g.x * pi, fy = (double)g.y * pi, fz = (double)g.z * pi;
    f[0] += fx;
    f[1] += fy;
    f[2] += fz;
    // virial + 1/2 sym(F x), x the vector of i's slot
    v[0] += fx * dv[0];
    v[1] += fy * dv[1];
    v[2] += fz * dv[2];
    v[3] += 0.5 * (fx * dv[1] + fy * dv[0]);
    v[4] += 0.5 * (fx * dv[2] + fz * dv[0]);
    v[5] += 0.5 * (fy * dv[2] + fz * dv[1]);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) f[k] = ts_wave_sum(f[k]);
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = ts_wave_sum(v[k]);
  if (lane == 0) {
    forces[3 * a] += f[0];
    forces[3 * a + 1] += f[1];
    forces[3 * a + 2] += f[2];
#pragma unroll
    for (int k = 0; k < 6; k++) vatom[6 * (size_t)a + k] += v[k];
  }
}

hipError_t launch_ts_volume(int n, const int* site_of_atom, const double* esite, double v0,
                            double* V, hipStream_t st) {
  hipLaunchKernelGGL(ts_volume_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     site_of_atom, esite, v0, V);
  return hipGetLastError();
}

hipError_t launch_ts_pairs(const TsDev& t, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           const double* V, double* eatom, double* forces, double* vatom,
                           double* pref, int do_forces, hipStream_t st) {
  const dim3 grid((n + 3) / 4), block(256);
  if (g.many)
    hipLaunchKernelGGL(ts_pair_kernel<CellMany>, grid, block, 0, st, t, pos, species, n,
                       CellMany{g.many, g.sid}, maxnb, count, list, V, eatom, forces, vatom,
                       pref, do_forces);
  else
    hipLaunchKernelGGL(ts_pair_kernel<CellOne>, grid, block, 0, st, t, pos, species, n,
                       CellOne{g.one}, maxnb, count, list, V, eatom, forces, vatom, pref,
                       do_forces);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ts_gather(const TsDev& t, const double* pos, int n, const Geom& g, int maxnb,
                            const int* count, const int* list, const int* rev, const T* gpair,
                            const double* pref, double* forces, double* vatom, hipStream_t st) {
  const dim3 grid((n + 3) / 4), block(256);
  if (g.many)
    hipLaunchKernelGGL((ts_gather_kernel<T, CellMany>), grid, block, 0, st, t, pos, n,
                       CellMany{g.many, g.sid}, maxnb, count, list, rev, gpair, pref, forces,
                       vatom);
  else
    hipLaunchKernelGGL((ts_gather_kernel<T, CellOne>), grid, block, 0, st, t, pos, n,
                       CellOne{g.one}, maxnb, count, list, rev, gpair, pref, forces, vatom);
  return hipGetLastError();
}
template hipError_t launch_ts_gather<float>(const TsDev&, const double*, int, const Geom&, int,
                                            const int*, const int*, const int*, const float*,
                                            const double*, double*, double*, hipStream_t);
template hipError_t launch_ts_gather<double>(const TsDev&, const double*, int, const Geom&, int,
                                             const int*, const int*, const int*, const double*,
                                             const double*, double*, double*, hipStream_t);

}  // namespace tgx
