// Device-resident NVE molecular dynamics: velocity Verlet (reference
// md/md_utils.f90:486-532), the skin test that triggers a neighbour-list
// rebuild (md/md_interface.f90:1112-1151) and the per-step energy history.
// Every kernel reads the step counter and flags from device memory so that a
// whole step can be captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

// v(t) = v(t-dt) + 1/2 (F(t) + F(t-dt))/m dt  (skipped on the first step),
// x(t+dt) = x(t) + v dt + 1/2 F/m dt^2, with the reference's operation order.
// ke_part[block] = kinetic energy of the block's atoms (fixed tree).  The
// skin test follows on the new positions: *flag is raised when an atom
// moved more than sqrt(lim2) from pref, its position at the last list build.
__global__ __launch_bounds__(256) void md_vv_kernel(int n, double* pos, double* vel,
                                                    const double* f, double* fprev,
                                                    const double* mass, double dt,
                                                    const int* step, double* ke_part,
                                                    const double* pref, double lim2,
                                                    int* flag) {
#pragma clang fp contract(off)
  __shared__ double sh[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double kei = 0.0;
  if (i < n) {
    const bool first = (*step == 0);
    const double m = mass[i];
    double v2 = 0.0;
    for (int k = 0; k < 3; k++) {
      const int a = 3 * i + k;
      double v = vel[a];
      if (!first) v = v + 0.5 * (f[a] + fprev[a]) / m * dt;
      vel[a] = v;
      v2 = v2 + v * v;
      fprev[a] = f[a];
      pos[a] = pos[a] + v * dt + 0.5 * f[a] / m * (dt * dt);
    }
    kei = 0.5 * m * v2;
    const double dx = pos[3 * i] - pref[3 * i], dy = pos[3 * i + 1] - pref[3 * i + 1],
                 dz = pos[3 * i + 2] - pref[3 * i + 2];
    if (dx * dx + dy * dy + dz * dz > lim2) atomicOr(flag, 1);
  }
  sh[threadIdx.x] = kei;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) ke_part[blockIdx.x] = sh[0];
}

// End of step: the final energy/virial sum, kinetic energy of v(t)
// (fixed-order tree), potential energy of x(t+dt), rebuild bookkeeping,
// step counter.
__global__ void md_finish_kernel(int n, const double* ke, double* out,
                                 int* flag, int* nreb, int* step,
                                 double* ep_hist, double* ek_hist, int cap) {
  __shared__ double sh[256];
  __shared__ double shr[TGX_RB][7];
  const int t = threadIdx.x;
  // the energy/virial partials of this step's reduction (out + 16)
  reduce_final_rows(out + 16, out, shr);
  double a = 0.0;
  const int nb = (n + 255) / 256;  // per-block partials written by md_vv_kernel
  for (int i = t; i < nb; i += 256) a += ke[i];
  sh[t] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) {
    const int s = *step;
    ek_hist[s % cap] = sh[0];
    ep_hist[(s + 1) % cap] = out[9];
    if (*flag) {
      *nreb += 1;
      *flag = 0;
    }
    *step = s + 1;
  }
}

// Velocities synchronous with the current positions (v(t) from v(t-dt)) and
// their kinetic energy; the integrator state itself is not changed.
__global__ void md_sync_vel_kernel(int n, const double* vel, const double* f,
                                   const double* fprev, const double* mass,
                                   double dt, const int* step, double* vout,
                                   double* ke) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool none = (*step == 0);
  const double m = mass[i];
  double v2 = 0.0;
  for (int k = 0; k < 3; k++) {
    const int a = 3 * i + k;
    double v = vel[a];
    if (!none) v = v + 0.5 * (f[a] + fprev[a]) / m * dt;
    vout[a] = v;
    v2 = v2 + v * v;
  }
  ke[i] = 0.5 * m * v2;
}

__global__ void md_sum_kernel(int n, const double* x, double* out) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double a = 0.0;
  for (int i = t; i < n; i += 256) a += x[i];
  sh[t] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) *out = sh[0];
}

hipError_t launch_md_vv(int n, double* pos, double* vel, const double* f, double* fprev,
                        const double* mass, double dt, const int* step, double* ke,
                        const double* pref, double lim2, int* flag, hipStream_t st) {
  hipLaunchKernelGGL(md_vv_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, pos, vel, f,
                     fprev, mass, dt, step, ke, pref, lim2, flag);
  return hipGetLastError();
}

hipError_t launch_md_finish(int n, const double* ke, double* out, int* flag, int* nreb,
                            int* step, double* ep_hist, double* ek_hist, int cap,
                            hipStream_t st) {
  hipLaunchKernelGGL(md_finish_kernel, dim3(1), dim3(256), 0, st, n, ke, out, flag, nreb, step,
                     ep_hist, ek_hist, cap);
  return hipGetLastError();
}

hipError_t launch_md_sync_vel(int n, const double* vel, const double* f, const double* fprev,
                              const double* mass, double dt, const int* step, double* vout,
                              double* ke, double* kesum, hipStream_t st) {
  hipLaunchKernelGGL(md_sync_vel_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, vel, f,
                     fprev, mass, dt, step, vout, ke);
  hipLaunchKernelGGL(md_sum_kernel, dim3(1), dim3(256), 0, st, n, ke, kesum);
  return hipGetLastError();
}

}  // namespace tgx
