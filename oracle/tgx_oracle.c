/*
 * tgx_oracle.c - FP64 CPU oracle of the GAP/SOAP energy-and-force path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared with.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product never calls it.  Built with -ffp-contract=off so that every
 * FP64 expression rounds as written.
 *
 * What each part restates and what pins it (tests/test_oracle_*.py):
 *   orc_mic / orc_cells_for_cutoff / orc_neighbors
 *       neighbors/neighbors_interface.f90:144-524 (cell list, self first),
 *       :547-581 (get_distance, gfortran MODULO = fmod + sign fix),
 *       :655-708 (cells for cutoff), xyz/read_xyz.f90:323-373 (supercell).
 *       Pinned by: linked cells == brute force == numpy brute force with
 *       explicit images; permutation/translation invariance.
 *   orc_poly3_overlap / orc_radial / orc_radial_raw / orc_sph / orc_soap
 *       the soap_turbo expansion (third-party, absent from the reference:
 *       libAtoms/soap_turbo, pinned version unknown; call sites
 *       gap/gap_interface.f90:175,301-306).  Restated from the published
 *       method (Caro, PRB 100, 024112 (2019)): poly3 radial basis
 *       orthonormalised by W = S^(-1/2), Gaussian radial smearing, heat-kernel
 *       angular smearing, polynomial amplitude scaling and soft cutoff, power
 *       spectrum over n<=n' with sqrt(2) multiplicity, L2 normalisation.
 *       Exact soap_turbo conventions: PARITY UNPINNED (no vector exists).
 *       Pinned against mathematics: S vs quadrature, W S W = I, raw radial
 *       integrals vs scipy quad, real harmonics vs scipy sph_harm_y, rotation
 *       and permutation invariance, |q| = 1, finite-difference derivatives,
 *       translation sum rule.
 *   orc_gap_kernel_step / orc_gap_soap_ef
 *       gap/gap.f90:39-244.  Pinned by numpy closed form (einsum) and by
 *       finite differences of the whole path (F = -dE/dr, virial =
 *       -dE/dstrain).
 *   orc_gap_2b_ef
 *       gap/gap.f90:329-460.  Pinned by the dimer closed form and FD.
 *   orc_velocity_verlet   md/md_utils.f90:486-532 (harmonic dimer test).
 *   orc_mc_accept         mc/mc_utils.f90:172-180, 215-257.
 *   orc_mc_uniform        the trial stream of tgx_mc_run (splitmix64 of
 *       seed + golden * (6t + k + 1)); pinned by splitmix64's published
 *       outputs for state 0.
 *   orc_mc_run            mc/mc_interface.f90:183-447 with single-atom
 *       moves (mc_utils.f90:259-277, 696-698): one full-system energy per
 *       trial, accept when min(1, p) > u.  Pinned by its limits (T -> inf
 *       accepts every move, T -> 0 only downhill ones) and by recomputing
 *       the final energy.
 *   orc_mlp               the surrogate network (no reference counterpart);
 *       its gradient pinned by central finite differences of the loss.
 *   orc_mc_run_da         delayed-acceptance Metropolis (Christen & Fox,
 *       J. Comput. Graph. Stat. 14, 795 (2005)) with a fixed surrogate;
 *       pinned by its limit: with a flat surrogate (zero network, no 2-body
 *       term) stage 1 passes every trial and the chain is plain Metropolis
 *       drawn with u_5.
 */
#include "tgx_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Neighbours                                                               */
/* ------------------------------------------------------------------------ */

/* Fortran MODULO(a,p) for reals as gfortran evaluates it: fmod, then add p
 * when a and p have opposite signs and the remainder is non-zero. */
static double f_modulo(double a, double p) {
  double r = fmod(a, p);
  if (r != 0.0) {
    if ((a < 0.0) != (p < 0.0)) r = r + p;
  } else {
    r = copysign(0.0, p);
  }
  return r;
}

void orc_mic(const double* posi, const double* posj, const double* L,
             double* dist, double* d) {
  double d2 = 0.0;
  for (int k = 0; k < 3; k++) {
    double t = f_modulo(posj[k] - posi[k], L[k]);
    if (t > L[k] / 2.0) t = t - L[k];
    dist[k] = t;
    d2 = d2 + t * t;
  }
  *d = sqrt(d2);
}

static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

void orc_cells_for_cutoff(const double* lat, double rcut, int* indices) {
  const double *a = lat, *b = lat + 3, *c = lat + 6;
  double axb[3], bxc[3], cxa[3], m[3][3], mi[3][3], md, v[3];
  cross3(a, b, axb);
  double t = sqrt(dot3(axb, axb));
  for (int k = 0; k < 3; k++) axb[k] /= t;
  cross3(c, a, cxa);
  t = sqrt(dot3(cxa, cxa));
  for (int k = 0; k < 3; k++) cxa[k] /= t;
  cross3(b, c, bxc);
  t = sqrt(dot3(bxc, bxc));
  for (int k = 0; k < 3; k++) bxc[k] /= t;
  const double* rows[3] = {axb, cxa, bxc};
  const double* cols[3] = {a, b, c};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = dot3(rows[i], cols[j]);
  md = -m[0][2] * m[1][1] * m[2][0] + m[0][1] * m[1][2] * m[2][0] +
       m[0][2] * m[1][0] * m[2][1] - m[0][0] * m[1][2] * m[2][1] -
       m[0][1] * m[1][0] * m[2][2] + m[0][0] * m[1][1] * m[2][2];
  mi[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  mi[0][1] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
  mi[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  mi[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  mi[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
  mi[1][2] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
  mi[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  mi[2][1] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
  mi[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) mi[i][j] = mi[i][j] / md;
  for (int i = 0; i < 3; i++) {
    v[i] = 0.0;
    for (int j = 0; j < 3; j++) v[i] += mi[i][j] * (2.0 * rcut);
    v[i] = fabs(v[i]);
    indices[i] = (int)v[i];
    if (v[i] > (double)indices[i]) indices[i] += 1;
  }
}

/* get_distance, general-cell branch (neighbors_interface.f90:582-629):
 * the matrix of lattice dot products and its explicit inverse (computed once
 * per lattice there; here by orc_general_mic_setup), the real-valued shift
 * from the normal equations, its nearest integer and the best of the 27
 * integer shifts around it (first minimum in i, j, k order).  a, b, c are
 * the supercell vectors.  Operation order as written in the reference. */
typedef struct {
  double a[3], b[3], c[3], mi[3][3];
} gen_cell;

static void orc_general_mic_setup(const double* a, const double* b, const double* c,
                                  gen_cell* g) {
  double m[3][3], md;
  for (int k = 0; k < 3; k++) {
    g->a[k] = a[k];
    g->b[k] = b[k];
    g->c[k] = c[k];
  }
  m[0][0] = dot3(a, a);
  m[0][1] = dot3(a, b);
  m[0][2] = dot3(a, c);
  m[1][0] = m[0][1];
  m[1][1] = dot3(b, b);
  m[1][2] = dot3(b, c);
  m[2][0] = m[0][2];
  m[2][1] = m[1][2];
  m[2][2] = dot3(c, c);
  md = -(m[0][2] * m[0][2]) * m[1][1] + 2.0 * m[0][1] * m[0][2] * m[1][2] -
       m[0][0] * (m[1][2] * m[1][2]) - (m[0][1] * m[0][1]) * m[2][2] +
       m[0][0] * m[1][1] * m[2][2];
  g->mi[0][0] = m[1][1] * m[2][2] - m[1][2] * m[1][2];
  g->mi[0][1] = m[0][2] * m[1][2] - m[0][1] * m[2][2];
  g->mi[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  g->mi[1][0] = g->mi[0][1];
  g->mi[1][1] = m[0][0] * m[2][2] - m[0][2] * m[0][2];
  g->mi[1][2] = m[0][1] * m[0][2] - m[0][0] * m[1][2];
  g->mi[2][0] = g->mi[0][2];
  g->mi[2][1] = g->mi[1][2];
  g->mi[2][2] = m[0][0] * m[1][1] - m[0][1] * m[0][1];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) g->mi[i][j] = g->mi[i][j] / md;
}

static void orc_general_mic(const double* posi, const double* posj, const gen_cell* g,
                            double* dist, double* d) {
  double dd[3], v[3], x[3], best[3];
  for (int k = 0; k < 3; k++) dd[k] = posj[k] - posi[k];
  v[0] = -1.0 * dot3(dd, g->a);
  v[1] = -1.0 * dot3(dd, g->b);
  v[2] = -1.0 * dot3(dd, g->c);
  for (int i = 0; i < 3; i++) x[i] = (g->mi[i][0] * v[0] + g->mi[i][1] * v[1]) + g->mi[i][2] * v[2];
  int nI = (int)lround(x[0]), nJ = (int)lround(x[1]), nK = (int)lround(x[2]);
  double res_opt = 1e10;
  best[0] = dd[0];
  best[1] = dd[1];
  best[2] = dd[2];
  for (int i = nI - 1; i <= nI + 1; i++)
    for (int j = nJ - 1; j <= nJ + 1; j++)
      for (int k = nK - 1; k <= nK + 1; k++) {
        double t[3];
        for (int q = 0; q < 3; q++)
          t[q] = ((dd[q] + (double)i * g->a[q]) + (double)j * g->b[q]) + (double)k * g->c[q];
        const double res = dot3(t, t);
        if (res < res_opt) {
          res_opt = res;
          best[0] = t[0];
          best[1] = t[1];
          best[2] = t[2];
        }
      }
  dist[0] = best[0];
  dist[1] = best[1];
  dist[2] = best[2];
  *d = sqrt(dot3(best, best));
}

/* The reference's neighbour list of a non-orthorhombic cell
 * (neighbors_interface.f90:390-421 with get_distance :582-629 and the second
 * get_distance pass :470-500): the supercell of read_xyz.f90:323-373 when a
 * plane spacing is below 2 rcut, then every supercell atom j != i tested
 * with the general minimum image against the supercell vectors; self first,
 * then ascending supercell id. */
static int orc_neighbors_general(int n, const double* pos, const double* lat, double rcut,
                                 int64_t cap, int* n_neigh, int* list, double* rjs,
                                 double* xyz, double* th, double* ph, int64_t* n_pairs,
                                 int* indices) {
  int idx[3];
  orc_cells_for_cutoff(lat, rcut, idx);
  if (indices) memcpy(indices, idx, sizeof(idx));
  const int nrep = idx[0] * idx[1] * idx[2], nsc = n * nrep;
  double* P = (double*)malloc(sizeof(double) * 3 * (size_t)nsc);
  int cnt0 = 0;
  for (int i2 = 0; i2 < idx[0]; i2++)
    for (int j2 = 0; j2 < idx[1]; j2++)
      for (int k2 = 0; k2 < idx[2]; k2++)
        for (int i = 0; i < n; i++) {
          for (int k = 0; k < 3; k++)
            P[3 * cnt0 + k] = ((pos[3 * i + k] + (double)i2 * lat[k]) + (double)j2 * lat[3 + k]) +
                              (double)k2 * lat[6 + k];
          cnt0++;
        }
  double A[3], B[3], Cv[3];
  for (int k = 0; k < 3; k++) {
    A[k] = (double)idx[0] * lat[k];
    B[k] = (double)idx[1] * lat[3 + k];
    Cv[k] = (double)idx[2] * lat[6 + k];
  }
  gen_cell g;
  orc_general_mic_setup(A, B, Cv, &g);
  int** nb = (int**)calloc((size_t)n, sizeof(int*));
  int* cnt = (int*)calloc((size_t)n, sizeof(int));
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    int capi = 64, ni = 0;
    int* buf = (int*)malloc(sizeof(int) * capi);
    for (int k = 0; k < nsc; k++) {
      if (k == i) continue;
      double dv[3], dk;
      orc_general_mic(P + 3 * i, P + 3 * k, &g, dv, &dk);
      if (dk < rcut) {
        if (ni == capi) buf = (int*)realloc(buf, sizeof(int) * (capi *= 2));
        buf[ni++] = k;
      }
    }
    nb[i] = buf;
    cnt[i] = ni;
  }
  int64_t tot = 0;
  for (int i = 0; i < n; i++) tot += cnt[i] + 1;
  *n_pairs = tot;
  int rc = TGX_OK;
  if (tot > cap) {
    rc = TGX_ERR_CAPACITY;
  } else {
    int64_t k2 = 0;
    for (int i = 0; i < n; i++) {
      n_neigh[i] = cnt[i] + 1;
      list[k2] = i + 1;
      if (rjs) rjs[k2] = 0.0;
      if (xyz) xyz[3 * k2] = xyz[3 * k2 + 1] = xyz[3 * k2 + 2] = 0.0;
      if (th) th[k2] = 0.0;
      if (ph) ph[k2] = 0.0;
      k2++;
      for (int q = 0; q < cnt[i]; q++, k2++) {
        const int j = nb[i][q];
        double dv[3], d;
        orc_general_mic(P + 3 * i, P + 3 * j, &g, dv, &d);
        list[k2] = j + 1;
        if (rjs) rjs[k2] = d;
        if (xyz) {
          xyz[3 * k2] = dv[0];
          xyz[3 * k2 + 1] = dv[1];
          xyz[3 * k2 + 2] = dv[2];
        }
        if (th) {
          if (fabs(dv[2]) >= d)
            th[k2] = dv[2] > 0.0 ? 0.0 : acos(-1.0);
          else
            th[k2] = acos(dv[2] / d);
        }
        if (ph) ph[k2] = atan2(dv[1], dv[0]);
      }
    }
  }
  for (int i = 0; i < n; i++) free(nb[i]);
  free(nb);
  free(cnt);
  free(P);
  return rc;
}

static int imod(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

int orc_neighbors(int n, const double* pos, const double* lat, double rcut,
                  int algo, int64_t cap, int* n_neigh, int* list, double* rjs,
                  double* xyz, double* th, double* ph, int64_t* n_pairs,
                  int* indices) {
  const double dtol = 1e-6;
  if (!(fabs(lat[1]) < dtol && fabs(lat[2]) < dtol && fabs(lat[3]) < dtol &&
        fabs(lat[5]) < dtol && fabs(lat[6]) < dtol && fabs(lat[7]) < dtol))
    return orc_neighbors_general(n, pos, lat, rcut, cap, n_neigh, list, rjs, xyz, th, ph,
                                 n_pairs, indices);
  double A[9] = {lat[0], 0, 0, 0, lat[4], 0, 0, 0, lat[8]};
  int idx[3];
  orc_cells_for_cutoff(A, rcut, idx);
  if (indices) memcpy(indices, idx, sizeof(idx));
  int small = idx[0] > 1 || idx[1] > 1 || idx[2] > 1;
  int nrep = idx[0] * idx[1] * idx[2];
  int nsc = n * nrep;
  double* P = (double*)malloc(sizeof(double) * 3 * (size_t)nsc);
  double L[3];
  if (small) {
    int cnt = 0;
    for (int i2 = 0; i2 < idx[0]; i2++)
      for (int j2 = 0; j2 < idx[1]; j2++)
        for (int k2 = 0; k2 < idx[2]; k2++)
          for (int i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++)
              P[3 * cnt + k] = pos[3 * i + k] + (double)i2 * A[k] +
                               (double)j2 * A[3 + k] + (double)k2 * A[6 + k];
            cnt++;
          }
    L[0] = (double)idx[0] * A[0];
    L[1] = (double)idx[1] * A[4];
    L[2] = (double)idx[2] * A[8];
  } else {
    memcpy(P, pos, sizeof(double) * 3 * (size_t)n);
    L[0] = A[0];
    L[1] = A[4];
    L[2] = A[8];
  }

  /* Candidate lists per atom (supercell ids, 0-based), sorted. */
  int** nb = (int**)calloc((size_t)n, sizeof(int*));
  int* cnt = (int*)calloc((size_t)n, sizeof(int));
  int use_cells = (algo == 0) && !small;
  int mx = 0, my = 0, mz = 0;
  int *head = NULL, *next = NULL;
  if (use_cells) {
    const double tol = 1e-10;
    mx = (int)(L[0] / rcut);
    my = (int)(L[1] / rcut);
    mz = (int)(L[2] / rcut);
    head = (int*)malloc(sizeof(int) * (size_t)mx * my * mz);
    next = (int*)malloc(sizeof(int) * (size_t)n);
    for (int c = 0; c < mx * my * mz; c++) head[c] = -1;
    double ctr[3] = {L[0] / 2.0, L[1] / 2.0, L[2] / 2.0};
    for (int i = 0; i < n; i++) {
      double dd[3], d;
      orc_mic(ctr, P + 3 * i, L, dd, &d);
      for (int k = 0; k < 3; k++) dd[k] = dd[k] + ctr[k];
      int cx = imod((int)(dd[0] / (L[0] + tol) * mx), mx);
      int cy = imod((int)(dd[1] / (L[1] + tol) * my), my);
      int cz = imod((int)(dd[2] / (L[2] + tol) * mz), mz);
      int c = cx + cy * mx + cz * mx * my;
      next[i] = head[c];
      head[c] = i;
    }
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    int capi = 64, ni = 0;
    int* buf = (int*)malloc(sizeof(int) * capi);
    if (use_cells) {
      const double tol = 1e-10;
      double ctr[3] = {L[0] / 2.0, L[1] / 2.0, L[2] / 2.0};
      double dd[3], d;
      orc_mic(ctr, P + 3 * i, L, dd, &d);
      for (int k = 0; k < 3; k++) dd[k] = dd[k] + ctr[k];
      int i2 = (int)(dd[0] / (L[0] + tol) * mx);
      int j2 = (int)(dd[1] / (L[1] + tol) * my);
      int k2 = (int)(dd[2] / (L[2] + tol) * mz);
      int seen[27], nseen = 0;
      for (int k3 = k2 - 1; k3 <= k2 + 1; k3++)
        for (int j3 = j2 - 1; j3 <= j2 + 1; j3++)
          for (int i3 = i2 - 1; i3 <= i2 + 1; i3++) {
            int c = imod(i3, mx) + imod(j3, my) * mx + imod(k3, mz) * mx * my;
            int dup = 0;
            for (int q = 0; q < nseen; q++) dup |= (seen[q] == c);
            if (dup) continue;
            seen[nseen++] = c;
            for (int k = head[c]; k >= 0; k = next[k]) {
              if (k == i) continue;
              double dv[3], dk;
              orc_mic(P + 3 * i, P + 3 * k, L, dv, &dk);
              if (dk < rcut) {
                if (ni == capi) buf = (int*)realloc(buf, sizeof(int) * (capi *= 2));
                buf[ni++] = k;
              }
            }
          }
    } else {
      for (int k = 0; k < nsc; k++) {
        if (k == i) continue;
        double dv[3], dk;
        orc_mic(P + 3 * i, P + 3 * k, L, dv, &dk);
        if (dk < rcut) {
          if (ni == capi) buf = (int*)realloc(buf, sizeof(int) * (capi *= 2));
          buf[ni++] = k;
        }
      }
    }
    qsort(buf, (size_t)ni, sizeof(int), cmp_int);
    nb[i] = buf;
    cnt[i] = ni;
  }
  int64_t tot = 0;
  for (int i = 0; i < n; i++) tot += cnt[i] + 1;
  *n_pairs = tot;
  int rc = TGX_OK;
  if (tot > cap) {
    rc = TGX_ERR_CAPACITY;
  } else {
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    off[0] = 0;
    for (int i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i] + 1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      int64_t k2 = off[i];
      n_neigh[i] = cnt[i] + 1;
      list[k2] = i + 1;
      if (rjs) rjs[k2] = 0.0;
      if (xyz) xyz[3 * k2] = xyz[3 * k2 + 1] = xyz[3 * k2 + 2] = 0.0;
      if (th) th[k2] = 0.0;
      if (ph) ph[k2] = 0.0;
      for (int q = 0; q < cnt[i]; q++) {
        int64_t kk = k2 + 1 + q;
        int j = nb[i][q];
        double dv[3], d;
        orc_mic(P + 3 * i, P + 3 * j, L, dv, &d);
        list[kk] = j + 1;
        if (rjs) rjs[kk] = d;
        if (xyz) {
          xyz[3 * kk] = dv[0];
          xyz[3 * kk + 1] = dv[1];
          xyz[3 * kk + 2] = dv[2];
        }
        if (th) {
          if (fabs(dv[2]) >= d)
            th[kk] = dv[2] > 0.0 ? 0.0 : acos(-1.0);
          else
            th[kk] = acos(dv[2] / d);
        }
        if (ph) ph[kk] = atan2(dv[1], dv[0]);
      }
    }
    free(off);
  }
  for (int i = 0; i < n; i++) free(nb[i]);
  free(nb);
  free(cnt);
  free(P);
  free(head);
  free(next);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Radial basis                                                             */
/* ------------------------------------------------------------------------ */

/* W = S^(-1/2) is needed to full FP64 accuracy although cond(S) ~ 1e13 for
 * n_max = 8: the eigen-decomposition runs in binary128 (gcc __float128) and
 * only the result is rounded to double. */
typedef __float128 q128;

static q128 q_sqrt(q128 v) {
  if (v <= 0) return 0;
  q128 x = (q128)sqrt((double)v);
  for (int it = 0; it < 4; it++) x = (x + v / x) / 2;
  return x;
}
static q128 q_abs(q128 v) { return v < 0 ? -v : v; }

static void jacobi_eig_q(int n, q128* a, q128* v, q128* w) {
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) v[i * n + j] = (i == j);
  for (int sweep = 0; sweep < 100; sweep++) {
    q128 off = 0, diag = 0;
    for (int p = 0; p < n; p++) {
      diag += a[p * n + p] * a[p * n + p];
      for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
    }
    if (off <= diag * (q128)1e-66) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        q128 apq = a[p * n + q];
        if (apq == 0) continue;
        q128 theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
        q128 t = (theta >= 0 ? 1 : -1) / (q_abs(theta) + q_sqrt(theta * theta + 1));
        q128 c = 1 / q_sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          q128 akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          q128 apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          q128 vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = a[i * n + i];
}

void orc_poly3_overlap(int nmax, double* S, double* W) {
  /* phi_a(r) = sqrt((2a+5)/rh^(2a+5)) (rh-r)^(a+2), a = 1..nmax, on [0,rh]:
   * S_ab = sqrt((2a+5)(2b+5)) / (a+b+5), independent of rh. */
  q128 A[TGX_MAX_NR * TGX_MAX_NR], V[TGX_MAX_NR * TGX_MAX_NR], w[TGX_MAX_NR];
  for (int a = 1; a <= nmax; a++)
    for (int b = 1; b <= nmax; b++) {
      q128 s = q_sqrt((q128)((2 * a + 5) * (2 * b + 5))) / (q128)(a + b + 5);
      S[(a - 1) * nmax + (b - 1)] = (double)s;
      A[(a - 1) * nmax + (b - 1)] = s;
    }
  jacobi_eig_q(nmax, A, V, w);
  for (int i = 0; i < nmax; i++)
    for (int j = 0; j < nmax; j++) {
      q128 acc = 0;
      for (int k = 0; k < nmax; k++) acc += V[i * nmax + k] * V[j * nmax + k] / q_sqrt(w[k]);
      W[i * nmax + j] = (double)acc;
    }
}

/* poly3gauss basis (soap_turbo's get_orthonormalization_matrix_poly3gauss,
 * called at gpu/gpu_interface.f90:431): phi_1 = N_g exp(-r^2/(2 sigma^2)),
 * the atom's Gaussian at the origin, then the first alpha_max - 1 poly3
 * functions N_k (rh - r)^(k+2), all normalised on [0, rh] with weight 1:
 *   S_11 = 1, S_1k = N_g N_k int_0^rh exp(-r^2/(2 sigma^2)) (rh - r)^(k+2) dr,
 *   S_kl = sqrt((2k+5)(2l+5)) / (k+l+5);
 * N_g^-2 = int_0^rh exp(-r^2/sigma^2) dr = sqrt(pi)/2 sigma erf(rh/sigma).
 * The Gaussian-polynomial integrals come from the recurrence of radial_J
 * at r = 0, all in binary128 (cond(S) is about 1e13, so W must come from an
 * S exact to far below FP64 rounding). */
static q128 q_expq(q128 x) { return expq(x); }

void orc_poly3gauss_overlap(int nmax, double rh_in, double sigma_in, double* S, double* W) {
  q128 A[TGX_MAX_NR * TGX_MAX_NR], V[TGX_MAX_NR * TGX_MAX_NR], w[TGX_MAX_NR];
  const q128 rh = rh_in, sg = sigma_in, s2 = sg * sg;
  const q128 pi = M_PIq;
  /* J_m = int_0^rh v^m exp(-(v - rh)^2/(2 s^2)) dv, v = rh - r */
  q128 J[TGX_MAX_NR + 8];
  const q128 er = 1, ec = q_expq(-rh * rh / (2 * s2));
  J[0] = sg * q_sqrt(pi / 2) * erfq(rh / (sg * q_sqrt((q128)2)));
  J[1] = rh * J[0] + s2 * (ec - er);
  q128 rhp = 1;
  for (int k = 2; k < nmax + 4; k++) {
    rhp *= rh;
    J[k] = rh * J[k - 1] + (k - 1) * s2 * J[k - 2] - s2 * rhp * er;
  }
  const q128 Ng = 1 / q_sqrt(q_sqrt(pi) / 2 * sg * erfq(rh / sg));
  for (int a = 0; a < nmax; a++)
    for (int b = 0; b < nmax; b++) {
      q128 v;
      if (a == 0 && b == 0) {
        v = 1;
      } else if (a == 0 || b == 0) {
        const int k = a + b; /* the polynomial's index */
        const q128 Nk = q_sqrt((q128)(2 * k + 5)) / powq(rh, (q128)(2 * k + 5) / 2);
        v = Ng * Nk * J[k + 2];
      } else {
        v = q_sqrt((q128)((2 * a + 5) * (2 * b + 5))) / (q128)(a + b + 5);
      }
      A[a * nmax + b] = v;
      S[a * nmax + b] = (double)v;
    }
  jacobi_eig_q(nmax, A, V, w);
  for (int i = 0; i < nmax; i++)
    for (int j = 0; j < nmax; j++) {
      q128 acc = 0;
      for (int k = 0; k < nmax; k++) acc += V[i * nmax + k] * V[j * nmax + k] / q_sqrt(w[k]);
      W[i * nmax + j] = (double)acc;
    }
}

/* J_k = int_0^rh u^k exp(-(u-c)^2/(2 s^2)) du, c = rh - r, k = 0..kmax. */
static void radial_J(double r, double rh, double s, int kmax, double* J) {
  double c = rh - r, s2 = s * s;
  double er = exp(-r * r / (2.0 * s2)), ec = exp(-c * c / (2.0 * s2));
  J[0] = s * sqrt(PI / 2.0) *
         (erf(r / (s * sqrt(2.0))) + erf(c / (s * sqrt(2.0))));
  if (kmax >= 1) J[1] = c * J[0] + s2 * (ec - er);
  double rhp = 1.0; /* rh^(k-1) */
  for (int k = 2; k <= kmax; k++) {
    rhp *= rh;
    J[k] = c * J[k - 1] + (k - 1) * s2 * J[k - 2] - s2 * rhp * er;
  }
}

/* T_k for poly degree d = k+3 and radial weight rho^p, from J. */
static double poly_T(const double* J, int d, int p, double rh) {
  if (p == 0) return J[d];
  if (p == 1) return rh * J[d] - J[d + 1];
  return rh * rh * J[d] - 2.0 * rh * J[d + 1] + J[d + 2];
}

void orc_radial_raw(double r, double rh, double sig, int p, int nk,
                    double* out) {
  double J[64];
  radial_J(r, rh, sig, nk + 2 + p + 2, J);
  double pre = 1.0 / (sqrt(2.0 * PI) * sig);
  for (int k = 0; k < nk; k++) out[k] = pre * poly_T(J, k + 3, p, rh);
}

/* Unnormalised (no W) basis integrals I_a(r) (a = 1..nmax) and dI_a/dr for
 * a Gaussian of width sig0 + sigs*r centred at r. */
static void radial_I(double r, double rh, double sig0, double sigs, int p,
                     int nmax, int gauss, double* I, double* dI) {
  double s = sig0 + sigs * r;
  double J[64];
  int kmax = nmax + 2 + p + 2;
  radial_J(r, rh, s, kmax, J);
  double c = rh - r, s2 = s * s;
  double pre = 1.0 / (sqrt(2.0 * PI) * s);
  double dJc[64], dJs[64];
  for (int k = 0; k + 2 <= kmax; k++) {
    dJc[k] = (J[k + 1] - c * J[k]) / s2;
    dJs[k] = (J[k + 2] - 2.0 * c * J[k + 1] + c * c * J[k]) / (s2 * s);
  }
  /* poly3gauss: slot 0 is the Gaussian, slots 1.. the poly3 functions 1.. */
  for (int a = 1; a <= nmax - gauss; a++) {
    int d = a + 2;
    double Na = sqrt((2.0 * a + 5.0) / pow(rh, 2.0 * a + 5.0));
    double T = poly_T(J, d, p, rh);
    double Tc = poly_T(dJc, d, p, rh);
    double Ts = poly_T(dJs, d, p, rh);
    I[a - 1 + gauss] = Na * pre * T;
    if (dI)
      dI[a - 1 + gauss] = Na * (pre * (-Tc + Ts * sigs) - pre / s * sigs * T);
  }
  if (gauss) {
    /* int_0^rh exp(-u^2/(2 sb^2)) u^p G(u; r, s) du with sb = sig0: the
     * product of the two Gaussians is exp(-r^2/(2 (sb^2 + s^2))) times a
     * Gaussian of centre mu = r sb^2/(sb^2+s^2) and width tau = sb s /
     * sqrt(sb^2+s^2); M_k = int_0^rh u^k exp(-(u-mu)^2/(2 tau^2)) du by
     * parts: M_k = mu M_(k-1) + (k-1) tau^2 M_(k-2) - tau^2 [u^(k-1) e]_0^rh. */
    const double sb2 = sig0 * sig0, Sg = sb2 + s * s;
    const double A = exp(-r * r / (2.0 * Sg));
    const double mu = r * sb2 / Sg, t2 = sb2 * s * s / Sg, t = sqrt(t2);
    const double e0 = exp(-mu * mu / (2.0 * t2)), eh = exp(-(rh - mu) * (rh - mu) / (2.0 * t2));
    double M[5];
    M[0] = t * sqrt(PI / 2.0) * (erf((rh - mu) / (t * sqrt(2.0))) + erf(mu / (t * sqrt(2.0))));
    M[1] = mu * M[0] - t2 * (eh - e0);
    double rp = 1.0;
    for (int k = 2; k < 5; k++) {
      rp *= rh;
      M[k] = mu * M[k - 1] + (k - 1) * t2 * M[k - 2] - t2 * rp * eh;
    }
    const double Ng = 1.0 / sqrt(sqrt(PI) / 2.0 * sig0 * erf(rh / sig0));
    I[0] = Ng * pre * A * M[p];
    if (dI) {
      const double dMmu = (M[p + 1] - mu * M[p]) / t2;
      const double dMt = (M[p + 2] - 2.0 * mu * M[p + 1] + mu * mu * M[p]) / (t2 * t);
      const double dmu = sb2 / Sg - r * sb2 * 2.0 * s * sigs / (Sg * Sg);
      const double dt = sigs * sig0 * sb2 / (Sg * sqrt(Sg));
      const double dA = A * (-r / Sg + r * r * s * sigs / (Sg * Sg));
      dI[0] = Ng * (-pre * sigs / s * A * M[p] + pre * dA * M[p] +
                    pre * A * (dMmu * dmu + dMt * dt));
    }
  }
}

/* W = S^(-1/2) of species s's basis (alpha_max x alpha_max, row-major). */
static void basis_W(const tgx_soap_hypers* h, int s, double* W) {
  double S[TGX_MAX_NR * TGX_MAX_NR];
  if (h->basis == 1)
    orc_poly3gauss_overlap(h->alpha_max[s], h->rcut_hard[s], h->atom_sigma_r[s], S, W);
  else
    orc_poly3_overlap(h->alpha_max[s], S, W);
}

/* Amplitude (polynomial scaling, global scaling) times soft cutoff. */
static void radial_envelope(const tgx_soap_hypers* h, int s, double r,
                            double* e, double* de) {
  double rh = h->rcut_hard[s], rs = h->rcut_soft[s];
  double x = r / rh;
  double poly = 1.0 + 2.0 * x * x * x - 3.0 * x * x;
  double dpoly = (6.0 * x * x - 6.0 * x) / rh;
  double as = h->amplitude_scaling[s];
  double amp, damp;
  if (as == 0.0) {
    amp = 1.0;
    damp = 0.0;
  } else if (as == 1.0) {
    amp = poly;
    damp = dpoly;
  } else {
    amp = pow(poly, as);
    damp = poly > 0.0 ? as * pow(poly, as - 1.0) * dpoly : 0.0;
  }
  double fc = 1.0, dfc = 0.0;
  if (r >= rh) {
    fc = 0.0;
  } else if (rs < rh && r > rs) {
    double y = (r - rs) / (rh - rs);
    fc = 1.0 - 3.0 * y * y + 2.0 * y * y * y;
    dfc = (-6.0 * y + 6.0 * y * y) / (rh - rs);
  }
  double g = h->global_scaling[s];
  *e = g * amp * fc;
  *de = g * (damp * fc + amp * dfc);
}

/* R_n, dR_n/dr with a given W (alpha_max x alpha_max, row n). */
static void radial_W(const tgx_soap_hypers* h, int s, const double* W, int nr,
                     const double* r, double* R, double* dR) {
  int nmax = h->alpha_max[s];
  for (int q = 0; q < nr; q++) {
    double I[TGX_MAX_NR], dI[TGX_MAX_NR], e, de;
    double rr = r[q];
    if (rr >= h->rcut_hard[s]) {
      for (int n = 0; n < nmax; n++) R[q * nmax + n] = dR[q * nmax + n] = 0.0;
      continue;
    }
    radial_I(rr, h->rcut_hard[s], h->atom_sigma_r[s],
             h->atom_sigma_r_scaling[s], h->radial_enhancement, nmax, h->basis == 1, I, dI);
    radial_envelope(h, s, rr, &e, &de);
    for (int n = 0; n < nmax; n++) {
      double v = 0.0, dv = 0.0;
      for (int a = 0; a < nmax; a++) {
        v += W[n * nmax + a] * I[a];
        dv += W[n * nmax + a] * dI[a];
      }
      R[q * nmax + n] = e * v;
      dR[q * nmax + n] = de * v + e * dv;
    }
  }
}

void orc_radial(const tgx_soap_hypers* h, int s, int nr, const double* r,
                double* R, double* dR) {
  double W[TGX_MAX_NR * TGX_MAX_NR];
  basis_W(h, s, W);
  radial_W(h, s, W, nr, r, R, dR);
}

/* Central-atom radial coefficients of species s at r = 0. */
static void radial_self(const tgx_soap_hypers* h, int s, const double* W,
                        double* R) {
  int nmax = h->alpha_max[s];
  double I[TGX_MAX_NR];
  radial_I(0.0, h->rcut_hard[s], h->atom_sigma_r[s], 0.0,
           h->radial_enhancement, nmax, h->basis == 1, I, NULL);
  for (int n = 0; n < nmax; n++) {
    double v = 0.0;
    for (int a = 0; a < nmax; a++) v += W[n * nmax + a] * I[a];
    R[n] = h->central_weight[s] * v;
  }
}

/* ------------------------------------------------------------------------ */
/* Real spherical harmonics                                                 */
/* ------------------------------------------------------------------------ */

void orc_sph(int L, const double* u, double* Y, double* dF) {
  double x = u[0], y = u[1], z = u[2];
  double C[TGX_MAX_LMAX + 2], S[TGX_MAX_LMAX + 2];
  double Q[TGX_MAX_LMAX + 1][TGX_MAX_LMAX + 1], Qd[TGX_MAX_LMAX + 1][TGX_MAX_LMAX + 1];
  C[0] = 1.0;
  S[0] = 0.0;
  for (int m = 0; m < L; m++) {
    C[m + 1] = x * C[m] - y * S[m];
    S[m + 1] = x * S[m] + y * C[m];
  }
  double dfact = 1.0; /* (2m-1)!! */
  for (int m = 0; m <= L; m++) {
    if (m > 0) dfact *= (2.0 * m - 1.0);
    Q[m][m] = dfact;
    Qd[m][m] = 0.0;
    if (m + 1 <= L) {
      Q[m + 1][m] = (2.0 * m + 1.0) * z * dfact;
      Qd[m + 1][m] = (2.0 * m + 1.0) * dfact;
    }
    for (int l = m + 2; l <= L; l++) {
      Q[l][m] = ((2.0 * l - 1.0) * z * Q[l - 1][m] - (l + m - 1.0) * Q[l - 2][m]) / (l - m);
      Qd[l][m] = ((2.0 * l - 1.0) * (Q[l - 1][m] + z * Qd[l - 1][m]) -
                  (l + m - 1.0) * Qd[l - 2][m]) / (l - m);
    }
  }
  for (int l = 0; l <= L; l++) {
    for (int m = 0; m <= l; m++) {
      double ratio = 1.0; /* (l-m)!/(l+m)! */
      for (int k = l - m + 1; k <= l + m; k++) ratio /= (double)k;
      double K = sqrt((2.0 * l + 1.0) / (4.0 * PI) * ratio);
      if (m > 0) K *= sqrt(2.0);
      double q = K * Q[l][m], qd = K * Qd[l][m];
      if (m == 0) {
        int i = l * l + l;
        Y[i] = q;
        if (dF) {
          dF[3 * i] = 0.0;
          dF[3 * i + 1] = 0.0;
          dF[3 * i + 2] = qd;
        }
      } else {
        int ic = l * l + l + m, is = l * l + l - m;
        Y[ic] = q * C[m];
        Y[is] = q * S[m];
        if (dF) {
          dF[3 * ic] = q * m * C[m - 1];
          dF[3 * ic + 1] = -q * m * S[m - 1];
          dF[3 * ic + 2] = qd * C[m];
          dF[3 * is] = q * m * S[m - 1];
          dF[3 * is + 1] = q * m * C[m - 1];
          dF[3 * is + 2] = qd * S[m];
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* SOAP                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nr, L, L2, npair, dim, dim_full; /* dim: descriptor, dim_full: power spectrum */
  int off[TGX_MAX_SPECIES], nmax[TGX_MAX_SPECIES];
  double W[TGX_MAX_SPECIES][TGX_MAX_NR * TGX_MAX_NR];
} soap_geom;

static void soap_geom_init(const tgx_soap_hypers* h, soap_geom* g) {
  g->nr = 0;
  for (int s = 0; s < h->n_species; s++) {
    g->off[s] = g->nr;
    g->nmax[s] = h->alpha_max[s];
    g->nr += h->alpha_max[s];
    basis_W(h, s, g->W[s]);
  }
  g->L = h->l_max;
  g->L2 = (h->l_max + 1) * (h->l_max + 1);
  g->npair = g->nr * (g->nr + 1) / 2;
  g->dim_full = g->npair * (h->l_max + 1);
  g->dim = g->dim_full;
  if (h->compress_P_nonzero > 0) {
    g->dim = 0;
    for (int k = 0; k < h->compress_P_nonzero; k++)
      g->dim = h->compress_P_i[k] > g->dim ? h->compress_P_i[k] : g->dim;
  }
}

/* compress_soap: out[P_i - 1] += P_el x[P_j - 1] over the entries, each of
 * `stride` components (1 for p, 3 for dp/dr) (soap_turbo's P
 * transformation, read_gap.f90:485-549 reads it). */
static void compress_apply(const tgx_soap_hypers* h, int dim_out, const double* x, int stride,
                           double* out) {
  memset(out, 0, sizeof(double) * (size_t)dim_out * stride);
  for (int k = 0; k < h->compress_P_nonzero; k++) {
    const int i = h->compress_P_i[k] - 1, j = h->compress_P_j[k] - 1;
    for (int t = 0; t < stride; t++) out[i * stride + t] += h->compress_P_el[k] * x[j * stride + t];
  }
}

/* Expansion coefficients c[(n*L2 + lm)] of one site (c zeroed by the
 * caller): the central atom's c_n00 plus, per neighbour j, R_n(r_j) A_l(r_j)
 * Y_lm(u_j) with the angular smearing factor A_l = exp(-l(l+1) st^2 /
 * (2 r_j^2)).  Keeps R, dR/dr, Y, dY, A, dA/dr per neighbour for the
 * derivative pass. */
static void site_expand(const tgx_soap_hypers* h, const soap_geom* g, int si, int np,
                        const double* x, const int* sj, double* c, double* Rv, double* dRv,
                        double* AY, double* Yv, double* dFv, double* Al, double* dAl) {
  int L = g->L, L2 = g->L2;
  /* central atom */
  {
    double Rs[TGX_MAX_NR];
    radial_self(h, si, g->W[si], Rs);
    for (int n = 0; n < g->nmax[si]; n++)
      c[(g->off[si] + n) * L2 + 0] += Rs[n] * (1.0 / sqrt(4.0 * PI));
  }
  for (int j = 0; j < np; j++) {
    const double* v = x + 3 * j;
    double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    int s = sj[j];
    radial_W(h, s, g->W[s], 1, &r, Rv + (size_t)j * TGX_MAX_NR, dRv + (size_t)j * TGX_MAX_NR);
    double u[3] = {v[0] / r, v[1] / r, v[2] / r};
    orc_sph(L, u, Yv + (size_t)j * L2, dFv + (size_t)j * L2 * 3);
    double st = h->atom_sigma_t[s] + h->atom_sigma_t_scaling[s] * r;
    double dst = h->atom_sigma_t_scaling[s];
    for (int l = 0; l <= L; l++) {
      double ex = -0.5 * l * (l + 1.0) * st * st / (r * r);
      double a = exp(ex);
      double dex = -0.5 * l * (l + 1.0) * (2.0 * st * dst / (r * r) - 2.0 * st * st / (r * r * r));
      Al[j * (L + 1) + l] = a;
      dAl[j * (L + 1) + l] = a * dex;
    }
    for (int l = 0; l <= L; l++)
      for (int m = -l; m <= l; m++) {
        int lm = l * l + l + m;
        AY[(size_t)j * L2 + lm] = Al[j * (L + 1) + l] * Yv[(size_t)j * L2 + lm];
      }
    for (int n = 0; n < g->nmax[s]; n++) {
      double Rn = Rv[(size_t)j * TGX_MAX_NR + n];
      double* cn = c + (size_t)(g->off[s] + n) * L2;
      for (int lm = 0; lm < L2; lm++) cn[lm] += Rn * AY[(size_t)j * L2 + lm];
    }
  }
}

/* One site: central species si, np neighbour vectors x (3 each) with species
 * sj (0-based).  Writes q (dim) and, if dq != NULL, dq[(j*dim + d)*3 + k] =
 * dq_d / dr_j,k for each neighbour j; the self derivative is -sum_j. */
static void soap_site(const tgx_soap_hypers* h, const soap_geom* g, int si,
                      int np, const double* x, const int* sj, double* q,
                      double* dq) {
  int NR = g->nr, L = g->L, L2 = g->L2, dim = g->dim, dimf = g->dim_full;
  const int cmp = h->compress_P_nonzero > 0;
  double* c = (double*)calloc((size_t)NR * L2, sizeof(double));
  double* Rv = (double*)malloc(sizeof(double) * (size_t)np * TGX_MAX_NR);
  double* dRv = (double*)malloc(sizeof(double) * (size_t)np * TGX_MAX_NR);
  double* AY = (double*)malloc(sizeof(double) * (size_t)np * L2);
  double* Yv = (double*)malloc(sizeof(double) * (size_t)np * L2);
  double* dFv = (double*)malloc(sizeof(double) * (size_t)np * L2 * 3);
  double* Al = (double*)malloc(sizeof(double) * (size_t)np * (L + 1));
  double* dAl = (double*)malloc(sizeof(double) * (size_t)np * (L + 1));
  site_expand(h, g, si, np, x, sj, c, Rv, dRv, AY, Yv, dFv, Al, dAl);
  /* power spectrum (compressed: pc = P p) */
  double* pf = (double*)calloc((size_t)dimf, sizeof(double));
  double* p = pf;
  int k = 0;
  for (int n = 0; n < NR; n++)
    for (int n2 = n; n2 < NR; n2++) {
      double w = (n == n2) ? 1.0 : sqrt(2.0);
      for (int l = 0; l <= L; l++) {
        double acc = 0.0;
        for (int m = -l; m <= l; m++) {
          int lm = l * l + l + m;
          acc += c[n * L2 + lm] * c[n2 * L2 + lm];
        }
        pf[k * (L + 1) + l] = w * acc;
      }
      k++;
    }
  if (cmp) {
    p = (double*)malloc(sizeof(double) * (size_t)dim);
    compress_apply(h, dim, pf, 1, p);
  }
  double nrm = 0.0;
  for (int d = 0; d < dim; d++) nrm += p[d] * p[d];
  nrm = sqrt(nrm);
  for (int d = 0; d < dim; d++) q[d] = p[d] / nrm;

  if (dq) {
    double* dc = (double*)malloc(sizeof(double) * TGX_MAX_NR * L2 * 3);
    double* dpf = (double*)malloc(sizeof(double) * (size_t)dimf * 3);
    double* dp = cmp ? (double*)malloc(sizeof(double) * (size_t)dim * 3) : dpf;
    for (int j = 0; j < np; j++) {
      const double* v = x + 3 * j;
      double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      double u[3] = {v[0] / r, v[1] / r, v[2] / r};
      int s = sj[j], nm = g->nmax[s], o = g->off[s];
      /* dc[(n_local*L2 + lm)*3 + k] */
      for (int n = 0; n < nm; n++) {
        double Rn = Rv[(size_t)j * TGX_MAX_NR + n], dRn = dRv[(size_t)j * TGX_MAX_NR + n];
        for (int l = 0; l <= L; l++) {
          double A = Al[j * (L + 1) + l], dA = dAl[j * (L + 1) + l];
          for (int m = -l; m <= l; m++) {
            int lm = l * l + l + m;
            double Y = Yv[(size_t)j * L2 + lm];
            const double* dF = dFv + ((size_t)j * L2 + lm) * 3;
            double udF = u[0] * dF[0] + u[1] * dF[1] + u[2] * dF[2];
            double radial = (dRn * A + Rn * dA) * Y;
            double tang = Rn * A / r;
            for (int kk = 0; kk < 3; kk++)
              dc[((size_t)n * L2 + lm) * 3 + kk] = radial * u[kk] + tang * (dF[kk] - udF * u[kk]);
          }
        }
      }
      memset(dpf, 0, sizeof(double) * (size_t)dimf * 3);
      int kp = 0;
      for (int n = 0; n < NR; n++)
        for (int n2 = n; n2 < NR; n2++) {
          double w = (n == n2) ? 1.0 : sqrt(2.0);
          int in1 = (n >= o && n < o + nm), in2 = (n2 >= o && n2 < o + nm);
          if (in1 || in2) {
            for (int l = 0; l <= L; l++) {
              double acc[3] = {0, 0, 0};
              for (int m = -l; m <= l; m++) {
                int lm = l * l + l + m;
                for (int kk = 0; kk < 3; kk++) {
                  if (in1) acc[kk] += dc[((size_t)(n - o) * L2 + lm) * 3 + kk] * c[n2 * L2 + lm];
                  if (in2) acc[kk] += c[n * L2 + lm] * dc[((size_t)(n2 - o) * L2 + lm) * 3 + kk];
                }
              }
              for (int kk = 0; kk < 3; kk++) dpf[(kp * (L + 1) + l) * 3 + kk] = w * acc[kk];
            }
          }
          kp++;
        }
      if (cmp) compress_apply(h, dim, dpf, 3, dp);
      double qdp[3] = {0, 0, 0};
      for (int d = 0; d < dim; d++)
        for (int kk = 0; kk < 3; kk++) qdp[kk] += q[d] * dp[d * 3 + kk];
      for (int d = 0; d < dim; d++)
        for (int kk = 0; kk < 3; kk++)
          dq[((size_t)j * dim + d) * 3 + kk] = (dp[d * 3 + kk] - q[d] * qdp[kk]) / nrm;
    }
    free(dc);
    if (cmp) free(dp);
    free(dpf);
  }
  if (cmp) free(p);
  free(pf);
  free(c);
  free(Rv);
  free(dRv);
  free(AY);
  free(Yv);
  free(dFv);
  free(Al);
  free(dAl);
}

/* Expansion coefficients c_nlm of one site (species si, np neighbour
 * vectors x with species sj, 0-based): c[(n*(l_max+1)^2 + l*l+l+m)], n
 * running over all species' radial channels.  Exposed so that the tests can
 * check them against a 3-D quadrature of the smeared density. */
int orc_soap_coeffs(const tgx_soap_hypers* h, int si, int np, const double* x, const int* sj,
                    double* c) {
  soap_geom g;
  soap_geom_init(h, &g);
  int L = g.L, L2 = g.L2, m = np > 0 ? np : 1;
  memset(c, 0, sizeof(double) * (size_t)g.nr * L2);
  double* Rv = (double*)malloc(sizeof(double) * (size_t)m * TGX_MAX_NR);
  double* dRv = (double*)malloc(sizeof(double) * (size_t)m * TGX_MAX_NR);
  double* AY = (double*)malloc(sizeof(double) * (size_t)m * L2);
  double* Yv = (double*)malloc(sizeof(double) * (size_t)m * L2);
  double* dFv = (double*)malloc(sizeof(double) * (size_t)m * L2 * 3);
  double* Al = (double*)malloc(sizeof(double) * (size_t)m * (L + 1));
  double* dAl = (double*)malloc(sizeof(double) * (size_t)m * (L + 1));
  site_expand(h, &g, si, np, x, sj, c, Rv, dRv, AY, Yv, dFv, Al, dAl);
  free(Rv);
  free(dRv);
  free(AY);
  free(Yv);
  free(dFv);
  free(Al);
  free(dAl);
  return TGX_OK;
}

/* nf enters soap_turbo's radial expansion (read_gap.f90:164-165,192-194
 * read it, default 4; gap_interface.f90:301-306 pass it to get_soap).  Its
 * role in soap_turbo's arithmetic is not in the reference tree, so only the
 * default is accepted: any other value is refused rather than silently
 * treated as 4. */
static int hypers_supported(const tgx_soap_hypers* h) {
  for (int s = 0; s < h->n_species; s++)
    if (h->nf[s] != 4.0) return TGX_ERR_UNSUPPORTED;
  if (h->basis != 0 && h->basis != 1) return TGX_ERR_UNSUPPORTED;
  if (h->compress_P_nonzero > 0) {
    int nr = 0;
    for (int s = 0; s < h->n_species; s++) nr += h->alpha_max[s];
    const int dimf = nr * (nr + 1) / 2 * (h->l_max + 1);
    for (int k = 0; k < h->compress_P_nonzero; k++)
      if (h->compress_P_i[k] < 1 || h->compress_P_j[k] < 1 || h->compress_P_j[k] > dimf)
        return TGX_ERR_ARG;
  }
  return TGX_OK;
}

static double soap_rcut_max(const tgx_soap_hypers* h) {
  double r = 0.0;
  for (int s = 0; s < h->n_species; s++) r = h->rcut_hard[s] > r ? h->rcut_hard[s] : r;
  return r;
}

/* Site/pair selection of gap_interface.f90:184-289 on a full neighbour list:
 * central sites by central_species, pairs with rjs < rcut_max. */
typedef struct {
  int n_sites;
  int64_t n_pairs;
  int* site_atom;     /* n_sites, 0-based */
  int* site_nn;       /* pairs per site (self included) */
  int64_t* site_off;  /* n_sites+1 */
  int* pair_atom;     /* folded 0-based */
  double* pair_xyz;   /* 3 per pair */
} site_pairs;

static void select_sites(const tgx_soap_hypers* h, int n, const int* species,
                         const int* n_neigh, const int* list,
                         const double* rjs, const double* xyz, site_pairs* sp) {
  double rc = soap_rcut_max(h);
  sp->n_sites = 0;
  sp->n_pairs = 0;
  sp->site_atom = (int*)malloc(sizeof(int) * (size_t)n);
  sp->site_nn = (int*)malloc(sizeof(int) * (size_t)n);
  sp->site_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t k = 0, tot = 0;
  for (int i = 0; i < n; i++) {
    if (h->central_species == 0 || species[i] == h->central_species) {
      int c = 0;
      for (int q = 0; q < n_neigh[i]; q++)
        if (rjs[k + q] < rc) c++;
      sp->site_atom[sp->n_sites] = i;
      sp->site_nn[sp->n_sites] = c;
      sp->site_off[sp->n_sites] = tot;
      tot += c;
      sp->n_sites++;
    }
    k += n_neigh[i];
  }
  sp->site_off[sp->n_sites] = tot;
  sp->n_pairs = tot;
  sp->pair_atom = (int*)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
  sp->pair_xyz = (double*)malloc(sizeof(double) * 3 * (size_t)(tot > 0 ? tot : 1));
  k = 0;
  int64_t t = 0;
  for (int i = 0; i < n; i++) {
    if (h->central_species == 0 || species[i] == h->central_species) {
      for (int q = 0; q < n_neigh[i]; q++)
        if (rjs[k + q] < rc) {
          sp->pair_atom[t] = (list[k + q] - 1) % n;
          for (int kk = 0; kk < 3; kk++) sp->pair_xyz[3 * t + kk] = xyz[3 * (k + q) + kk];
          t++;
        }
    }
    k += n_neigh[i];
  }
}

static void free_sites(site_pairs* sp) {
  free(sp->site_atom);
  free(sp->site_nn);
  free(sp->site_off);
  free(sp->pair_atom);
  free(sp->pair_xyz);
}

/* Descriptors and derivatives for sites [s0, s1) of sp. */
static void soap_chunk(const tgx_soap_hypers* h, const soap_geom* g,
                       const site_pairs* sp, const int* species, int s0,
                       int s1, double* soap, double* soap_der) {
  int dim = g->dim;
  int64_t pbase = sp->site_off[s0];
#pragma omp parallel for schedule(dynamic, 4)
  for (int st = s0; st < s1; st++) {
    int64_t o = sp->site_off[st];
    int nn = sp->site_nn[st];
    int np = nn - 1; /* exclude self */
    int si = species[sp->site_atom[st]] - 1;
    int* sj = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    for (int j = 0; j < np; j++) sj[j] = species[sp->pair_atom[o + 1 + j]] - 1;
    double* q = soap + (size_t)(st - s0) * dim;
    double* dq = NULL;
    if (soap_der) dq = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1) * dim * 3);
    soap_site(h, g, si, np, sp->pair_xyz + 3 * (o + 1), sj, q, dq);
    if (soap_der) {
      double* self = soap_der + (size_t)(o - pbase) * dim * 3;
      memset(self, 0, sizeof(double) * (size_t)dim * 3);
      for (int j = 0; j < np; j++) {
        double* dst = soap_der + (size_t)(o - pbase + 1 + j) * dim * 3;
        memcpy(dst, dq + (size_t)j * dim * 3, sizeof(double) * (size_t)dim * 3);
        for (int d = 0; d < dim * 3; d++) self[d] -= dst[d];
      }
      free(dq);
    }
    free(sj);
  }
}

int orc_soap(const tgx_soap_hypers* h, int n, const double* pos,
             const double* lat, const int* species, double* soap,
             double* soap_der, int64_t pair_cap, int* n_sites,
             int64_t* n_pairs, int* site_atom, int* pair_atom,
             double* pair_xyz, int* site_n_neigh) {
  if (hypers_supported(h) != TGX_OK) return TGX_ERR_UNSUPPORTED;
  double rc = soap_rcut_max(h);
  int64_t np = 0;
  int rc0 = orc_neighbors(n, pos, lat, rc, 0, 0, NULL, NULL, NULL, NULL, NULL, NULL, &np, NULL);
  if (rc0 != TGX_OK && rc0 != TGX_ERR_CAPACITY) return rc0;
  int* nn = (int*)malloc(sizeof(int) * (size_t)n);
  int* list = (int*)malloc(sizeof(int) * (size_t)np);
  double* rjs = (double*)malloc(sizeof(double) * (size_t)np);
  double* xyz = (double*)malloc(sizeof(double) * 3 * (size_t)np);
  orc_neighbors(n, pos, lat, rc, 0, np, nn, list, rjs, xyz, NULL, NULL, &np, NULL);
  site_pairs sp;
  select_sites(h, n, species, nn, list, rjs, xyz, &sp);
  soap_geom g;
  soap_geom_init(h, &g);
  *n_sites = sp.n_sites;
  *n_pairs = sp.n_pairs;
  int rcode = TGX_OK;
  if (soap_der && sp.n_pairs > pair_cap) {
    rcode = TGX_ERR_CAPACITY;
  } else {
    soap_chunk(h, &g, &sp, species, 0, sp.n_sites, soap, soap_der);
    if (site_atom) memcpy(site_atom, sp.site_atom, sizeof(int) * (size_t)sp.n_sites);
    if (site_n_neigh) memcpy(site_n_neigh, sp.site_nn, sizeof(int) * (size_t)sp.n_sites);
    if (pair_atom && sp.n_pairs <= pair_cap)
      memcpy(pair_atom, sp.pair_atom, sizeof(int) * (size_t)sp.n_pairs);
    if (pair_xyz && sp.n_pairs <= pair_cap)
      memcpy(pair_xyz, sp.pair_xyz, sizeof(double) * 3 * (size_t)sp.n_pairs);
  }
  free_sites(&sp);
  free(nn);
  free(list);
  free(rjs);
  free(xyz);
  return rcode;
}

/* ------------------------------------------------------------------------ */
/* GAP kernel step                                                          */
/* ------------------------------------------------------------------------ */

static double ipow(double x, int e) {
  double r = 1.0;
  for (int i = 0; i < e; i++) r *= x;
  return r;
}

void orc_gap_kernel_step(int n_sites, int dim, int n_sparse,
                         const double* soap, const double* Qs,
                         const double* alphas, double delta, double zeta0,
                         double* energies, double* qss) {
  int is_int = fabs(zeta0 - (double)(int)zeta0) < 1e-5;
  int zi = (int)zeta0;
  double zeta = is_int ? (double)zi : zeta0;
  /* Blocks of 4 sites x 4 sparse points share their row loads; every K(i,s)
   * is still the sum over d in ascending order, every energy the sum over s
   * in ascending order, as the unblocked loops give. */
#pragma omp parallel for schedule(dynamic, 1)
  for (int i0 = 0; i0 < n_sites; i0 += 4) {
    const int ni = n_sites - i0 < 4 ? n_sites - i0 : 4;
    double* kr = (double*)malloc(sizeof(double) * 4 * (size_t)(n_sparse > 0 ? n_sparse : 1));
    const double* qi[4];
    for (int a = 0; a < 4; a++) qi[a] = soap + (size_t)(i0 + (a < ni ? a : ni - 1)) * dim;
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s0 = 0; s0 < n_sparse; s0 += 4) {
      const int nsb = n_sparse - s0 < 4 ? n_sparse - s0 : 4;
      const double* qs[4];
      for (int b = 0; b < 4; b++) qs[b] = Qs + (size_t)(s0 + (b < nsb ? b : nsb - 1)) * dim;
      double k[4][4];
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) k[a][b] = 0.0;
      for (int d = 0; d < dim; d++) {
        const double x0 = qi[0][d], x1 = qi[1][d], x2 = qi[2][d], x3 = qi[3][d];
        for (int b = 0; b < 4; b++) {
          const double y = qs[b][d];
          k[0][b] += x0 * y;
          k[1][b] += x1 * y;
          k[2][b] += x2 * y;
          k[3][b] += x3 * y;
        }
      }
      for (int a = 0; a < ni; a++)
        for (int b = 0; b < nsb; b++) {
          const int s = s0 + b;
          const double kk = k[a][b];
          double kz = is_int ? ipow(kk, zi) : pow(kk, zeta);
          e[a] += kz * alphas[s];
          kr[(size_t)a * n_sparse + s] = (is_int ? ipow(kk, zi - 1) : pow(kk, zeta - 1.0)) * alphas[s];
        }
    }
    for (int a = 0; a < ni; a++) energies[i0 + a] = delta * delta * e[a];
    if (qss) {
      /* acc[d] = sum over s (ascending) of kr[s] Qs[s][d], Qs read row by row */
      double* acc = (double*)calloc((size_t)4 * dim, sizeof(double));
      for (int s = 0; s < n_sparse; s++) {
        const double* qr = Qs + (size_t)s * dim;
        for (int a = 0; a < ni; a++) {
          const double ks = kr[(size_t)a * n_sparse + s];
          double* ac = acc + (size_t)a * dim;
          for (int d = 0; d < dim; d++) ac[d] += ks * qr[d];
        }
      }
      for (int a = 0; a < ni; a++)
        for (int d = 0; d < dim; d++)
          qss[(size_t)d * n_sites + i0 + a] = -zeta * delta * delta * acc[(size_t)a * dim + d];
      free(acc);
    }
    free(kr);
  }
}

void orc_gap_soap_ef(int n_sites, int dim, int n_sparse, const double* soap,
                     const double* soap_der, const double* alphas,
                     double delta, double zeta, double e0, const double* Qs,
                     const int* n_neigh, const int* pair_atom,
                     const double* xyz, int n_atoms, int do_forces,
                     double* energies, double* forces, double* virial) {
  double* qss = do_forces ? (double*)malloc(sizeof(double) * (size_t)n_sites * dim) : NULL;
  orc_gap_kernel_step(n_sites, dim, n_sparse, soap, Qs, alphas, delta, zeta, energies, qss);
  for (int i = 0; i < n_sites; i++) energies[i] += e0;
  if (!do_forces) return;
  memset(forces, 0, sizeof(double) * 3 * (size_t)n_atoms);
  memset(virial, 0, sizeof(double) * 9);
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_sites + 1));
  off[0] = 0;
  for (int i = 0; i < n_sites; i++) off[i + 1] = off[i] + n_neigh[i];
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  double* fth = (double*)calloc((size_t)nth * 3 * n_atoms, sizeof(double));
  double* vth = (double*)calloc((size_t)nth * 9, sizeof(double));
#pragma omp parallel
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double* F = fth + (size_t)t * 3 * n_atoms;
    double* V = vth + (size_t)t * 9;
    double* qrow = (double*)malloc(sizeof(double) * (size_t)dim);
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < n_sites; i++) {
      for (int d = 0; d < dim; d++) qrow[d] = qss[(size_t)d * n_sites + i];
      for (int64_t l = off[i]; l < off[i + 1]; l++) {
        int j2 = pair_atom[l];
        double f[3];
        for (int k = 0; k < 3; k++) {
          double acc = 0.0;
          for (int d = 0; d < dim; d++) acc += qrow[d] * soap_der[((size_t)l * dim + d) * 3 + k];
          f[k] = acc;
          F[3 * j2 + k] += acc;
        }
        for (int k1 = 0; k1 < 3; k1++)
          for (int k2 = 0; k2 < 3; k2++)
            V[k2 * 3 + k1] += 0.5 * (f[k1] * xyz[3 * l + k2] + f[k2] * xyz[3 * l + k1]);
      }
    }
    free(qrow);
  }
  for (int t = 0; t < nth; t++) {
    for (size_t a = 0; a < (size_t)3 * n_atoms; a++) forces[a] += fth[(size_t)t * 3 * n_atoms + a];
    for (int a = 0; a < 9; a++) virial[a] += vth[t * 9 + a];
  }
  free(fth);
  free(vth);
  free(off);
  free(qss);
}

/* ------------------------------------------------------------------------ */
/* 2-body                                                                   */
/* ------------------------------------------------------------------------ */

void orc_gap_2b_ef(const orc_2b* b, int n, const int* species,
                   const int* n_neigh, const int* list, const double* rjs,
                   const double* xyz, int do_forces, double* energies,
                   double* forces, double* virial) {
  const double pi = acos(-1.0);
  double rcut = b->rcut, buffer = b->buffer, delta = b->delta, sigma = b->sigma;
  int sp1 = b->sp1, sp2 = b->sp2;
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  off[0] = 0;
  for (int i = 0; i < n; i++) off[i + 1] = off[i] + n_neigh[i];
  if (do_forces) memset(virial, 0, sizeof(double) * 9);
  double vir[9] = {0};
  for (int i = 0; i < n; i++) {
    energies[i] = 0.0;
    if (do_forces) forces[3 * i] = forces[3 * i + 1] = forces[3 * i + 2] = 0.0;
    if (species[i] != sp1 && species[i] != sp2) continue;
    for (int64_t k = off[i] + 1; k < off[i + 1]; k++) {
      int sj = species[(list[k] - 1) % n];
      if (!((species[i] == sp1 && sj == sp2) || (species[i] == sp2 && sj == sp1))) continue;
      double r = rjs[k];
      if (!(r < rcut)) continue;
      double fcut, dfcut;
      if (r < rcut - buffer) {
        fcut = 1.0;
        dfcut = 0.0;
      } else {
        fcut = (cos(pi * (r - rcut + buffer) / buffer) + 1.0) / 2.0;
        dfcut = pi / 2.0 / buffer * sin(pi * (r - rcut + buffer) / buffer);
      }
      for (int s = 0; s < b->n_sparse; s++) {
        double ex = exp(-0.5 * (r - b->Qs[s]) * (r - b->Qs[s]) / (sigma * sigma));
        energies[i] += delta * delta * b->alphas[s] * b->cutoff[s] * fcut * ex;
        if (do_forces) {
          double f[3];
          for (int kk = 0; kk < 3; kk++)
            f[kk] = -2.0 * delta * delta * b->alphas[s] * b->cutoff[s] * ex * xyz[3 * k + kk] / r *
                    ((r - b->Qs[s]) / (sigma * sigma) * fcut + dfcut);
          for (int kk = 0; kk < 3; kk++) forces[3 * i + kk] += f[kk];
          for (int k1 = 0; k1 < 3; k1++)
            for (int k2 = 0; k2 < 3; k2++)
              vir[k2 * 3 + k1] -= 0.5 * (f[k1] * xyz[3 * k + k2] + f[k2] * xyz[3 * k + k1]);
        }
      }
    }
  }
  if (do_forces)
    for (int a = 0; a < 9; a++) virial[a] = 0.5 * vir[a];
  free(off);
}

/* ------------------------------------------------------------------------ */
/* core_pot and angle_3b terms                                              */
/* ------------------------------------------------------------------------ */

/* read_gap.f90:701-752: the tridiagonal sweep for the spline's second
 * derivatives (natural end where the slope exceeds 0.99e30, else clamped),
 * then natural-end slopes replaced by spline_der at the end points.  At
 * x(n) the reference's interval search runs past the table; the last
 * interval is used here. */
void orc_core_pot_setup(int n, const double* x, const double* V, double* yp1, double* ypn,
                        double* d2) {
  double* u = (double*)malloc(sizeof(double) * (size_t)n);
  if (*yp1 > 0.99e30) {
    d2[0] = 0.0;
    u[0] = 0.0;
  } else {
    d2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((V[1] - V[0]) / (x[1] - x[0]) - *yp1);
  }
  for (int i = 1; i < n - 1; i++) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * d2[i - 1] + 2.0;
    d2[i] = (sig - 1.0) / p;
    u[i] = (V[i + 1] - V[i]) / (x[i + 1] - x[i]) - (V[i] - V[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  double qn, un;
  if (*ypn > 0.99e30) {
    qn = 0.0;
    un = 0.0;
  } else {
    qn = 0.5;
    un = (3.0 / (x[n - 1] - x[n - 2])) * (*ypn - (V[n - 1] - V[n - 2]) / (x[n - 1] - x[n - 2]));
  }
  d2[n - 1] = (un - qn * u[n - 2]) / (qn * d2[n - 2] + 1.0);
  for (int i = n - 2; i >= 0; i--) d2[i] = d2[i] * d2[i + 1] + u[i];
  if (*yp1 > 0.99e30) *yp1 = orc_spline(n, x, V, d2, *yp1, *ypn, x[0], 1e100, 1);
  if (*ypn > 0.99e30) *ypn = orc_spline(n, x, V, d2, *yp1, *ypn, x[n - 1], 1e100, 1);
  free(u);
}

/* splines.f90:33-156: linear continuation outside [x(1), x(n)], else the
 * cubic on the interval with x(j) <= r < x(j+1) (the first j whose x(j+1)
 * exceeds r; r = x(n) takes the last interval). */
double orc_spline(int n, const double* x, const double* y, const double* y2, double yp1,
                  double ypn, double r, double rcut, int derivative) {
  if (!(r < rcut)) return 0.0;
  if (r < x[0]) return derivative ? yp1 : y[0] + (r - x[0]) * yp1;
  if (r > x[n - 1]) return derivative ? ypn : y[n - 1] + (r - x[n - 1]) * ypn;
  int j = 0;
  while (j < n - 2 && !(r < x[j + 1])) j++;
  double h = x[j + 1] - x[j];
  double A = (x[j + 1] - r) / h;
  double B = 1.0 - A;
  if (!derivative) {
    double h26 = h * h / 6.0;
    double C = (A * A * A - A) * h26;
    double D = (B * B * B - B) * h26;
    return A * y[j] + B * y[j + 1] + C * y2[j] + D * y2[j + 1];
  }
  double h6 = h / 6.0;
  double dA = -1.0 / h, dB = -dA;
  double dC = (1.0 - 3.0 * A * A) * h6;
  double dD = (3.0 * B * B - 1.0) * h6;
  return dA * y[j] + dB * y[j + 1] + dC * y2[j] + dD * y2[j + 1];
}

/* gap.f90:552-686: for every pair of the right species within max(x),
 * E_i += V(r)/2; F_i += V'(r) xyz/r; virial -= sym(F x xyz)/2, then x 1/2. */
void orc_core_pot_ef(const orc_core* b, int n, const int* species, const int* n_neigh,
                     const int* list, const double* rjs, const double* xyz, int do_forces,
                     double* energies, double* forces, double* virial) {
  double rcut = b->x[0];
  for (int a = 1; a < b->n; a++) rcut = b->x[a] > rcut ? b->x[a] : rcut;
  int sp1 = b->sp1, sp2 = b->sp2;
  double vir[9] = {0};
  int64_t k = 0;
  for (int i = 0; i < n; i++) {
    energies[i] = 0.0;
    if (do_forces) forces[3 * i] = forces[3 * i + 1] = forces[3 * i + 2] = 0.0;
  }
  for (int i = 0; i < n; i++) {
    if (species[i] != sp1 && species[i] != sp2) {
      k += n_neigh[i];
      continue;
    }
    k++; /* self */
    for (int j = 1; j < n_neigh[i]; j++, k++) {
      int sj = species[(list[k] - 1) % n];
      if (!((species[i] == sp1 && sj == sp2) || (species[i] == sp2 && sj == sp1))) continue;
      double r = rjs[k];
      if (!(r < rcut)) continue;
      energies[i] += 0.5 * orc_spline(b->n, b->x, b->V, b->d2, b->yp1, b->ypn, r, rcut, 0);
      if (do_forces) {
        double dV = orc_spline(b->n, b->x, b->V, b->d2, b->yp1, b->ypn, r, rcut, 1);
        double f[3];
        for (int a = 0; a < 3; a++) f[a] = dV * xyz[3 * k + a] / r;
        for (int a = 0; a < 3; a++) forces[3 * i + a] += f[a];
        for (int k1 = 0; k1 < 3; k1++)
          for (int k2 = 0; k2 < 3; k2++)
            vir[k2 * 3 + k1] -= 0.5 * (f[k1] * xyz[3 * k + k2] + f[k2] * xyz[3 * k + k1]);
      }
    }
  }
  if (do_forces)
    for (int a = 0; a < 9; a++) virial[a] = 0.5 * vir[a];
}

/* cov_pp / cov_pp_der with d = 3, q = 1 (gap.f90:1022-1081): j = 3,
 * (1 - r)^4 (4 r + 1) on (0, 1), 1 below, 0 above; the derivative holds the
 * 1/r factor. */
static double cov_pp31(double r) {
  if (r >= 1.0) return 0.0;
  if (r <= 0.0) return 1.0;
  return pow(1.0 - r, 4) * (4.0 * r + 1.0);
}
static double cov_pp31_der(double r) {
  if (r >= 1.0 || r <= 0.0) return 0.0;
  return (-4.0 * pow(1.0 - r, 3) * (4.0 * r + 1.0) + pow(1.0 - r, 4) * 4.0) / r;
}

/* gap.f90:756-1015 with buffer 0.5 and e0 0 (the call at :730-751): for a
 * centre of species sp0 and every pair of its neighbours (list order) of
 * species (sp1, sp2) in either order within rcut, q = (r12 + r13,
 * (r12 - r13)^2, r23), kernel pref_s k(|q - Q_s|_sigma), pref = 2 delta^2
 * alpha c, both cosine cutoffs; forces on the three atoms and the virial
 * sym(F2 x12 + F3 x13)/2. */
void orc_angle_3b_ef(const orc_3b* b, int n, const int* species, const int* n_neigh,
                     const int* list, const double* rjs, const double* xyz, int do_forces,
                     double* energies, double* forces, double* virial) {
  const double pi = acos(-1.0), buffer = 0.5, rcut = b->rcut;
  const int S = b->n_sparse;
  const double* Q1 = b->Qs;
  const double* Q2 = b->Qs + S;
  const double* Q3 = b->Qs + 2 * S;
  double* pref = (double*)malloc(sizeof(double) * (size_t)S);
  double* kern = (double*)malloc(sizeof(double) * (size_t)S);
  double* kder = (double*)malloc(sizeof(double) * (size_t)S);
  double* rr = (double*)malloc(sizeof(double) * (size_t)S);
  double* dq = (double*)malloc(sizeof(double) * 3 * (size_t)S);
  for (int s = 0; s < S; s++) pref[s] = 2.0 * b->delta * b->delta * b->alphas[s] * b->cutoff[s];
  for (int i = 0; i < n; i++) energies[i] = 0.0;
  if (do_forces) {
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(virial, 0, sizeof(double) * 9);
  }
  int64_t k = 0;
  for (int i = 0; i < n; i++) {
    if (species[i] != b->sp0) {
      k += n_neigh[i];
      continue;
    }
    int i3 = (list[k] - 1) % n;
    k++;
    for (int j = 1; j < n_neigh[i]; j++) {
      int64_t kj = k + j - 1;
      double r12 = rjs[kj];
      int sj = species[(list[kj] - 1) % n];
      if ((sj != b->sp1 && sj != b->sp2) || r12 > rcut) continue;
      int j3 = (list[kj] - 1) % n;
      const double* x12 = xyz + 3 * kj;
      double u12[3] = {x12[0] / r12, x12[1] / r12, x12[2] / r12};
      for (int j2 = j + 1; j2 < n_neigh[i]; j2++) {
        int64_t kk = k + j2 - 1;
        double r13 = rjs[kk];
        int sk = species[(list[kk] - 1) % n];
        if (!(((sj == b->sp1 && sk == b->sp2) || (sj == b->sp2 && sk == b->sp1)) && r13 < rcut))
          continue;
        int k3 = (list[kk] - 1) % n;
        const double* x13 = xyz + 3 * kk;
        double u13[3] = {x13[0] / r13, x13[1] / r13, x13[2] / r13};
        double x23[3] = {x13[0] - x12[0], x13[1] - x12[1], x13[2] - x12[2]};
        double r23 = sqrt(x23[0] * x23[0] + x23[1] * x23[1] + x23[2] * x23[2]);
        double u23[3] = {x23[0] / r23, x23[1] / r23, x23[2] / r23};
        if (!(r12 < rcut && r13 < rcut)) continue;
        double fc12, fc13, df12[3], df13[3];
        if (r12 < rcut - buffer) {
          fc12 = 1.0;
          df12[0] = df12[1] = df12[2] = 0.0;
        } else {
          fc12 = (cos(pi * (r12 - rcut + buffer) / buffer) + 1.0) / 2.0;
          for (int a = 0; a < 3; a++)
            df12[a] = sin(pi * (r12 - rcut + buffer) / buffer) / 2.0 * pi / buffer * u12[a];
        }
        if (r13 < rcut - buffer) {
          fc13 = 1.0;
          df13[0] = df13[1] = df13[2] = 0.0;
        } else {
          fc13 = (cos(pi * (r13 - rcut + buffer) / buffer) + 1.0) / 2.0;
          for (int a = 0; a < 3; a++)
            df13[a] = sin(pi * (r13 - rcut + buffer) / buffer) / 2.0 * pi / buffer * u13[a];
        }
        double fcut = fc12 * fc13, dfc[3];
        for (int a = 0; a < 3; a++) dfc[a] = fc12 * df13[a] + fc13 * df12[a];
        double q[3] = {r12 + r13, (r12 - r13) * (r12 - r13), r23};
        for (int s = 0; s < S; s++) {
          double d1 = q[0] - Q1[s], d2 = q[1] - Q2[s], d3 = q[2] - Q3[s];
          double t = d1 * d1 / (b->sigma[0] * b->sigma[0]);
          t = t + d2 * d2 / (b->sigma[1] * b->sigma[1]);
          t = t + d3 * d3 / (b->sigma[2] * b->sigma[2]);
          rr[s] = sqrt(t);
          dq[3 * s] = d1 / (b->sigma[0] * b->sigma[0]);
          dq[3 * s + 1] = d2 / (b->sigma[1] * b->sigma[1]);
          dq[3 * s + 2] = d3 / (b->sigma[2] * b->sigma[2]);
          if (b->kernel == 0) {
            kern[s] = pref[s] * exp(-0.5 * rr[s] * rr[s]);
          } else {
            kern[s] = pref[s] * cov_pp31(rr[s]);
            kder[s] = pref[s] * cov_pp31_der(rr[s]);
          }
        }
        double ks = 0.0;
        for (int s = 0; s < S; s++) ks += kern[s];
        energies[i] += fcut * ks;
        if (!do_forces) continue;
        double F1[3], F2[3], F3[3];
        for (int l = 0; l < 3; l++) {
          double s1 = 0.0, s2 = 0.0, s3 = 0.0;
          for (int s = 0; s < S; s++) {
            double dx1 = dq[3 * s] * (u12[l] + u13[l]) +
                         dq[3 * s + 1] * 2.0 * (r12 - r13) * (u12[l] - u13[l]);
            double dx2 = -dq[3 * s] * u12[l] - dq[3 * s + 1] * 2.0 * (r12 - r13) * u12[l] +
                         dq[3 * s + 2] * u23[l];
            double dx3 = -dq[3 * s] * u13[l] - dq[3 * s + 1] * 2.0 * (r13 - r12) * u13[l] -
                         dq[3 * s + 2] * u23[l];
            if (b->kernel == 0) {
              s1 += kern[s] * (dfc[l] + fcut * dx1);
              s2 += kern[s] * (-fc13 * df12[l] + fcut * dx2);
              s3 += kern[s] * (-fc12 * df13[l] + fcut * dx3);
            } else {
              s1 += kern[s] * dfc[l] - kder[s] * fcut * dx1;
              s2 += -kern[s] * fc13 * df12[l] - kder[s] * fcut * dx2;
              s3 += -kern[s] * fc12 * df13[l] - kder[s] * fcut * dx3;
            }
          }
          F1[l] = -s1;
          F2[l] = -s2;
          F3[l] = -s3;
        }
        for (int a = 0; a < 3; a++) {
          forces[3 * i3 + a] += F1[a];
          forces[3 * j3 + a] += F2[a];
          forces[3 * k3 + a] += F3[a];
        }
        for (int k1 = 0; k1 < 3; k1++)
          for (int k4 = 0; k4 < 3; k4++) {
            virial[k4 * 3 + k1] += 0.5 * (F2[k1] * x12[k4] + F2[k4] * x12[k1]);
            virial[k4 * 3 + k1] += 0.5 * (F3[k1] * x13[k4] + F3[k4] * x13[k1]);
          }
      }
    }
    k += n_neigh[i] - 1;
  }
  free(pref);
  free(kern);
  free(kder);
  free(rr);
  free(dq);
}

/* ------------------------------------------------------------------------ */
/* Tkatchenko-Scheffler dispersion                                          */
/* ------------------------------------------------------------------------ */

/* get_ts_energy_and_forces (vdw/vdw_interface.f90:127-385) with its
 * conventions kept:
 *  - per pair k of row i (j >= 2): C6_ij = 2 C6_i C6_j / (a_j/a_i C6_i +
 *    a_i/a_j C6_j) from C6 = c6_ref V^2, a = alpha0_ref V, R0_ij = R0_i + R0_j
 *    with R0 = r0_ref V^(1/3) (:235-276); zero when either C6 is zero;
 *  - r^-6 times the cubic switch 1 - 3 x^2 + 2 x^3 in the outer buffer
 *    (rcut - buffer, rcut) and the inner one (rcut_inner, rcut_inner +
 *    buffer_inner) (:212-226); its derivative is r6' of the unswitched
 *    r^-6 plus r^-6 times the switch's slope (:199-211);
 *  - E_i = -1/2 sum_j C6_ij r6 f_damp over rcut_inner < r < rcut, f_damp =
 *    1/(1 + exp(-d (r/(sR R0_ij) - 1))) (:287-300);
 *  - the Hirshfeld term: dE_i/dV_i prefactors (:314-339) times dV_i/dr_j
 *    for every entry of row i with rcut_inner < r < rcut, so the self entry
 *    (r = 0) takes no part (:345-363), virial + 1/2 sym(F x);
 *  - the pair force on i from each entry (j not i itself), virial - 1/4
 *    sym(F x) (:365-382).
 * The pair parameters are formed for pairs where they take part only, so
 * species with zero references yield zeros where the reference's 0/0 would
 * give NaN. */
void orc_ts_ef(const orc_ts* t, int n, const int* species, const int* n_neigh, const int* list,
               const double* rjs, const double* xyz, const double* V, const double* V_der,
               int do_forces, double* energies, double* forces, double* virial) {
  const double rc = t->rcut, b = t->buffer, rci = t->rcut_inner, bi = t->buffer_inner;
  double* pref = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  if (do_forces) {
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(virial, 0, sizeof(double) * 9);
  }
  int64_t k = 0;
  for (int i = 0; i < n; i++) {
    const int si = species[i] - 1;
    const double vi = V[i];
    const double c6i = t->c6_ref[si] * vi * vi;
    const double r0i = t->r0_ref[si] * pow(vi, 1.0 / 3.0);
    const double ai = t->alpha0_ref[si] * vi;
    double e = 0.0, p1 = 0.0, p2 = 0.0;
    for (int jj = 1; jj < n_neigh[i]; jj++) {
      const int64_t kk = k + jj;
      const double r = rjs[kk];
      if (!(r > rci && r < rc)) continue;
      const int j = (list[kk] - 1) % n;
      const int sj = species[j] - 1;
      const double vj = V[j];
      const double c6j = t->c6_ref[sj] * vj * vj;
      const double aj = t->alpha0_ref[sj] * vj;
      const double c6ij =
          (c6i == 0.0 || c6j == 0.0) ? 0.0 : 2.0 * c6i * c6j / (aj / ai * c6i + ai / aj * c6j);
      const double r0ij = r0i + t->r0_ref[sj] * pow(vj, 1.0 / 3.0);
      const double r2 = r * r;
      double r6 = 1.0 / (r2 * r2 * r2);
      double r6d = -6.0 / r * r6;
      if ((b > 0.0 || bi > 0.0) && ((r > rc - b && r < rc) || (r < rci + bi && r > rci))) {
        double x, dsw;
        if (r <= rci + bi) {
          x = (rci + bi - r) / bi;
          dsw = 6.0 * r6 / (-bi) * (-x + x * x);
        } else {
          x = (r - rc + b) / b;
          dsw = 6.0 * r6 / b * (-x + x * x);
        }
        r6d += dsw;
        r6 *= 1.0 - 3.0 * x * x + 2.0 * x * x * x;
      }
      if (c6ij == 0.0) continue;
      const double ed = exp(-t->d * (r / (t->sR * r0ij) - 1.0));
      const double fd = 1.0 / (1.0 + ed);
      e += c6ij * r6 * fd;
      if (!do_forces) continue;
      p1 += c6ij * fd * r6;
      p2 -= c6ij * fd * fd * r * r6 * ed * t->d / t->sR / (r0ij * r0ij);
      const double s = c6ij / r * fd * (-r6d - r6 * fd * ed * t->d / t->sR / r0ij);
      const double f[3] = {s * xyz[3 * kk], s * xyz[3 * kk + 1], s * xyz[3 * kk + 2]};
      if (j != i) {
        forces[3 * i] += f[0];
        forces[3 * i + 1] += f[1];
        forces[3 * i + 2] += f[2];
      }
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++)
          virial[3 * c + a] -= 0.25 * (f[a] * xyz[3 * kk + c] + f[c] * xyz[3 * kk + a]);
    }
    energies[i] = -0.5 * e;
    if (do_forces)
      pref[i] = vi == 0.0 ? 0.0 : p1 / vi + p2 * t->r0_ref[si] / 3.0 / pow(vi, 2.0 / 3.0);
    k += n_neigh[i];
  }
  if (do_forces) {
    k = 0;
    for (int i = 0; i < n; i++) {
      for (int jj = 0; jj < n_neigh[i]; jj++) {
        const int64_t kk = k + jj;
        const double r = rjs[kk];
        const double* g = V_der + 3 * kk;
        if (!(r > rci && r < rc) || pref[i] == 0.0) continue;
        if (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0) continue;
        const int j = (list[kk] - 1) % n;
        const double f[3] = {g[0] * pref[i], g[1] * pref[i], g[2] * pref[i]};
        forces[3 * j] += f[0];
        forces[3 * j + 1] += f[1];
        forces[3 * j + 2] += f[2];
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++)
            virial[3 * c + a] += 0.5 * (f[a] * xyz[3 * kk + c] + f[c] * xyz[3 * kk + a]);
      }
      k += n_neigh[i];
    }
  }
  free(pref);
}

/* ------------------------------------------------------------------------ */
/* Whole path                                                               */
/* ------------------------------------------------------------------------ */

/* The SOAP term of one soap_turbo block on a full neighbour list, added to
 * e_atom, forces and virial (get_gap_soap + get_soap_energy_and_forces as
 * calculate_gap_soap.f90:359-447 call them; one chunk of chunk_sites sites
 * at a time, the reference's memory split). */
static void soap_block_ef(const orc_model* m, int n, const int* species, const int* nn,
                          const int* list, const double* rjs, const double* xyz,
                          int do_forces, int chunk_sites, double* e_atom, double* forces,
                          double* virial, double* tf) {
  double tv[9];
  site_pairs sp;
  select_sites(&m->h, n, species, nn, list, rjs, xyz, &sp);
  soap_geom g;
  soap_geom_init(&m->h, &g);
  int dim = g.dim;
  int cs = chunk_sites > 0 ? chunk_sites : (sp.n_sites > 0 ? sp.n_sites : 1);
  for (int s0 = 0; s0 < sp.n_sites; s0 += cs) {
    int s1 = s0 + cs < sp.n_sites ? s0 + cs : sp.n_sites;
    int ns = s1 - s0;
    int64_t p0 = sp.site_off[s0], p1 = sp.site_off[s1];
    double* soap = (double*)malloc(sizeof(double) * (size_t)ns * dim);
    double* der = do_forces ? (double*)malloc(sizeof(double) * (size_t)(p1 - p0) * dim * 3) : NULL;
    soap_chunk(&m->h, &g, &sp, species, s0, s1, soap, der);
    double* en = (double*)malloc(sizeof(double) * (size_t)ns);
    orc_gap_soap_ef(ns, dim, m->n_sparse, soap, der, m->alphas, m->delta, m->zeta, 0.0,
                    m->Qs, sp.site_nn + s0, sp.pair_atom + p0, sp.pair_xyz + 3 * p0, n,
                    do_forces, en, tf, tv);
    for (int i = 0; i < ns; i++) e_atom[sp.site_atom[s0 + i]] += en[i];
    if (do_forces) {
      for (size_t a = 0; a < (size_t)3 * n; a++) forces[a] += tf[a];
      for (int a = 0; a < 9; a++) virial[a] += tv[a];
    }
    free(en);
    free(soap);
    free(der);
  }
  free_sites(&sp);
}

int orc_energy_forces(const orc_model* m, int n, const double* pos,
                      const int* species, const double* lat, int do_forces,
                      int chunk_sites, double* e_atom, double* forces,
                      double* virial) {
  double rc = 0.0;
  if (m->has_soap && hypers_supported(&m->h) != TGX_OK) return TGX_ERR_UNSUPPORTED;
  if (m->has_soap) rc = soap_rcut_max(&m->h);
  for (int k = 0; m->has_soap && k < m->n_more; k++) {
    if (hypers_supported(&m->more[k].h) != TGX_OK) return TGX_ERR_UNSUPPORTED;
    const double r = soap_rcut_max(&m->more[k].h);
    rc = r > rc ? r : rc;
  }
  for (int b = 0; b < m->n2b; b++) rc = m->b2[b].rcut > rc ? m->b2[b].rcut : rc;
  for (int b = 0; b < m->n3b; b++) rc = m->b3[b].rcut > rc ? m->b3[b].rcut : rc;
  for (int b = 0; b < m->ncore; b++)
    for (int a = 0; a < m->core[b].n; a++) rc = m->core[b].x[a] > rc ? m->core[b].x[a] : rc;
  int64_t np = 0;
  int r0 = orc_neighbors(n, pos, lat, rc, 0, 0, NULL, NULL, NULL, NULL, NULL, NULL, &np, NULL);
  if (r0 != TGX_OK && r0 != TGX_ERR_CAPACITY) return r0;
  int* nn = (int*)malloc(sizeof(int) * (size_t)n);
  int* list = (int*)malloc(sizeof(int) * (size_t)np);
  double* rjs = (double*)malloc(sizeof(double) * (size_t)np);
  double* xyz = (double*)malloc(sizeof(double) * 3 * (size_t)np);
  orc_neighbors(n, pos, lat, rc, 0, np, nn, list, rjs, xyz, NULL, NULL, &np, NULL);

  for (int i = 0; i < n; i++) e_atom[i] = m->e0[species[i] - 1];
  if (do_forces) {
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(virial, 0, sizeof(double) * 9);
  }
  double* tf = (double*)malloc(sizeof(double) * 3 * (size_t)n);
  double tv[9];
  if (m->has_soap) {
    /* the soap_turbo blocks in turn (calculate_gap_soap.f90:324-447) */
    soap_block_ef(m, n, species, nn, list, rjs, xyz, do_forces, chunk_sites, e_atom, forces,
                  virial, tf);
    for (int k = 0; k < m->n_more; k++)
      soap_block_ef(&m->more[k], n, species, nn, list, rjs, xyz, do_forces, chunk_sites, e_atom,
                    forces, virial, tf);
  }
  double* te = (double*)malloc(sizeof(double) * (size_t)n);
  for (int b = 0; b < m->n2b; b++) {
    orc_gap_2b_ef(&m->b2[b], n, species, nn, list, rjs, xyz, do_forces, te, tf, tv);
    for (int i = 0; i < n; i++) e_atom[i] += te[i];
    if (do_forces) {
      for (size_t a = 0; a < (size_t)3 * n; a++) forces[a] += tf[a];
      for (int a = 0; a < 9; a++) virial[a] += tv[a];
    }
  }
  /* angle_3b then core_pot, as turbogap_main.f90:809-840 calls them */
  for (int b = 0; b < m->n3b + m->ncore; b++) {
    if (b < m->n3b)
      orc_angle_3b_ef(&m->b3[b], n, species, nn, list, rjs, xyz, do_forces, te, tf, tv);
    else
      orc_core_pot_ef(&m->core[b - m->n3b], n, species, nn, list, rjs, xyz, do_forces, te, tf,
                      tv);
    for (int i = 0; i < n; i++) e_atom[i] += te[i];
    if (do_forces) {
      for (size_t a = 0; a < (size_t)3 * n; a++) forces[a] += tf[a];
      for (int a = 0; a < 9; a++) virial[a] += tv[a];
    }
  }
  free(te);
  free(tf);
  free(nn);
  free(list);
  free(rjs);
  free(xyz);
  return TGX_OK;
}

/* ------------------------------------------------------------------------ */
/* MD / MC                                                                  */
/* ------------------------------------------------------------------------ */

void orc_velocity_verlet(int n, double* pos, double* pos_prev, double* vel,
                         const double* forces, double* forces_prev,
                         const double* masses, double dt, double* dt_prev,
                         int first_step) {
  if (!first_step)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < 3; j++)
        vel[3 * i + j] = vel[3 * i + j] + 0.5 * (forces[3 * i + j] + forces_prev[3 * i + j]) / masses[i] * *dt_prev;
  memcpy(pos_prev, pos, sizeof(double) * 3 * (size_t)n);
  memcpy(forces_prev, forces, sizeof(double) * 3 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 3; j++)
      pos[3 * i + j] = pos[3 * i + j] + vel[3 * i + j] * dt + 0.5 * forces[3 * i + j] / masses[i] * (dt * dt);
  *dt_prev = dt;
}

double orc_mc_accept(double e_new, double e_prev, double temp) {
  /* kB is a default-real literal in the reference (mc_utils.f90:176). */
  const double kB = (double)8.617333262e-5f;
  double p = exp(-(e_new - e_prev) / (kB * temp));
  return p < 1.0 ? p : 1.0;
}

/* u(seed, t, k) in [0, 1): the top 53 bits of the splitmix64 finaliser
 * applied to seed + 0x9E3779B97F4A7C15 * (6t + k + 1). */
double orc_mc_uniform(uint64_t seed, int64_t t, int k) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(6 * t + k + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static double orc_total_energy(const orc_model* m, int n, const double* pos,
                               const int* species, const double* lat,
                               double* e_atom, int* rc) {
  double e = 0.0;
  *rc = orc_energy_forces(m, n, pos, species, lat, 0, 0, e_atom, NULL, NULL);
  for (int i = 0; i < n; i++) e += e_atom[i];
  return e;
}

/* One Markov chain of single-atom moves on any energy function: the part of
 * orc_mc_run / orc_mc_run_da that decides moves, shared with the toy-
 * potential chains that pin it by sampling statistics.
 *   trial t: atom floor(u0 n), disp = move_max (u1..u3 - 1/2), positions +
 *   disp (mc_utils.f90:267-273, 696-698).
 *   plain (sur == NULL): E of the whole trial structure, p = min(1,
 *   exp(-(E_trial - E_current)/(kB T))), accepted when p > u4
 *   (mc_interface.f90:245-299, mc_utils.f90:172-180).
 *   delayed acceptance (Christen & Fox 2005): stage 1 accepts with
 *   min(1, exp(-(S_trial - S_current)/(kB T))) > u4, stage 2 with
 *   min(1, exp(-((E_trial - E_current) - (S_trial - S_current))/(kB T)))
 *   > u5.  Each stage-2 trial updates ema = 0.98 ema + 0.02 d^2 with
 *   d = (E_trial - E_current) - (S_trial - S_current); once ema exceeds
 *   max_rmse^2 (max_rmse >= 0) the following trials are plain Metropolis
 *   (u4), and *fallback_at = the first such trial (-1: none).
 *   variant 1 (tests only): stage 2 with the plain ratio
 *   min(1, exp(-(E_trial - E_current)/(kB T))), which does not preserve the
 *   Boltzmann distribution -- the sampling tests must reject it.
 * e_trace[0] = initial energy, e_trace[t+1] = current energy after trial t. */
typedef double (*orc_efn)(void* ctx, int n, const double* pos, int* rc);

static int mc_chain(int n, double* pos, int n_trials, double temp, double move_max,
                    uint64_t seed, orc_efn E, void* ectx, orc_efn S, void* sctx,
                    double max_rmse, int variant, double* e_trace, int* n_accepted,
                    int* n_rej1, int* fallback_at) {
  int rc = 0, acc = 0, r1 = 0, fb = -1;
  const double kB = (double)8.617333262e-5f;
  double ema = 0.0;
  double e_cur = E(ectx, n, pos, &rc);
  double s_cur = (rc == 0 && S) ? S(sctx, n, pos, &rc) : 0.0;
  if (e_trace) e_trace[0] = e_cur;
  for (int t = 0; t < n_trials && rc == 0; t++) {
    int idx = (int)floor(orc_mc_uniform(seed, t, 0) * n);
    if (idx > n - 1) idx = n - 1;
    double old[3];
    for (int k = 0; k < 3; k++) {
      const double disp = move_max * (orc_mc_uniform(seed, t, 1 + k) - 0.5);
      old[k] = pos[3 * idx + k];
      pos[3 * idx + k] = old[k] + disp;
    }
    const int da = S != NULL && fb < 0;
    int accept = 0;
    double e_new = e_cur, s_new = s_cur;
    if (da) {
      s_new = S(sctx, n, pos, &rc);
      double p1 = exp(-(s_new - s_cur) / (kB * temp));
      p1 = p1 < 1.0 ? p1 : 1.0;
      if (p1 > orc_mc_uniform(seed, t, 4)) {
        e_new = E(ectx, n, pos, &rc);
        const double d = (e_new - e_cur) - (s_new - s_cur);
        double p2 = exp(-(variant == 1 ? e_new - e_cur : d) / (kB * temp));
        p2 = p2 < 1.0 ? p2 : 1.0;
        accept = p2 > orc_mc_uniform(seed, t, 5);
        ema = 0.98 * ema + 0.02 * d * d;
        if (max_rmse >= 0.0 && ema > max_rmse * max_rmse) fb = t + 1;
      } else {
        r1++;
      }
    } else {
      e_new = E(ectx, n, pos, &rc);
      accept = orc_mc_accept(e_new, e_cur, temp) > orc_mc_uniform(seed, t, 4);
    }
    if (accept) {
      e_cur = e_new;
      acc++;
      if (da) s_cur = s_new;
    } else {
      for (int k = 0; k < 3; k++) pos[3 * idx + k] = old[k];
    }
    /* after the switch the surrogate is not evaluated; re-seed its current
     * value if it is ever needed again (it is not: the switch is final) */
    if (e_trace) e_trace[t + 1] = e_cur;
  }
  if (n_accepted) *n_accepted = acc;
  if (n_rej1) *n_rej1 = r1;
  if (fallback_at) *fallback_at = fb;
  return rc;
}

typedef struct {
  const orc_model* m;
  const int* species;
  const double* lat;
  double* ea;
  const double* P;
  double mu, sigma;
} gap_chain_ctx;

static double gap_chain_energy(void* ctx, int n, const double* pos, int* rc) {
  gap_chain_ctx* g = (gap_chain_ctx*)ctx;
  return orc_total_energy(g->m, n, pos, g->species, g->lat, g->ea, rc);
}

static double orc_surrogate_energy(const orc_model* m, int n, const double* pos,
                                   const int* species, const double* lat,
                                   const double* P, double mu, double sigma,
                                   int* rc);

static double gap_chain_surrogate(void* ctx, int n, const double* pos, int* rc) {
  gap_chain_ctx* g = (gap_chain_ctx*)ctx;
  return orc_surrogate_energy(g->m, n, pos, g->species, g->lat, g->P, g->mu, g->sigma, rc);
}

/* Sequential Metropolis with one whole-structure GAP energy per trial
 * (mc_interface.f90:183-447): mc_chain without a surrogate. */
int orc_mc_run(const orc_model* m, int n, double* pos, const int* species,
               const double* lat, int n_trials, double temp, double move_max,
               uint64_t seed, double* e_trace, int* n_accepted) {
  gap_chain_ctx g = {m, species, lat, (double*)malloc(sizeof(double) * (size_t)n), NULL, 0, 1};
  int rc = mc_chain(n, pos, n_trials, temp, move_max, seed, gap_chain_energy, &g, NULL, NULL,
                    -1.0, 0, e_trace, n_accepted, NULL, NULL);
  free(g.ea);
  return rc;
}

/* Toy potentials for the sampling tests of mc_chain: n (1 or 2) particles,
 * only their x coordinates enter,
 *   V = sum_i a (x_i^2 - 1)^2 + c x_0 x_1 (n = 2),
 * and the surrogate V + eps sum_i sin(w x_i) (inexact, not flat).
 * prm = {a, c, eps, w}. */
static double toy_v(void* ctx, int n, const double* pos, int* rc) {
  const double* p = (const double*)ctx;
  double v = 0.0;
  for (int i = 0; i < n; i++) {
    const double x = pos[3 * i];
    v += p[0] * (x * x - 1.0) * (x * x - 1.0);
  }
  if (n == 2) v += p[1] * pos[0] * pos[3];
  *rc = 0;
  return v;
}
static double toy_s(void* ctx, int n, const double* pos, int* rc) {
  const double* p = (const double*)ctx;
  double v = toy_v(ctx, n, pos, rc);
  for (int i = 0; i < n; i++) v += p[2] * sin(p[3] * pos[3 * i]);
  return v;
}

/* n_chains independent chains of the toy potential, chain k from x0[k*n ..]
 * with seed + k, n_steps trials each; x_end receives the final x
 * coordinates.  use_sur: delayed acceptance with the toy surrogate
 * (no fallback); variant as mc_chain. */
int orc_mc_toy(int n, const double* prm, int n_chains, int n_steps, double temp,
               double move_max, uint64_t seed, int use_sur, int variant, const double* x0,
               double* x_end, int64_t* n_accepted, int64_t* n_rej1) {
  if (n < 1 || n > 2) return TGX_ERR_ARG;
  int64_t acc = 0, r1 = 0;
  int rc = 0;
#pragma omp parallel for reduction(+ : acc, r1) schedule(static)
  for (int k = 0; k < n_chains; k++) {
    double pos[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) pos[3 * i] = x0[(size_t)k * n + i];
    int a = 0, b = 0;
    int r = mc_chain(n, pos, n_steps, temp, move_max, seed + (uint64_t)k, toy_v, (void*)prm,
                     use_sur ? toy_s : NULL, (void*)prm, -1.0, variant, NULL, &a, &b, NULL);
    if (r) rc = r;
    for (int i = 0; i < n; i++) x_end[(size_t)k * n + i] = pos[3 * i];
    acc += a;
    r1 += b;
  }
  if (n_accepted) *n_accepted = acc;
  if (n_rej1) *n_rej1 = r1;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Surrogate MLP                                                            */
/* ------------------------------------------------------------------------ */

static double orc_bf16(double x, int round) {
  if (!round) return x;
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}
static double orc_sig(double x) { return 1.0 / (1.0 + exp(-x)); }

/* e(q) = mu + sigma (w3 . s(W2 s(W1 q + b1) + b2) + b3), s(x) = x sigmoid(x);
 * P = [W1 (H x D), b1, W2 (H x H), b2, w3, b3].  round_bf16: q, W1, W2 enter
 * rounded to bfloat16 (round to nearest even), as the device kernels use
 * them.  grad (may be NULL): d/dP of L = (1/n) sum (e - y)^2 / sigma^2 with
 * the rounding taken as the identity. */
double orc_mlp(int n, int D, int H, const double* P, const double* q, double mu,
               double sigma, int round_bf16, const double* y, double* e,
               double* grad) {
  const double* W1 = P;
  const double* b1 = W1 + (size_t)H * D;
  const double* W2 = b1 + H;
  const double* b2 = W2 + (size_t)H * H;
  const double* w3 = b2 + H;
  const double b3 = w3[H];
  const size_t NP = (size_t)H * D + (size_t)H * H + 3 * (size_t)H + 1;
  double* z1 = (double*)malloc(sizeof(double) * H);
  double* z2 = (double*)malloc(sizeof(double) * H);
  double* h1 = (double*)malloc(sizeof(double) * H);
  double* d2 = (double*)malloc(sizeof(double) * H);
  double* qr = (double*)malloc(sizeof(double) * D);
  double loss = 0.0;
  if (grad) memset(grad, 0, sizeof(double) * NP);
  for (int s = 0; s < n; s++) {
    for (int k = 0; k < D; k++) qr[k] = orc_bf16(q[(size_t)s * D + k], round_bf16);
    for (int u = 0; u < H; u++) {
      double a = 0.0;
      for (int k = 0; k < D; k++) a += orc_bf16(W1[(size_t)u * D + k], round_bf16) * qr[k];
      z1[u] = a + b1[u];
      h1[u] = z1[u] * orc_sig(z1[u]);
    }
    double out = b3;
    for (int u = 0; u < H; u++) {
      double a = 0.0;
      for (int k = 0; k < H; k++) a += orc_bf16(W2[(size_t)u * H + k], round_bf16) * h1[k];
      z2[u] = a + b2[u];
      out += w3[u] * z2[u] * orc_sig(z2[u]);
    }
    const double es = mu + sigma * out;
    if (e) e[s] = es;
    if (y) {
      const double r = (es - y[s]) / sigma;
      loss += r * r / n;
      if (grad) {
        const double g = 2.0 * r / sigma / n * sigma; /* dL/d(out) */
        double* gW1 = grad;
        double* gb1 = gW1 + (size_t)H * D;
        double* gW2 = gb1 + H;
        double* gb2 = gW2 + (size_t)H * H;
        double* gw3 = gb2 + H;
        gw3[H] += g;
        for (int u = 0; u < H; u++) {
          const double sg = orc_sig(z2[u]);
          gw3[u] += g * z2[u] * sg;
          d2[u] = g * w3[u] * sg * (1.0 + z2[u] * (1.0 - sg));
          gb2[u] += d2[u];
          for (int k = 0; k < H; k++) gW2[(size_t)u * H + k] += d2[u] * h1[k];
        }
        for (int k = 0; k < H; k++) {
          double a = 0.0;
          for (int u = 0; u < H; u++) a += orc_bf16(W2[(size_t)u * H + k], round_bf16) * d2[u];
          const double sg = orc_sig(z1[k]);
          const double d1 = a * sg * (1.0 + z1[k] * (1.0 - sg));
          gb1[k] += d1;
          for (int j = 0; j < D; j++) gW1[(size_t)k * D + j] += d1 * qr[j];
        }
      }
    }
  }
  free(z1);
  free(z2);
  free(h1);
  free(d2);
  free(qr);
  return loss;
}

/* Surrogate energy of a whole structure: the MLP on every site's SOAP plus
 * the exact 2-body terms (the parts delayed-acceptance stage 1 sees). */
static double orc_surrogate_energy(const orc_model* m, int n, const double* pos,
                                   const int* species, const double* lat,
                                   const double* P, double mu, double sigma,
                                   int* rc) {
  const int H = 128;
  soap_geom g;
  soap_geom_init(&m->h, &g);
  double* q = (double*)malloc(sizeof(double) * (size_t)n * g.dim);
  int n_sites = 0;
  int64_t n_pairs = 0;
  *rc = orc_soap(&m->h, n, pos, lat, species, q, NULL, 0, &n_sites, &n_pairs, NULL, NULL,
                 NULL, NULL);
  double e = 0.0;
  if (*rc == 0) {
    double* es = (double*)malloc(sizeof(double) * (size_t)n_sites);
    orc_mlp(n_sites, g.dim, H, P, q, mu, sigma, 1, NULL, es, NULL);
    for (int i = 0; i < n_sites; i++) e += es[i];
    free(es);
    /* 2-body part: the whole-path energy without the SOAP term (minus e0) */
    if (m->n2b > 0) {
      orc_model m2 = *m;
      m2.has_soap = 0;
      double* ea = (double*)malloc(sizeof(double) * (size_t)n);
      *rc = orc_energy_forces(&m2, n, pos, species, lat, 0, 0, ea, NULL, NULL);
      for (int i = 0; i < n; i++) e += ea[i] - m->e0[species[i] - 1];
      free(ea);
    }
  }
  free(q);
  return e;
}

/* Delayed-acceptance Metropolis with a fixed surrogate (P, mu, sigma): the
 * MLP on every site's SOAP plus the exact 2-body terms as stage 1, the GAP
 * as stage 2, whole-structure energies per trial; mc_chain with the fallback
 * to plain Metropolis once the running mean square of the stage-2 energy
 * difference exceeds max_rmse^2 (max_rmse < 0: never). */
int orc_mc_run_da(const orc_model* m, int n, double* pos, const int* species,
                  const double* lat, int n_trials, double temp, double move_max,
                  uint64_t seed, const double* P, double mu, double sigma,
                  double max_rmse, double* e_trace, int* n_accepted, int* n_rej1,
                  int* fallback_at) {
  gap_chain_ctx g = {m, species, lat, (double*)malloc(sizeof(double) * (size_t)n), P, mu, sigma};
  int rc = mc_chain(n, pos, n_trials, temp, move_max, seed, gap_chain_energy, &g,
                    gap_chain_surrogate, &g, max_rmse, 0, e_trace, n_accepted, n_rej1,
                    fallback_at);
  free(g.ea);
  return rc;
}
