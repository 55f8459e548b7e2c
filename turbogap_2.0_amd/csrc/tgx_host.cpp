// Host-side model math of the tgx library: the poly3 orthonormalisation
// matrix W = S^(-1/2) (computed in binary128 and rounded once, because
// cond(S) ~ 1e13 at n_max = 8), the FP64 radial closed form used for the
// central-atom term, spherical-harmonic recurrence constants and the cell.
#include <quadmath.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "tgx_internal.h"

namespace tgx {

namespace {

typedef __float128 qd;

qd qsqrt(qd v) {
  if (v <= 0) return 0;
  qd x = (qd)std::sqrt((double)v);
  for (int i = 0; i < 5; i++) x = (x + v / x) / 2;
  return x;
}

// Symmetric eigen-decomposition A = V diag(w) V^T by cyclic Jacobi, binary128.
void eig_sym(int n, qd* a, qd* v, qd* w) {
  for (int i = 0; i < n * n; i++) v[i] = 0;
  for (int i = 0; i < n; i++) v[i * n + i] = 1;
  for (int sweep = 0; sweep < 64; sweep++) {
    qd off = 0, dg = 0;
    for (int i = 0; i < n; i++) {
      dg += a[i * n + i] * a[i * n + i];
      for (int j = i + 1; j < n; j++) off += a[i * n + j] * a[i * n + j];
    }
    if (off <= dg * (qd)1e-60 * (qd)1e-6) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const qd apq = a[p * n + q];
        if (apq == 0) continue;
        const qd th = (a[q * n + q] - a[p * n + p]) / (2 * apq);
        const qd t = (th >= 0 ? 1 : -1) / ((th >= 0 ? th : -th) + qsqrt(th * th + 1));
        const qd c = 1 / qsqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          const qd x = a[k * n + p], y = a[k * n + q];
          a[k * n + p] = c * x - s * y;
          a[k * n + q] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          const qd x = a[p * n + k], y = a[q * n + k];
          a[p * n + k] = c * x - s * y;
          a[q * n + k] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          const qd x = v[k * n + p], y = v[k * n + q];
          v[k * n + p] = c * x - s * y;
          v[k * n + q] = s * x + c * y;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = a[i * n + i];
}

}  // namespace

// Overlap of phi_a(r) = sqrt((2a+5)/rh^(2a+5)) (rh - r)^(a+2) on [0, rh]:
// S_ab = sqrt((2a+5)(2b+5))/(a+b+5) for a, b = 1..nmax; W = S^(-1/2).
void poly3_W(int nmax, double* W) {
  qd S[TGX_MAX_NR * TGX_MAX_NR], V[TGX_MAX_NR * TGX_MAX_NR], w[TGX_MAX_NR];
  for (int a = 0; a < nmax; a++)
    for (int b = 0; b < nmax; b++)
      S[a * nmax + b] = qsqrt((qd)((2 * a + 7) * (2 * b + 7))) / (qd)(a + b + 7);
  eig_sym(nmax, S, V, w);
  for (int i = 0; i < nmax; i++)
    for (int j = 0; j < nmax; j++) {
      qd acc = 0;
      for (int k = 0; k < nmax; k++) acc += V[i * nmax + k] * V[j * nmax + k] / qsqrt(w[k]);
      W[i * TGX_MAX_NR + j] = (double)acc;
    }
}

// The poly3gauss basis on [0, rh], every function normalised (weight 1):
// slot 0 the Gaussian g(r) = exp(-r^2/(2 sigma^2)), slot k = 1..nmax-1 the
// poly3 function (rh - r)^(k+2).  S_00 = 1, S_kl as poly3, and
// S_0k = N_g N_k int_0^rh g(r) (rh - r)^(k+2) dr, the integrals by the
// recurrence J_m = rh J_(m-1) + (m-1) sigma^2 J_(m-2) - sigma^2 rh^(m-1) of
// J_m = int_0^rh v^m exp(-(v - rh)^2/(2 sigma^2)) dv (v = rh - r).  All in
// binary128: cond(S) is about 1e13, so S must be exact far below FP64
// rounding for W = S^(-1/2) to be exact to FP64 rounding.
void poly3gauss_W(int nmax, double rh_in, double sigma_in, double* Sout, double* W) {
  qd S[TGX_MAX_NR * TGX_MAX_NR], V[TGX_MAX_NR * TGX_MAX_NR], w[TGX_MAX_NR];
  const qd rh = rh_in, sg = sigma_in, s2 = sg * sg, pi = acosq((qd)-1);
  qd J[TGX_MAX_NR + 8];
  J[0] = sg * sqrtq(pi / 2) * erfq(rh / (sg * sqrtq((qd)2)));
  J[1] = rh * J[0] + s2 * (expq(-rh * rh / (2 * s2)) - 1);
  qd rhp = 1;
  for (int m = 2; m < nmax + 4; m++) {
    rhp *= rh;
    J[m] = rh * J[m - 1] + (m - 1) * s2 * J[m - 2] - s2 * rhp;
  }
  const qd ng = 1 / sqrtq(sqrtq(pi) / 2 * sg * erfq(rh / sg));
  for (int a = 0; a < nmax; a++)
    for (int b = 0; b < nmax; b++) {
      qd v;
      if (a == 0 && b == 0) {
        v = 1;
      } else if (a == 0 || b == 0) {
        const int k = a + b;
        v = ng * sqrtq((qd)(2 * k + 5)) / powq(rh, (qd)(2 * k + 5) / 2) * J[k + 2];
      } else {
        v = sqrtq((qd)((2 * a + 5) * (2 * b + 5))) / (qd)(a + b + 5);
      }
      S[a * nmax + b] = v;
      if (Sout) Sout[a * nmax + b] = (double)v;
    }
  eig_sym(nmax, S, V, w);
  for (int i = 0; i < nmax; i++)
    for (int j = 0; j < nmax; j++) {
      qd acc = 0;
      for (int k = 0; k < nmax; k++) acc += V[i * nmax + k] * V[j * nmax + k] / sqrtq(w[k]);
      W[i * TGX_MAX_NR + j] = (double)acc;
    }
}

// FP64 closed form of R_n(r) (no envelope) for species s: see the device
// version in tgx_device.h; used on the host for the central-atom term.
void radial_closed_form(const SoapDev& sd, const double* W, int s, double r, double* R,
                        double* /*dR*/) {
  const int nmax = sd.nmax[s], p = sd.radial_enhancement;
  const double rh = sd.rh[s];
  const double sg = sd.sig_r[s] + sd.sig_rs[s] * r;
  const double c = rh - r, s2 = sg * sg;
  const double er = std::exp(-r * r / (2 * s2)), ec = std::exp(-c * c / (2 * s2));
  double J[TGX_MAX_NR + 8];
  J[0] = sg * std::sqrt(M_PI / 2) * (std::erf(r / (sg * std::sqrt(2.0))) + std::erf(c / (sg * std::sqrt(2.0))));
  J[1] = c * J[0] + s2 * (ec - er);
  double rhp = 1;
  for (int k = 2; k < nmax + 8; k++) {
    rhp *= rh;
    J[k] = c * J[k - 1] + (k - 1) * s2 * J[k - 2] - s2 * rhp * er;
  }
  const double pre = 1.0 / (std::sqrt(2 * M_PI) * sg);
  double I[TGX_MAX_NR];
  const int g = sd.basis == 1 ? 1 : 0;  // poly3gauss: slot 0 is the Gaussian
  for (int a = 1; a <= nmax - g; a++) {
    const int d = a + 2;
    const double Na = std::sqrt((2.0 * a + 5.0) / std::pow(rh, 2.0 * a + 5.0));
    double T;
    if (p == 0) T = J[d];
    else if (p == 1) T = rh * J[d] - J[d + 1];
    else T = rh * rh * J[d] - 2 * rh * J[d + 1] + J[d + 2];
    I[a - 1 + g] = Na * pre * T;
  }
  if (g) {
    // the Gaussian times the smeared density: exp(-r^2/(2(sb^2+s^2))) times
    // M_p = int_0^rh u^p exp(-(u-mu)^2/(2 t^2)) du (see the device version)
    const double sb2 = sd.sig_r[s] * sd.sig_r[s], Sg = sb2 + s2;
    const double mu = r * sb2 / Sg, t2 = sb2 * s2 / Sg, t = std::sqrt(t2);
    const double e0 = std::exp(-mu * mu / (2 * t2)), eh = std::exp(-(rh - mu) * (rh - mu) / (2 * t2));
    double M[5];
    M[0] = t * std::sqrt(M_PI / 2) *
           (std::erf((rh - mu) / (t * std::sqrt(2.0))) + std::erf(mu / (t * std::sqrt(2.0))));
    M[1] = mu * M[0] - t2 * (eh - e0);
    double rp = 1;
    for (int k = 2; k < 5; k++) {
      rp *= rh;
      M[k] = mu * M[k - 1] + (k - 1) * t2 * M[k - 2] - t2 * rp * eh;
    }
    I[0] = sd.gnorm[s] * pre * std::exp(-r * r / (2 * Sg)) * M[p];
  }
  const double* Ws = W + s * TGX_MAX_NR * TGX_MAX_NR;
  for (int n = 0; n < nmax; n++) {
    double v = 0;
    for (int a = 0; a < nmax; a++) v += Ws[n * TGX_MAX_NR + a] * I[a];
    R[n] = v;
  }
}

// Orthorhombic cell, supercell repetitions as number_of_unit_cells_for_given_
// cutoff (neighbors_interface.f90:655-708) and the linked-cell grid
// mx = int(L/rcut) (:299-301).
// number_of_unit_cells_for_given_cutoff (neighbors_interface.f90:655-708)
// in the reference's arithmetic: unit normals of the cell faces, the 3x3
// matrix of normal.edge, its explicit inverse, times 2 rcut, rounded up.
static void cells_for_cutoff(const double* a, const double* b, const double* cc, double rcut,
                             int* idx) {
  auto cross = [](const double* x, const double* y, double* o) {
    o[0] = x[1] * y[2] - x[2] * y[1];
    o[1] = x[2] * y[0] - x[0] * y[2];
    o[2] = x[0] * y[1] - x[1] * y[0];
  };
  auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  double axb[3], cxa[3], bxc[3];
  cross(a, b, axb);
  double t = std::sqrt(dot(axb, axb));
  for (int k = 0; k < 3; k++) axb[k] /= t;
  cross(cc, a, cxa);
  t = std::sqrt(dot(cxa, cxa));
  for (int k = 0; k < 3; k++) cxa[k] /= t;
  cross(b, cc, bxc);
  t = std::sqrt(dot(bxc, bxc));
  for (int k = 0; k < 3; k++) bxc[k] /= t;
  const double* rows[3] = {axb, cxa, bxc};
  const double* cols[3] = {a, b, cc};
  double m[3][3], mi[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = dot(rows[i], cols[j]);
  const double md = -m[0][2] * m[1][1] * m[2][0] + m[0][1] * m[1][2] * m[2][0] +
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
  for (int i = 0; i < 3; i++) {
    double v = 0;
    for (int j = 0; j < 3; j++) v += (mi[i][j] / md) * (2.0 * rcut);
    v = std::fabs(v);
    idx[i] = (int)v;
    if (v > (double)idx[i]) idx[i] += 1;
    if (idx[i] < 1) idx[i] = 1;
  }
}

bool is_orthorhombic(const double* lat) {
  const double tol = 1e-6;  // get_distance's d_tol
  return std::fabs(lat[1]) < tol && std::fabs(lat[2]) < tol && std::fabs(lat[3]) < tol &&
         std::fabs(lat[5]) < tol && std::fabs(lat[6]) < tol && std::fabs(lat[7]) < tol;
}

int make_general_cell(const double* lat, double rcut, GenCell* g, std::string* err) {
  const double* a = lat;
  const double* b = lat + 3;
  const double* cc = lat + 6;
  const double vol = a[0] * (b[1] * cc[2] - b[2] * cc[1]) - a[1] * (b[0] * cc[2] - b[2] * cc[0]) +
                     a[2] * (b[0] * cc[1] - b[1] * cc[0]);
  if (!(std::fabs(vol) > 0) || !(rcut > 0)) {
    *err = "the cell vectors must span a volume and the cutoff must be positive";
    return TGX_ERR_ARG;
  }
  int idx[3];
  cells_for_cutoff(a, b, cc, rcut, idx);
  g->ia = idx[0];
  g->ib = idx[1];
  g->ic = idx[2];
  g->nrep = idx[0] * idx[1] * idx[2];
  if (g->nrep > (1 << (31 - TGX_IMG_SHIFT))) {
    *err = "cell too small for the cutoff (too many images)";
    return TGX_ERR_UNSUPPORTED;
  }
  for (int k = 0; k < 9; k++) g->u[k] = lat[k];
  // supercell vectors (set_supercell, read_xyz.f90:370-372)
  for (int k = 0; k < 3; k++) {
    g->s[k] = (double)idx[0] * a[k];
    g->s[3 + k] = (double)idx[1] * b[k];
    g->s[6 + k] = (double)idx[2] * cc[k];
  }
  // get_distance's matrix of supercell-vector dot products and its explicit
  // inverse (neighbors_interface.f90:593-613), in the reference's order
  auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  const double *A = g->s, *B = g->s + 3, *C = g->s + 6;
  double m[3][3];
  m[0][0] = dot(A, A);
  m[0][1] = dot(A, B);
  m[0][2] = dot(A, C);
  m[1][0] = m[0][1];
  m[1][1] = dot(B, B);
  m[1][2] = dot(B, C);
  m[2][0] = m[0][2];
  m[2][1] = m[1][2];
  m[2][2] = dot(C, C);
  const double md = -(m[0][2] * m[0][2]) * m[1][1] + 2.0 * m[0][1] * m[0][2] * m[1][2] -
                    m[0][0] * (m[1][2] * m[1][2]) - (m[0][1] * m[0][1]) * m[2][2] +
                    m[0][0] * m[1][1] * m[2][2];
  double* mi = g->mi;
  mi[0] = m[1][1] * m[2][2] - m[1][2] * m[1][2];
  mi[1] = m[0][2] * m[1][2] - m[0][1] * m[2][2];
  mi[2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  mi[3] = mi[1];
  mi[4] = m[0][0] * m[2][2] - m[0][2] * m[0][2];
  mi[5] = m[0][1] * m[0][2] - m[0][0] * m[1][2];
  mi[6] = mi[2];
  mi[7] = mi[5];
  mi[8] = m[0][0] * m[1][1] - m[0][1] * m[0][1];
  for (int k = 0; k < 9; k++) mi[k] = mi[k] / md;
  return TGX_OK;
}

int make_image_cell(const double* lat, double rcut, ImgCell* g, std::string* err) {
  const double* a = lat;
  const double* b = lat + 3;
  const double* c = lat + 6;
  auto cross = [](const double* x, const double* y, double* z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
  };
  double bc[3], ca[3], ab[3];
  cross(b, c, bc);
  cross(c, a, ca);
  cross(a, b, ab);
  const double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
  if (!(std::fabs(det) > 0) || !(rcut > 0)) {
    *err = "the cell vectors must span a volume and the cutoff must be positive";
    return TGX_ERR_ARG;
  }
  // inverse of the row matrix (rows a, b, c): columns are the reciprocal
  // vectors b x c, c x a, a x b over det
  for (int k = 0; k < 3; k++) {
    g->ui[3 * k + 0] = bc[k] / det;
    g->ui[3 * k + 1] = ca[k] / det;
    g->ui[3 * k + 2] = ab[k] / det;
  }
  for (int k = 0; k < 9; k++) g->u[k] = lat[k];
  g->vol = std::fabs(det);
  // a pair within rcut of wrapped atoms (fractional coordinates in [0, 1])
  // needs |t_k| <= rcut / w_k + 1, w_k the cell's width across axis k
  const double* rec[3] = {bc, ca, ab};
  long long nt = 1;
  for (int k = 0; k < 3; k++) {
    const double w = g->vol / std::sqrt(rec[k][0] * rec[k][0] + rec[k][1] * rec[k][1] +
                                        rec[k][2] * rec[k][2]);
    const double mk = std::floor(rcut / w) + 1.0;
    if (!(mk < 64)) {
      *err = "cell too thin for the cutoff";
      return TGX_ERR_UNSUPPORTED;
    }
    g->m[k] = (int)mk;
    nt *= 2 * g->m[k] + 1;
  }
  g->nt = (int)nt;
  return TGX_OK;
}

Cell open_cell() {
  Cell c{};
  c.ia = c.ib = c.ic = 1;
  c.nrep = 1;
  c.small = 0;
  c.mx = c.my = c.mz = 1;
  return c;  // edges and inverse edges 0: no minimum image
}

int make_cell(const double* lat, double rcut, Cell* c, std::string* err) {
  if (!is_orthorhombic(lat)) {
    *err = "non-orthorhombic cells are supported by the single-structure calls (neighbours, "
           "descriptors, energies and forces, local-property values, TS dispersion) only";
    return TGX_ERR_UNSUPPORTED;
  }
  if (!(lat[0] > 0 && lat[4] > 0 && lat[8] > 0 && rcut > 0)) {
    *err = "cell edges and cutoff must be positive";
    return TGX_ERR_ARG;
  }
  // the (zeroed off-diagonal) lattice, as get_distance reads it
  const double a[3] = {lat[0], 0, 0}, b[3] = {0, lat[4], 0}, cc[3] = {0, 0, lat[8]};
  int idx[3];
  cells_for_cutoff(a, b, cc, rcut, idx);
  c->a = lat[0];
  c->b = lat[4];
  c->c = lat[8];
  c->ia = idx[0];
  c->ib = idx[1];
  c->ic = idx[2];
  c->nrep = idx[0] * idx[1] * idx[2];
  c->small = c->nrep > 1;
  c->La = (double)idx[0] * lat[0];
  c->Lb = (double)idx[1] * lat[4];
  c->Lc = (double)idx[2] * lat[8];
  c->iLa = 1.0 / c->La;
  c->iLb = 1.0 / c->Lb;
  c->iLc = 1.0 / c->Lc;
  c->mx = (int)(c->La / rcut);
  c->my = (int)(c->Lb / rcut);
  c->mz = (int)(c->Lc / rcut);
  if (c->nrep > (1 << (31 - TGX_IMG_SHIFT))) {
    *err = "cell too small for the cutoff (too many images)";
    return TGX_ERR_UNSUPPORTED;
  }
  return TGX_OK;
}

}  // namespace tgx
