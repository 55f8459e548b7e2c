/*
 * tgx_oracle.h - FP64 CPU oracle of the GAP/SOAP energy-and-force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never calls it.
 *
 * Parity status of each function is stated in tgx_oracle.c.
 */
#ifndef TGX_ORACLE_H
#define TGX_ORACLE_H

#include <stdint.h>

#include "../include/tgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One distance_2b block (gap_2b_t, reference types/types.f90:314-321). */
typedef struct {
  int32_t sp1, sp2; /* 1-based */
  int32_t n_sparse;
  int32_t pad_;
  double rcut, buffer, delta, sigma;
  const double* Qs;
  const double* alphas;
  const double* cutoff;
} orc_2b;

/* One angle_3b block (gap_3b_t; read_gap.f90:599-643, sparse set
 * :850-872).  Qs is n_sparse x 3 column-major as the reference stores it;
 * kernel 0 = "exp", 1 = "pol" (cov_pp with d = 3, q = 1). */
typedef struct {
  int32_t sp0, sp1, sp2; /* 1-based: centre and the two neighbour species */
  int32_t n_sparse;
  int32_t kernel;
  int32_t pad_;
  double rcut, delta;
  double sigma[3];
  const double* Qs;
  const double* alphas;
  const double* cutoff;
} orc_3b;

/* One core_pot block (gap_core_pot_t): tabulated V(x) with its spline
 * second derivatives d2 and end slopes as orc_core_pot_setup leaves them. */
typedef struct {
  int32_t sp1, sp2, n, pad_;
  double yp1, ypn;
  const double* x;
  const double* V;
  const double* d2;
} orc_core;

/* Whole model: a soap_turbo block, optional further soap_turbo blocks
 * (more[0..n_more-1], of which only the SOAP fields are read), optional 2b
 * blocks, e0 per species. */
typedef struct orc_model {
  tgx_soap_hypers h;
  int32_t has_soap;
  int32_t dim, n_sparse;
  int32_t n2b;
  const double* Qs;     /* dim x n_sparse, column-major */
  const double* alphas; /* n_sparse */
  double delta, zeta;
  double e0[TGX_MAX_SPECIES];
  const orc_2b* b2;
  int32_t n_more;
  int32_t pad_;
  const struct orc_model* more;
  int32_t n3b, ncore;
  const orc_3b* b3;
  const orc_core* core;
} orc_model;

/* get_distance, orthorhombic branch (neighbors_interface.f90:547-581). */
void orc_mic(const double* posi, const double* posj, const double* L,
             double* dist, double* d);

/* number_of_unit_cells_for_given_cutoff (neighbors_interface.f90:655-708). */
void orc_cells_for_cutoff(const double* lat, double rcut, int* indices);

/* Neighbour lists; see tgx_neighbors_build for the output contract.
 * algo: 0 = linked cells when the box allows it, 1 = brute force. */
int orc_neighbors(int n, const double* pos, const double* lat, double rcut,
                  int algo, int64_t cap, int* n_neigh, int* list, double* rjs,
                  double* xyz, double* th, double* ph, int64_t* n_pairs,
                  int* indices);

/* Overlap S and W = S^(-1/2) of the poly3 basis (alpha_max x alpha_max). */
void orc_poly3_overlap(int alpha_max, double* S, double* W);
/* The poly3gauss basis' overlap S and W = S^(-1/2) (alpha_max x alpha_max,
 * row-major): a Gaussian of width sigma at the origin, then alpha_max - 1
 * poly3 functions, normalised on [0, rh]. */
void orc_poly3gauss_overlap(int alpha_max, double rh, double sigma, double* S, double* W);

/* Radial coefficients R_n(r) and dR_n/dr of neighbour species s (0-based)
 * for nr distances; R and dR are nr x alpha_max[s] (n fastest). */
void orc_radial(const tgx_soap_hypers* h, int s, int nr, const double* r,
                double* R, double* dR);

/* Raw radial integrals of the unnormalised polynomials (test hook):
 * out[k] = int_0^rh rho^p (rh-rho)^(k+3) exp(-(rho-r)^2/(2 sig^2)) drho /
 * (sqrt(2 pi) sig), k = 0..nk-1. */
void orc_radial_raw(double r, double rh, double sig, int p, int nk,
                    double* out);

/* Real spherical harmonics Y[l*l+l+m] at unit vector u and the gradient
 * (3 x (lmax+1)^2, component fastest) of their polynomial extension. */
void orc_sph(int lmax, const double* u, double* Y, double* dF);
/* c_nlm of one site: c[(n*(l_max+1)^2 + l*l+l+m)] over all radial channels. */
int orc_soap_coeffs(const tgx_soap_hypers* h, int si, int np, const double* x, const int* sj,
                    double* c);

/* SOAP vectors (dim x n_sites) and, when soap_der != NULL, the Cartesian
 * derivatives soap_cart_der(3,dim,n_pairs) over the site pair list (pairs
 * within the descriptor cutoff, self first), as the reference's get_soap.
 * Uses the neighbour list of orc_neighbors with rcut = max rcut_hard.
 * pair_atom (0-based folded atom of each pair) and site_atom (0-based atom
 * of each site) are returned when not NULL.  Returns sites / pairs counts. */
int orc_soap(const tgx_soap_hypers* h, int n, const double* pos,
             const double* lat, const int* species, double* soap,
             double* soap_der, int64_t pair_cap, int* n_sites,
             int64_t* n_pairs, int* site_atom, int* pair_atom,
             double* pair_xyz, int* site_n_neigh);

/* get_soap_energy_and_forces (gap/gap.f90:39-244) on given arrays.
 * n_neigh(n_sites), pair_atom(n_pairs) 0-based folded targets,
 * xyz(3,n_pairs); forces(3,n_atoms). */
void orc_gap_soap_ef(int n_sites, int dim, int n_sparse, const double* soap,
                     const double* soap_der, const double* alphas,
                     double delta, double zeta, double e0, const double* Qs,
                     const int* n_neigh, const int* pair_atom,
                     const double* xyz, int n_atoms, int do_forces,
                     double* energies, double* forces, double* virial);

/* The dgemm/pow/dgemm part of gap.f90:118-171: energies(n_sites) and
 * Qss(n_sites,dim) column-major (qss may be NULL). */
void orc_gap_kernel_step(int n_sites, int dim, int n_sparse,
                         const double* soap, const double* Qs,
                         const double* alphas, double delta, double zeta,
                         double* energies, double* qss);

/* get_2b_energy_and_forces (gap/gap.f90:329-460) on a neighbour list. */
void orc_gap_2b_ef(const orc_2b* b, int n, const int* species,
                   const int* n_neigh, const int* list, const double* rjs,
                   const double* xyz, int do_forces, double* energies,
                   double* forces, double* virial);

/* Whole path for one structure: neighbours, SOAP term (sites processed in
 * chunks of at most chunk_sites, 0 = all), 2b terms, e0. */
/* Spline second derivatives of a core_pot table and the end slopes
 * (read_gap.f90:701-752): a slope above 0.99e30 means a natural end and is
 * then replaced by the spline's own derivative there. */
void orc_core_pot_setup(int n, const double* x, const double* V, double* yp1, double* ypn,
                        double* d2);
/* spline / spline_der (utils/splines.f90:33-156) at r < rcut, else 0. */
double orc_spline(int n, const double* x, const double* y, const double* y2, double yp1,
                  double ypn, double r, double rcut, int derivative);
/* get_core_pot_energy_and_forces (gap.f90:552-686) on a full list. */
void orc_core_pot_ef(const orc_core* b, int n, const int* species, const int* n_neigh,
                     const int* list, const double* rjs, const double* xyz, int do_forces,
                     double* energies, double* forces, double* virial);
/* get_3b_energy_and_forces (gap.f90:756-1015) on a full list. */
void orc_angle_3b_ef(const orc_3b* b, int n, const int* species, const int* n_neigh,
                     const int* list, const double* rjs, const double* xyz, int do_forces,
                     double* energies, double* forces, double* virial);

/* Tkatchenko-Scheffler dispersion parameters (vdw_t, vdw/vdw_types.f90:32-48):
 * per-species free-atom references (1-based species index s -> [s-1]). */
typedef struct {
  int n_species;
  const double* c6_ref;
  const double* r0_ref;
  const double* alpha0_ref;
  double rcut, buffer, rcut_inner, buffer_inner, sR, d;
} orc_ts;

/* get_ts_energy_and_forces (vdw/vdw_interface.f90:127-385) on a full list in
 * the reference's layout (self first, 1-based supercell ids folded modulo n):
 * V (n) the Hirshfeld volumes per atom, V_der (3 per list entry) dV_i/dr_j.
 * energies (n) are overwritten; forces (3n) and the virial are overwritten
 * when do_forces. */
void orc_ts_ef(const orc_ts* t, int n, const int* species, const int* n_neigh, const int* list,
               const double* rjs, const double* xyz, const double* V, const double* V_der,
               int do_forces, double* energies, double* forces, double* virial);

int orc_energy_forces(const orc_model* m, int n, const double* pos,
                      const int* species, const double* lat, int do_forces,
                      int chunk_sites, double* e_atom, double* forces,
                      double* virial);

/* Velocity Verlet step (md/md_utils.f90:486-532); masses in eV fs^2/A^2. */
void orc_velocity_verlet(int n, double* pos, double* pos_prev, double* vel,
                         const double* forces, double* forces_prev,
                         const double* masses, double dt, double* dt_prev,
                         int first_step);

/* Metropolis acceptance of mc_utils.f90:172-180,215-257 ("move"). */
double orc_mc_accept(double e_new, double e_prev, double temp);

/* Trial random stream of single-atom-move MC: u(seed, t, k) in [0, 1). */
double orc_mc_uniform(uint64_t seed, int64_t t, int k);

/* Sequential Metropolis MC with single-atom moves and one full-system energy
 * per trial (mc/mc_interface.f90:183-447); pos updated in place. */
int orc_mc_run(const orc_model* m, int n, double* pos, const int* species,
               const double* lat, int n_trials, double temp, double move_max,
               uint64_t seed, double* e_trace, int* n_accepted);

/* Surrogate MLP e(q) and, with y and grad, the gradient of the mean squared
 * error in normalised units; returns the loss (0 without y). */
double orc_mlp(int n, int D, int H, const double* P, const double* q, double mu,
               double sigma, int round_bf16, const double* y, double* e,
               double* grad);

/* Delayed-acceptance MC with a fixed surrogate (P, mu, sigma). */
int orc_mc_run_da(const orc_model* m, int n, double* pos, const int* species,
                  const double* lat, int n_trials, double temp, double move_max,
                  uint64_t seed, const double* P, double mu, double sigma,
                  double max_rmse, double* e_trace, int* n_accepted, int* n_rej1,
                  int* fallback_at);

/* Independent chains of the shared Metropolis / delayed-acceptance driver on
 * a toy potential of 1 or 2 particles (sampling tests). */
int orc_mc_toy(int n, const double* prm, int n_chains, int n_steps, double temp,
               double move_max, uint64_t seed, int use_sur, int variant, const double* x0,
               double* x_end, int64_t* n_accepted, int64_t* n_rej1);

#ifdef __cplusplus
}
#endif

#endif
