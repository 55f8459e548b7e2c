/*
 * tgx.h - C-ABI of the MI355X GAP/SOAP energy-and-force engine.
 *
 * This is the drop-in boundary behind TurboGAP's Fortran drivers.  Every
 * entry point takes plain pointers, sizes and scalars (no HIP or torch types)
 * so that Fortran binds it with iso_c_binding exactly as the reference binds
 * its GPU layer (reference: src/gpu/gpu_interface.f90:34-703).
 *
 * Conventions (same as the reference's Fortran arrays, so buffers pass
 * straight through):
 *   - reals are double, integers are int32 (Fortran default integer);
 *   - positions(3,N), forces(3,N), xyz(3,n_pairs) are column-major, i.e.
 *     x,y,z of one atom are contiguous;
 *   - soap(dim,n_sites) and Qs(dim,n_sparse) are column-major;
 *   - the lattice is the 9 doubles a(1:3), b(1:3), c(1:3)
 *     (state%a_box, b_box, c_box, reference types/types.f90:147-211);
 *   - species are 1-based indices into the model's species list
 *     (state%species), neighbour ids are 1-based supercell ids
 *     (neighbors%neighbors_list, reference types/types.f90:83-111);
 *   - virial(3,3) is symmetric, column-major.
 * Every function returns TGX_OK (0) or an error code; tgx_last_error() gives
 * the message.  The library never stops the process (the reference prints and
 * calls stop; the Fortran shim keeps that behaviour on a non-zero code).
 * Calls on one context are stream-ordered and not thread-safe; separate
 * contexts are independent.
 */
#ifndef TGX_H
#define TGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TGX_MAX_SPECIES 8
#define TGX_MAX_LMAX 10  /* (l_max+1)^2 <= 121 real harmonics      */
#define TGX_MAX_NR 32    /* sum over species of alpha_max (n_max)  */
#define TGX_MAX_SOAP_BLOCKS 8  /* soap_turbo blocks of one model        */

enum {
  TGX_OK = 0,
  TGX_ERR_ARG = 1,         /* bad argument or model/shape mismatch           */
  TGX_ERR_HIP = 2,         /* HIP runtime failure                            */
  TGX_ERR_CAPACITY = 3,    /* caller buffer too small (size is reported)     */
  TGX_ERR_UNSUPPORTED = 4, /* e.g. non-orthorhombic cell for MD / MC        */
  TGX_ERR_STATE = 5        /* no model loaded, wrong call order              */
};

/* Arithmetic of the SOAP expansion and of the kernel GEMM. */
enum {
  TGX_FP64 = 0,  /* SOAP FP64, kernel GEMM FP64                           */
  TGX_FP32 = 1,  /* SOAP FP32, kernel GEMM FP32                           */
  TGX_MIXED = 2  /* SOAP FP32, kernel GEMM BF16 on MFMA, FP32 accumulate   */
};

typedef struct tgx_ctx tgx_ctx;

/*
 * Hyper-parameters of one soap_turbo descriptor block.  Field meaning and
 * defaults follow the reference's soap_turbo type (types/types.f90:252-312)
 * and its .gap parser (gap/read_gap.f90:156-299, 425-432).  Per-species
 * arrays are indexed by species-1.  rcut_soft is the soft radius
 * (rcut_hard - buffer), as read_gap.f90:425-432 converts it.
 */
typedef struct {
  int32_t n_species;
  int32_t l_max;
  int32_t radial_enhancement; /* 0, 1 or 2                                  */
  int32_t central_species;    /* 0 = all atoms are sites, else 1-based       */
  int32_t basis;              /* 0 = poly3, 1 = poly3gauss                   */
  int32_t pad_;
  int32_t alpha_max[TGX_MAX_SPECIES];
  double rcut_hard[TGX_MAX_SPECIES];
  double rcut_soft[TGX_MAX_SPECIES];
  double atom_sigma_r[TGX_MAX_SPECIES];
  double atom_sigma_t[TGX_MAX_SPECIES];
  double atom_sigma_r_scaling[TGX_MAX_SPECIES];
  double atom_sigma_t_scaling[TGX_MAX_SPECIES];
  double amplitude_scaling[TGX_MAX_SPECIES];
  double central_weight[TGX_MAX_SPECIES];
  double global_scaling[TGX_MAX_SPECIES];
  double nf[TGX_MAX_SPECIES]; /* only the default 4 is accepted (else
                                 TGX_ERR_UNSUPPORTED)                    */
  /* compress_soap (types.f90:268-305, read_gap.f90:485-549): with
   * compress_P_nonzero > 0 the descriptor is q = P p / |P p|, p the
   * uncompressed power spectrum and P the sparse matrix of entries
   * (compress_P_i[k], compress_P_j[k], compress_P_el[k]), 1-based row in
   * 1..dim (the model's descriptor dimension) and column in 1..n_max
   * (n_max+1)/2 (l_max+1); 0 = no compression.  The arrays are copied when
   * the model is set. */
  int32_t compress_P_nonzero;
  int32_t pad2_;
  const int32_t* compress_P_i;
  const int32_t* compress_P_j;
  const double* compress_P_el;
} tgx_soap_hypers;

/* Context ------------------------------------------------------------------ */

/* Replaces the empty _GPU init block (turbogap_main.f90:320-327) and the
 * handle/stream creation of gpu_interface.f90:682-700. */
int tgx_create(int device, int precision, tgx_ctx** ctx);
int tgx_destroy(tgx_ctx* ctx);
const char* tgx_last_error(const tgx_ctx* ctx);
int tgx_set_precision(tgx_ctx* ctx, int precision);

/* Autotuner (no counterpart in the reference; the engine's per-shape choice
 * of launch shapes): on = 1 times the candidate launch shapes of the kernel step
 * (single kernel or separate GEMMs, GEMM tile shapes) and of the SOAP
 * kernels (waves per workgroup) on the first call of each shape and keeps
 * the fastest; on = 0 uses stored decisions or the built-in shapes.
 * cache_path (may be NULL): decisions are read from this text file now and
 * appended to it as they are made.  TGX_AUTOTUNE=1 in the environment turns
 * tuning on for contexts that never call this.  Decisions are never made
 * inside graph capture or for guarded (Monte Carlo) steps. */
int tgx_set_autotune(tgx_ctx* ctx, int on, const char* cache_path);

/* Models ------------------------------------------------------------------- */

/* One soap_turbo GAP block: hypers + sparse set Qs(dim,n_sparse) + alphas,
 * delta, zeta (read_gap.f90:45-771, 773-875).  dim must equal
 * N_R(N_R+1)/2*(l_max+1) with N_R = sum(alpha_max) (read_gap.f90:556-558). */
int tgx_set_soap_model(tgx_ctx* ctx, const tgx_soap_hypers* h, int dim,
                       int n_sparse, const double* Qs, const double* alphas,
                       double delta, double zeta);

/* Further soap_turbo blocks of the same model.  The reference reads any
 * number of soap_turbo blocks from the .gap file (read_gap.f90:150-299,
 * n_soap_turbo) and calculate_gap_soap.f90:324-447 evaluates them in turn,
 * adding each block's energies, forces and virial.  tgx_set_soap_model
 * replaces the model by one block; tgx_add_soap_model appends a block
 * (arguments as tgx_set_soap_model; with no block loaded it acts as
 * tgx_set_soap_model).  Every block uses the model's species list (the same
 * n_species and species order, h->n_species of the first block); its
 * central_species selects its sites.  At most TGX_MAX_SOAP_BLOCKS blocks.
 * tgx_clear_soap_models removes all blocks; tgx_n_soap_models returns how
 * many are loaded.  tgx_soap_descriptors, tgx_gap_kernel_step and the Monte
 * Carlo entry points work on the first block (Monte Carlo refuses models of
 * more than one block with TGX_ERR_UNSUPPORTED). */
int tgx_add_soap_model(tgx_ctx* ctx, const tgx_soap_hypers* h, int dim,
                       int n_sparse, const double* Qs, const double* alphas,
                       double delta, double zeta);
int tgx_clear_soap_models(tgx_ctx* ctx);
int tgx_n_soap_models(const tgx_ctx* ctx);

/* Per-species isolated-atom energies added to every atom of that species
 * (calculate_e0, gap/calculate_gap_soap.f90:135-169). */
int tgx_set_e0(tgx_ctx* ctx, int n_species, const double* e0);

/* One distance_2b block (gap_2b_t, types.f90:314-321; alphas/cutoff/Qs read
 * by read_gap.f90:829-848).  sp1/sp2 are 1-based species. */
int tgx_add_2b_model(tgx_ctx* ctx, int sp1, int sp2, double rcut,
                     double buffer, double delta, double sigma, int n_sparse,
                     const double* Qs, const double* alphas,
                     const double* cutoff);
int tgx_clear_2b_models(tgx_ctx* ctx);

/* One angle_3b block (gap_3b_t, types.f90; read_gap.f90:599-643, sparse set
 * :850-872; evaluated as get_3b_energy_and_forces, gap.f90:756-1015, with
 * buffer 0.5 as calculate_gap_3b passes it, :730-751).  sp_center, sp1, sp2
 * 1-based; sigma[3]; kernel 0 = "exp", 1 = "pol"; Qs(n_sparse,3)
 * column-major as read_alphas_and_descriptors stores it. */
int tgx_add_angle_3b(tgx_ctx* ctx, int sp_center, int sp1, int sp2, double rcut,
                     double delta, const double* sigma, int kernel, int n_sparse,
                     const double* Qs, const double* alphas, const double* cutoff);

/* One core_pot block (gap_core_pot_t; read_gap.f90:644-752; evaluated as
 * get_core_pot_energy_and_forces, gap.f90:552-686): V tabulated at n >= 2
 * strictly ascending x (after any core_pot_cutoff taper, which the reader
 * applies), end slopes yp1/ypn; a slope above 0.99e30 makes that end of the
 * cubic spline natural.  The cutoff is the largest x. */
int tgx_add_core_pot(tgx_ctx* ctx, int sp1, int sp2, int n, const double* x,
                     const double* V, double yp1, double ypn);

/* Drops every angle_3b and core_pot block. */
int tgx_clear_extra_terms(tgx_ctx* ctx);

/* A local-property model (local_property_models of a soap_turbo block,
 * read_gap.f90:296-380, e.g. Hirshfeld volumes) on SOAP block `block`
 * (0-based, in the order the blocks were set and added), predicted as
 * local_property_predict does (gap/local_properties.f90:46-146): V_i =
 * delta^2 sum_s alphas_s K(i,s)^zeta + V0 with K = q_i . Qs_s, clipped at 0.
 * Qs(dim, n_sparse) column-major.  *index_out (may be NULL) receives its
 * index.  Integer zeta follows the GAP kernel's rule (|zeta - int(zeta)| <
 * 1e-5), where the reference's property code takes |zeta - nint(zeta)| <
 * 1e-10. */
int tgx_add_local_property(tgx_ctx* ctx, int block, int dim, int n_sparse,
                           const double* Qs, const double* alphas, double delta,
                           double zeta, double V0, int* index_out);
int tgx_clear_local_properties(tgx_ctx* ctx);

/* Local property `index` of one structure (get_local_properties,
 * gap/gap_interface.f90:300-476): V(n), 0 for atoms that are no site of its
 * block; with V_der (may be NULL; orthorhombic cells) dV_i/dr_j (3, n_pairs)
 * for every entry of the list tgx_neighbors_build returns at rcut_list (at
 * least the block's cutoff), the self entry being dV_i/dr_i, all zero where
 * V_i = 0 or atom i is no site (local_properties.f90:125-137).  cap is the
 * capacity of V_der in pairs; TGX_ERR_CAPACITY leaves the size needed in
 * *n_pairs. */
int tgx_local_property(tgx_ctx* ctx, int index, int n, const double* pos,
                       const int* species, const double* lattice, double rcut_list,
                       double* V, int64_t cap, double* V_der, int64_t* n_pairs);

/* Tkatchenko-Scheffler dispersion (vdw_type "ts"; get_vdw_energy_and_forces
 * and get_ts_energy_and_forces, vdw/vdw_interface.f90:40-385) on the
 * Hirshfeld volumes of local property `local_property`.  Per species s
 * (1-based, s-1 in the arrays): free-atom C6, R0 and alpha0 (eV A^6, A,
 * A^3; get_vdw_ref_params :388-466 holds the reference's defaults for H, C
 * and P).  The cutoffs and damping are vdw_t's (vdw/vdw_types.f90:32-48;
 * defaults rcut 10, buffer 1, rcut_inner 0.5, buffer_inner 0.5, sR 0.94,
 * d 20).  Replaces an earlier setting; tgx_clear_local_properties clears
 * it. */
int tgx_set_vdw_ts(tgx_ctx* ctx, int local_property, int n_species, const double* c6_ref,
                   const double* r0_ref, const double* alpha0_ref, double rcut,
                   double buffer, double rcut_inner, double buffer_inner, double sR,
                   double d);
int tgx_clear_vdw(tgx_ctx* ctx);

/* The TS term of one structure, as get_vdw_energy_and_forces returns it:
 * energies(n) = -1/2 sum_j C6_ij f_damp r^-6 (switched in the buffers) over
 * rcut_inner < r_ij < rcut; with do_forces, forces(3,n) and virial(3,3)
 * (overwritten; not touched otherwise), the pair force plus the Hirshfeld
 * term dE/dV_i dV_i/dr_j over the list entries of i with rcut_inner < r <
 * rcut (so not its own entry, r = 0, as in the reference, :345-363).  Cells
 * of any shape.  Species with zero references contribute nothing (the
 * reference's 0/0 there would be NaN). */
int tgx_vdw_energy_forces(tgx_ctx* ctx, int n, const double* pos, const int* species,
                          const double* lattice, int do_forces, double* energies,
                          double* forces, double* virial);

/* Monte Carlo trials per speculative batch (tgx_mc_run, tgx_mc_run_da after
 * the surrogate's training trials): 0 = automatic (8), 1 = one trial per
 * launch sequence, at most 64.  No counterpart in the reference (its trials
 * run one by one, mc/mc_interface.f90:183-447). */
int tgx_set_mc_batch(tgx_ctx* ctx, int trials_per_batch);

/* Hot path ----------------------------------------------------------------- */

/*
 * Full neighbour list within rcut (build_neighbors_list,
 * neighbors/neighbors_interface.f90:144-524, with get_distance :547-643).
 * pos(3,n) is the unit cell; when the cell is thinner than 2*rcut it is
 * replicated exactly as read_xyz.f90:133-155 / set_supercell :323-373 do and
 * neighbour ids are supercell ids.  Each atom's list holds itself first
 * (rjs=0, xyz=0, theta=phi=0) and then its neighbours in ascending supercell
 * id (the reference's order is cell-traversal order; the set is identical).
 * rjs/xyz are bit-identical to the reference's FP64 arithmetic.  cap is the
 * length of list/rjs/thetas/phis (xyz holds 3*cap); when too small,
 * TGX_ERR_CAPACITY is returned and *n_pairs holds the required size.
 * indices_out (may be NULL) receives the supercell repetitions.
 * Orthorhombic cells take the reference's linked-cell path
 * (neighbors_interface.f90:291-386); cells of any other shape its brute-force
 * path (:390-421) with get_distance's general minimum image (:582-629): every
 * supercell atom tested against every atom, O(n^2 images), lists in
 * ascending supercell id, rjs/xyz bit-identical to that arithmetic.
 * tgx_energy_forces(_noalloc) and tgx_soap_descriptors take cells of any
 * shape too; MD, Monte Carlo and batched structures take orthorhombic cells
 * only (TGX_ERR_UNSUPPORTED otherwise).
 */
int tgx_neighbors_build(tgx_ctx* ctx, int n, const double* pos,
                        const double* lattice, double rcut, int64_t cap,
                        int* n_neigh, int* list, double* rjs, double* xyz,
                        double* thetas, double* phis, int64_t* n_pairs,
                        int* indices_out);

/*
 * Normalised SOAP vectors soap(dim,n_sites) of the central sites, in site
 * order (the get_soap call at gap/gap_interface.f90:301-306, forward part).
 * n_sites_out receives the number of sites (atoms whose species matches
 * central_species, or all atoms).  site_index (may be NULL, length n) gets
 * the 1-based atom id of each site.
 */
int tgx_soap_descriptors(tgx_ctx* ctx, int n, const double* pos,
                         const int* species, const double* lattice,
                         double* soap, int* n_sites_out, int* site_index);

/*
 * The GAP kernel step on given descriptors (gap/gap.f90:118-171):
 * energies(i) = delta^2 * sum_s K(i,s)^zeta alphas(s), K = soap^T Qs, and,
 * when qss is not NULL, Qss(n_sites,dim) = -zeta delta^2 (K^(zeta-1) o alphas)
 * Qs^T (column-major, as the reference's Qss).  Runs in the context's
 * precision mode.  e0 is not added.
 */
int tgx_gap_kernel_step(tgx_ctx* ctx, int n_sites, const double* soap,
                        double* energies, double* qss);

/*
 * Energies, forces and virial of one periodic structure: neighbour list, SOAP
 * term, 2-body terms and e0 (calculate_gap_soap.f90:171-481 +
 * calculate_gap_2b gap.f90:246-327 + calculate_e0, summed as
 * turbogap_main.f90:993-1109 does).  e_atom(n), forces(3,n), virial(3,3).
 * With do_forces = 0 only e_atom is written (MC mode,
 * control/read_control.f90:47-51).  A cell of any shape is accepted: its
 * atoms are wrapped into the cell and every lattice translation within reach
 * of the cutoff becomes a set of image atoms (the pairs of the reference's
 * supercell list, neighbors_interface.f90:390-421), searched brute force.
 */
int tgx_energy_forces(tgx_ctx* ctx, int n, const double* pos,
                      const int* species, const double* lattice,
                      int do_forces, double* e_atom, double* forces,
                      double* virial);

/*
 * tgx_energy_forces without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, results and errors as tgx_energy_forces (host buffers in
 * and out; the model set beforehand stays where it is).  Every work buffer
 * of the call (structure copy, neighbour list and cells, descriptors,
 * kernel-step and adjoint work, outputs before the copy back) comes from one
 * 64 MiB array per GPU that the library holds until the process ends; the
 * call makes no hipMalloc/hipFree.  A call that would need more (roughly
 * more than 3,500 atoms of a 324-dim, 500-sparse model in FP64, 7,000 in
 * mixed precision) returns TGX_ERR_CAPACITY and leaves the context usable.
 * All calls on one GPU share that array, across contexts: calls that
 * overlap on one GPU (other host threads) race and must be serialised by
 * the caller.  The autotuner does not time candidates inside this call.
 */
int tgx_energy_forces_noalloc(tgx_ctx* ctx, int n, const double* pos,
                              const int* species, const double* lattice,
                              int do_forces, double* e_atom, double* forces,
                              double* virial);

/*
 * Many independent structures in one call (structures s = 0..n_struct-1 own
 * atoms offsets[s]..offsets[s+1]-1 of pos/species/e_atom/forces; lattices
 * holds 9 doubles per structure, virials 9 per structure).
 */
int tgx_energy_forces_batched(tgx_ctx* ctx, int n_struct, const int* offsets,
                              const double* pos, const int* species,
                              const double* lattices, int do_forces,
                              double* e_atom, double* forces,
                              double* virials);

/*
 * NVE molecular dynamics, n_steps velocity-Verlet steps of dt fs
 * (md/md_utils.f90:486-532) with the neighbour list built with skin and
 * rebuilt when any atom moved more than skin/2 since the last build
 * (md/md_interface.f90:1112-1151).  Masses in amu (converted to
 * eV fs^2/A^2 with 103.6426965268, read/read_files.f90:333).  pos and vel
 * (3,n) are updated in place; forces(3,n) receives the final forces.
 * e_pot and e_kin (length n_steps+1, may be NULL) receive total potential
 * and kinetic energy before the first and after every step.  The step is
 * captured once in a hipGraph and replayed.  n_rebuilds (may be NULL)
 * receives the number of list rebuilds.
 */
int tgx_md_run(tgx_ctx* ctx, int n, int n_steps, double dt, double skin,
               const int* species, const double* lattice, const double* mass,
               double* pos, double* vel, double* forces, double* e_pot,
               double* e_kin, int* n_rebuilds);

/*
 * Metropolis Monte Carlo with single-atom moves, energies only (replaces the
 * "move" branch of calculate_mc_step, mc/mc_interface.f90:183-447, with
 * mc_get_atom_disp mc/mc_utils.f90:259-277 and the acceptance of
 * get_mc_acceptance / monte_carlo_move :172-180, 215-257).  Trial t
 * (0-based) draws u_k = u(seed, t, k), k = 0..4, where u is the top 53 bits
 * of the splitmix64 finaliser of seed + 0x9E3779B97F4A7C15 * (6t + k + 1)
 * divided by 2^53; it moves atom floor(u_0 n) by move_max (u_1..u_3 - 1/2)
 * (positions + disp, mc_utils.f90:696-698) and accepts when
 * min(1, exp(-(E_trial - E)/(kB T))) > u_4, kB = 8.617333262e-5 rounded to
 * single precision as the reference's literal is.  Each trial sees the state
 * left by the previous one; the energy change is computed from the sites
 * whose environment changed (the reference recomputes the whole system).
 * The neighbour list carries a skin (> sqrt(3) move_max) and is rebuilt when
 * an atom leaves it; the box must be at least 2 (rcut + skin).  pos (3,n) is
 * updated in place.  e_trace (n_trials + 1, may be NULL): energy before the
 * first trial and after every trial.  n_accepted / n_rebuilds may be NULL.
 * Trials are proposed and evaluated in speculative batches
 * (tgx_set_mc_batch) and decided in order; a trial whose energy change an
 * earlier accepted trial of its batch could have altered (an atom moved
 * within 2 rcut + sqrt(3) move_max) is evaluated again in the next batch,
 * so the decisions are those of the one-trial sequence.
 */
int tgx_mc_run(tgx_ctx* ctx, int n, const int* species, const double* lattice,
               double* pos, int n_trials, double temp, double move_max,
               double skin, uint64_t seed, double* e_trace, int* n_accepted,
               int* n_rebuilds);

/*
 * Delayed-acceptance Metropolis MC (no counterpart in the reference; the
 * chain samples the same distribution as tgx_mc_run).  A per-atom MLP
 * surrogate e_s(q) screens every trial first: stage 1 accepts with
 * min(1, exp(-dE_s/(kB T))) > u_4, dE_s = surrogate change of the affected
 * sites + the exact 2-body change; only stage-1 survivors are evaluated with
 * the GAP and accepted with min(1, exp(-(dE - dE_s)/(kB T))) > u_5 (Christen
 * and Fox 2005).  The surrogate (tgx_surrogate_init first, input dimension =
 * SOAP dimension) is trained on the GPU on the GAP site energies of the
 * first n_train_trials trials, which run as plain Metropolis; training
 * rounds every train_every trials.  When the running RMS of dE - dE_s over
 * GAP-evaluated trials exceeds max_rmse the run falls back to plain
 * Metropolis (u_4) for the remaining trials.
 */
typedef struct {
  int32_t n_train_trials; /* plain-MC trials whose GAP site energies train  */
  int32_t train_every;    /* trials between training rounds (>= 64)         */
  int32_t epochs;         /* Adam passes over the samples per round         */
  int32_t batch;          /* sites per Adam step                            */
  double lr;              /* Adam learning rate                             */
  double max_rmse;        /* eV; fall back to plain Metropolis above this   */
} tgx_da_options;

typedef struct {
  int64_t n_accepted;
  int64_t n_stage1_rejected; /* trials the surrogate rejected (no GAP)     */
  int64_t n_gap_evaluations; /* trials evaluated with the GAP              */
  int64_t n_rebuilds;
  double train_rmse;         /* eV per site, last training round, all samples */
  double da_rmse;            /* eV, running RMS of dE - dE_s (stage 2)       */
  int32_t fallback;          /* 1: the run fell back to plain Metropolis    */
  int32_t pad_;
} tgx_mc_stats;

int tgx_mc_run_da(tgx_ctx* ctx, int n, const int* species, const double* lattice,
                  double* pos, int n_trials, double temp, double move_max,
                  double skin, uint64_t seed, const tgx_da_options* opt,
                  double* e_trace, tgx_mc_stats* stats);

/*
 * The surrogate network e_s(q) = mu + sigma (w3 . s(W2 s(W1 q + b1) + b2) + b3),
 * s(x) = x / (1 + exp(-x)), 128 hidden units; q, W1 and W2 enter as BF16,
 * products accumulate in FP32.  Parameters P (FP32 on the device, doubles at
 * the boundary): W1 (128 x dim, row-major), b1, W2 (128 x 128), b2, w3, b3.
 *   tgx_surrogate_init      Xavier-uniform weights from seed, dim <= 448;
 *   tgx_surrogate_set/get   parameters and the output scaling (mu, sigma);
 *   tgx_surrogate_predict   e_s(q) for n descriptors (n x dim, row-major);
 *   tgx_surrogate_gradient  d/dP of L = (1/n) sum ((e_s - y)/sigma)^2;
 *   tgx_surrogate_fit       Adam epochs on (q, y), rmse (eV) afterwards; the
 *                           first fit sets mu, sigma to the mean and spread of y.
 * None has a counterpart in the reference.
 */
int tgx_surrogate_init(tgx_ctx* ctx, int dim, uint64_t seed);
int tgx_surrogate_set(tgx_ctx* ctx, int n_params, const double* P, double mu,
                      double sigma);
int tgx_surrogate_get(tgx_ctx* ctx, int n_params, double* P, double* mu,
                      double* sigma);
int tgx_surrogate_predict(tgx_ctx* ctx, int n, const double* q, double* e);
int tgx_surrogate_gradient(tgx_ctx* ctx, int n, const double* q, const double* y,
                           double* grad, double* loss);
int tgx_surrogate_fit(tgx_ctx* ctx, int n, const double* q, const double* y,
                      int epochs, int batch, double lr, uint64_t seed,
                      double* rmse);

/* Overlap S(alpha_max,alpha_max) and orthonormalisation W = S^(-1/2) of the
 * poly3 radial basis, both column-major (symmetric); replaces
 * get_orthonormalization_matrix_poly3 (gpu/gpu_interface.f90:424-429).
 * Host-only (no device needed).  W is exact to FP64 rounding although
 * cond(S) ~ 1e13 at alpha_max = 8. */
int tgx_poly3_orthonormalization(int alpha_max, double* S, double* W);

/* The same for the poly3gauss basis (a normalised Gaussian of width
 * atom_sigma at the origin, then alpha_max - 1 normalised poly3 functions on
 * [0, rcut_hard]); replaces get_orthonormalization_matrix_poly3gauss
 * (gpu/gpu_interface.f90:431-437, implemented in soap_turbo, which is not
 * in the reference tree: this basis is the engine's reading of it).
 * Host-only.  S is formed in binary128, so W is exact to FP64 rounding. */
int tgx_poly3gauss_orthonormalization(int alpha_max, double atom_sigma, double rcut_hard,
                                      double* S, double* W);

/*
 * The same dynamics split for callers that keep the state on the device
 * between calls (the Fortran md driver, benchmarks):
 *   tgx_md_begin  uploads the state, builds the list with skin, computes the
 *                 initial forces and captures one step as a hipGraph;
 *                 history = length of the e_pot/e_kin ring kept on device;
 *   tgx_md_steps  enqueues n_steps steps (asynchronous, no host sync);
 *   tgx_md_read   waits and copies back positions, velocities synchronous
 *                 with them, forces and the energies of the last
 *                 min(steps+1, history) states (any pointer may be NULL);
 *   tgx_md_end    releases the graph.
 */
int tgx_md_begin(tgx_ctx* ctx, int n, const int* species, const double* lattice,
                 const double* mass, const double* pos, const double* vel,
                 double dt, double skin, int history);
int tgx_md_steps(tgx_ctx* ctx, int n_steps);
int tgx_md_read(tgx_ctx* ctx, double* pos, double* vel, double* forces,
                double* e_pot, double* e_kin, int* n_steps_done,
                int* n_rebuilds);
int tgx_md_end(tgx_ctx* ctx);

/* Wait for all work enqueued on the context's stream. */
int tgx_sync(tgx_ctx* ctx);

/* Neighbour-list buffer of tgx_energy_forces and tgx_energy_forces_batched:
 * lists are built to the largest model cutoff + buffer (neighbors%buffer,
 * utils/misc.f90:79-81; default 0).  Results do not depend on it: pairs
 * beyond a term's cutoff contribute nothing. */
int tgx_set_list_buffer(tgx_ctx* ctx, double buffer);

/* Timing of the library's own launches (the reference's do_timing switch,
 * gap/gap.f90:230-242): on, tgx_energy_forces and tgx_energy_forces_batched
 * bracket their stages and their main kernels with HIP events for
 * tgx_stage_times and tgx_kernel_times.  Off by default: the events cost the
 * GPU several microseconds of idle time each; off, those calls report 0. */
int tgx_set_timing(tgx_ctx* ctx, int on);

/* Per-stage timing of the last call, milliseconds (HIP events, with
 * tgx_set_timing on):
 * [0] neighbours, [1] SOAP forward, [2] kernel GEMMs, [3] SOAP backward and
 * force gather, [4] 2-body, [5] integrator, [6] total. */
int tgx_stage_times(const tgx_ctx* ctx, double* ms7);

/* Device time of single kernels in the last tgx_energy_forces call,
 * milliseconds, each bracketed by HIP events on the context's stream (with
 * tgx_set_timing on):
 * [0] SOAP forward, [1] kernel GEMMs (both), [2] SOAP adjoint, [3] force
 * gather.  0 for a kernel that did not run.  (Instrumentation; the
 * reference's equivalent is the do_timing wall-clock printout,
 * gap/gap.f90:230-242.) */
int tgx_kernel_times(const tgx_ctx* ctx, double* ms4);

#ifdef __cplusplus
}
#endif

#endif /* TGX_H */
