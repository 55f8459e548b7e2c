// Internal declarations of the tgx HIP library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tgx.h"

#define TGX_L2MAX ((TGX_MAX_LMAX + 1) * (TGX_MAX_LMAX + 1))

// Number of intervals of the FP32 radial lookup table (per species, over
// [0, rcut_hard]); cubic Hermite interpolation of value and derivative.
#define TGX_RTAB_N 1024

// Packed neighbour entry: atom id in the low 26 bits, supercell image in the
// high bits.  Ascending packed value == ascending reference supercell id.
#define TGX_IMG_SHIFT 26
#define TGX_ATOM_MASK ((1 << TGX_IMG_SHIFT) - 1)

namespace tgx {

// Orthorhombic periodic cell as the kernels see it.
struct Cell {
  double a, b, c;        // unit-cell edges (state%a_box(1), b_box(2), c_box(3))
  double La, Lb, Lc;     // supercell edges (indices * a ...)
  int ia, ib, ic;        // supercell repetitions
  int nrep;              // ia*ib*ic
  int small;             // any repetition > 1
  int mx, my, mz;        // linked-cell grid (large boxes only)
  double iLa, iLb, iLc;  // 1 / supercell edges
};

// A periodic cell of any shape, for the reference's brute-force neighbour
// search (neighbors_interface.f90:390-421): unit-cell vectors u (rows a, b,
// c), the supercell vectors s = indices * u (read_xyz.f90:370-372) and
// get_distance's inverse of their dot-product matrix mi (row-major).
struct GenCell {
  double u[9], s[9], mi[9];
  int ia, ib, ic, nrep;
};

// A periodic cell of any shape for energies, forces and descriptors: the
// atoms are wrapped into the cell and every lattice translation t with
// |t_k| <= m[k] of them becomes an image atom J = img * n + j.  img 0 is
// t = 0 (the atoms themselves); the image of -t is (nt - img) % nt.
struct ImgCell {
  double u[9];   // rows a, b, c
  double ui[9];  // inverse of the row matrix: fractional f = p * ui
  int m[3];      // translation range per axis
  int nt;        // (2 m0 + 1)(2 m1 + 1)(2 m2 + 1)
  double vol;    // |det|
};

// Bytes of the linked-cell atom buffer for n atoms: the atom ids in cell
// order, then (16-byte aligned) their wrapped FP32 positions with the id in
// the fourth component, read by the one-pass list build.
inline size_t cell_atoms_bytes(int n) {
  return 4 * (size_t)((n + 3) & ~3) + 16 * (size_t)n;
}

// Pair geometry handed to the launchers: one periodic cell for every atom,
// or (batched structures) a device array of cells with sid[atom] = the
// atom's structure.
struct Geom {
  Cell one;
  const Cell* many = nullptr;
  const int* sid = nullptr;
};

// SOAP model constants in device-friendly form.
struct SoapDev {
  int n_species, l_max, L2, nr, dim, npair, radial_enhancement, central_species;
  int nmax[TGX_MAX_SPECIES], off[TGX_MAX_SPECIES];
  double rh[TGX_MAX_SPECIES], rs[TGX_MAX_SPECIES];
  double sig_r[TGX_MAX_SPECIES], sig_rs[TGX_MAX_SPECIES];
  double sig_t[TGX_MAX_SPECIES], sig_ts[TGX_MAX_SPECIES];
  double amp_scal[TGX_MAX_SPECIES], glob[TGX_MAX_SPECIES];
  double rself[TGX_MAX_SPECIES * TGX_MAX_NR];  // central weight * R_n(0) * Y00
  double rcut_max;
  // Kernel shape: nmx = per-species n_max bound (8 or 16); nrp = radial
  // channels padded so that off[s] + nmx <= nrp for every species (rows of
  // width nrp can then be read nmx wide from any species offset).
  int nmx, nrp;
  // radial basis: 0 poly3, 1 poly3gauss (slot 0 the normalised Gaussian
  // gnorm[s] exp(-r^2/(2 sig_r[s]^2)), then the poly3 functions)
  int basis;
  double gnorm[TGX_MAX_SPECIES];
  // The adjoint's dE/dc rows in LDS (nrp wide): species sub-rows at offsets
  // offa[s], multiples of four with offa[s] + nmx <= nrp, so that the
  // nmx-wide reads of a pair are 16-byte aligned; chana[n] = LDS slot of
  // channel n (-1: no species' channel); lds_gap = some slot of a row is
  // filled by no channel (and is cleared).
  int lds_gap;
  int offa[TGX_MAX_SPECIES];
  signed char chana[TGX_MAX_NR];
};

struct TwoBodyDev {
  int sp1, sp2, n_sparse;
  double rcut, buffer, delta, sigma;
  double* Qs;      // device
  double* alphas;  // device (alpha * cutoff folded in)
  float2* tab;     // device: (g(r_k), h g'(r_k)), r_k = k rcut/nt, k = 0..nt
  int nt;
};

// One angle_3b block on the device: per sparse point (q1, q2, q3, pref),
// pref = 2 delta^2 alpha c; is2 = 1 / sigma^2 per descriptor component.
struct Angle3bDev {
  int sp0, sp1, sp2, n_sparse, kernel;  // kernel 0 "exp", 1 "pol"
  double rcut, buffer;
  double is2[3];
  double* Qp;  // device, 4 doubles per sparse point
};

// One core_pot block on the device: the table, its spline second
// derivatives and end slopes; rcut = largest x.
struct CorePotDev {
  int sp1, sp2, n;
  double rcut, yp1, ypn;
  double* x;   // device
  double* V;   // device
  double* d2;  // device
};

// Tkatchenko-Scheffler dispersion: per-species free-atom references and the
// cutoffs of the pair sum (vdw_t, vdw/vdw_types.f90:32-48).
struct TsDev {
  double c6[TGX_MAX_SPECIES], r0[TGX_MAX_SPECIES], a0[TGX_MAX_SPECIES];
  double rcut, buffer, rcut_inner, buffer_inner, sR, d;
  int n_species;
};

// Monte Carlo control words on the device (int and double arrays).
enum {
  MC_REBUILD = 0,  // list rebuild requested by this trial
  MC_STEP,         // trials done
  MC_END,          // trials to run
  MC_LIVE,         // this trial runs
  MC_ATOM,         // moved atom
  MC_NREB,         // list rebuilds
  MC_NACC,         // accepted trials
  MC_GATE,         // this trial reaches the GAP evaluation
  MC_DA,           // delayed acceptance active (surrogate first stage)
  MC_FALLBACK,     // surrogate drifted: plain Metropolis from here on
  MC_COLLECT,      // store (q, GAP site energy) samples for training
  MC_NREJ1,        // trials rejected by the surrogate stage
  MC_NGAP,         // trials evaluated with the GAP
  MC_CURSOR,       // sample ring cursor (total samples written)
  MC_NCTL = 16
};
enum {
  MC_OLD = 0,   // 0..2 position of the moved atom before the move
  MC_ETOT = 3,  // current total energy
  MC_DESUR,     // surrogate energy change of this trial (stage 1)
  MC_DE2B,      // exact 2-body energy change of this trial
  MC_EMA,       // running mean of (dE_gap - dE_sur)^2 over GAP-evaluated trials
  MC_MAXMS,     // fall back when MC_EMA exceeds this
  MC_NDCTL = 8
};

// Growable device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want);
  void release();
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace tgx

namespace tgx {

// One soap_turbo descriptor block of the model (the reference keeps one
// soap_turbo_hypers entry per block and loops over them,
// gap/calculate_gap_soap.f90:324-447): its hyper-parameters, sparse set and
// device tables, and the per-call buffers of its sites.  Every block uses
// the model's species list; its own central species picks its sites.
struct SoapBlock {
  tgx_soap_hypers hyp{};
  SoapDev sd{};
  int dim = 0, n_sparse = 0, dim_pad = 0, sp_pad = 0;
  double delta = 1.0, zeta = 2.0;
  std::vector<double> Qs_host;      // n_sparse x dim
  std::vector<double> alphas_host;  // n_sparse
  // W matrices per species (n_max x n_max, row-major), FP64
  std::vector<double> W_host;       // TGX_MAX_SPECIES * NR * NR
  DevBuf W_dev, rtab, Qs64, QsT64, Qs32, QsT32, Qs16, QsT16, alph64, alph32, sd_dev, pairn;
  // per call: sites (atoms of the central species, or all), descriptors,
  // kernel-step and adjoint work
  DevBuf site_of_atom, site_atom, q, qb, cbuf, pnorm, wmat, gmat, epart, esite, vsite;
  // every atom is a site, so site s is atom s (the SOAP kernels then take no
  // site_atom lookup)
  bool site_ident = false;
  int n_sites = 0;
  // compress_soap (cmp = number of entries of P, 0 = none): the entries
  // (1-based, as given), P by rows and by columns (0-based) on the device,
  // and per call the uncompressed descriptor q_f, its norm and dE/dq_f, in
  // rows of dimf_pad = roundup(power-spectrum dimension, 64)
  int cmp = 0, dimf_pad = 0;
  std::vector<int32_t> cP_i, cP_j;
  std::vector<double> cP_el;
  DevBuf c_roff, c_rcol, c_rval, c_coff, c_crow, c_cval, qf, pnf, gf;
  void release() {
    DevBuf* b[] = {&W_dev, &rtab, &Qs64, &QsT64, &Qs32, &QsT32, &Qs16, &QsT16, &alph64,
                   &alph32, &sd_dev, &pairn, &site_of_atom, &site_atom, &q, &qb, &cbuf,
                   &pnorm, &wmat, &gmat, &epart, &esite, &vsite, &c_roff, &c_rcol,
                   &c_rval, &c_coff, &c_crow, &c_cval, &qf, &pnf, &gf};
    for (DevBuf* d : b) d->release();
  }
};

}  // namespace tgx

struct tgx_ctx {
  int device = 0;
  int precision = TGX_FP64;
  hipStream_t stream = nullptr;
  std::string err;

  // model: SOAP blocks (at least one object; has_soap = a model is loaded),
  // blk = the block the SOAP launchers work on
  bool has_soap = false;
  std::vector<std::unique_ptr<tgx::SoapBlock>> blocks;
  tgx::SoapBlock* blk = nullptr;
  double e0[TGX_MAX_SPECIES] = {0};
  double list_buffer = 0.0;  // list cutoff = largest model cutoff + list_buffer
  std::vector<tgx::TwoBodyDev> twob;
  std::vector<tgx::Angle3bDev> a3b;
  std::vector<tgx::CorePotDev> corep;

  // per-call work buffers
  tgx::DevBuf pos, species, nb_count, nb_list, nb_rev, nb_flag, cell_of, cell_start,
      cell_atoms;
  tgx::DevBuf ksum, gpair, forces, eatom, virial, scratch, csr_off, csr_out;
  int maxnb = 0;
  double stage_ms[7] = {0};
  double kernel_ms[4] = {0};  // SOAP forward, kernel GEMMs, SOAP adjoint, force gather

  // MD state
  tgx::DevBuf md_vel, md_fprev, md_mass, md_pos0, md_ep, md_ek, md_reb, md_disp;
};

namespace tgx {

int set_err(tgx_ctx* c, int code, const char* fmt, ...);
#define TGX_HIP(ctx, call)                                                   \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return tgx::set_err(ctx, TGX_ERR_HIP, "%s failed: %s (%s:%d)", #call,  \
                          hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while (0)

// Host helpers (tgx_host.cpp)
void poly3_W(int nmax, double* W);                 // S^(-1/2), binary128
// poly3gauss: S and W = S^(-1/2) (binary128, rounded once); S may be NULL
void poly3gauss_W(int nmax, double rh, double sigma, double* S, double* W);
void radial_closed_form(const SoapDev& sd, const double* W, int s, double r,
                        double* R, double* dR);   // FP64 host evaluation
int make_cell(const double* lattice, double rcut, Cell* cell, std::string* err);
bool is_orthorhombic(const double* lattice);
int make_general_cell(const double* lattice, double rcut, GenCell* cell, std::string* err);
int make_image_cell(const double* lattice, double rcut, ImgCell* cell, std::string* err);
// Orthorhombic Cell with no periodicity (pair vectors taken as they are).
Cell open_cell();

// Launch plan of the kernel step, chosen by the autotuner (or all zero: the
// built-in choice).  path: 0 built-in, 2 separate GEMMs.  t0: tile of the
// kernel-epilogue GEMM on the separate path (0: 64x64, 1: 128x64, 2:
// 64x128).  t1: tile of the store GEMM (0: the type's default, 1: the
// alternative, 64x64 for BF16 and 128x128 for FP32).
struct GemmPlan {
  int path, t0, t1;
};

// Kernel launchers are declared next to their callers in tgx_api.hip.
// launch_gap_kernel: eatom/e0/species (every atom a site, in order) also
// write eatom[i] = e0[species[i] - 1] + esite[i].  few_parts (energies only):
// when the few-site kernel runs, its row-sum launch is left out and
// *few_parts = the number of FP32 partials per site in epart[part][site]
// (E_i = delta^2 times their sum in part order, in FP64); otherwise
// *few_parts = 0 and esite is written.
hipError_t launch_gap_kernel(int mode, int n_sites, int dim, int dim_pad,
                             int n_sparse, int sp_pad, const void* q,
                             const void* qs, const void* qsT,
                             const void* alphas, double delta, double zeta,
                             void* wmat, void* epart, double* esite, void* G,
                             int want_grad, const int* guard, hipStream_t st,
                             double* eatom = nullptr, const double* e0 = nullptr,
                             const int* species = nullptr, int* few_parts = nullptr,
                             const GemmPlan* plan = nullptr);
// overwrite: forces[k] = gathered sum (otherwise added to forces[k])
// The partial rows of the energy/virial reduction (reduce_part_rows), done
// by extra workgroups of the same launch when nothing later in the
// evaluation writes its inputs.
struct ReduceArgs {
  int n_sites, n;
  const double* vsite;
  const double* vatom;
  const double* eatom;
  double* part;
};
template <typename T>
hipError_t launch_force_gather(int n, int maxnb, const int* count,
                               const int* list, const int* rev,
                               const int* site_of_atom, int all_sites, const T* gpair,
                               double* forces, hipStream_t st, int overwrite = 0,
                               const ReduceArgs* red = nullptr);

}  // namespace tgx
