// C-ABI entry points of the tgx library (declared in include/tgx.h) and the
// device-resident pipeline behind them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "tgx_internal.h"

namespace tgx {

// launchers defined in tgx_kernels.hip (not in the shared header because they
// take the device SoapDev pointer)
hipError_t launch_build_rtab(const SoapDev* sd_dev, int n_species, const double* W, float* tab,
                             hipStream_t st);
hipError_t launch_neighbors(const double* pos, int n, const Cell& c, double rcut, int* cell_of,
                            int* cell_count, int* cell_start, int* cell_fill, int* cell_atoms,
                            int* count, int maxnb, int* list, int* rev, int* flag,
                            hipStream_t st, const int* guard, double* pref = nullptr);
hipError_t launch_csr_out(const double* pos, int n, const Cell& c, int maxnb, const int* count,
                          const int* list, const long long* off, int* out_list, double* rjs,
                          double* xyz, double* th, double* ph, hipStream_t st);
template <typename T>
hipError_t launch_soap_fwd(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, T* q, int q_ld,
                           unsigned short* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st,
                           int split = 0, int wpb = 0);
template <typename T>
hipError_t launch_soap_bwd(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, const T* q,
                           int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                           T* gpair, double* vsite, const TwoBodyDev* twob, int n2b,
                           double* eatom, hipStream_t st, int wpb = 0);
hipError_t launch_two_body(const TwoBodyDev& b, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           double* eatom, double* forces, double* vatom, int do_forces,
                           int exact, hipStream_t st);
hipError_t launch_neighbors_many(const double* pos, int n, const Cell* cells, const int* sid,
                                 const int* boff, const int* aoff, int n_bins, double rcut,
                                 int* cell_of, int* cell_count, int* cell_start, int* cell_fill,
                                 int* cell_atoms, int* count, int maxnb, int* list, int* rev,
                                 int* flag, hipStream_t st);
hipError_t launch_reduce_many(int n_struct, const int* aoff, const int* site_of_atom,
                              const double* vsite, const double* vatom, const double* eatom,
                              double* out, hipStream_t st);
hipError_t launch_mc_propose(int n, double* pos, const double* pref, unsigned long long seed,
                             double move_max, double lim2, int* ctl, double* dctl,
                             hipStream_t st);
hipError_t launch_mc_affected(int* ctl, double* dctl, const double* pos, const Cell& cell,
                              double rsoap, const int* count, const int* list, int maxnb,
                              const int* species, int central, int* aff, int cap, int* n_aff,
                              const TwoBodyDev* twob, int n2b, int exact, int with_2b,
                              int open_gate, hipStream_t st);
hipError_t launch_mc_stage1(double* pos, const int* species, const Cell& cell, int maxnb,
                            const int* count, const int* list, const int* aff, const int* n_aff,
                            const double* esur, const double* esur_cur, const TwoBodyDev* twob,
                            int n2b, int exact, int have_2b, unsigned long long seed, double kT,
                            int* ctl, double* dctl, double* trace, hipStream_t st);
hipError_t launch_mc_stage2(double* pos, const int* aff, const int* n_aff, const double* esite,
                            const float* parts, int nparts, int part_ld, double d2, double* e_cur, const double* esur, double* esur_cur,
                            unsigned long long seed, double kT, int* ctl, double* dctl,
                            double* trace, hipStream_t st);
template <typename T>
hipError_t launch_mc_collect(const int* ctl, const int* n_aff, const T* q, int q_ld, int D,
                             const double* esite, float* sq, double* sy, int cap_s, int* cursor,
                             hipStream_t st);
size_t mlp_nparams(int D);
size_t mlp_w1_elems(int D);
int mlp_hidden();
int mlp_max_dim();
template <typename T>
hipError_t launch_mlp_forward(int n, int D, const T* q, int q_ld, const int* rows,
                              const float* P, const unsigned short* W1t,
                              const unsigned short* W2t, const float* norm, double* e,
                              float* z1, float* z2, const int* guard, hipStream_t st);
template <typename T>
hipError_t launch_mlp_gradient(int n, int D, const T* q, int q_ld, const int* rows,
                               const double* y, const double* e, const float* z1,
                               const float* z2, const float* P, const unsigned short* W2t,
                               const float* norm, float* part, float* grad, hipStream_t st);
hipError_t launch_mlp_adam(int D, float* P, const float* grad, float* m, float* v, float lr,
                           int step, unsigned short* W1t, unsigned short* W2t, hipStream_t st);
hipError_t launch_mlp_cast(int D, const float* P, unsigned short* W1t, unsigned short* W2t,
                           hipStream_t st);
hipError_t launch_mcb_propose(int n, int nb, double* pos, int* species, const double* pref,
                              unsigned long long seed, double move_max, double lim2,
                              int* ctl, int* mi, double* md, hipStream_t st);
hipError_t launch_mcb_affected(int n, int nb, int cap, int vb, double* pos, int* species,
                               int* count, int* list, int maxnb, const Cell& cell, double rsoap,
                               int central, int* site_atom, int* vatom, int* mi, double* md,
                               const TwoBodyDev* twob, int n2b, int exact, hipStream_t st);
hipError_t launch_mcb_walk(int nb, int cap, double* pos, const Cell& cell, double rdep,
                           const int* vatom, const double* esite, double* e_cur,
                           const double* esur, double* esur_cur, unsigned long long seed,
                           double kT, int* ctl, double* dctl, const int* mi, const double* md,
                           double* trace, hipStream_t st);
hipError_t launch_mc_init(int n, const int* site_of_atom, const double* esite, double* e_cur,
                          hipStream_t st);
hipError_t launch_neighbors_general(const double* pos, int n, const GenCell& g, double rcut,
                                    int maxnb, int* count, int* list, int* flag,
                                    hipStream_t st);
hipError_t launch_csr_out_general(const double* pos, int n, const GenCell& g, int maxnb,
                                  const int* count, const int* list, const long long* off,
                                  int* out_list, double* rjs, double* xyz, double* th,
                                  double* ph, hipStream_t st);
hipError_t launch_image_atoms(const double* pos0, const int* sp0, int n, const ImgCell& g,
                              double* pos, int* sp, hipStream_t st);
hipError_t launch_neighbors_image(const double* pos, int n, const ImgCell& g, double rcut,
                                  int maxnb, int* count, int* list, int* rev, int* flag,
                                  hipStream_t st);
hipError_t launch_core_pot(const CorePotDev& b, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           double* eatom, double* forces, double* vatom, int do_forces,
                           hipStream_t st);
hipError_t launch_angle_3b(const Angle3bDev& b, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           double* eatom, double* vatom, double* g3, int do_forces,
                           hipStream_t st);
template <typename T>
hipError_t launch_compress_fwd(int n_sites, int dimf, int dimf_ld, const T* qf, const T* pnf,
                               int dim, int dim_ld, const int* roff, const int* rcol,
                               const double* rval, T* q, unsigned short* qb, int qb_ld, T* pn,
                               hipStream_t st);
template <typename T>
hipError_t launch_compress_bwd(int n_sites, int dim, int dim_ld, const T* q, const T* pn,
                               const T* G, int g_ld, int dimf, int dimf_ld, const T* pnf,
                               const int* coff, const int* crow, const double* cval, T* Gf,
                               hipStream_t st);
hipError_t launch_ts_volume(int n, const int* site_of_atom, const double* esite, double v0,
                            double* V, hipStream_t st);
hipError_t launch_ts_pairs(const TsDev& t, const double* pos, const int* species, int n,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           const double* V, double* eatom, double* forces, double* vatom,
                           double* pref, int do_forces, hipStream_t st);
template <typename T>
hipError_t launch_ts_gather(const TsDev& t, const double* pos, int n, const Geom& g, int maxnb,
                            const int* count, const int* list, const int* rev, const T* gpair,
                            const double* pref, double* forces, double* vatom, hipStream_t st);
hipError_t launch_site_add(int n_sites, const int* site_atom, const double* esite,
                           const double* vsite, double* eatom, double* vatom, hipStream_t st);
hipError_t launch_energy_init(int n, const int* species, const double* e0,
                              const int* site_of_atom, const double* esite, double* eatom,
                              double* forces, double* vatom, hipStream_t st);
hipError_t launch_reduce(int n_sites, const double* vsite, int n, const double* vatom,
                         const double* eatom, double* out, hipStream_t st,
                         bool with_final = true);
hipError_t launch_reduce_final(double* out, hipStream_t st);

hipError_t launch_md_vv(int n, double* pos, double* vel, const double* f, double* fprev,
                        const double* mass, double dt, const int* step, double* ke,
                        const double* pref, double lim2, int* flag, hipStream_t st);
hipError_t launch_md_finish(int n, const double* ke, double* out, int* flag, int* nreb,
                            int* step, double* ep_hist, double* ek_hist, int cap,
                            hipStream_t st);
hipError_t launch_md_sync_vel(int n, const double* vel, const double* f, const double* fprev,
                              const double* mass, double dt, const int* step, double* vout,
                              double* ke, double* kesum, hipStream_t st);

int set_err(tgx_ctx* c, int code, const char* fmt, ...) {
  if (!c) return code;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  c->err = buf;
  return code;
}

int DevBuf::ensure(size_t want) {
  if (want <= bytes && p) return 0;
  release();
  size_t sz = want + want / 4 + 256;
  if (hipMalloc(&p, sz) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return 1;
  }
  bytes = sz;
  return 0;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

}  // namespace tgx

using namespace tgx;

namespace {

#define ENSURE(ctx, buf, nbytes)                                                     \
  do {                                                                               \
    if (grow((ctx), (ctx)->buf, (nbytes))) return grow_err((ctx), (nbytes), #buf);   \
  } while (0)

// Bump region of the per-device scratch array that the no-allocation entry
// point hands its call: while `on`, buffers are carved from it instead of
// allocated, and the context's own buffers are restored afterwards.
struct Arena {
  bool on = false, overflow = false;
  char* base = nullptr;
  size_t size = 0, off = 0;
  std::vector<std::pair<DevBuf*, DevBuf>> saved;  // carved buffer, its own value
};

struct MdState {
  bool active = false;
  int n = 0, n_sites = 0, cap = 0;
  double dt = 1.0, skin = 0.0, rcut_list = 0.0;
  Cell cell{};
  DevBuf pref, ke, ctl /* flag, nreb, step, pad */, ep, ek, vout, kesum;
  // internal atom k is the caller's atom perm[k] (a Z-order of the atoms'
  // positions at tgx_md_begin, so that neighbours sit close in memory)
  std::vector<int> perm;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
};

struct Extra {
  DevBuf e0_dev, iota, cell_count, cell_fill, out, stage;
  DevBuf b_cells, b_sid, b_boff, b_aoff, b_out;  // batched structures
  DevBuf mc_pref, mc_ctl, mc_dctl, mc_ecur, mc_naff, mc_trace, mc_twob;  // Monte Carlo
  DevBuf mc_esur, mc_esur_cur, mc_sq, mc_sy;                          // delayed acceptance
  DevBuf mcb_mi, mcb_md, mcb_vatom;  // speculative batches: trial data, virtual sites' atoms
  // local-property models: a block each (the descriptor block's hypers, the
  // property's sparse set), its descriptor block and V0
  std::vector<std::unique_ptr<SoapBlock>> lp;
  std::vector<int> lp_block;
  std::vector<double> lp_v0;
  // Tkatchenko-Scheffler dispersion on local property ts_lp (-1: not set)
  int ts_lp = -1;
  TsDev ts{};
  DevBuf ts_V, ts_pref, ts_vatom;
  int mc_batch = 0;                  // trials per batch (0: automatic, 1: one at a time)
  // surrogate MLP: FP32 master parameters, Adam moments, BF16 copies, work
  DevBuf sg_P, sg_m, sg_v, sg_grad, sg_W1t, sg_W2t, sg_part, sg_z1, sg_z2, sg_e, sg_q, sg_y,
      sg_rows, sg_norm /* mu, sigma on the device */;
  int sg_dim = 0, sg_step = 0;
  float sg_mu = 0.f, sg_sigma = 1.f;
  bool sg_norm_set = false;
  DevBuf twob_store;
  DevBuf g_pos, g_sp;  // the structure itself while its image atoms fill pos / species
  DevBuf g3;           // angle_3b forces per list slot (double4)
  std::vector<int> species_host;
  // autotuner: on (-1 = not yet read from TGX_AUTOTUNE), cache file, the
  // decisions by shape key, and a copy of the atom energies (the adjoint
  // with the fused 2-body term adds to them, so its candidates restore them)
  int tune_on = -1;
  std::string tune_path;
  std::map<std::string, int> tune_choice;
  DevBuf tune_save;
  Arena arena;
  hipEvent_t ev[8];
  hipEvent_t kev[8];  // brackets of single kernels: fwd, GEMMs, adjoint, gather
  bool ev_ready = false;
  bool ev_on = true;
  // stage and kernel timing events (tgx_set_timing; off by default: every
  // event between two kernels leaves the GPU idle for several microseconds)
  bool timing = false;
  MdState md;
};

}  // namespace

// The context keeps an Extra block next to it.
struct tgx_ctx_full {
  tgx_ctx c;
  Extra x;
};

static Extra& X(tgx_ctx* c) { return reinterpret_cast<tgx_ctx_full*>(c)->x; }

// Makes b hold at least want bytes: DevBuf::ensure (which may allocate), or,
// while the context's arena is on, a region of the scratch array (a buffer
// carved earlier in the same call is reused when large enough).  Non-zero on
// failure.
static int grow(tgx_ctx* c, DevBuf& b, size_t want) {
  Arena& a = X(c).arena;
  if (!a.on) return b.ensure(want);
  bool carved = false;
  for (const auto& e : a.saved)
    if (e.first == &b) carved = true;
  if (carved && b.bytes >= want) return 0;
  const size_t off = (a.off + 255) & ~(size_t)255;
  if (want > a.size || off > a.size - want) {
    a.overflow = true;
    return 1;
  }
  if (!carved) a.saved.emplace_back(&b, b);
  b.p = a.base + off;
  b.bytes = want;
  a.off = off + want;
  return 0;
}

static int grow_err(tgx_ctx* c, size_t want, const char* what) {
  if (X(c).arena.on && X(c).arena.overflow)
    return set_err(c, TGX_ERR_CAPACITY,
                   "the call needs more than the %zu-byte scratch array (%s: %zu bytes at "
                   "offset %zu)",
                   X(c).arena.size, what, want, X(c).arena.off);
  return set_err(c, TGX_ERR_HIP, "device allocation of %zu bytes failed (%s)", want, what);
}

static size_t tsize(int prec) { return prec == TGX_FP64 ? 8 : 4; }

// Neighbour-list capacity per atom for n atoms in a box at cutoff rcut.
static int guess_maxnb(int n_total_images, double volume, double rcut) {
  double rho = n_total_images / volume;
  double est = 4.0 / 3.0 * M_PI * rcut * rcut * rcut * rho;
  int m = (int)(est * 1.6) + 16;
  return (m + 7) / 8 * 8;
}

// Build the padded, sorted neighbour list (and reverse index) on the device.
static int build_list(tgx_ctx* c, int n, const Cell& cell, double rcut, bool want_rev) {
  Extra& x = X(c);
  const int ncell = cell.small ? 1 : cell.mx * cell.my * cell.mz;
  ENSURE(c, cell_of, sizeof(int) * n);
  ENSURE(c, cell_atoms, cell_atoms_bytes(n));
  ENSURE(c, cell_start, sizeof(int) * (ncell + 1));
  if (grow(c, x.cell_count, sizeof(int) * ncell) || grow(c, x.cell_fill, sizeof(int) * ncell))
    return grow_err(c, sizeof(int) * ncell, "cells");
  ENSURE(c, nb_count, sizeof(int) * n);
  ENSURE(c, nb_flag, sizeof(int) * 4);
  int maxnb = c->maxnb;
  if (maxnb <= 0) maxnb = guess_maxnb(n * cell.nrep, cell.La * cell.Lb * cell.Lc, rcut);
  for (int attempt = 0; attempt < 3; attempt++) {
    ENSURE(c, nb_list, sizeof(int) * (size_t)n * maxnb);
    if (want_rev) ENSURE(c, nb_rev, sizeof(int) * (size_t)n * maxnb);
    TGX_HIP(c, hipMemsetAsync(c->nb_flag.p, 0, sizeof(int) * 4, c->stream));
    TGX_HIP(c, launch_neighbors(c->pos.as<double>(), n, cell, rcut, c->cell_of.as<int>(),
                                x.cell_count.as<int>(), c->cell_start.as<int>(),
                                x.cell_fill.as<int>(), c->cell_atoms.as<int>(),
                                c->nb_count.as<int>(), maxnb, c->nb_list.as<int>(),
                                want_rev ? c->nb_rev.as<int>() : nullptr, c->nb_flag.as<int>(),
                                c->stream, nullptr));
    int flag[4];
    TGX_HIP(c, hipMemcpyAsync(flag, c->nb_flag.p, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    if (flag[0] > maxnb) {
      maxnb = (flag[0] + flag[0] / 4 + 8 + 7) / 8 * 8;
      continue;
    }
    if (flag[1]) return set_err(c, TGX_ERR_STATE, "neighbour list is not symmetric");
    c->maxnb = maxnb;
    return TGX_OK;
  }
  return set_err(c, TGX_ERR_CAPACITY, "neighbour capacity could not be sized");
}

// Image atoms and their list for a cell of any shape (energies, forces,
// descriptors).  The uploaded structure moves to g_pos / g_sp, the n * nt
// image atoms (wrapped atoms plus each lattice translation) fill pos /
// species, and the rows over them (with the reverse index when forces are
// wanted) are built; the pair kernels then take the vectors as they are
// (*open_out has no minimum image).  Brute force over all images, as the
// reference's search for such cells.
static int prepare_image_cell(tgx_ctx* c, int n, const double* lattice, double rcut,
                              bool want_rev, Cell* open_out, int* n_ext_out) {
  Extra& x = X(c);
  ImgCell g;
  std::string err;
  int rcode = make_image_cell(lattice, rcut, &g, &err);
  if (rcode) return set_err(c, rcode, "%s", err.c_str());
  const long long next = (long long)n * g.nt;
  if (next >= (1LL << TGX_IMG_SHIFT))
    return set_err(c, TGX_ERR_UNSUPPORTED, "%d atoms x %d lattice translations: too many image atoms",
                   n, g.nt);
  if (grow(c, x.g_pos, 24 * (size_t)n) || grow(c, x.g_sp, 4 * (size_t)n))
    return grow_err(c, 28 * (size_t)n, "structure copy");
  TGX_HIP(c, hipMemcpyAsync(x.g_pos.p, c->pos.p, 24 * (size_t)n, hipMemcpyDeviceToDevice,
                            c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.g_sp.p, c->species.p, 4 * (size_t)n, hipMemcpyDeviceToDevice,
                            c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));  // before pos / species may be replaced
  ENSURE(c, pos, 24 * (size_t)next);
  ENSURE(c, species, 4 * (size_t)next);
  TGX_HIP(c, launch_image_atoms(x.g_pos.as<double>(), x.g_sp.as<int>(), n, g, c->pos.as<double>(),
                                c->species.as<int>(), c->stream));
  ENSURE(c, nb_count, sizeof(int) * n);
  ENSURE(c, nb_flag, sizeof(int) * 4);
  int maxnb = c->maxnb > 0 ? c->maxnb : guess_maxnb(n, g.vol, rcut);
  for (int attempt = 0; attempt < 3; attempt++) {
    ENSURE(c, nb_list, sizeof(int) * (size_t)n * maxnb);
    if (want_rev) ENSURE(c, nb_rev, sizeof(int) * (size_t)n * maxnb);
    TGX_HIP(c, hipMemsetAsync(c->nb_flag.p, 0, sizeof(int) * 4, c->stream));
    TGX_HIP(c, launch_neighbors_image(c->pos.as<double>(), n, g, rcut, maxnb,
                                      c->nb_count.as<int>(), c->nb_list.as<int>(),
                                      want_rev ? c->nb_rev.as<int>() : nullptr,
                                      c->nb_flag.as<int>(), c->stream));
    int flag[4];
    TGX_HIP(c, hipMemcpyAsync(flag, c->nb_flag.p, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    if (flag[0] > maxnb) {
      maxnb = (flag[0] + flag[0] / 4 + 8 + 7) / 8 * 8;
      continue;
    }
    if (flag[1]) return set_err(c, TGX_ERR_STATE, "neighbour list is not symmetric");
    c->maxnb = maxnb;
    *open_out = open_cell();
    *n_ext_out = (int)next;
    return TGX_OK;
  }
  return set_err(c, TGX_ERR_CAPACITY, "neighbour capacity could not be sized");
}

static double rcut_all(const tgx_ctx* c) {
  double r = 0.0;
  if (c->has_soap)
    for (const auto& b : c->blocks) r = std::max(r, b->sd.rcut_max);
  for (const auto& b : c->twob) r = std::max(r, b.rcut);
  for (const auto& b : c->a3b) r = std::max(r, b.rcut);
  for (const auto& b : c->corep) r = std::max(r, b.rcut);
  return r;
}

static bool has_model(const tgx_ctx* c) {
  return c->has_soap || !c->twob.empty() || !c->a3b.empty() || !c->corep.empty();
}

// angle_3b and core_pot blocks in turn (turbogap_main.f90:809-840), after
// the 2-body terms: energies, forces and per-atom virials added.
static int launch_extra_terms(tgx_ctx* c, int n, const Geom& cell, int do_forces) {
  Extra& x = X(c);
  for (const auto& b : c->a3b) {
    if (do_forces) {
      if (grow(c, x.g3, 32 * (size_t)n * c->maxnb))
        return grow_err(c, 32 * (size_t)n * c->maxnb, "angle_3b slots");
    }
    TGX_HIP(c, launch_angle_3b(b, c->pos.as<double>(), c->species.as<int>(), n, cell, c->maxnb,
                               c->nb_count.as<int>(), c->nb_list.as<int>(), c->eatom.as<double>(),
                               c->virial.as<double>(), do_forces ? x.g3.as<double>() : nullptr,
                               do_forces, c->stream));
    if (do_forces)
      TGX_HIP(c, launch_force_gather<double>(n, c->maxnb, c->nb_count.as<int>(),
                                             c->nb_list.as<int>(), c->nb_rev.as<int>(), nullptr,
                                             1, x.g3.as<double>(), c->forces.as<double>(),
                                             c->stream, 0));
  }
  for (const auto& b : c->corep)
    TGX_HIP(c, launch_core_pot(b, c->pos.as<double>(), c->species.as<int>(), n, cell, c->maxnb,
                               c->nb_count.as<int>(), c->nb_list.as<int>(), c->eatom.as<double>(),
                               c->forces.as<double>(), c->virial.as<double>(), do_forces,
                               c->stream));
  return TGX_OK;
}

// Sites of every SOAP block: all atoms, or those of its central_species.
// *n_sites receives the first block's count (each block keeps its own).
// With image atoms (n_ext > n) site_of_atom also covers them: the site of
// image J is that of atom J % n.
static int setup_block_sites(tgx_ctx* c, SoapBlock* bp, int n, const int* species_host,
                             int n_ext);

static int setup_sites(tgx_ctx* c, int n, const int* species_host, int* n_sites,
                       int n_ext = 0) {
  for (auto& bp : c->blocks) {
    const int rcode = setup_block_sites(c, bp.get(), n, species_host, n_ext);
    if (rcode) return rcode;
  }
  *n_sites = c->blocks[0]->n_sites;
  return TGX_OK;
}

// Sites of one block (see setup_sites).
static int setup_block_sites(tgx_ctx* c, SoapBlock* bp, int n, const int* species_host,
                             int n_ext) {
  SoapBlock* const keep = c->blk;
  if (n_ext < n) n_ext = n;
  {
    c->blk = bp;
    SoapBlock& b = *bp;
    ENSURE(c, blk->site_of_atom, sizeof(int) * n_ext);
    ENSURE(c, blk->site_atom, sizeof(int) * n);
    std::vector<int> soa(n_ext), sa;
    sa.reserve(n);
    for (int i = 0; i < n; i++) {
      if (c->has_soap && (b.sd.central_species == 0 || species_host[i] == b.sd.central_species)) {
        soa[i] = (int)sa.size();
        sa.push_back(i);
      } else {
        soa[i] = -1;
      }
    }
    for (int J = n; J < n_ext; J++) soa[J] = soa[J % n];
    TGX_HIP(c, hipMemcpyAsync(b.site_of_atom.p, soa.data(), sizeof(int) * n_ext,
                              hipMemcpyHostToDevice, c->stream));
    if (!sa.empty())
      TGX_HIP(c, hipMemcpyAsync(b.site_atom.p, sa.data(), sizeof(int) * sa.size(),
                                hipMemcpyHostToDevice, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    b.n_sites = (int)sa.size();
    b.site_ident = b.n_sites == n;
  }
  c->blk = keep;
  return TGX_OK;
}

static void kev_rec(tgx_ctx* c, int k);

// ---- autotuner -------------------------------------------------------------
// Launch shapes with more than one good choice (the kernel-step path and its
// GEMM tiles, the SOAP kernels' waves per workgroup) are picked by timing the
// candidates on the call's own buffers, once per shape key, when tuning is
// on (tgx_set_autotune, or TGX_AUTOTUNE=1).  Decisions are kept per context
// and, with a cache file, appended to it and read back by later contexts.
// With tuning off the stored decisions or else the built-in choices apply,
// so that results never depend on timing noise.
static bool tune_enabled(tgx_ctx* c) {
  Extra& x = X(c);
  if (x.tune_on < 0) {
    const char* e = getenv("TGX_AUTOTUNE");
    x.tune_on = (e && e[0] == '1') ? 1 : 0;
  }
  return x.tune_on == 1;
}

static void tune_load(Extra& x) {
  if (x.tune_path.empty()) return;
  FILE* f = fopen(x.tune_path.c_str(), "r");
  if (!f) return;
  char line[512];
  while (fgets(line, sizeof line, f)) {
    char* sp = strrchr(line, ' ');
    if (!sp) continue;
    *sp = 0;
    x.tune_choice[line] = atoi(sp + 1);
  }
  fclose(f);
}

static std::string bucket_key(const char* kind, int a, int b, int c2, int d, int n) {
  int m = 64;
  while (m < n) m <<= 1;
  char buf[160];
  snprintf(buf, sizeof buf, "%s/%d/%d/%d/%d/%d", kind, a, b, c2, d, m);
  return buf;
}

// The decision for key among n candidates: the stored one; else, when tuning
// is on and the stream is not being captured, the fastest candidate (one
// warm-up launch, then five timed with events; launch(k) enqueues candidate
// k and must give the same results when repeated); else dflt.  The caller
// launches the decision afterwards, so its results stand.
template <typename F>
static int tune_pick(tgx_ctx* c, const std::string& key, int n, int dflt, F&& launch) {
  Extra& x = X(c);
  auto it = x.tune_choice.find(key);
  if (it != x.tune_choice.end() && it->second >= 0 && it->second < n) return it->second;
  if (n < 2 || !tune_enabled(c)) return dflt;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return dflt;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return dflt;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return dflt;
  }
  // another candidate replaces the built-in one only when it is at least 3 %
  // faster, so that timing noise between equal shapes changes nothing
  int best = dflt;
  float best_ms = 1e30f, dflt_ms = 1e30f;
  for (int k = 0; k < n; k++) {
    bool ok = launch(k) == hipSuccess;
    ok = ok && hipEventRecord(e0, c->stream) == hipSuccess;
    for (int r = 0; r < 5 && ok; r++) ok = launch(k) == hipSuccess;
    ok = ok && hipEventRecord(e1, c->stream) == hipSuccess &&
         hipEventSynchronize(e1) == hipSuccess;
    float ms = 0.f;
    if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
      if (k == dflt) dflt_ms = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = k;
      }
    }
    if (!ok) (void)hipGetLastError();  // a candidate that does not launch is skipped
  }
  if (dflt_ms < 1e30f && best_ms > 0.97f * dflt_ms) best = dflt;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  x.tune_choice[key] = best;
  if (!x.tune_path.empty()) {
    FILE* f = fopen(x.tune_path.c_str(), "a");
    if (f) {
      fprintf(f, "%s %d\n", key.c_str(), best);
      fclose(f);
    }
  }
  return best;
}

// Waves per workgroup of the SOAP kernels: 0 = the built-in choice.
static const int kTuneWpb[4] = {0, 1, 2, 4};

// SOAP forward for the current list into c->blk->q (and c->blk->qb in mixed mode).
// few: a few sites (a Monte Carlo trial) take the workgroup-per-site FP32
// kernel; TGX_FWD_SPLIT=1 asks for it on every FP32 call of up to 256 sites.
// ident_ok: the sites are the block's own (setup_block_sites), not a Monte
// Carlo trial's list, so an all-atom block may skip the site_atom lookup.
static int soap_forward(tgx_ctx* c, int n, const Geom& cell, int n_sites, bool few = false,
                        bool ident_ok = false) {
  Extra& x = X(c);
  const int prec = c->precision;
  const size_t ts = tsize(prec);
  const int NRP = c->blk->sd.nrp;
  SoapBlock* const b = c->blk;
  const bool cmp = b->cmp > 0;
  if (cmp && few)
    return set_err(c, TGX_ERR_UNSUPPORTED, "compress_soap blocks are not implemented for Monte Carlo");
  ENSURE(c, blk->q, ts * (size_t)std::max(n_sites, 1) * c->blk->dim_pad);
  ENSURE(c, blk->cbuf, ts * (size_t)std::max(n_sites, 1) * NRP * c->blk->sd.L2);
  ENSURE(c, blk->pnorm, ts * (size_t)std::max(n_sites, 1));
  unsigned short* qb = nullptr;
  if (prec == TGX_MIXED) {
    ENSURE(c, blk->qb, 2 * (size_t)std::max(n_sites, 1) * c->blk->dim_pad);
    qb = c->blk->qb.as<unsigned short>();
  }
  // with compress_soap the expansion writes the uncompressed descriptor and
  // its norm (q_f, |p|); the compression pass below writes q (and its BF16
  // copy) and |P p|
  if (cmp) {
    ENSURE(c, blk->qf, ts * (size_t)std::max(n_sites, 1) * b->dimf_pad);
    ENSURE(c, blk->pnf, ts * (size_t)std::max(n_sites, 1));
  }
  void* const fq = cmp ? b->qf.p : b->q.p;
  void* const fpn = cmp ? b->pnf.p : b->pnorm.p;
  const int fld = cmp ? b->dimf_pad : b->dim_pad;
  unsigned short* const fqb = cmp ? nullptr : qb;
  const char* fs = getenv("TGX_FWD_SPLIT");
  const int split = (few || (fs && fs[0] == '1')) ? 1 : 0;
  const int* sat = (ident_ok && !few && b->site_ident) ? nullptr : b->site_atom.as<int>();
  auto fwd = [&](int wpb) -> hipError_t {
    if (prec == TGX_FP64)
      return launch_soap_fwd<double>(c->blk->sd, c->blk->sd_dev.as<SoapDev>(), c->blk->W_dev.as<double>(),
                                     c->blk->rtab.as<float>(), c->pos.as<double>(),
                                     c->species.as<int>(), cell, c->maxnb,
                                     c->nb_count.as<int>(), c->nb_list.as<int>(),
                                     sat, n_sites, c->blk->pairn.as<int>(),
                                     static_cast<double*>(fq), fld, nullptr, 0,
                                     c->blk->cbuf.as<double>(), static_cast<double*>(fpn), c->stream, 0,
                                     wpb);
    return launch_soap_fwd<float>(c->blk->sd, c->blk->sd_dev.as<SoapDev>(), c->blk->W_dev.as<double>(),
                                  c->blk->rtab.as<float>(), c->pos.as<double>(), c->species.as<int>(),
                                  cell, c->maxnb, c->nb_count.as<int>(), c->nb_list.as<int>(),
                                  sat, n_sites, c->blk->pairn.as<int>(),
                                  static_cast<float*>(fq), fld, fqb, c->blk->dim_pad,
                                  c->blk->cbuf.as<float>(), static_cast<float*>(fpn), c->stream, split,
                                  wpb);
  };
  // TGX_SOAP_WPB=1|2|4 forces the waves per workgroup (for tests)
  const char* fw = getenv("TGX_SOAP_WPB");
  int wpb = fw ? atoi(fw) : 0;
  if (!split && !fw) {
    const std::string key =
        bucket_key("fwd", prec, c->blk->sd.nrp, c->blk->sd.L2, cell.many ? 1 : 0, n_sites);
    wpb = kTuneWpb[tune_pick(c, key, 4, 0, [&](int k) { return fwd(kTuneWpb[k]); })];
  }
  kev_rec(c, 0);
  TGX_HIP(c, fwd(wpb));
  kev_rec(c, 1);
  if (cmp && n_sites > 0) {
    if (prec == TGX_FP64)
      TGX_HIP(c, launch_compress_fwd<double>(n_sites, b->sd.dim, b->dimf_pad, b->qf.as<double>(),
                                             b->pnf.as<double>(), b->dim, b->dim_pad,
                                             b->c_roff.as<int>(), b->c_rcol.as<int>(),
                                             b->c_rval.as<double>(), b->q.as<double>(), nullptr,
                                             0, b->pnorm.as<double>(), c->stream));
    else
      TGX_HIP(c, launch_compress_fwd<float>(n_sites, b->sd.dim, b->dimf_pad, b->qf.as<float>(),
                                            b->pnf.as<float>(), b->dim, b->dim_pad,
                                            b->c_roff.as<int>(), b->c_rcol.as<int>(),
                                            b->c_rval.as<double>(), b->q.as<float>(), qb,
                                            b->dim_pad, b->pnorm.as<float>(), c->stream));
  }
  (void)n;
  return TGX_OK;
}

// Inputs of the SOAP adjoint of block c->blk: the descriptor, its norm and
// dE/dq, each of the power spectrum.  With compress_soap the kernel step's
// dE/dq of the compressed descriptor is first taken back to the power
// spectrum (after the kernel step with the gradient).
struct AdjIn {
  void* q;
  int q_ld;
  void* pn;
  void* G;
  int g_ld;
};

static int adjoint_inputs(tgx_ctx* c, int n_sites, AdjIn* a) {
  SoapBlock* const b = c->blk;
  if (!b->cmp) {
    *a = AdjIn{b->q.p, b->dim_pad, b->pnorm.p, b->gmat.p, b->dim_pad};
    return TGX_OK;
  }
  ENSURE(c, blk->gf, tsize(c->precision) * (size_t)std::max(n_sites, 1) * b->dimf_pad);
  if (n_sites > 0) {
    if (c->precision == TGX_FP64)
      TGX_HIP(c, launch_compress_bwd<double>(n_sites, b->dim, b->dim_pad, b->q.as<double>(),
                                             b->pnorm.as<double>(), b->gmat.as<double>(),
                                             b->dim_pad, b->sd.dim, b->dimf_pad,
                                             b->pnf.as<double>(), b->c_coff.as<int>(),
                                             b->c_crow.as<int>(), b->c_cval.as<double>(),
                                             b->gf.as<double>(), c->stream));
    else
      TGX_HIP(c, launch_compress_bwd<float>(n_sites, b->dim, b->dim_pad, b->q.as<float>(),
                                            b->pnorm.as<float>(), b->gmat.as<float>(),
                                            b->dim_pad, b->sd.dim, b->dimf_pad,
                                            b->pnf.as<float>(), b->c_coff.as<int>(),
                                            b->c_crow.as<int>(), b->c_cval.as<double>(),
                                            b->gf.as<float>(), c->stream));
  }
  *a = AdjIn{b->qf.p, b->dimf_pad, b->pnf.p, b->gf.p, b->dimf_pad};
  return TGX_OK;
}

// Kernel step on the descriptors in c->blk->q / c->blk->qb.
// With eatom_out, every atom is a site (in atom order) and the kernel step
// also writes eatom_out[i] = e0[species[i] - 1] + esite[i].
static int kernel_step(tgx_ctx* c, int n_sites, bool want_grad, const int* guard = nullptr,
                       double* eatom_out = nullptr, int* few_parts = nullptr) {
  const int prec = c->precision;
  const size_t tin = prec == TGX_FP64 ? 8 : (prec == TGX_FP32 ? 4 : 2);
  const size_t tacc = prec == TGX_FP64 ? 8 : 4;
  const int ns = std::max(n_sites, 1);
  ENSURE(c, blk->wmat, tin * (size_t)ns * c->blk->sp_pad);
  ENSURE(c, blk->epart, tacc * (size_t)ns * (c->blk->sp_pad / 32));  // per 32 sparse points at most
  ENSURE(c, blk->esite, 8 * (size_t)ns);
  if (want_grad) ENSURE(c, blk->gmat, tacc * (size_t)ns * c->blk->dim_pad);
  const void *qa, *qs, *qsT, *al;
  if (prec == TGX_FP64) {
    qa = c->blk->q.p; qs = c->blk->Qs64.p; qsT = c->blk->QsT64.p; al = c->blk->alph64.p;
  } else if (prec == TGX_FP32) {
    qa = c->blk->q.p; qs = c->blk->Qs32.p; qsT = c->blk->QsT32.p; al = c->blk->alph32.p;
  } else {
    qa = c->blk->qb.p; qs = c->blk->Qs16.p; qsT = c->blk->QsT16.p; al = c->blk->alph32.p;
  }
  auto step = [&](const GemmPlan* plan) {
    return launch_gap_kernel(prec, n_sites, c->blk->dim, c->blk->dim_pad, c->blk->n_sparse, c->blk->sp_pad, qa, qs,
                             qsT, al, c->blk->delta, c->blk->zeta, c->blk->wmat.p, c->blk->epart.p,
                             c->blk->esite.as<double>(), want_grad ? c->blk->gmat.p : nullptr,
                             want_grad ? 1 : 0, guard, c->stream, eatom_out,
                             eatom_out ? X(c).e0_dev.as<double>() : nullptr,
                             eatom_out ? c->species.as<int>() : nullptr, few_parts, plan);
  };
  // candidates: the built-in path, then the separate GEMMs with each tile
  GemmPlan plans[5] = {{0, 0, 0}};
  int np = 1;
  if (prec == TGX_MIXED) plans[np++] = GemmPlan{2, 0, 0};
  plans[np++] = GemmPlan{2, 1, 0};
  plans[np++] = GemmPlan{2, 2, 0};
  if (want_grad && prec != TGX_FP64) plans[np++] = GemmPlan{2, 0, 1};
  int pk = 0;
  // TGX_GEMM_PLAN="path t0 t1" forces one plan (for tests of every shape)
  const char* fp = getenv("TGX_GEMM_PLAN");
  GemmPlan forced{0, 0, 0};
  const bool force = fp && sscanf(fp, "%d %d %d", &forced.path, &forced.t0, &forced.t1) == 3;
  if (!guard && !force) {  // a guarded (Monte Carlo) step may not run: nothing to time
    const std::string key =
        bucket_key("step", prec, want_grad ? 1 : 0, c->blk->sp_pad, c->blk->dim_pad, n_sites);
    pk = tune_pick(c, key, np, 0, [&](int k) { return step(&plans[k]); });
  }
  kev_rec(c, 2);
  TGX_HIP(c, step(force ? &forced : &plans[pk]));
  kev_rec(c, 3);
  return TGX_OK;
}

static void ev_rec(tgx_ctx* c, int k) {
  Extra& x = X(c);
  if (x.ev_ready && x.ev_on && x.timing) (void)hipEventRecord(x.ev[k], c->stream);
}
static void kev_rec(tgx_ctx* c, int k) {
  Extra& x = X(c);
  if (x.ev_ready && x.ev_on && x.timing) (void)hipEventRecord(x.kev[k], c->stream);
}

static int ef_launch(tgx_ctx* c, int n, int n_sites, const Geom& cell, int do_forces,
                     bool events, int n_struct = 0, bool defer_final = false);

// Energies/forces/virial of the structure whose positions and species are
// already in c->pos / c->species.  Results stay on the device:
// c->eatom, c->forces, X(c).out[0..8] virial, [9] total energy.
static int run_ef(tgx_ctx* c, int n, const double* lattice, const int* species_host,
                  int do_forces) {
  Extra& x = X(c);
  if (!has_model(c)) return set_err(c, TGX_ERR_STATE, "no model loaded");
  Cell cell;
  std::string err;
  const double rc = rcut_all(c) + c->list_buffer;
  if (!is_orthorhombic(lattice)) {
    x.ev_on = true;
    ev_rec(c, 0);
    int n_ext = n;
    int rcode = prepare_image_cell(c, n, lattice, rc, do_forces != 0, &cell, &n_ext);
    if (rcode) return rcode;
    ev_rec(c, 1);
    int n_sites = 0;
    rcode = setup_sites(c, n, species_host, &n_sites, n_ext);
    if (rcode) return rcode;
    return ef_launch(c, n, n_sites, Geom{cell}, do_forces, true);
  }
  int rcode = make_cell(lattice, rc, &cell, &err);
  if (rcode) return set_err(c, rcode, "%s", err.c_str());
  if ((long long)n * cell.nrep >= (1LL << TGX_IMG_SHIFT))
    return set_err(c, TGX_ERR_ARG, "too many atoms");
  if (cell.nrep >= (1 << (31 - TGX_IMG_SHIFT)))
    return set_err(c, TGX_ERR_UNSUPPORTED, "box too small for the cutoff (%d images)", cell.nrep);
  x.ev_on = true;
  ev_rec(c, 0);
  rcode = build_list(c, n, cell, rc, do_forces != 0);
  if (rcode) return rcode;
  ev_rec(c, 1);
  int n_sites = 0;
  rcode = setup_sites(c, n, species_host, &n_sites);
  if (rcode) return rcode;
  return ef_launch(c, n, n_sites, Geom{cell}, do_forces, true);
}

// The SOAP adjoint adds the distance_2b terms itself (one pass over the
// pairs instead of two) when every atom is a SOAP site and every 2b block's
// cutoff lies within the SOAP cutoff of both its species.
static bool fuse_two_body(const tgx_ctx* c, int n, int n_sites, int do_forces) {
  if (!do_forces || !c->has_soap || c->twob.empty() || n_sites != n || c->blocks.size() != 1 ||
      !c->a3b.empty() || !c->corep.empty())
    return false;
  for (const auto& b : c->twob) {
    if (b.sp1 > c->blk->sd.n_species || b.sp2 > c->blk->sd.n_species) return false;
    if (b.rcut > c->blk->sd.rh[b.sp1 - 1] || b.rcut > c->blk->sd.rh[b.sp2 - 1]) return false;
  }
  return true;
}

// Energy/force launch sequence on the current list and sites.  Asynchronous:
// no host synchronisation, no allocation once the buffers are sized (the MD
// step graph captures it).
static int ef_launch_multi(tgx_ctx* c, int n, const Geom& cell, int do_forces, bool events,
                           int n_struct, bool defer_final);

static int ef_launch(tgx_ctx* c, int n, int n_sites, const Geom& cell, int do_forces,
                     bool events, int n_struct, bool defer_final) {
  if (c->has_soap && c->blocks.size() > 1)
    return ef_launch_multi(c, n, cell, do_forces, events, n_struct, defer_final);
  Extra& x = X(c);
  int rcode;
  x.ev_on = events;
  const int prec = c->precision;
  ENSURE(c, eatom, 8 * (size_t)n);
  ENSURE(c, forces, 24 * (size_t)n);
  ENSURE(c, blk->vsite, 48 * (size_t)std::max(n_sites, 1));
  ENSURE(c, virial, 48 * (size_t)n);  // per-atom 2b virial
  if (grow(c, x.out, (16 + 8 * 128) * sizeof(double)))
    return grow_err(c, (16 + 8 * 128) * sizeof(double), "out");
  // fused 2b: every atom is a site, so the kernel step writes the atom
  // energies, the force gather overwrites the forces and no per-atom virial
  // is used -- nothing to initialise
  const bool fuse = fuse_two_body(c, n, n_sites, do_forces);
  bool red_in_gather = false;  // the reduction's partial rows done by the force gather
  if (c->has_soap && n_sites > 0) {
    rcode = soap_forward(c, n, cell, n_sites, false, true);
    if (rcode) return rcode;
    ev_rec(c, 2);
    rcode = kernel_step(c, n_sites, do_forces != 0, nullptr,
                        fuse ? c->eatom.as<double>() : nullptr);
    if (rcode) return rcode;
    ev_rec(c, 3);
  } else {
    ev_rec(c, 2);
    ev_rec(c, 3);
  }
  if (!fuse)
    TGX_HIP(c, launch_energy_init(n, c->species.as<int>(), x.e0_dev.as<double>(),
                                  c->blk->site_of_atom.as<int>(),
                                  (c->has_soap && n_sites > 0) ? c->blk->esite.as<double>() : nullptr,
                                  c->eatom.as<double>(), c->forces.as<double>(),
                                  c->virial.as<double>(), c->stream));
  if (do_forces && c->has_soap && n_sites > 0) {
    const size_t ts = tsize(prec);
    ENSURE(c, gpair, ts * 4 * (size_t)n * c->maxnb);  // (g, 0) per list slot
    AdjIn ai;
    int arc = adjoint_inputs(c, n_sites, &ai);
    if (arc) return arc;
    auto bwd = [&](int wpb) -> hipError_t {
      if (prec == TGX_FP64)
        return launch_soap_bwd<double>(
            c->blk->sd, c->blk->sd_dev.as<SoapDev>(), c->blk->W_dev.as<double>(), c->blk->rtab.as<float>(),
            c->pos.as<double>(), c->species.as<int>(), cell, c->maxnb, c->nb_count.as<int>(),
            c->nb_list.as<int>(), (c->blk->site_ident ? nullptr : c->blk->site_atom.as<int>()), n_sites, c->blk->pairn.as<int>(),
            static_cast<const double*>(ai.q), ai.q_ld, c->blk->cbuf.as<double>(),
            static_cast<const double*>(ai.pn), static_cast<const double*>(ai.G), ai.g_ld,
            c->gpair.as<double>(), c->blk->vsite.as<double>(),
            fuse ? x.twob_store.as<TwoBodyDev>() : nullptr, fuse ? (int)c->twob.size() : 0,
            c->eatom.as<double>(), c->stream, wpb);
      return launch_soap_bwd<float>(
          c->blk->sd, c->blk->sd_dev.as<SoapDev>(), c->blk->W_dev.as<double>(), c->blk->rtab.as<float>(),
          c->pos.as<double>(), c->species.as<int>(), cell, c->maxnb, c->nb_count.as<int>(),
          c->nb_list.as<int>(), (c->blk->site_ident ? nullptr : c->blk->site_atom.as<int>()), n_sites, c->blk->pairn.as<int>(),
          static_cast<const float*>(ai.q), ai.q_ld, c->blk->cbuf.as<float>(),
          static_cast<const float*>(ai.pn), static_cast<const float*>(ai.G), ai.g_ld,
          c->gpair.as<float>(), c->blk->vsite.as<double>(),
          fuse ? x.twob_store.as<TwoBodyDev>() : nullptr, fuse ? (int)c->twob.size() : 0,
          c->eatom.as<double>(), c->stream, wpb);
    };
    // the fused 2-body term adds to the atom energies: while candidates are
    // timed, every launch starts from a copy of them
    const std::string bkey = bucket_key("bwd", prec, c->blk->sd.nrp, c->blk->sd.L2,
                                        (fuse ? 2 : 0) + (cell.many ? 1 : 0), n_sites);
    bool snap = false;
    if (fuse && tune_enabled(c) && !x.tune_choice.count(bkey)) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(c->stream, &cs) == hipSuccess &&
          cs == hipStreamCaptureStatusNone) {
        if (grow(c, x.tune_save, 8 * (size_t)n)) return grow_err(c, 8 * (size_t)n, "tune_save");
        TGX_HIP(c, hipMemcpyAsync(x.tune_save.p, c->eatom.p, 8 * (size_t)n,
                                  hipMemcpyDeviceToDevice, c->stream));
        snap = true;
      }
    }
    auto restore = [&]() -> hipError_t {
      return snap ? hipMemcpyAsync(c->eatom.p, x.tune_save.p, 8 * (size_t)n,
                                   hipMemcpyDeviceToDevice, c->stream)
                  : hipSuccess;
    };
    const char* fw = getenv("TGX_SOAP_WPB");
    const int wk = fw ? 0 : tune_pick(c, bkey, 4, 0, [&](int k) {
      const hipError_t r = restore();
      return r == hipSuccess ? bwd(kTuneWpb[k]) : r;
    });
    TGX_HIP(c, restore());
    kev_rec(c, 4);
    hipError_t e = bwd(fw ? atoi(fw) : kTuneWpb[wk]);
    kev_rec(c, 5);
    kev_rec(c, 6);
    // with the 2-body term fused and no further terms, nothing after the
    // gather writes the energies or virials: the same launch computes the
    // partial rows of their reduction
    red_in_gather = fuse && c->a3b.empty() && c->corep.empty() && !cell.many;
    const ReduceArgs ra{n_sites, n, c->blk->vsite.as<double>(), nullptr, c->eatom.as<double>(),
                        x.out.as<double>() + 16};
    if (e == hipSuccess) {
      if (prec == TGX_FP64)
        e = launch_force_gather<double>(n, c->maxnb, c->nb_count.as<int>(),
                                        c->nb_list.as<int>(), c->nb_rev.as<int>(),
                                        c->blk->site_of_atom.as<int>(), n_sites == n ? 1 : 0,
                                        c->gpair.as<double>(), c->forces.as<double>(),
                                        c->stream, fuse ? 1 : 0, red_in_gather ? &ra : nullptr);
      else
        e = launch_force_gather<float>(n, c->maxnb, c->nb_count.as<int>(),
                                       c->nb_list.as<int>(), c->nb_rev.as<int>(),
                                       c->blk->site_of_atom.as<int>(), n_sites == n ? 1 : 0,
                                       c->gpair.as<float>(), c->forces.as<double>(), c->stream,
                                       fuse ? 1 : 0, red_in_gather ? &ra : nullptr);
    }
    TGX_HIP(c, e);
    kev_rec(c, 7);
  }
  ev_rec(c, 4);
  for (const auto& b : c->twob)
    if (!fuse)
      TGX_HIP(c, launch_two_body(b, c->pos.as<double>(), c->species.as<int>(), n, cell, c->maxnb,
                               c->nb_count.as<int>(), c->nb_list.as<int>(), c->eatom.as<double>(),
                               c->forces.as<double>(), c->virial.as<double>(), do_forces,
                               prec == TGX_FP64, c->stream));
  rcode = launch_extra_terms(c, n, cell, do_forces);
  if (rcode) return rcode;
  ev_rec(c, 5);
  if (cell.many) {
    if (x.b_out.ensure(sizeof(double) * 10 * (size_t)n_struct))
      return set_err(c, TGX_ERR_HIP, "allocation failed");
    TGX_HIP(c, launch_reduce_many(n_struct, x.b_aoff.as<int>(), c->blk->site_of_atom.as<int>(),
                                  (do_forces && c->has_soap && n_sites > 0) ? c->blk->vsite.as<double>()
                                                                            : nullptr,
                                  (do_forces && !fuse) ? c->virial.as<double>() : nullptr,
                                  c->eatom.as<double>(), x.b_out.as<double>(), c->stream));
  } else if (red_in_gather) {
    if (!defer_final) TGX_HIP(c, launch_reduce_final(x.out.as<double>(), c->stream));
  } else {
    TGX_HIP(c, launch_reduce((do_forces && c->has_soap) ? n_sites : 0,
                             (do_forces && c->has_soap) ? c->blk->vsite.as<double>() : nullptr, n,
                             (do_forces && !fuse) ? c->virial.as<double>() : nullptr,
                             c->eatom.as<double>(), x.out.as<double>(), c->stream, !defer_final));
  }
  ev_rec(c, 6);
  x.ev_on = true;
  return TGX_OK;
}

static void collect_times(tgx_ctx* c) {
  Extra& x = X(c);
  if (!x.ev_ready || !x.ev_on) return;
  if (!x.timing) {  // no figures of this call
    for (int k = 0; k < 7; k++) c->stage_ms[k] = 0.0;
    for (int k = 0; k < 4; k++) c->kernel_ms[k] = 0.0;
    return;
  }
  float ms[6];
  for (int k = 0; k < 6; k++) {
    ms[k] = 0.f;
    (void)hipEventElapsedTime(&ms[k], x.ev[k], x.ev[k + 1]);
  }
  c->stage_ms[0] = ms[0];  // neighbour list
  c->stage_ms[1] = ms[1];  // site setup + SOAP forward
  c->stage_ms[2] = ms[2];  // kernel GEMMs
  c->stage_ms[3] = ms[3];  // SOAP adjoint + force gather
  c->stage_ms[4] = ms[4];  // 2-body
  c->stage_ms[5] = ms[5];  // reductions
  float tot = 0.f;
  (void)hipEventElapsedTime(&tot, x.ev[0], x.ev[6]);
  c->stage_ms[6] = tot;
  for (int k = 0; k < 4; k++) {
    float t = 0.f;
    c->kernel_ms[k] = hipEventElapsedTime(&t, x.kev[2 * k], x.kev[2 * k + 1]) == hipSuccess ? t : 0.0;
  }
  (void)hipGetLastError();
}

// The same launch sequence for a model of several soap_turbo blocks
// (calculate_gap_soap.f90:324-447 evaluates the blocks in turn and adds
// their energies, forces and virials): e0 first, then per block the SOAP
// forward, the kernel step and its site energies added to the atoms, the
// adjoint with its forces gathered onto the atoms and its site virials added
// to the per-atom virials; then the 2-body terms and the reductions.  Every
// sum runs in a fixed order.
static int ef_launch_multi(tgx_ctx* c, int n, const Geom& cell, int do_forces, bool events,
                           int n_struct, bool defer_final) {
  Extra& x = X(c);
  x.ev_on = events;
  const int prec = c->precision;
  ENSURE(c, eatom, 8 * (size_t)n);
  ENSURE(c, forces, 24 * (size_t)n);
  ENSURE(c, virial, 48 * (size_t)n);
  if (grow(c, x.out, (16 + 8 * 128) * sizeof(double)))
    return grow_err(c, (16 + 8 * 128) * sizeof(double), "out");
  TGX_HIP(c, launch_energy_init(n, c->species.as<int>(), x.e0_dev.as<double>(),
                                c->blocks[0]->site_of_atom.as<int>(), nullptr,
                                c->eatom.as<double>(), c->forces.as<double>(),
                                c->virial.as<double>(), c->stream));
  ev_rec(c, 2);
  int rcode = TGX_OK;
  for (auto& bp : c->blocks) {
    c->blk = bp.get();
    const int ns = c->blk->n_sites;
    if (ns == 0) continue;
    rcode = soap_forward(c, n, cell, ns, false, true);
    if (!rcode) rcode = kernel_step(c, ns, do_forces != 0);
    if (rcode) break;
    TGX_HIP(c, launch_site_add(ns, c->blk->site_atom.as<int>(), c->blk->esite.as<double>(),
                               nullptr, c->eatom.as<double>(), nullptr, c->stream));
    if (!do_forces) continue;
    ENSURE(c, blk->vsite, 48 * (size_t)ns);
    const size_t ts = tsize(prec);
    ENSURE(c, gpair, ts * 4 * (size_t)n * c->maxnb);
    AdjIn ai;
    rcode = adjoint_inputs(c, ns, &ai);
    if (rcode) break;
    kev_rec(c, 4);
    hipError_t e;
    if (prec == TGX_FP64)
      e = launch_soap_bwd<double>(
          c->blk->sd, c->blk->sd_dev.as<SoapDev>(), c->blk->W_dev.as<double>(),
          c->blk->rtab.as<float>(), c->pos.as<double>(), c->species.as<int>(), cell, c->maxnb,
          c->nb_count.as<int>(), c->nb_list.as<int>(), (c->blk->site_ident ? nullptr : c->blk->site_atom.as<int>()), ns,
          c->blk->pairn.as<int>(), static_cast<const double*>(ai.q), ai.q_ld,
          c->blk->cbuf.as<double>(), static_cast<const double*>(ai.pn),
          static_cast<const double*>(ai.G), ai.g_ld, c->gpair.as<double>(),
          c->blk->vsite.as<double>(), nullptr, 0, c->eatom.as<double>(), c->stream, 0);
    else
      e = launch_soap_bwd<float>(
          c->blk->sd, c->blk->sd_dev.as<SoapDev>(), c->blk->W_dev.as<double>(),
          c->blk->rtab.as<float>(), c->pos.as<double>(), c->species.as<int>(), cell, c->maxnb,
          c->nb_count.as<int>(), c->nb_list.as<int>(), (c->blk->site_ident ? nullptr : c->blk->site_atom.as<int>()), ns,
          c->blk->pairn.as<int>(), static_cast<const float*>(ai.q), ai.q_ld,
          c->blk->cbuf.as<float>(), static_cast<const float*>(ai.pn),
          static_cast<const float*>(ai.G), ai.g_ld, c->gpair.as<float>(),
          c->blk->vsite.as<double>(), nullptr, 0, c->eatom.as<double>(), c->stream, 0);
    kev_rec(c, 5);
    kev_rec(c, 6);
    if (e == hipSuccess) {
      if (prec == TGX_FP64)
        e = launch_force_gather<double>(n, c->maxnb, c->nb_count.as<int>(),
                                        c->nb_list.as<int>(), c->nb_rev.as<int>(),
                                        c->blk->site_of_atom.as<int>(), ns == n ? 1 : 0,
                                        c->gpair.as<double>(), c->forces.as<double>(),
                                        c->stream, 0);
      else
        e = launch_force_gather<float>(n, c->maxnb, c->nb_count.as<int>(), c->nb_list.as<int>(),
                                       c->nb_rev.as<int>(), c->blk->site_of_atom.as<int>(),
                                       ns == n ? 1 : 0, c->gpair.as<float>(),
                                       c->forces.as<double>(), c->stream, 0);
    }
    if (e == hipSuccess)
      e = launch_site_add(ns, c->blk->site_atom.as<int>(), nullptr, c->blk->vsite.as<double>(),
                          nullptr, c->virial.as<double>(), c->stream);
    kev_rec(c, 7);
    if (e != hipSuccess) {
      c->blk = c->blocks[0].get();
      TGX_HIP(c, e);
    }
  }
  c->blk = c->blocks[0].get();
  if (rcode) return rcode;
  ev_rec(c, 3);
  ev_rec(c, 4);
  for (const auto& b : c->twob)
    TGX_HIP(c, launch_two_body(b, c->pos.as<double>(), c->species.as<int>(), n, cell, c->maxnb,
                               c->nb_count.as<int>(), c->nb_list.as<int>(), c->eatom.as<double>(),
                               c->forces.as<double>(), c->virial.as<double>(), do_forces,
                               prec == TGX_FP64, c->stream));
  rcode = launch_extra_terms(c, n, cell, do_forces);
  if (rcode) return rcode;
  ev_rec(c, 5);
  if (cell.many) {
    if (x.b_out.ensure(sizeof(double) * 10 * (size_t)n_struct))
      return set_err(c, TGX_ERR_HIP, "allocation failed");
    TGX_HIP(c, launch_reduce_many(n_struct, x.b_aoff.as<int>(), c->blocks[0]->site_of_atom.as<int>(),
                                  nullptr, do_forces ? c->virial.as<double>() : nullptr,
                                  c->eatom.as<double>(), x.b_out.as<double>(), c->stream));
  } else {
    TGX_HIP(c, launch_reduce(0, nullptr, n, do_forces ? c->virial.as<double>() : nullptr,
                             c->eatom.as<double>(), x.out.as<double>(), c->stream, !defer_final));
  }
  ev_rec(c, 6);
  x.ev_on = true;
  return TGX_OK;
}

static int upload_structure(tgx_ctx* c, int n, const double* pos, const int* species) {
  ENSURE(c, pos, 24 * (size_t)n);
  ENSURE(c, species, 4 * (size_t)n);
  const int ns = c->has_soap ? c->blk->sd.n_species : TGX_MAX_SPECIES;
  for (int i = 0; i < n; i++) {
    if (species[i] < 1 || species[i] > ns)
      return set_err(c, TGX_ERR_ARG, "species[%d] = %d out of range 1..%d", i, species[i], ns);
    if (!std::isfinite(pos[3 * i]) || !std::isfinite(pos[3 * i + 1]) || !std::isfinite(pos[3 * i + 2]))
      return set_err(c, TGX_ERR_ARG, "position of atom %d is not finite", i);
  }
  TGX_HIP(c, hipMemcpyAsync(c->pos.p, pos, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(c->species.p, species, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  return TGX_OK;
}

// tgx_neighbors_build for a cell of any shape: the reference's brute-force
// search (neighbors_interface.f90:390-421) over the supercell with
// get_distance's general minimum image, padded rows on the device, then the
// CSR export.  O(n^2 images), as in the reference.
static int neighbors_general(tgx_ctx* c, int n, const double* pos, const double* lattice,
                             double rcut, int64_t cap, int* n_neigh, int* list, double* rjs,
                             double* xyz, double* thetas, double* phis, int64_t* n_pairs,
                             int* indices_out) {
  GenCell g;
  std::string err;
  int rcode = make_general_cell(lattice, rcut, &g, &err);
  if (rcode) return set_err(c, rcode, "%s", err.c_str());
  if ((long long)n * g.nrep >= (1LL << TGX_IMG_SHIFT))
    return set_err(c, TGX_ERR_ARG, "too many atoms");
  if (indices_out) {
    indices_out[0] = g.ia;
    indices_out[1] = g.ib;
    indices_out[2] = g.ic;
  }
  for (int i = 0; i < 3 * n; i++)
    if (!std::isfinite(pos[i])) return set_err(c, TGX_ERR_ARG, "non-finite position");
  ENSURE(c, pos, 24 * (size_t)n);
  ENSURE(c, nb_count, sizeof(int) * n);
  ENSURE(c, nb_flag, sizeof(int) * 4);
  TGX_HIP(c, hipMemcpyAsync(c->pos.p, pos, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  const double vol = std::fabs(g.u[0] * (g.u[4] * g.u[8] - g.u[5] * g.u[7]) -
                               g.u[1] * (g.u[3] * g.u[8] - g.u[5] * g.u[6]) +
                               g.u[2] * (g.u[3] * g.u[7] - g.u[4] * g.u[6]));
  int maxnb = guess_maxnb(n, vol, rcut);
  bool sized = false;
  for (int attempt = 0; attempt < 3 && !sized; attempt++) {
    ENSURE(c, nb_list, sizeof(int) * (size_t)n * maxnb);
    TGX_HIP(c, hipMemsetAsync(c->nb_flag.p, 0, sizeof(int) * 4, c->stream));
    TGX_HIP(c, launch_neighbors_general(c->pos.as<double>(), n, g, rcut, maxnb,
                                        c->nb_count.as<int>(), c->nb_list.as<int>(),
                                        c->nb_flag.as<int>(), c->stream));
    int flag[4];
    TGX_HIP(c, hipMemcpyAsync(flag, c->nb_flag.p, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    if (flag[0] > maxnb)
      maxnb = (flag[0] + 7) / 8 * 8;
    else
      sized = true;
  }
  if (!sized) return set_err(c, TGX_ERR_CAPACITY, "neighbour capacity could not be sized");
  std::vector<int> cnt(n);
  TGX_HIP(c, hipMemcpy(cnt.data(), c->nb_count.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
  std::vector<long long> off(n + 1, 0);
  for (int i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i] + 1;
  *n_pairs = off[n];
  if (off[n] > cap)
    return set_err(c, TGX_ERR_CAPACITY, "need %lld pairs, cap %lld", off[n], (long long)cap);
  if (!n_neigh || !list) return set_err(c, TGX_ERR_ARG, "null output arrays");
  const size_t np = (size_t)off[n];
  ENSURE(c, csr_off, 8 * (size_t)(n + 1));
  ENSURE(c, csr_out, (4 + 8 + 24 + 8 + 8) * np + 8);
  char* base = c->csr_out.as<char>();
  int* d_list = reinterpret_cast<int*>(base);
  double* d_rjs = reinterpret_cast<double*>(base + ((4 * np + 7) / 8 * 8));
  double* d_xyz = d_rjs + np;
  double* d_th = d_xyz + 3 * np;
  double* d_ph = d_th + np;
  TGX_HIP(c, hipMemcpyAsync(c->csr_off.p, off.data(), 8 * (size_t)(n + 1), hipMemcpyHostToDevice,
                            c->stream));
  TGX_HIP(c, launch_csr_out_general(c->pos.as<double>(), n, g, maxnb, c->nb_count.as<int>(),
                                    c->nb_list.as<int>(), c->csr_off.as<long long>(), d_list,
                                    d_rjs, d_xyz, d_th, d_ph, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) n_neigh[i] = cnt[i] + 1;
  TGX_HIP(c, hipMemcpy(list, d_list, 4 * np, hipMemcpyDeviceToHost));
  if (rjs) TGX_HIP(c, hipMemcpy(rjs, d_rjs, 8 * np, hipMemcpyDeviceToHost));
  if (xyz) TGX_HIP(c, hipMemcpy(xyz, d_xyz, 24 * np, hipMemcpyDeviceToHost));
  if (thetas) TGX_HIP(c, hipMemcpy(thetas, d_th, 8 * np, hipMemcpyDeviceToHost));
  if (phis) TGX_HIP(c, hipMemcpy(phis, d_ph, 8 * np, hipMemcpyDeviceToHost));
  return TGX_OK;
}

// Scratch array of tgx_energy_forces_noalloc: every device buffer that call
// needs beyond the model (structure, list, cells, descriptors, kernel-step
// and adjoint work, outputs) is carved from it, so the call allocates
// nothing.  Each GPU holds its own copy for the life of the process; host
// code uses only the per-device address taken with hipGetSymbolAddress at
// tgx_create, never the array's name.
#define TGX_SCRATCH_BYTES ((size_t)64 << 20)
#define TGX_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char scratch[TGX_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[TGX_SCRATCH_DEVICES];

// Address of this device's scratch array (the current device must be
// `device`); taken once per device.
static hipError_t scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= TGX_SCRATCH_DEVICES) return hipErrorInvalidDevice;
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

static void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < TGX_SCRATCH_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

// Entry points that write the context's structure, list or model buffers
// refuse to run while an MD run holds them: the captured step graph reads
// those buffers at fixed addresses.
static void md_release_graph(MdState& m);
static int md_busy(tgx_ctx* c) {
  if (X(c).md.active)
    return set_err(c, TGX_ERR_STATE,
                   "an MD run is active on this context (call tgx_md_end first)");
  return TGX_OK;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int tgx_create(int device, int precision, tgx_ctx** out) {
  if (!out) return TGX_ERR_ARG;
  *out = nullptr;
  if (precision < TGX_FP64 || precision > TGX_MIXED) return TGX_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return TGX_ERR_HIP;
  tgx_ctx_full* f = new (std::nothrow) tgx_ctx_full();
  if (!f) return TGX_ERR_ARG;
  tgx_ctx* c = &f->c;
  c->device = device;
  c->precision = precision;
  c->blocks.emplace_back(new (std::nothrow) SoapBlock());
  if (!c->blocks.back()) {
    delete f;
    return TGX_ERR_ARG;
  }
  c->blk = c->blocks[0].get();
  if (hipSetDevice(device) != hipSuccess || scratch_init(device) != hipSuccess ||
      hipStreamCreate(&c->stream) != hipSuccess) {
    delete f;
    return TGX_ERR_HIP;
  }
  for (int k = 0; k < 8; k++) (void)hipEventCreate(&f->x.ev[k]);
  for (int k = 0; k < 8; k++) (void)hipEventCreate(&f->x.kev[k]);
  f->x.ev_ready = true;
  (void)hipStreamSynchronize(c->stream);
  *out = c;
  return TGX_OK;
}

int tgx_destroy(tgx_ctx* c) {
  if (!c) return TGX_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  tgx_ctx_full* f = reinterpret_cast<tgx_ctx_full*>(c);
  md_release_graph(f->x.md);
  f->x.md.active = false;
  for (auto& b : c->blocks) b->release();
  DevBuf* bufs[] = {&c->pos, &c->species, &c->nb_count, &c->nb_list, &c->nb_rev, &c->nb_flag,
                    &c->cell_of, &c->cell_start, &c->cell_atoms, &c->ksum, &c->gpair,
                    &c->forces, &c->eatom, &c->virial, &c->scratch, &c->csr_off,
                    &c->csr_out, &c->md_vel, &c->md_fprev, &c->md_mass, &c->md_pos0,
                    &c->md_ep, &c->md_ek, &c->md_reb, &c->md_disp,
                    &f->x.e0_dev, &f->x.iota, &f->x.cell_count, &f->x.cell_fill, &f->x.out,
                    &f->x.stage, &f->x.twob_store, &f->x.b_cells, &f->x.b_sid,
                    &f->x.b_boff, &f->x.b_aoff, &f->x.b_out, &f->x.mc_pref, &f->x.mc_ctl,
                    &f->x.mc_dctl, &f->x.mc_ecur, &f->x.mc_naff, &f->x.mc_trace,
                    &f->x.mc_twob, &f->x.mc_esur, &f->x.mc_esur_cur, &f->x.mc_sq,
                    &f->x.mc_sy, &f->x.sg_P, &f->x.sg_m, &f->x.sg_v, &f->x.sg_grad,
                    &f->x.sg_W1t, &f->x.sg_W2t, &f->x.sg_part, &f->x.sg_z1, &f->x.sg_z2,
                    &f->x.sg_e, &f->x.sg_q, &f->x.sg_y, &f->x.sg_rows, &f->x.sg_norm,
                    &f->x.tune_save, &f->x.md.pref, &f->x.md.ke, &f->x.md.ctl, &f->x.md.ep,
                    &f->x.md.ek, &f->x.md.vout, &f->x.md.kesum, &f->x.g_pos, &f->x.g_sp, &f->x.g3,
                    &f->x.mcb_mi, &f->x.mcb_md, &f->x.mcb_vatom, &f->x.ts_V, &f->x.ts_pref,
                    &f->x.ts_vatom};
  for (DevBuf* b : bufs) b->release();
  for (auto& b : c->twob) {
    if (b.Qs) (void)hipFree(b.Qs);
    if (b.alphas) (void)hipFree(b.alphas);
    if (b.tab) (void)hipFree(b.tab);
  }
  (void)tgx_clear_extra_terms(c);
  for (auto& b : f->x.lp) b->release();
  if (f->x.ev_ready) {
    for (int k = 0; k < 8; k++) (void)hipEventDestroy(f->x.ev[k]);
    for (int k = 0; k < 8; k++) (void)hipEventDestroy(f->x.kev[k]);
  }
  (void)hipStreamDestroy(c->stream);
  delete f;
  return TGX_OK;
}

const char* tgx_last_error(const tgx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int tgx_set_precision(tgx_ctx* c, int precision) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (precision < TGX_FP64 || precision > TGX_MIXED)
    return set_err(c, TGX_ERR_ARG, "precision %d not in 0..2", precision);
  c->precision = precision;
  return TGX_OK;
}

int tgx_set_autotune(tgx_ctx* c, int on, const char* cache_path) {
  if (!c) return TGX_ERR_ARG;
  if (on != 0 && on != 1) return set_err(c, TGX_ERR_ARG, "autotune %d not 0 or 1", on);
  Extra& x = X(c);
  x.tune_on = on;
  if (cache_path) {
    x.tune_path = cache_path;
    tune_load(x);
  }
  return TGX_OK;
}

}  // extern "C"

// Loads one soap_turbo block into c->blk (hypers, sparse set, device tables).
static int load_soap_block(tgx_ctx* c, const tgx_soap_hypers* h, int dim, int n_sparse,
                           const double* Qs, const double* alphas, double delta, double zeta) {
  if (!h || !Qs || !alphas) return set_err(c, TGX_ERR_ARG, "null argument");
  (void)hipSetDevice(c->device);
  if (h->n_species < 1 || h->n_species > TGX_MAX_SPECIES)
    return set_err(c, TGX_ERR_ARG, "n_species %d not in 1..%d", h->n_species, TGX_MAX_SPECIES);
  if (h->l_max < 0 || h->l_max > TGX_MAX_LMAX)
    return set_err(c, TGX_ERR_ARG, "l_max %d not in 0..%d", h->l_max, TGX_MAX_LMAX);
  if (h->basis != 0 && h->basis != 1)
    return set_err(c, TGX_ERR_ARG, "basis %d: 0 (poly3) or 1 (poly3gauss)", h->basis);
  if (h->radial_enhancement < 0 || h->radial_enhancement > 2)
    return set_err(c, TGX_ERR_ARG, "radial_enhancement must be 0, 1 or 2");
  SoapDev sd{};
  sd.n_species = h->n_species;
  sd.l_max = h->l_max;
  sd.L2 = (h->l_max + 1) * (h->l_max + 1);
  sd.radial_enhancement = h->radial_enhancement;
  sd.central_species = h->central_species;
  sd.basis = h->basis;
  int nr = 0;
  double rcm = 0;
  for (int s = 0; s < h->n_species; s++) {
    if (h->alpha_max[s] < 1 || h->alpha_max[s] > 16)
      return set_err(c, TGX_ERR_ARG, "alpha_max[%d] = %d not in 1..16", s, h->alpha_max[s]);
    if (!(h->rcut_hard[s] > 0) || h->rcut_soft[s] > h->rcut_hard[s] || h->rcut_soft[s] < 0)
      return set_err(c, TGX_ERR_ARG, "need 0 <= rcut_soft <= rcut_hard (species %d)", s + 1);
    if (!(h->atom_sigma_r[s] > 0) || !(h->atom_sigma_t[s] > 0))
      return set_err(c, TGX_ERR_ARG, "atom_sigma must be positive (species %d)", s + 1);
    // nf (read_gap.f90:164-165,192-194, default 4) is passed to soap_turbo's
    // get_soap; its role there is not in the reference tree, so only the
    // default is accepted rather than silently treating another value as 4
    if (h->nf[s] != 4.0)
      return set_err(c, TGX_ERR_UNSUPPORTED, "nf = %g (species %d): only nf = 4 is implemented",
                     h->nf[s], s + 1);
    sd.nmax[s] = h->alpha_max[s];
    sd.off[s] = nr;
    nr += h->alpha_max[s];
    sd.rh[s] = h->rcut_hard[s];
    sd.rs[s] = h->rcut_soft[s];
    sd.sig_r[s] = h->atom_sigma_r[s];
    sd.sig_rs[s] = h->atom_sigma_r_scaling[s];
    sd.sig_t[s] = h->atom_sigma_t[s];
    sd.sig_ts[s] = h->atom_sigma_t_scaling[s];
    sd.amp_scal[s] = h->amplitude_scaling[s];
    sd.glob[s] = h->global_scaling[s];
    // the poly3gauss Gaussian's normalisation: (int_0^rh exp(-r^2/sig^2) dr)^(-1/2)
    sd.gnorm[s] = 1.0 / std::sqrt(std::sqrt(M_PI) / 2.0 * h->atom_sigma_r[s] *
                                  std::erf(h->rcut_hard[s] / h->atom_sigma_r[s]));
    rcm = std::max(rcm, h->rcut_hard[s]);
  }
  if (nr > TGX_MAX_NR) return set_err(c, TGX_ERR_ARG, "sum of alpha_max %d > %d", nr, TGX_MAX_NR);
  sd.nr = nr;
  sd.npair = nr * (nr + 1) / 2;
  {
    int nmx = 0, need = 0;
    for (int s = 0; s < h->n_species; s++) nmx = std::max(nmx, h->alpha_max[s]);
    sd.nmx = nmx <= 8 ? 8 : 16;
    for (int s = 0; s < h->n_species; s++) need = std::max(need, sd.off[s] + sd.nmx);
    // the adjoint's species sub-rows at 4-aligned offsets must fit as well
    int oa = 0;
    for (int s = 0; s < h->n_species; s++) {
      sd.offa[s] = oa;
      need = std::max(need, oa + sd.nmx);
      oa += (h->alpha_max[s] + 3) / 4 * 4;
    }
    sd.nrp = (std::max(need, nr) + 7) / 8 * 8;
    if (sd.nrp > TGX_MAX_NR)
      return set_err(c, TGX_ERR_UNSUPPORTED, "radial channels (padded %d) exceed %d", sd.nrp,
                     TGX_MAX_NR);
    for (int k = 0; k < TGX_MAX_NR; k++) sd.chana[k] = -1;
    for (int s = 0; s < h->n_species; s++)
      for (int k = 0; k < h->alpha_max[s]; k++)
        sd.chana[sd.off[s] + k] = (signed char)(sd.offa[s] + k);
    sd.lds_gap = sd.nrp != nr;
  }
  sd.dim = sd.npair * (h->l_max + 1);
  sd.rcut_max = rcm;
  if (h->central_species < 0 || h->central_species > h->n_species)
    return set_err(c, TGX_ERR_ARG, "central_species out of range");
  const int ncmp = h->compress_P_nonzero;
  if (ncmp < 0) return set_err(c, TGX_ERR_ARG, "compress_P_nonzero < 0");
  if (ncmp > 0) {
    if (!h->compress_P_i || !h->compress_P_j || !h->compress_P_el)
      return set_err(c, TGX_ERR_ARG, "compress_soap: P arrays missing");
    if (dim < 1 || dim > 6144)
      return set_err(c, TGX_ERR_ARG, "compressed dim %d not in 1..6144", dim);
    for (int k = 0; k < ncmp; k++)
      if (h->compress_P_i[k] < 1 || h->compress_P_i[k] > dim || h->compress_P_j[k] < 1 ||
          h->compress_P_j[k] > sd.dim || !std::isfinite(h->compress_P_el[k]))
        return set_err(c, TGX_ERR_ARG,
                       "compress_soap entry %d (%d, %d) outside %d x %d or not finite", k + 1,
                       h->compress_P_i[k], h->compress_P_j[k], dim, sd.dim);
  } else if (dim != sd.dim) {
    return set_err(c, TGX_ERR_ARG, "dim %d != n_max(n_max+1)/2*(l_max+1) = %d", dim, sd.dim);
  }
  if (n_sparse < 1) return set_err(c, TGX_ERR_ARG, "n_sparse must be >= 1");
  // W per species and the central-atom coefficients
  std::vector<double> W((size_t)TGX_MAX_SPECIES * TGX_MAX_NR * TGX_MAX_NR, 0.0);
  for (int s = 0; s < h->n_species; s++) {
    double* Ws = W.data() + (size_t)s * TGX_MAX_NR * TGX_MAX_NR;
    if (h->basis == 1)
      poly3gauss_W(h->alpha_max[s], h->rcut_hard[s], h->atom_sigma_r[s], nullptr, Ws);
    else
      poly3_W(h->alpha_max[s], Ws);
  }
  for (int s = 0; s < h->n_species; s++) {
    double R[TGX_MAX_NR];
    SoapDev tmp = sd;
    tmp.sig_rs[s] = 0.0;
    radial_closed_form(tmp, W.data(), s, 0.0, R, nullptr);
    for (int k = 0; k < h->alpha_max[s]; k++)
      sd.rself[s * TGX_MAX_NR + k] = h->central_weight[s] * R[k] * (1.0 / std::sqrt(4.0 * M_PI));
  }
  c->blk->W_host = W;
  c->blk->hyp = *h;
  // the block keeps its own copy of P (its hypers point there, so that
  // blocks loaded from them later, e.g. local properties, find it)
  c->blk->cmp = ncmp;
  c->blk->cP_i.assign(h->compress_P_i, h->compress_P_i + ncmp);
  c->blk->cP_j.assign(h->compress_P_j, h->compress_P_j + ncmp);
  c->blk->cP_el.assign(h->compress_P_el, h->compress_P_el + ncmp);
  c->blk->hyp.compress_P_i = ncmp ? c->blk->cP_i.data() : nullptr;
  c->blk->hyp.compress_P_j = ncmp ? c->blk->cP_j.data() : nullptr;
  c->blk->hyp.compress_P_el = ncmp ? c->blk->cP_el.data() : nullptr;
  c->blk->dimf_pad = (sd.dim + 63) / 64 * 64;
  if (ncmp) {
    // P by rows (forward) and by columns (adjoint), entries in input order
    std::vector<int> roff(dim + 1, 0), coff(sd.dim + 1, 0), rcol(ncmp), crow(ncmp);
    std::vector<double> rval(ncmp), cval(ncmp);
    for (int k = 0; k < ncmp; k++) {
      roff[h->compress_P_i[k]]++;
      coff[h->compress_P_j[k]]++;
    }
    for (int i = 0; i < dim; i++) roff[i + 1] += roff[i];
    for (int j = 0; j < sd.dim; j++) coff[j + 1] += coff[j];
    std::vector<int> rf(roff.begin(), roff.end() - 1), cf(coff.begin(), coff.end() - 1);
    for (int k = 0; k < ncmp; k++) {
      const int i = h->compress_P_i[k] - 1, j = h->compress_P_j[k] - 1;
      rcol[rf[i]] = j;
      rval[rf[i]++] = h->compress_P_el[k];
      crow[cf[j]] = i;
      cval[cf[j]++] = h->compress_P_el[k];
    }
    ENSURE(c, blk->c_roff, 4 * roff.size());
    ENSURE(c, blk->c_coff, 4 * coff.size());
    ENSURE(c, blk->c_rcol, 4 * (size_t)ncmp);
    ENSURE(c, blk->c_crow, 4 * (size_t)ncmp);
    ENSURE(c, blk->c_rval, 8 * (size_t)ncmp);
    ENSURE(c, blk->c_cval, 8 * (size_t)ncmp);
    TGX_HIP(c, hipMemcpy(c->blk->c_roff.p, roff.data(), 4 * roff.size(), hipMemcpyHostToDevice));
    TGX_HIP(c, hipMemcpy(c->blk->c_coff.p, coff.data(), 4 * coff.size(), hipMemcpyHostToDevice));
    TGX_HIP(c, hipMemcpy(c->blk->c_rcol.p, rcol.data(), 4 * (size_t)ncmp, hipMemcpyHostToDevice));
    TGX_HIP(c, hipMemcpy(c->blk->c_crow.p, crow.data(), 4 * (size_t)ncmp, hipMemcpyHostToDevice));
    TGX_HIP(c, hipMemcpy(c->blk->c_rval.p, rval.data(), 8 * (size_t)ncmp, hipMemcpyHostToDevice));
    TGX_HIP(c, hipMemcpy(c->blk->c_cval.p, cval.data(), 8 * (size_t)ncmp, hipMemcpyHostToDevice));
  }
  c->blk->sd = sd;
  c->blk->dim = dim;
  c->blk->n_sparse = n_sparse;
  c->blk->dim_pad = (dim + 63) / 64 * 64;  // whole 64-deep k-slabs of the BF16 GEMM
  c->blk->sp_pad = (n_sparse + 63) / 64 * 64;
  c->blk->delta = delta;
  c->blk->zeta = zeta;
  c->blk->Qs_host.assign(Qs, Qs + (size_t)dim * n_sparse);
  c->blk->alphas_host.assign(alphas, alphas + n_sparse);
  Extra& x = X(c);
  // padded layouts of Qs for the three precisions
  const int Dp = c->blk->dim_pad, Sp = c->blk->sp_pad;
  std::vector<double> q64((size_t)Sp * Dp, 0.0), qT64((size_t)Dp * Sp, 0.0), a64(Sp, 0.0);
  std::vector<float> q32((size_t)Sp * Dp, 0.f), qT32((size_t)Dp * Sp, 0.f), a32(Sp, 0.f);
  std::vector<unsigned short> q16((size_t)Sp * Dp, 0), qT16((size_t)Dp * Sp, 0);
  auto bf = [](float f) {
    unsigned int u;
    std::memcpy(&u, &f, 4);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  };
  for (int s = 0; s < n_sparse; s++) {
    a64[s] = alphas[s];
    a32[s] = (float)alphas[s];
    for (int d = 0; d < dim; d++) {
      const double v = Qs[(size_t)s * dim + d];
      q64[(size_t)s * Dp + d] = v;
      qT64[(size_t)d * Sp + s] = v;
      q32[(size_t)s * Dp + d] = (float)v;
      qT32[(size_t)d * Sp + s] = (float)v;
      q16[(size_t)s * Dp + d] = bf((float)v);
      qT16[(size_t)d * Sp + s] = bf((float)v);
    }
  }
  ENSURE(c, blk->Qs64, 8 * q64.size());
  ENSURE(c, blk->QsT64, 8 * qT64.size());
  ENSURE(c, blk->Qs32, 4 * q32.size());
  ENSURE(c, blk->QsT32, 4 * qT32.size());
  ENSURE(c, blk->Qs16, 2 * q16.size());
  ENSURE(c, blk->QsT16, 2 * qT16.size());
  ENSURE(c, blk->alph64, 8 * (size_t)Sp);
  ENSURE(c, blk->alph32, 4 * (size_t)Sp);
  ENSURE(c, blk->W_dev, 8 * W.size());
  ENSURE(c, blk->rtab, sizeof(float) * 2 * (size_t)TGX_MAX_SPECIES * (TGX_RTAB_N + 1) * TGX_MAX_NR);
  if (c->blk->sd_dev.ensure(sizeof(SoapDev))) return set_err(c, TGX_ERR_HIP, "allocation failed");
  // component d of the power spectrum -> (n, n', l), packed n | n' << 8 | l << 16
  std::vector<int> pn(sd.dim);
  {
    int k = 0;
    for (int a = 0; a < nr; a++)
      for (int b = a; b < nr; b++)
        for (int l = 0; l <= sd.l_max; l++) pn[k++] = a | (b << 8) | (l << 16);
  }
  if (c->blk->pairn.ensure(sizeof(int) * pn.size())) return set_err(c, TGX_ERR_HIP, "allocation failed");
  hipStream_t st = c->stream;
  TGX_HIP(c, hipMemcpyAsync(c->blk->Qs64.p, q64.data(), 8 * q64.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->QsT64.p, qT64.data(), 8 * qT64.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->Qs32.p, q32.data(), 4 * q32.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->QsT32.p, qT32.data(), 4 * qT32.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->Qs16.p, q16.data(), 2 * q16.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->QsT16.p, qT16.data(), 2 * qT16.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->alph64.p, a64.data(), 8 * (size_t)Sp, hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->alph32.p, a32.data(), 4 * (size_t)Sp, hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->W_dev.p, W.data(), 8 * W.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->sd_dev.p, &c->blk->sd, sizeof(SoapDev), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemcpyAsync(c->blk->pairn.p, pn.data(), sizeof(int) * pn.size(), hipMemcpyHostToDevice, st));
  TGX_HIP(c, hipMemsetAsync(c->blk->rtab.p, 0, c->blk->rtab.bytes, st));
  TGX_HIP(c, launch_build_rtab(c->blk->sd_dev.as<SoapDev>(), sd.n_species, c->blk->W_dev.as<double>(),
                               c->blk->rtab.as<float>(), st));
  TGX_HIP(c, hipStreamSynchronize(st));
  c->has_soap = true;
  c->maxnb = 0;
  if (!x.e0_dev.p) {
    if (x.e0_dev.ensure(8 * TGX_MAX_SPECIES)) return set_err(c, TGX_ERR_HIP, "allocation failed");
    TGX_HIP(c, hipMemcpy(x.e0_dev.p, c->e0, 8 * TGX_MAX_SPECIES, hipMemcpyHostToDevice));
  }
  return TGX_OK;
}

// Releases every SOAP block but the first object (kept empty).
static void drop_soap_blocks(tgx_ctx* c) {
  for (auto& b : c->blocks) b->release();
  c->blocks.resize(1);
  *c->blocks[0] = SoapBlock();
  c->blk = c->blocks[0].get();
  c->has_soap = false;
  c->maxnb = 0;
}

extern "C" {

int tgx_set_soap_model(tgx_ctx* c, const tgx_soap_hypers* h, int dim, int n_sparse,
                       const double* Qs, const double* alphas, double delta, double zeta) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  drop_soap_blocks(c);
  const int rcode = load_soap_block(c, h, dim, n_sparse, Qs, alphas, delta, zeta);
  if (rcode) drop_soap_blocks(c);
  return rcode;
}

int tgx_add_soap_model(tgx_ctx* c, const tgx_soap_hypers* h, int dim, int n_sparse,
                       const double* Qs, const double* alphas, double delta, double zeta) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (!c->has_soap) return tgx_set_soap_model(c, h, dim, n_sparse, Qs, alphas, delta, zeta);
  if (!h) return set_err(c, TGX_ERR_ARG, "null argument");
  if (h->n_species != c->blocks[0]->hyp.n_species)
    return set_err(c, TGX_ERR_ARG,
                   "every soap_turbo block uses the model's species list (%d species, got %d)",
                   c->blocks[0]->hyp.n_species, h->n_species);
  if (c->blocks.size() >= TGX_MAX_SOAP_BLOCKS)
    return set_err(c, TGX_ERR_ARG, "at most %d soap_turbo blocks", TGX_MAX_SOAP_BLOCKS);
  (void)hipSetDevice(c->device);
  c->blocks.emplace_back(new (std::nothrow) SoapBlock());
  if (!c->blocks.back()) {
    c->blocks.pop_back();
    return set_err(c, TGX_ERR_ARG, "out of host memory");
  }
  c->blk = c->blocks.back().get();
  const int rcode = load_soap_block(c, h, dim, n_sparse, Qs, alphas, delta, zeta);
  if (rcode) {
    c->blocks.back()->release();
    c->blocks.pop_back();
  }
  c->blk = c->blocks[0].get();
  c->has_soap = true;
  c->maxnb = 0;
  return rcode;
}

int tgx_clear_soap_models(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  drop_soap_blocks(c);
  return TGX_OK;
}

int tgx_n_soap_models(const tgx_ctx* c) {
  if (!c) return -TGX_ERR_ARG;
  return c->has_soap ? (int)c->blocks.size() : 0;
}

int tgx_set_e0(tgx_ctx* c, int n_species, const double* e0) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n_species < 0 || n_species > TGX_MAX_SPECIES || (n_species > 0 && !e0))
    return set_err(c, TGX_ERR_ARG, "bad e0 arguments");
  (void)hipSetDevice(c->device);
  for (int s = 0; s < TGX_MAX_SPECIES; s++) c->e0[s] = s < n_species ? e0[s] : 0.0;
  Extra& x = X(c);
  if (x.e0_dev.ensure(8 * TGX_MAX_SPECIES)) return set_err(c, TGX_ERR_HIP, "allocation failed");
  TGX_HIP(c, hipMemcpy(x.e0_dev.p, c->e0, 8 * TGX_MAX_SPECIES, hipMemcpyHostToDevice));
  return TGX_OK;
}

int tgx_add_2b_model(tgx_ctx* c, int sp1, int sp2, double rcut, double buffer, double delta,
                     double sigma, int n_sparse, const double* Qs, const double* alphas,
                     const double* cutoff) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (sp1 < 1 || sp2 < 1 || sp1 > TGX_MAX_SPECIES || sp2 > TGX_MAX_SPECIES || n_sparse < 1 ||
      !Qs || !alphas || !cutoff || !(rcut > 0) || !(sigma > 0) || buffer < 0)
    return set_err(c, TGX_ERR_ARG, "bad distance_2b arguments");
  (void)hipSetDevice(c->device);
  TwoBodyDev b{};
  b.sp1 = sp1;
  b.sp2 = sp2;
  b.n_sparse = n_sparse;
  b.rcut = rcut;
  b.buffer = buffer;
  b.delta = delta;
  b.sigma = sigma;
  std::vector<double> ac(n_sparse);
  for (int k = 0; k < n_sparse; k++) ac[k] = alphas[k] * cutoff[k];
  // FP32 table of g(r) = delta^2 sum_s alpha_s c_s exp(-(r-Q_s)^2/(2 sigma^2))
  // and h g'(r) on [0, rcut], evaluated in FP64 (the cutoff is applied
  // analytically by the kernel).
  b.nt = 4096;
  std::vector<float> tab(2 * (size_t)(b.nt + 1));
  const double hh = rcut / b.nt, is2 = 1.0 / (sigma * sigma);
  for (int k = 0; k <= b.nt; k++) {
    const double r = k * hh;
    double g = 0.0, dg = 0.0;
    for (int s = 0; s < n_sparse; s++) {
      const double dr = r - Qs[s];
      const double ex = std::exp(-0.5 * dr * dr * is2);
      g += ac[s] * ex;
      dg -= ac[s] * ex * dr * is2;
    }
    tab[2 * k] = (float)(delta * delta * g);
    tab[2 * k + 1] = (float)(hh * delta * delta * dg);
  }
  TGX_HIP(c, hipMalloc(&b.Qs, 8 * (size_t)n_sparse));
  TGX_HIP(c, hipMalloc(&b.alphas, 8 * (size_t)n_sparse));
  TGX_HIP(c, hipMalloc(&b.tab, sizeof(float) * tab.size()));
  TGX_HIP(c, hipMemcpy(b.Qs, Qs, 8 * (size_t)n_sparse, hipMemcpyHostToDevice));
  TGX_HIP(c, hipMemcpy(b.alphas, ac.data(), 8 * (size_t)n_sparse, hipMemcpyHostToDevice));
  TGX_HIP(c, hipMemcpy(b.tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
  c->twob.push_back(b);
  c->maxnb = 0;
  Extra& x = X(c);
  // device copy of all blocks, read by the SOAP adjoint when it adds the
  // pair terms itself
  if (x.twob_store.ensure(sizeof(TwoBodyDev) * c->twob.size()))
    return set_err(c, TGX_ERR_HIP, "allocation failed");
  TGX_HIP(c, hipMemcpy(x.twob_store.p, c->twob.data(), sizeof(TwoBodyDev) * c->twob.size(),
                       hipMemcpyHostToDevice));
  if (!x.e0_dev.p) {
    if (x.e0_dev.ensure(8 * TGX_MAX_SPECIES)) return set_err(c, TGX_ERR_HIP, "allocation failed");
    TGX_HIP(c, hipMemcpy(x.e0_dev.p, c->e0, 8 * TGX_MAX_SPECIES, hipMemcpyHostToDevice));
  }
  return TGX_OK;
}

int tgx_clear_2b_models(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  (void)hipSetDevice(c->device);
  for (auto& b : c->twob) {
    (void)hipFree(b.Qs);
    (void)hipFree(b.alphas);
    (void)hipFree(b.tab);
  }
  c->twob.clear();
  c->maxnb = 0;
  return TGX_OK;
}

int tgx_add_angle_3b(tgx_ctx* c, int sp_center, int sp1, int sp2, double rcut, double delta,
                     const double* sigma, int kernel, int n_sparse, const double* Qs,
                     const double* alphas, const double* cutoff) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  auto sp_ok = [](int s) { return s >= 1 && s <= TGX_MAX_SPECIES; };
  if (!sp_ok(sp_center) || !sp_ok(sp1) || !sp_ok(sp2) || !(rcut > 0) || !sigma ||
      !(sigma[0] > 0 && sigma[1] > 0 && sigma[2] > 0) || (kernel != 0 && kernel != 1) ||
      n_sparse < 1 || !Qs || !alphas || !cutoff)
    return set_err(c, TGX_ERR_ARG, "bad angle_3b arguments");
  (void)hipSetDevice(c->device);
  Angle3bDev b{};
  b.sp0 = sp_center;
  b.sp1 = sp1;
  b.sp2 = sp2;
  b.n_sparse = n_sparse;
  b.kernel = kernel;
  b.rcut = rcut;
  b.buffer = 0.5;  // calculate_gap_3b passes 0.5 (gap.f90:738)
  for (int k = 0; k < 3; k++) b.is2[k] = 1.0 / (sigma[k] * sigma[k]);
  std::vector<double> qp(4 * (size_t)n_sparse);
  for (int s = 0; s < n_sparse; s++) {
    for (int k = 0; k < 3; k++) qp[4 * s + k] = Qs[(size_t)k * n_sparse + s];
    qp[4 * s + 3] = 2.0 * delta * delta * alphas[s] * cutoff[s];
  }
  TGX_HIP(c, hipMalloc(&b.Qp, 8 * qp.size()));
  TGX_HIP(c, hipMemcpy(b.Qp, qp.data(), 8 * qp.size(), hipMemcpyHostToDevice));
  c->a3b.push_back(b);
  c->maxnb = 0;
  return TGX_OK;
}

// Second derivatives M of the cubic spline through (x, V) by one tridiagonal
// solve: interior rows h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1}
// = 6 (s_i - s_{i-1}) (s_i the slope of interval i); an end is clamped to
// its slope, or natural (M = 0) when the slope exceeds 0.99e30, and then
// takes the spline's own end derivative.
static void core_pot_spline(int n, const double* x, const double* V, double* yp1, double* ypn,
                            double* M) {
  std::vector<double> h(n - 1), sl(n - 1), lo(n), di(n), up(n), rhs(n);
  for (int i = 0; i + 1 < n; i++) {
    h[i] = x[i + 1] - x[i];
    sl[i] = (V[i + 1] - V[i]) / h[i];
  }
  const bool nat0 = *yp1 > 0.99e30, natn = *ypn > 0.99e30;
  lo[0] = 0.0;
  di[0] = nat0 ? 1.0 : 2.0 * h[0];
  up[0] = nat0 ? 0.0 : h[0];
  rhs[0] = nat0 ? 0.0 : 6.0 * (sl[0] - *yp1);
  for (int i = 1; i + 1 < n; i++) {
    lo[i] = h[i - 1];
    di[i] = 2.0 * (h[i - 1] + h[i]);
    up[i] = h[i];
    rhs[i] = 6.0 * (sl[i] - sl[i - 1]);
  }
  lo[n - 1] = natn ? 0.0 : h[n - 2];
  di[n - 1] = natn ? 1.0 : 2.0 * h[n - 2];
  up[n - 1] = 0.0;
  rhs[n - 1] = natn ? 0.0 : 6.0 * (*ypn - sl[n - 2]);
  // Thomas algorithm (the matrix is diagonally dominant)
  for (int i = 1; i < n; i++) {
    const double w = lo[i] / di[i - 1];
    di[i] -= w * up[i - 1];
    rhs[i] -= w * rhs[i - 1];
  }
  M[n - 1] = rhs[n - 1] / di[n - 1];
  for (int i = n - 2; i >= 0; i--) M[i] = (rhs[i] - up[i] * M[i + 1]) / di[i];
  if (nat0) *yp1 = sl[0] - h[0] * (2.0 * M[0] + M[1]) / 6.0;
  if (natn) *ypn = sl[n - 2] + h[n - 2] * (M[n - 2] + 2.0 * M[n - 1]) / 6.0;
}

int tgx_add_core_pot(tgx_ctx* c, int sp1, int sp2, int n, const double* x, const double* V,
                     double yp1, double ypn) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (sp1 < 1 || sp2 < 1 || sp1 > TGX_MAX_SPECIES || sp2 > TGX_MAX_SPECIES || n < 2 || !x || !V)
    return set_err(c, TGX_ERR_ARG, "bad core_pot arguments");
  for (int i = 0; i + 1 < n; i++)
    if (!(x[i + 1] > x[i])) return set_err(c, TGX_ERR_ARG, "core_pot x must ascend strictly");
  (void)hipSetDevice(c->device);
  CorePotDev b{};
  b.sp1 = sp1;
  b.sp2 = sp2;
  b.n = n;
  b.rcut = x[n - 1];
  b.yp1 = yp1;
  b.ypn = ypn;
  std::vector<double> M(n);
  core_pot_spline(n, x, V, &b.yp1, &b.ypn, M.data());
  TGX_HIP(c, hipMalloc(&b.x, 8 * (size_t)n));
  TGX_HIP(c, hipMalloc(&b.V, 8 * (size_t)n));
  TGX_HIP(c, hipMalloc(&b.d2, 8 * (size_t)n));
  TGX_HIP(c, hipMemcpy(b.x, x, 8 * (size_t)n, hipMemcpyHostToDevice));
  TGX_HIP(c, hipMemcpy(b.V, V, 8 * (size_t)n, hipMemcpyHostToDevice));
  TGX_HIP(c, hipMemcpy(b.d2, M.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
  c->corep.push_back(b);
  c->maxnb = 0;
  return TGX_OK;
}

int tgx_clear_extra_terms(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  (void)hipSetDevice(c->device);
  for (auto& b : c->a3b) (void)hipFree(b.Qp);
  for (auto& b : c->corep) {
    (void)hipFree(b.x);
    (void)hipFree(b.V);
    (void)hipFree(b.d2);
  }
  c->a3b.clear();
  c->corep.clear();
  c->maxnb = 0;
  return TGX_OK;
}

int tgx_neighbors_build(tgx_ctx* c, int n, const double* pos, const double* lattice, double rcut,
                        int64_t cap, int* n_neigh, int* list, double* rjs, double* xyz,
                        double* thetas, double* phis, int64_t* n_pairs, int* indices_out) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n < 1 || !pos || !lattice || !n_pairs) return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  if (!is_orthorhombic(lattice))
    return neighbors_general(c, n, pos, lattice, rcut, cap, n_neigh, list, rjs, xyz, thetas,
                             phis, n_pairs, indices_out);
  Cell cell;
  std::string err;
  int rcode = make_cell(lattice, rcut, &cell, &err);
  if (rcode) return set_err(c, rcode, "%s", err.c_str());
  if (indices_out) {
    indices_out[0] = cell.ia;
    indices_out[1] = cell.ib;
    indices_out[2] = cell.ic;
  }
  ENSURE(c, pos, 24 * (size_t)n);
  for (int i = 0; i < 3 * n; i++)
    if (!std::isfinite(pos[i])) return set_err(c, TGX_ERR_ARG, "non-finite position");
  TGX_HIP(c, hipMemcpyAsync(c->pos.p, pos, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  const int saved = c->maxnb;
  c->maxnb = 0;
  rcode = build_list(c, n, cell, rcut, false);
  const int maxnb = c->maxnb;
  c->maxnb = saved;
  if (rcode) return rcode;
  std::vector<int> cnt(n);
  TGX_HIP(c, hipMemcpy(cnt.data(), c->nb_count.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
  std::vector<long long> off(n + 1, 0);
  for (int i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i] + 1;
  *n_pairs = off[n];
  if (off[n] > cap) return set_err(c, TGX_ERR_CAPACITY, "need %lld pairs, cap %lld", off[n], (long long)cap);
  if (!n_neigh || !list) return set_err(c, TGX_ERR_ARG, "null output arrays");
  const size_t np = (size_t)off[n];
  ENSURE(c, csr_off, 8 * (size_t)(n + 1));
  ENSURE(c, csr_out, (4 + 8 + 24 + 8 + 8) * np);
  char* base = c->csr_out.as<char>();
  int* d_list = reinterpret_cast<int*>(base);
  double* d_rjs = reinterpret_cast<double*>(base + ((4 * np + 7) / 8 * 8));
  double* d_xyz = d_rjs + np;
  double* d_th = d_xyz + 3 * np;
  double* d_ph = d_th + np;
  TGX_HIP(c, hipMemcpyAsync(c->csr_off.p, off.data(), 8 * (size_t)(n + 1), hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, launch_csr_out(c->pos.as<double>(), n, cell, maxnb, c->nb_count.as<int>(),
                            c->nb_list.as<int>(), c->csr_off.as<long long>(), d_list, d_rjs, d_xyz,
                            d_th, d_ph, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) n_neigh[i] = cnt[i] + 1;
  TGX_HIP(c, hipMemcpy(list, d_list, 4 * np, hipMemcpyDeviceToHost));
  if (rjs) TGX_HIP(c, hipMemcpy(rjs, d_rjs, 8 * np, hipMemcpyDeviceToHost));
  if (xyz) TGX_HIP(c, hipMemcpy(xyz, d_xyz, 24 * np, hipMemcpyDeviceToHost));
  if (thetas) TGX_HIP(c, hipMemcpy(thetas, d_th, 8 * np, hipMemcpyDeviceToHost));
  if (phis) TGX_HIP(c, hipMemcpy(phis, d_ph, 8 * np, hipMemcpyDeviceToHost));
  return TGX_OK;
}

int tgx_soap_descriptors(tgx_ctx* c, int n, const double* pos, const int* species,
                         const double* lattice, double* soap, int* n_sites_out, int* site_index) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n < 1 || !pos || !species || !lattice || !soap || !n_sites_out)
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  if (!c->has_soap) return set_err(c, TGX_ERR_STATE, "no SOAP model loaded");
  (void)hipSetDevice(c->device);
  int rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  Cell cell;
  std::string err;
  const bool general = !is_orthorhombic(lattice);
  if (!general) {
    rcode = make_cell(lattice, c->blk->sd.rcut_max, &cell, &err);
    if (rcode) return set_err(c, rcode, "%s", err.c_str());
  }
  const int saved = c->maxnb;
  c->maxnb = 0;
  int n_ext = n;
  rcode = general ? prepare_image_cell(c, n, lattice, c->blk->sd.rcut_max, false, &cell, &n_ext)
                  : build_list(c, n, cell, c->blk->sd.rcut_max, false);
  if (rcode) {
    c->maxnb = saved;
    return rcode;
  }
  int n_sites = 0;
  rcode = setup_sites(c, n, species, &n_sites, n_ext);
  if (!rcode && n_sites > 0) rcode = soap_forward(c, n, Geom{cell}, n_sites, false, true);
  c->maxnb = saved;
  if (rcode) return rcode;
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  const int Dp = c->blk->dim_pad, dim = c->blk->dim;
  if (c->precision == TGX_FP64) {
    std::vector<double> tmp((size_t)n_sites * Dp);
    TGX_HIP(c, hipMemcpy(tmp.data(), c->blk->q.p, 8 * tmp.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < n_sites; s++)
      std::memcpy(soap + (size_t)s * dim, tmp.data() + (size_t)s * Dp, 8 * (size_t)dim);
  } else {
    std::vector<float> tmp((size_t)n_sites * Dp);
    TGX_HIP(c, hipMemcpy(tmp.data(), c->blk->q.p, 4 * tmp.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < n_sites; s++)
      for (int d = 0; d < dim; d++) soap[(size_t)s * dim + d] = tmp[(size_t)s * Dp + d];
  }
  *n_sites_out = n_sites;
  if (site_index) {
    std::vector<int> sa(n_sites);
    if (n_sites)
      TGX_HIP(c, hipMemcpy(sa.data(), c->blk->site_atom.p, 4 * (size_t)n_sites, hipMemcpyDeviceToHost));
    for (int s = 0; s < n_sites; s++) site_index[s] = sa[s] + 1;
  }
  return TGX_OK;
}

int tgx_gap_kernel_step(tgx_ctx* c, int n_sites, const double* soap, double* energies,
                        double* qss) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n_sites < 0 || (n_sites > 0 && (!soap || !energies)))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  if (!c->has_soap) return set_err(c, TGX_ERR_STATE, "no SOAP model loaded");
  if (n_sites == 0) return TGX_OK;
  (void)hipSetDevice(c->device);
  const int prec = c->precision, Dp = c->blk->dim_pad, dim = c->blk->dim;
  const size_t ts = prec == TGX_FP64 ? 8 : (prec == TGX_FP32 ? 4 : 2);
  if (prec == TGX_MIXED) {
    ENSURE(c, blk->qb, 2 * (size_t)n_sites * Dp);
    std::vector<unsigned short> t((size_t)n_sites * Dp, 0);
    for (int s = 0; s < n_sites; s++)
      for (int d = 0; d < dim; d++) {
        float f = (float)soap[(size_t)s * dim + d];
        unsigned int u;
        std::memcpy(&u, &f, 4);
        t[(size_t)s * Dp + d] = (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
      }
    TGX_HIP(c, hipMemcpy(c->blk->qb.p, t.data(), 2 * t.size(), hipMemcpyHostToDevice));
  } else {
    ENSURE(c, blk->q, ts * (size_t)n_sites * Dp);
    if (prec == TGX_FP64) {
      std::vector<double> t((size_t)n_sites * Dp, 0.0);
      for (int s = 0; s < n_sites; s++)
        std::memcpy(t.data() + (size_t)s * Dp, soap + (size_t)s * dim, 8 * (size_t)dim);
      TGX_HIP(c, hipMemcpy(c->blk->q.p, t.data(), 8 * t.size(), hipMemcpyHostToDevice));
    } else {
      std::vector<float> t((size_t)n_sites * Dp, 0.f);
      for (int s = 0; s < n_sites; s++)
        for (int d = 0; d < dim; d++) t[(size_t)s * Dp + d] = (float)soap[(size_t)s * dim + d];
      TGX_HIP(c, hipMemcpy(c->blk->q.p, t.data(), 4 * t.size(), hipMemcpyHostToDevice));
    }
  }
  int rcode = kernel_step(c, n_sites, qss != nullptr);
  if (rcode) return rcode;
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  TGX_HIP(c, hipMemcpy(energies, c->blk->esite.p, 8 * (size_t)n_sites, hipMemcpyDeviceToHost));
  if (qss) {
    if (prec == TGX_FP64) {
      std::vector<double> g((size_t)n_sites * Dp);
      TGX_HIP(c, hipMemcpy(g.data(), c->blk->gmat.p, 8 * g.size(), hipMemcpyDeviceToHost));
      for (int s = 0; s < n_sites; s++)
        for (int d = 0; d < dim; d++) qss[(size_t)d * n_sites + s] = -g[(size_t)s * Dp + d];
    } else {
      std::vector<float> g((size_t)n_sites * Dp);
      TGX_HIP(c, hipMemcpy(g.data(), c->blk->gmat.p, 4 * g.size(), hipMemcpyDeviceToHost));
      for (int s = 0; s < n_sites; s++)
        for (int d = 0; d < dim; d++) qss[(size_t)d * n_sites + s] = -(double)g[(size_t)s * Dp + d];
    }
  }
  return TGX_OK;
}

int tgx_energy_forces(tgx_ctx* c, int n, const double* pos, const int* species,
                      const double* lattice, int do_forces, double* e_atom, double* forces,
                      double* virial) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n < 1 || !pos || !species || !lattice || !e_atom || (do_forces && (!forces || !virial)))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  int rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  rcode = run_ef(c, n, lattice, species, do_forces);
  if (rcode) return rcode;
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  collect_times(c);
  TGX_HIP(c, hipMemcpy(e_atom, c->eatom.p, 8 * (size_t)n, hipMemcpyDeviceToHost));
  if (do_forces) {
    TGX_HIP(c, hipMemcpy(forces, c->forces.p, 24 * (size_t)n, hipMemcpyDeviceToHost));
    TGX_HIP(c, hipMemcpy(virial, X(c).out.p, 72, hipMemcpyDeviceToHost));
  }
  return TGX_OK;
}

int tgx_energy_forces_noalloc(tgx_ctx* c, int n, const double* pos, const int* species,
                              const double* lattice, int do_forces, double* e_atom,
                              double* forces, double* virial) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n < 1 || !pos || !species || !lattice || !e_atom || (do_forces && (!forces || !virial)))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  void* base = scratch_ptr(c->device);
  if (!base) return set_err(c, TGX_ERR_STATE, "no scratch array on device %d", c->device);
  Extra& x = X(c);
  Arena& a = x.arena;
  a.on = true;
  a.overflow = false;
  a.base = static_cast<char*>(base);
  a.size = TGX_SCRATCH_BYTES;
  a.off = 0;
  a.saved.clear();
  // no autotuner timing here: its candidates would need buffers of their own
  const int tune = x.tune_on;
  x.tune_on = 0;
  int rcode = upload_structure(c, n, pos, species);
  if (!rcode) rcode = run_ef(c, n, lattice, species, do_forces);
  if (!rcode) {
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      rcode = set_err(c, TGX_ERR_HIP, "stream synchronisation failed");
  }
  if (!rcode) {
    collect_times(c);
    if (hipMemcpy(e_atom, c->eatom.p, 8 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess ||
        (do_forces &&
         (hipMemcpy(forces, c->forces.p, 24 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(virial, x.out.p, 72, hipMemcpyDeviceToHost) != hipSuccess)))
      rcode = set_err(c, TGX_ERR_HIP, "copy of the results failed");
  } else {
    (void)hipStreamSynchronize(c->stream);  // nothing of this call still runs
  }
  // the context gets its own buffers back
  a.on = false;
  for (auto it = a.saved.rbegin(); it != a.saved.rend(); ++it) *it->first = it->second;
  a.saved.clear();
  x.tune_on = tune;
  return rcode;
}

int tgx_energy_forces_batched(tgx_ctx* c, int n_struct, const int* offsets, const double* pos,
                              const int* species, const double* lattices, int do_forces,
                              double* e_atom, double* forces, double* virials) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (n_struct < 1 || !offsets || !pos || !species || !lattices || !e_atom ||
      (do_forces && (!forces || !virials)))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  if (offsets[0] != 0) return set_err(c, TGX_ERR_ARG, "offsets[0] must be 0");
  for (int s = 0; s < n_struct; s++)
    if (offsets[s + 1] <= offsets[s]) return set_err(c, TGX_ERR_ARG, "structure %d is empty", s);
  const int n = offsets[n_struct];
  if (n >= (1 << TGX_IMG_SHIFT)) return set_err(c, TGX_ERR_ARG, "too many atoms");
  if (!has_model(c)) return set_err(c, TGX_ERR_STATE, "no model loaded");
  (void)hipSetDevice(c->device);
  Extra& x = X(c);
  int rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  // per-structure cells, bin offsets and the structure of every atom
  const double rc = rcut_all(c) + c->list_buffer;
  std::vector<Cell> cells(n_struct);
  std::vector<int> boff(n_struct + 1), sid(n);
  int maxnb = c->maxnb;
  int guess = 0;
  for (int s = 0; s < n_struct; s++) {
    std::string err;
    rcode = make_cell(lattices + 9 * (size_t)s, rc, &cells[s], &err);
    if (rcode) return set_err(c, rcode, "structure %d: %s", s, err.c_str());
    if (cells[s].nrep >= (1 << (31 - TGX_IMG_SHIFT)))
      return set_err(c, TGX_ERR_UNSUPPORTED, "structure %d: box too small for the cutoff", s);
    const int ns = offsets[s + 1] - offsets[s];
    boff[s + 1] = boff[s] + (cells[s].small ? 1 : cells[s].mx * cells[s].my * cells[s].mz);
    for (int i = offsets[s]; i < offsets[s + 1]; i++) sid[i] = s;
    guess = std::max(guess, guess_maxnb(ns * cells[s].nrep,
                                        cells[s].La * cells[s].Lb * cells[s].Lc, rc));
  }
  if (maxnb <= 0 || maxnb < guess) maxnb = guess;
  const int n_bins = boff[n_struct];
  if (x.b_cells.ensure(sizeof(Cell) * n_struct) || x.b_sid.ensure(sizeof(int) * n) ||
      x.b_boff.ensure(sizeof(int) * (n_struct + 1)) ||
      x.b_aoff.ensure(sizeof(int) * (n_struct + 1)) || x.cell_count.ensure(sizeof(int) * n_bins) ||
      x.cell_fill.ensure(sizeof(int) * n_bins))
    return set_err(c, TGX_ERR_HIP, "device allocation failed");
  ENSURE(c, cell_of, sizeof(int) * n);
  ENSURE(c, cell_atoms, sizeof(int) * n);
  ENSURE(c, cell_start, sizeof(int) * (n_bins + 1));
  ENSURE(c, nb_count, sizeof(int) * n);
  ENSURE(c, nb_flag, sizeof(int) * 4);
  TGX_HIP(c, hipMemcpyAsync(x.b_cells.p, cells.data(), sizeof(Cell) * n_struct,
                            hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.b_sid.p, sid.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                            c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.b_boff.p, boff.data(), sizeof(int) * (n_struct + 1),
                            hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.b_aoff.p, offsets, sizeof(int) * (n_struct + 1),
                            hipMemcpyHostToDevice, c->stream));
  x.ev_on = true;
  ev_rec(c, 0);
  bool sized = false;
  for (int attempt = 0; attempt < 3 && !sized; attempt++) {
    ENSURE(c, nb_list, sizeof(int) * (size_t)n * maxnb);
    if (do_forces) ENSURE(c, nb_rev, sizeof(int) * (size_t)n * maxnb);
    TGX_HIP(c, hipMemsetAsync(c->nb_flag.p, 0, sizeof(int) * 4, c->stream));
    TGX_HIP(c, launch_neighbors_many(c->pos.as<double>(), n, x.b_cells.as<Cell>(),
                                     x.b_sid.as<int>(), x.b_boff.as<int>(), x.b_aoff.as<int>(),
                                     n_bins, rc, c->cell_of.as<int>(), x.cell_count.as<int>(),
                                     c->cell_start.as<int>(), x.cell_fill.as<int>(),
                                     c->cell_atoms.as<int>(), c->nb_count.as<int>(), maxnb,
                                     c->nb_list.as<int>(), do_forces ? c->nb_rev.as<int>() : nullptr,
                                     c->nb_flag.as<int>(), c->stream));
    int flag[4];
    TGX_HIP(c, hipMemcpyAsync(flag, c->nb_flag.p, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    if (flag[0] > maxnb) {
      maxnb = (flag[0] + flag[0] / 4 + 8 + 7) / 8 * 8;
      continue;
    }
    if (flag[1]) return set_err(c, TGX_ERR_STATE, "neighbour list is not symmetric");
    sized = true;
  }
  if (!sized) return set_err(c, TGX_ERR_CAPACITY, "neighbour capacity could not be sized");
  c->maxnb = maxnb;
  ev_rec(c, 1);
  int n_sites = 0;
  rcode = setup_sites(c, n, species, &n_sites);
  if (rcode) return rcode;
  rcode = ef_launch(c, n, n_sites, Geom{cells[0], x.b_cells.as<Cell>(), x.b_sid.as<int>()},
                    do_forces, true, n_struct);
  if (rcode) return rcode;
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  collect_times(c);
  TGX_HIP(c, hipMemcpy(e_atom, c->eatom.p, 8 * (size_t)n, hipMemcpyDeviceToHost));
  if (do_forces) {
    TGX_HIP(c, hipMemcpy(forces, c->forces.p, 24 * (size_t)n, hipMemcpyDeviceToHost));
    std::vector<double> o(10 * (size_t)n_struct);
    TGX_HIP(c, hipMemcpy(o.data(), x.b_out.p, sizeof(double) * o.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < n_struct; s++)
      for (int k = 0; k < 9; k++) virials[9 * (size_t)s + k] = o[10 * (size_t)s + k];
  }
  return TGX_OK;
}

// ---- surrogate MLP -----------------------------------------------------------

// Xavier-uniform weights from a splitmix64 stream, zero biases.
static void sg_init_params(int D, int H, uint64_t seed, std::vector<float>* P) {
  const size_t NP = mlp_nparams(D);
  P->assign(NP, 0.f);
  uint64_t st = seed;
  auto uni = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const double a1 = std::sqrt(6.0 / (D + H)), a2 = std::sqrt(6.0 / (2.0 * H)),
               a3 = std::sqrt(6.0 / (H + 1.0));
  size_t k = 0;
  for (size_t i = 0; i < (size_t)H * D; i++) (*P)[k++] = (float)(a1 * (2 * uni() - 1));
  k += H;
  for (size_t i = 0; i < (size_t)H * H; i++) (*P)[k++] = (float)(a2 * (2 * uni() - 1));
  k += H;
  for (int i = 0; i < H; i++) (*P)[k++] = (float)(a3 * (2 * uni() - 1));
}

// mu, sigma to the device copy the kernels read.
static int sg_push_norm(tgx_ctx* c) {
  Extra& x = X(c);
  const float nm[2] = {x.sg_mu, x.sg_sigma};
  if (x.sg_norm.ensure(sizeof(nm))) return set_err(c, TGX_ERR_HIP, "allocation failed");
  TGX_HIP(c, hipMemcpyAsync(x.sg_norm.p, nm, sizeof(nm), hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  return TGX_OK;
}

static int sg_ready(tgx_ctx* c) {
  if (X(c).sg_dim <= 0) return set_err(c, TGX_ERR_STATE, "surrogate not initialised");
  return TGX_OK;
}

// Work buffers for n training/inference sites.
static int sg_work(tgx_ctx* c, int n) {
  Extra& x = X(c);
  const int H = mlp_hidden(), D = x.sg_dim;
  const int nb = (n + 15) / 16;
  if (x.sg_part.ensure(sizeof(float) * (size_t)nb * mlp_nparams(D)) ||
      x.sg_z1.ensure(sizeof(float) * (size_t)n * H) ||
      x.sg_z2.ensure(sizeof(float) * (size_t)n * H) || x.sg_e.ensure(8 * (size_t)n))
    return set_err(c, TGX_ERR_HIP, "allocation failed (surrogate)");
  return TGX_OK;
}

// One Adam step on the sites rows[0..n) of the sample matrix sq (FP32, D
// wide) with targets sy.
static int sg_adam_step(tgx_ctx* c, int n, const float* sq, const double* sy, const int* rows,
                        float lr) {
  Extra& x = X(c);
  const int D = x.sg_dim;
  int rc = sg_work(c, n);
  if (rc) return rc;
  // rows index both the descriptors and the targets: gather the targets
  TGX_HIP(c, launch_mlp_forward<float>(n, D, sq, D, rows, x.sg_P.as<float>(),
                                       x.sg_W1t.as<unsigned short>(),
                                       x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                       x.sg_e.as<double>(), x.sg_z1.as<float>(),
                                       x.sg_z2.as<float>(), nullptr, c->stream));
  TGX_HIP(c, launch_mlp_gradient<float>(n, D, sq, D, rows, sy, x.sg_e.as<double>(),
                                        x.sg_z1.as<float>(), x.sg_z2.as<float>(),
                                        x.sg_P.as<float>(), x.sg_W2t.as<unsigned short>(),
                                        x.sg_norm.as<float>(), x.sg_part.as<float>(),
                                        x.sg_grad.as<float>(), c->stream));
  x.sg_step += 1;
  TGX_HIP(c, launch_mlp_adam(D, x.sg_P.as<float>(), x.sg_grad.as<float>(), x.sg_m.as<float>(),
                             x.sg_v.as<float>(), lr, x.sg_step, x.sg_W1t.as<unsigned short>(),
                             x.sg_W2t.as<unsigned short>(), c->stream));
  return TGX_OK;
}

// Targets in row order for a minibatch (the gradient kernel reads y[site]).
__global__ void sg_gather_y_kernel(int n, const int* rows, const double* sy, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = sy[rows[i]];
}

// epochs passes over samples [0, n_s) of (sq, sy) in minibatches of `batch`
// in a seeded random order; returns the RMS error (eV) of the final weights
// on all samples.
static int sg_train(tgx_ctx* c, int n_s, const float* sq, const double* sy, int epochs,
                    int batch, float lr, uint64_t seed, double* rmse) {
  Extra& x = X(c);
  if (n_s <= 0) return set_err(c, TGX_ERR_ARG, "no samples");
  batch = std::max(16, std::min(batch, n_s));
  if (x.sg_rows.ensure(sizeof(int) * (size_t)n_s) || x.sg_y.ensure(8 * (size_t)n_s))
    return set_err(c, TGX_ERR_HIP, "allocation failed (surrogate)");
  if (!x.sg_norm_set) {
    std::vector<double> y(n_s);
    TGX_HIP(c, hipMemcpyAsync(y.data(), sy, 8 * (size_t)n_s, hipMemcpyDeviceToHost, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    double m = 0, v = 0;
    for (double t : y) m += t;
    m /= n_s;
    for (double t : y) v += (t - m) * (t - m);
    v = std::sqrt(v / n_s);
    x.sg_mu = (float)m;
    x.sg_sigma = (float)(v > 1e-6 ? v : 1e-6);
    x.sg_norm_set = true;
    int rc = sg_push_norm(c);
    if (rc) return rc;
  }
  std::vector<int> perm(n_s);
  uint64_t st = seed;
  for (int e = 0; e < epochs; e++) {
    for (int i = 0; i < n_s; i++) perm[i] = i;
    for (int i = n_s - 1; i > 0; i--) {  // Fisher-Yates with a splitmix64 stream
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      std::swap(perm[i], perm[(size_t)(z % (uint64_t)(i + 1))]);
    }
    TGX_HIP(c, hipMemcpyAsync(x.sg_rows.p, perm.data(), sizeof(int) * (size_t)n_s,
                              hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(sg_gather_y_kernel, dim3((n_s + 255) / 256), dim3(256), 0, c->stream, n_s,
                       x.sg_rows.as<int>(), sy, x.sg_y.as<double>());
    for (int b0 = 0; b0 + batch <= n_s; b0 += batch) {
      int rc = sg_adam_step(c, batch, sq, x.sg_y.as<double>() + b0, x.sg_rows.as<int>() + b0, lr);
      if (rc) return rc;
    }
    // the ordered copy must not be overwritten before the queued steps ran
    TGX_HIP(c, hipStreamSynchronize(c->stream));
  }
  if (rmse) {
    int rc = sg_work(c, n_s);
    if (rc) return rc;
    TGX_HIP(c, launch_mlp_forward<float>(n_s, x.sg_dim, sq, x.sg_dim, nullptr,
                                         x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                         x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                         x.sg_e.as<double>(), nullptr, nullptr, nullptr,
                                         c->stream));
    std::vector<double> e(n_s), y(n_s);
    TGX_HIP(c, hipMemcpyAsync(e.data(), x.sg_e.p, 8 * (size_t)n_s, hipMemcpyDeviceToHost,
                              c->stream));
    TGX_HIP(c, hipMemcpyAsync(y.data(), sy, 8 * (size_t)n_s, hipMemcpyDeviceToHost, c->stream));
    TGX_HIP(c, hipStreamSynchronize(c->stream));
    double a = 0;
    for (int i = 0; i < n_s; i++) a += (e[i] - y[i]) * (e[i] - y[i]);
    *rmse = std::sqrt(a / n_s);
  }
  return TGX_OK;
}

// ---- Monte Carlo -----------------------------------------------------------

namespace {
struct McRun {
  int n, cap, central, n2b;
  int nb;  // trials per speculative batch (1: the one-trial sequence)
  Cell cell;
  double rcut_list, rsoap, move_max, lim2, kT;
  double rdep;  // range of atoms an energy change depends on
  unsigned long long seed;
  bool da;      // surrogate stage in the trial sequence
  int cap_s;    // sample ring capacity (0: no collection)
};
}  // namespace

// One trial: propose, guarded list rebuild, affected sites, SOAP forward,
// surrogate stage, guarded GAP energy GEMM, sample collection, decision.
// Fixed launch shapes (capturable).
static int mc_trial_launch(tgx_ctx* c, const McRun& r) {
  Extra& x = X(c);
  int* ctl = x.mc_ctl.as<int>();
  double* dctl = x.mc_dctl.as<double>();
  TGX_HIP(c, launch_mc_propose(r.n, c->pos.as<double>(), x.mc_pref.as<double>(), r.seed,
                               r.move_max, r.lim2, ctl, dctl, c->stream));
  TGX_HIP(c, launch_neighbors(c->pos.as<double>(), r.n, r.cell, r.rcut_list,
                              c->cell_of.as<int>(), x.cell_count.as<int>(),
                              c->cell_start.as<int>(), x.cell_fill.as<int>(),
                              c->cell_atoms.as<int>(), c->nb_count.as<int>(), c->maxnb,
                              c->nb_list.as<int>(), nullptr, c->nb_flag.as<int>(), c->stream,
                              ctl + MC_REBUILD, x.mc_pref.as<double>()));
  const int D = c->blk->dim;
  const int exact = c->precision == TGX_FP64 ? 1 : 0;
  if (c->has_soap) {
    // the affected-site scan also computes the 2-body change; for plain
    // Metropolis that is all of stage 1, which then needs no launch
    TGX_HIP(c, launch_mc_affected(ctl, dctl, c->pos.as<double>(), r.cell, r.rsoap,
                                  c->nb_count.as<int>(), c->nb_list.as<int>(), c->maxnb,
                                  c->species.as<int>(), r.central, c->blk->site_atom.as<int>(), r.cap,
                                  x.mc_naff.as<int>(), x.mc_twob.as<TwoBodyDev>(), r.n2b, exact,
                                  1, r.da ? 0 : 1, c->stream));
    int rcode = soap_forward(c, r.n, Geom{r.cell}, r.cap, true);
    if (rcode) return rcode;
    if (r.da) {
      hipError_t e =
          c->precision == TGX_FP64
              ? launch_mlp_forward<double>(r.cap, D, c->blk->q.as<double>(), c->blk->dim_pad, nullptr,
                                           x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                           x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                           x.mc_esur.as<double>(), nullptr, nullptr,
                                           ctl + MC_DA, c->stream)
              : launch_mlp_forward<float>(r.cap, D, c->blk->q.as<float>(), c->blk->dim_pad, nullptr,
                                          x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                          x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                          x.mc_esur.as<double>(), nullptr, nullptr, ctl + MC_DA,
                                          c->stream);
      TGX_HIP(c, e);
    }
  }
  if (r.da || !c->has_soap)
    TGX_HIP(c, launch_mc_stage1(c->pos.as<double>(), c->species.as<int>(), r.cell, c->maxnb,
                                c->nb_count.as<int>(), c->nb_list.as<int>(),
                                c->blk->site_atom.as<int>(), x.mc_naff.as<int>(),
                                x.mc_esur.as<double>(), x.mc_esur_cur.as<double>(),
                                x.mc_twob.as<TwoBodyDev>(), r.n2b, exact, c->has_soap ? 1 : 0,
                                r.seed, r.kT, ctl, dctl, x.mc_trace.as<double>(), c->stream));
  // without sample collection the decision kernel sums the few-site
  // kernel's partial energies itself (one launch less per trial)
  int nparts = 0;
  if (c->has_soap) {
    int rcode = kernel_step(c, r.cap, false, ctl + MC_GATE, nullptr,
                            r.cap_s > 0 ? nullptr : &nparts);
    if (rcode) return rcode;
    if (r.cap_s > 0) {
      hipError_t e =
          c->precision == TGX_FP64
              ? launch_mc_collect<double>(ctl, x.mc_naff.as<int>(), c->blk->q.as<double>(), c->blk->dim_pad,
                                          D, c->blk->esite.as<double>(), x.mc_sq.as<float>(),
                                          x.mc_sy.as<double>(), r.cap_s, ctl + MC_CURSOR,
                                          c->stream)
              : launch_mc_collect<float>(ctl, x.mc_naff.as<int>(), c->blk->q.as<float>(), c->blk->dim_pad,
                                         D, c->blk->esite.as<double>(), x.mc_sq.as<float>(),
                                         x.mc_sy.as<double>(), r.cap_s, ctl + MC_CURSOR,
                                         c->stream);
      TGX_HIP(c, e);
    }
  }
  TGX_HIP(c, launch_mc_stage2(c->pos.as<double>(), c->blk->site_atom.as<int>(), x.mc_naff.as<int>(),
                              c->has_soap ? c->blk->esite.as<double>() : nullptr,
                              nparts ? c->blk->epart.as<float>() : nullptr, nparts, r.cap,
                              c->blk->delta * c->blk->delta,
                              x.mc_ecur.as<double>(), x.mc_esur.as<double>(),
                              x.mc_esur_cur.as<double>(), r.seed, r.kT, ctl, dctl,
                              x.mc_trace.as<double>(), c->stream));
  return TGX_OK;
}

// One speculative batch of r.nb trials: the rebuild requested by the last
// walk, proposals, affected sites as virtual atoms with the 2-body changes,
// SOAP and (delayed acceptance) surrogate energies and the kernel step of
// all virtual sites, then the walk that decides the trials in order.
static int mc_batch_launch(tgx_ctx* c, const McRun& r) {
  Extra& x = X(c);
  int* ctl = x.mc_ctl.as<int>();
  double* dctl = x.mc_dctl.as<double>();
  const int vb = r.n + r.nb, ns = r.nb * r.cap;
  TGX_HIP(c, launch_neighbors(c->pos.as<double>(), r.n, r.cell, r.rcut_list,
                              c->cell_of.as<int>(), x.cell_count.as<int>(),
                              c->cell_start.as<int>(), x.cell_fill.as<int>(),
                              c->cell_atoms.as<int>(), c->nb_count.as<int>(), c->maxnb,
                              c->nb_list.as<int>(), nullptr, c->nb_flag.as<int>(), c->stream,
                              ctl + MC_REBUILD, x.mc_pref.as<double>()));
  TGX_HIP(c, launch_mcb_propose(r.n, r.nb, c->pos.as<double>(), c->species.as<int>(),
                                x.mc_pref.as<double>(), r.seed, r.move_max, r.lim2, ctl,
                                x.mcb_mi.as<int>(), x.mcb_md.as<double>(), c->stream));
  const int exact = c->precision == TGX_FP64 ? 1 : 0;
  TGX_HIP(c, launch_mcb_affected(r.n, r.nb, r.cap, vb, c->pos.as<double>(), c->species.as<int>(),
                                 c->nb_count.as<int>(), c->nb_list.as<int>(), c->maxnb, r.cell,
                                 r.rsoap, r.central, c->blk->site_atom.as<int>(),
                                 x.mcb_vatom.as<int>(), x.mcb_mi.as<int>(),
                                 x.mcb_md.as<double>(), x.mc_twob.as<TwoBodyDev>(), r.n2b, exact,
                                 c->stream));
  int rcode = soap_forward(c, r.n, Geom{r.cell}, ns);
  if (rcode) return rcode;
  const int D = c->blk->dim;
  if (r.da) {
    hipError_t e =
        c->precision == TGX_FP64
            ? launch_mlp_forward<double>(ns, D, c->blk->q.as<double>(), c->blk->dim_pad, nullptr,
                                         x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                         x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                         x.mc_esur.as<double>(), nullptr, nullptr, ctl + MC_DA,
                                         c->stream)
            : launch_mlp_forward<float>(ns, D, c->blk->q.as<float>(), c->blk->dim_pad, nullptr,
                                        x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                        x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                        x.mc_esur.as<double>(), nullptr, nullptr, ctl + MC_DA,
                                        c->stream);
    TGX_HIP(c, e);
  }
  rcode = kernel_step(c, ns, false);
  if (rcode) return rcode;
  TGX_HIP(c, launch_mcb_walk(r.nb, r.cap, c->pos.as<double>(), r.cell, r.rdep,
                             x.mcb_vatom.as<int>(), c->blk->esite.as<double>(),
                             x.mc_ecur.as<double>(), x.mc_esur.as<double>(),
                             x.mc_esur_cur.as<double>(), r.seed, r.kT, ctl, dctl,
                             x.mcb_mi.as<int>(), x.mcb_md.as<double>(), x.mc_trace.as<double>(),
                             c->stream));
  return TGX_OK;
}

// Surrogate energies of every atom of the current state: SOAP of all sites
// and the MLP (the trial sequence then keeps them current).
static int mc_refresh_esur(tgx_ctx* c, const McRun& r, int n_sites) {
  Extra& x = X(c);
  // all sites: site_atom is the trial's site list, so the iota list is used
  if (x.iota.ensure(sizeof(int) * (size_t)r.n)) return set_err(c, TGX_ERR_HIP, "allocation failed");
  std::vector<int> io(r.n);
  for (int i = 0; i < r.n; i++) io[i] = i;
  TGX_HIP(c, hipMemcpyAsync(x.iota.p, io.data(), sizeof(int) * r.n, hipMemcpyHostToDevice,
                            c->stream));
  TGX_HIP(c, hipMemcpyAsync(c->blk->site_atom.p, x.iota.p, sizeof(int) * r.n,
                            hipMemcpyDeviceToDevice, c->stream));
  int rcode = soap_forward(c, r.n, Geom{r.cell}, n_sites);
  if (rcode) return rcode;
  const int D = c->blk->dim;
  hipError_t e =
      c->precision == TGX_FP64
          ? launch_mlp_forward<double>(n_sites, D, c->blk->q.as<double>(), c->blk->dim_pad, nullptr,
                                       x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                       x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                       x.mc_esur_cur.as<double>(), nullptr, nullptr, nullptr,
                                       c->stream)
          : launch_mlp_forward<float>(n_sites, D, c->blk->q.as<float>(), c->blk->dim_pad, nullptr,
                                      x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                      x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                      x.mc_esur_cur.as<double>(), nullptr, nullptr, nullptr,
                                      c->stream);
  TGX_HIP(c, e);
  return TGX_OK;
}

static int mc_run_impl(tgx_ctx* c, int n, const int* species, const double* lattice,
                       double* pos, int n_trials, double temp, double move_max, double skin,
                       uint64_t seed, const tgx_da_options* da, double* e_trace,
                       tgx_mc_stats* stats) {
  if (!c) return TGX_ERR_ARG;
  if (n < 2 || !species || !lattice || !pos || n_trials < 0 || !(temp > 0) ||
      !(move_max > 0))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  if (!(skin > std::sqrt(3.0) * move_max))
    return set_err(c, TGX_ERR_ARG, "skin must exceed sqrt(3) * move_max (%g)",
                   std::sqrt(3.0) * move_max);
  if (!has_model(c)) return set_err(c, TGX_ERR_STATE, "no model loaded");
  if (X(c).md.active) return set_err(c, TGX_ERR_STATE, "end the MD run first (tgx_md_end)");
  if (!c->a3b.empty() || !c->corep.empty())
    return set_err(c, TGX_ERR_UNSUPPORTED, "Monte Carlo takes no angle_3b or core_pot blocks");
  if (c->has_soap && c->blocks[0]->cmp > 0)
    return set_err(c, TGX_ERR_UNSUPPORTED, "Monte Carlo takes no compress_soap block");
  if (c->has_soap && c->blocks.size() > 1)
    return set_err(c, TGX_ERR_UNSUPPORTED,
                   "Monte Carlo takes models of one soap_turbo block (this one has %d)",
                   (int)c->blocks.size());
  Extra& x = X(c);
  if (da) {
    if (!c->has_soap) return set_err(c, TGX_ERR_STATE, "delayed acceptance needs a SOAP model");
    int rc = sg_ready(c);
    if (rc) return rc;
    if (x.sg_dim != c->blk->dim)
      return set_err(c, TGX_ERR_ARG, "surrogate dimension %d != SOAP dimension %d", x.sg_dim,
                     c->blk->dim);
    if (c->blk->sd.central_species != 0)
      return set_err(c, TGX_ERR_UNSUPPORTED, "delayed acceptance needs every atom as a site");
    if (da->n_train_trials < 0 || da->train_every < 64 || da->epochs < 1 || da->batch < 16 ||
        !(da->lr > 0) || !(da->max_rmse > 0))
      return set_err(c, TGX_ERR_ARG, "bad delayed-acceptance options");
  }
  (void)hipSetDevice(c->device);
  int rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  McRun r{};
  r.n = n;
  r.move_max = move_max;
  r.lim2 = 0.25 * skin * skin;
  r.kT = (double)8.617333262e-5f * temp;  // the reference's default-real kB
  r.seed = seed;
  r.n2b = (int)c->twob.size();
  r.central = c->has_soap ? c->blk->sd.central_species : 0;
  r.rsoap = c->has_soap ? c->blk->sd.rcut_max : 0.0;
  r.da = da != nullptr;
  r.rcut_list = rcut_all(c) + skin;
  std::string err;
  rcode = make_cell(lattice, r.rcut_list, &r.cell, &err);
  if (rcode) return set_err(c, rcode, "%s", err.c_str());
  if (r.cell.small)
    return set_err(c, TGX_ERR_UNSUPPORTED,
                   "Monte Carlo needs a box of at least 2 (rcut + skin) = %g A",
                   2 * r.rcut_list);
  c->maxnb = 0;
  rcode = build_list(c, n, r.cell, r.rcut_list, false);
  if (rcode) return rcode;
  {
    std::vector<int> cnt(n);
    TGX_HIP(c, hipMemcpy(cnt.data(), c->nb_count.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    const int mx = *std::max_element(cnt.begin(), cnt.end());
    const int want = ((mx * 3) / 2 + 16 + 7) / 8 * 8;
    if (c->maxnb < want) {
      c->maxnb = want;
      rcode = build_list(c, n, r.cell, r.rcut_list, false);
      if (rcode) return rcode;
    }
  }
  r.cap = c->maxnb + 1;
  if (r.cap > 1024) return set_err(c, TGX_ERR_UNSUPPORTED, "neighbour lists too long for MC");
  // speculative batches need a SOAP model; the virtual atoms follow the
  // structure in pos / species / the list, which grow to hold them
  r.nb = c->has_soap ? (x.mc_batch > 0 ? x.mc_batch : 8) : 1;
  r.nb = std::min(r.nb, 64);
  r.rdep = 2.0 * rcut_all(c) + 1e-9;  // the walk adds each trial's displacement
  if (r.nb > 1) {
    const size_t n_ext = (size_t)n + r.nb + (size_t)r.nb * r.cap;
    if (n_ext >= ((size_t)1 << TGX_IMG_SHIFT)) return set_err(c, TGX_ERR_ARG, "too many atoms");
    ENSURE(c, pos, 24 * n_ext);
    ENSURE(c, species, 4 * n_ext);
    ENSURE(c, nb_count, 4 * n_ext);
    ENSURE(c, nb_list, 4 * n_ext * c->maxnb);
    if (x.mcb_mi.ensure(sizeof(int) * 4 * r.nb) || x.mcb_md.ensure(sizeof(double) * 8 * r.nb) ||
        x.mcb_vatom.ensure(sizeof(int) * (size_t)r.nb * r.cap))
      return set_err(c, TGX_ERR_HIP, "device allocation failed (mc batches)");
    TGX_HIP(c, hipMemcpyAsync(c->pos.p, pos, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    TGX_HIP(c, hipMemcpyAsync(c->species.p, species, 4 * (size_t)n, hipMemcpyHostToDevice,
                              c->stream));
    rcode = build_list(c, n, r.cell, r.rcut_list, false);
    if (rcode) return rcode;
  }
  const int nsb = r.nb > 1 ? r.nb * r.cap : r.cap;  // sites of one launch sequence
  const int n_train = da ? std::min(da->n_train_trials, n_trials) : 0;
  r.cap_s = (da && n_train > 0) ? 65536 : 0;
  int n_sites = 0;
  rcode = setup_sites(c, n, species, &n_sites);
  if (rcode) return rcode;
  // energy of the initial state; per-atom SOAP site energies
  x.ev_on = false;
  rcode = ef_launch(c, n, n_sites, Geom{r.cell}, 0, false);
  if (rcode) return rcode;
  x.ev_on = false;
  if (x.mc_pref.ensure(24 * (size_t)n) || x.mc_ctl.ensure(sizeof(int) * MC_NCTL) ||
      x.mc_dctl.ensure(sizeof(double) * MC_NDCTL) || x.mc_ecur.ensure(8 * (size_t)n) ||
      x.mc_naff.ensure(sizeof(int) * 4) || x.mc_trace.ensure(8 * ((size_t)n_trials + 1)) ||
      x.mc_twob.ensure(sizeof(TwoBodyDev) * std::max(1, r.n2b)) ||
      x.mc_esur.ensure(8 * (size_t)nsb) || x.mc_esur_cur.ensure(8 * (size_t)n) ||
      (r.cap_s && (x.mc_sq.ensure(sizeof(float) * (size_t)r.cap_s * c->blk->dim) ||
                   x.mc_sy.ensure(8 * (size_t)r.cap_s))))
    return set_err(c, TGX_ERR_HIP, "device allocation failed (mc)");
  ENSURE(c, blk->site_atom, sizeof(int) * (size_t)std::max(n, nsb));
  if (c->has_soap) {
    // descriptor buffers for the larger of a launch sequence's sites and
    // all sites, so that refreshing the surrogate energies never moves a
    // buffer the captured trial graphs use
    const size_t rows = (size_t)std::max(n_sites, nsb), ts = tsize(c->precision);
    ENSURE(c, blk->q, ts * rows * c->blk->dim_pad);
    ENSURE(c, blk->cbuf, ts * rows * c->blk->sd.nrp * c->blk->sd.L2);
    ENSURE(c, blk->pnorm, ts * rows);
    if (c->precision == TGX_MIXED) ENSURE(c, blk->qb, 2 * rows * c->blk->dim_pad);
  }
  TGX_HIP(c, launch_mc_init(n, c->blk->site_of_atom.as<int>(),
                            (c->has_soap && n_sites > 0) ? c->blk->esite.as<double>() : nullptr,
                            x.mc_ecur.as<double>(), c->stream));
  if (r.n2b)
    TGX_HIP(c, hipMemcpyAsync(x.mc_twob.p, c->twob.data(), sizeof(TwoBodyDev) * r.n2b,
                              hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.mc_pref.p, c->pos.p, 24 * (size_t)n, hipMemcpyDeviceToDevice,
                            c->stream));
  double e0;
  TGX_HIP(c, hipMemcpyAsync(&e0, x.out.as<double>() + 9, 8, hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipMemsetAsync(x.mc_ctl.p, 0, sizeof(int) * MC_NCTL, c->stream));
  TGX_HIP(c, hipMemsetAsync(c->nb_flag.p, 0, sizeof(int) * 4, c->stream));
  TGX_HIP(c, hipMemsetAsync(x.mc_naff.p, 0, sizeof(int) * 4, c->stream));
  TGX_HIP(c, hipMemsetAsync(c->blk->site_atom.p, 0, sizeof(int) * (size_t)r.cap, c->stream));
  TGX_HIP(c, hipMemsetAsync(x.mc_esur.p, 0, 8 * (size_t)nsb, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  // a dry trial and batch (MC_END = 0: every control kernel returns at
  // once) size the SOAP and GEMM buffers before anything is captured
  if (r.nb > 1) {
    rcode = mc_batch_launch(c, r);
    if (rcode) return rcode;
  }
  rcode = mc_trial_launch(c, r);
  if (rcode) return rcode;
  if (r.da) {
    rcode = sg_work(c, std::max(n, nsb));
    if (rcode) return rcode;
  }
  std::vector<double> d(MC_NDCTL, 0.0);
  d[MC_ETOT] = e0;
  d[MC_MAXMS] = da ? da->max_rmse * da->max_rmse : 0.0;
  int ctl_h[MC_NCTL] = {0};
  ctl_h[MC_END] = n_trials;
  ctl_h[MC_COLLECT] = n_train > 0 ? 1 : 0;
  ctl_h[MC_DA] = (da && n_train == 0) ? 1 : 0;
  if (ctl_h[MC_DA]) {
    rcode = mc_refresh_esur(c, r, n_sites);
    if (rcode) return rcode;
  }
  TGX_HIP(c, hipMemcpyAsync(x.mc_dctl.p, d.data(), sizeof(double) * MC_NDCTL,
                            hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.mc_ctl.p, ctl_h, sizeof(ctl_h), hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.mc_trace.p, &e0, 8, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  const int TB = std::max(1, std::min(64, n_trials));
  const int KB = 8;  // speculative batches per replay
  const char* ng = getenv("TGX_NO_GRAPH");
  const bool use_graph = !(ng && ng[0] == '1') && n_trials > 0;
  const bool batched = r.nb > 1;
  hipGraph_t g = nullptr, gb = nullptr;
  hipGraphExec_t ge = nullptr, gbe = nullptr;
  auto cleanup = [&]() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (gbe) (void)hipGraphExecDestroy(gbe);
    if (gb) (void)hipGraphDestroy(gb);
    ge = gbe = nullptr;
    g = gb = nullptr;
    x.ev_on = true;
  };
  auto capture = [&](bool batch, hipGraph_t* gr, hipGraphExec_t* ex) -> int {
    TGX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = TGX_OK;
    for (int b = 0; b < (batch ? KB : TB) && !rc; b++)
      rc = batch ? mc_batch_launch(c, r) : mc_trial_launch(c, r);
    hipError_t e = hipStreamEndCapture(c->stream, gr);
    if (rc) return rc;
    if (e != hipSuccess)
      return set_err(c, TGX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(ex, *gr, nullptr, nullptr, 0);
    if (e != hipSuccess)
      return set_err(c, TGX_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    return TGX_OK;
  };
  if (use_graph) {
    if (!batched || n_train > 0) rcode = capture(false, &g, &ge);
    if (!rcode && batched) rcode = capture(true, &gb, &gbe);
    if (rcode) {
      cleanup();
      return rcode;
    }
  }
  // trials in blocks of TB (one replay each) while training data are
  // collected (the surrogate is trained every train_every trials and its
  // stage switched on after n_train trials), else in speculative batches,
  // KB per replay, until the device's trial counter reaches n_trials
  double train_rmse = 0.0;
  const int every = da ? std::max(TB, da->train_every / TB * TB) : 0;
  int done = 0, last_train = 0;
  while (done < n_trials && !rcode) {
    const bool collecting = n_train > 0 && last_train < n_train;
    if (batched && !collecting) {
      if (gbe) {
        hipError_t le = hipGraphLaunch(gbe, c->stream);
        if (le != hipSuccess)
          rcode = set_err(c, TGX_ERR_HIP, "graph launch: %s", hipGetErrorString(le));
      } else {
        for (int b = 0; b < KB && !rcode; b++) rcode = mc_batch_launch(c, r);
      }
      if (rcode) break;
      int step = 0;
      TGX_HIP(c, hipMemcpyAsync(&step, x.mc_ctl.as<int>() + MC_STEP, sizeof(int),
                                hipMemcpyDeviceToHost, c->stream));
      TGX_HIP(c, hipStreamSynchronize(c->stream));
      if (step <= done) {
        rcode = set_err(c, TGX_ERR_STATE, "Monte Carlo batches made no progress at trial %d", step);
        break;
      }
      done = step;
      continue;
    }
    if (ge) {
      hipError_t le = hipGraphLaunch(ge, c->stream);
      if (le != hipSuccess) rcode = set_err(c, TGX_ERR_HIP, "graph launch: %s", hipGetErrorString(le));
    } else {
      for (int b = 0; b < TB && !rcode; b++) rcode = mc_trial_launch(c, r);
    }
    done = std::min(n_trials, done + TB);
    if (rcode || n_train == 0 || last_train >= n_train) continue;
    if (done - last_train >= every || done >= n_train) {
      int cur = 0;
      TGX_HIP(c, hipMemcpyAsync(&cur, x.mc_ctl.as<int>() + MC_CURSOR, sizeof(int),
                                hipMemcpyDeviceToHost, c->stream));
      TGX_HIP(c, hipStreamSynchronize(c->stream));
      const int n_s = std::min(cur, r.cap_s);
      if (n_s >= 16) {
        rcode = sg_train(c, n_s, x.mc_sq.as<float>(), x.mc_sy.as<double>(), da->epochs,
                         da->batch, (float)da->lr, seed ^ (uint64_t)done, &train_rmse);
        if (rcode) break;
      }
      last_train = done;
      if (done >= n_train) {
        // switch to delayed acceptance from the next trial on
        rcode = mc_refresh_esur(c, r, n_sites);
        if (rcode) break;
        int on[2] = {1, 0};
        TGX_HIP(c, hipMemcpyAsync(x.mc_ctl.as<int>() + MC_DA, &on[0], sizeof(int),
                                  hipMemcpyHostToDevice, c->stream));
        TGX_HIP(c, hipMemcpyAsync(x.mc_ctl.as<int>() + MC_COLLECT, &on[1], sizeof(int),
                                  hipMemcpyHostToDevice, c->stream));
        TGX_HIP(c, hipStreamSynchronize(c->stream));
        last_train = n_train;
      }
    }
  }
  hipError_t se = hipStreamSynchronize(c->stream);
  cleanup();
  if (rcode) return rcode;
  if (se != hipSuccess) return set_err(c, TGX_ERR_HIP, "%s", hipGetErrorString(se));
  int flag[4];
  TGX_HIP(c, hipMemcpy(flag, c->nb_flag.p, sizeof(flag), hipMemcpyDeviceToHost));
  if (flag[0] > c->maxnb)
    return set_err(c, TGX_ERR_CAPACITY, "neighbour capacity exceeded during the run (%d > %d)",
                   flag[0], c->maxnb);
  TGX_HIP(c, hipMemcpy(ctl_h, x.mc_ctl.p, sizeof(ctl_h), hipMemcpyDeviceToHost));
  TGX_HIP(c, hipMemcpy(d.data(), x.mc_dctl.p, sizeof(double) * MC_NDCTL, hipMemcpyDeviceToHost));
  TGX_HIP(c, hipMemcpy(pos, c->pos.p, 24 * (size_t)n, hipMemcpyDeviceToHost));
  if (e_trace)
    TGX_HIP(c, hipMemcpy(e_trace, x.mc_trace.p, 8 * ((size_t)n_trials + 1),
                         hipMemcpyDeviceToHost));
  if (stats) {
    stats->n_accepted = ctl_h[MC_NACC];
    stats->n_stage1_rejected = ctl_h[MC_NREJ1];
    stats->n_gap_evaluations = ctl_h[MC_NGAP];
    stats->n_rebuilds = ctl_h[MC_NREB];
    stats->train_rmse = train_rmse;
    stats->da_rmse = std::sqrt(d[MC_EMA]);
    stats->fallback = ctl_h[MC_FALLBACK];
    stats->pad_ = 0;
  }
  if (ctl_h[MC_STEP] != n_trials)
    return set_err(c, TGX_ERR_STATE, "ran %d of %d trials", ctl_h[MC_STEP], n_trials);
  return TGX_OK;
}

int tgx_mc_run(tgx_ctx* c, int n, const int* species, const double* lattice, double* pos,
               int n_trials, double temp, double move_max, double skin, uint64_t seed,
               double* e_trace, int* n_accepted, int* n_rebuilds) {
  tgx_mc_stats st{};
  const int rc = mc_run_impl(c, n, species, lattice, pos, n_trials, temp, move_max, skin, seed,
                             nullptr, e_trace, &st);
  if (rc == TGX_OK) {
    if (n_accepted) *n_accepted = (int)st.n_accepted;
    if (n_rebuilds) *n_rebuilds = (int)st.n_rebuilds;
  }
  return rc;
}

int tgx_mc_run_da(tgx_ctx* c, int n, const int* species, const double* lattice, double* pos,
                  int n_trials, double temp, double move_max, double skin, uint64_t seed,
                  const tgx_da_options* opt, double* e_trace, tgx_mc_stats* stats) {
  if (!c) return TGX_ERR_ARG;
  if (!opt) return set_err(c, TGX_ERR_ARG, "options required");
  return mc_run_impl(c, n, species, lattice, pos, n_trials, temp, move_max, skin, seed, opt,
                     e_trace, stats);
}

int tgx_surrogate_init(tgx_ctx* c, int dim, uint64_t seed) {
  if (!c) return TGX_ERR_ARG;
  if (dim < 1 || dim > mlp_max_dim())
    return set_err(c, TGX_ERR_UNSUPPORTED, "surrogate input dimension %d not in 1..%d", dim,
                   mlp_max_dim());
  (void)hipSetDevice(c->device);
  Extra& x = X(c);
  const int H = mlp_hidden();
  const size_t NP = mlp_nparams(dim);
  std::vector<float> P;
  sg_init_params(dim, H, seed, &P);
  if (x.sg_P.ensure(4 * NP) || x.sg_m.ensure(4 * NP) || x.sg_v.ensure(4 * NP) ||
      x.sg_grad.ensure(4 * NP) || x.sg_W1t.ensure(2 * mlp_w1_elems(dim)) ||
      x.sg_W2t.ensure(2 * (size_t)H * H))
    return set_err(c, TGX_ERR_HIP, "allocation failed (surrogate)");
  // the zero padding of the unit-major W1 copy
  TGX_HIP(c, hipMemsetAsync(x.sg_W1t.p, 0, 2 * mlp_w1_elems(dim), c->stream));
  TGX_HIP(c, hipMemcpyAsync(x.sg_P.p, P.data(), 4 * NP, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemsetAsync(x.sg_m.p, 0, 4 * NP, c->stream));
  TGX_HIP(c, hipMemsetAsync(x.sg_v.p, 0, 4 * NP, c->stream));
  TGX_HIP(c, launch_mlp_cast(dim, x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                             x.sg_W2t.as<unsigned short>(), c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  x.sg_dim = dim;
  x.sg_step = 0;
  x.sg_mu = 0.f;
  x.sg_sigma = 1.f;
  x.sg_norm_set = false;
  return sg_push_norm(c);
}

int tgx_surrogate_set(tgx_ctx* c, int n_params, const double* P, double mu, double sigma) {
  if (!c) return TGX_ERR_ARG;
  int rc = sg_ready(c);
  if (rc) return rc;
  Extra& x = X(c);
  const size_t NP = mlp_nparams(x.sg_dim);
  if (!P || n_params != (int)NP || !(sigma > 0))
    return set_err(c, TGX_ERR_ARG, "expected %zu parameters and sigma > 0", NP);
  std::vector<float> f(P, P + NP);
  TGX_HIP(c, hipMemcpyAsync(x.sg_P.p, f.data(), 4 * NP, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, launch_mlp_cast(x.sg_dim, x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                             x.sg_W2t.as<unsigned short>(), c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  x.sg_mu = (float)mu;
  x.sg_sigma = (float)sigma;
  x.sg_norm_set = true;
  return sg_push_norm(c);
}

int tgx_surrogate_get(tgx_ctx* c, int n_params, double* P, double* mu, double* sigma) {
  if (!c) return TGX_ERR_ARG;
  int rc = sg_ready(c);
  if (rc) return rc;
  Extra& x = X(c);
  const size_t NP = mlp_nparams(x.sg_dim);
  if (n_params != (int)NP || !P) return set_err(c, TGX_ERR_ARG, "expected %zu parameters", NP);
  std::vector<float> f(NP);
  TGX_HIP(c, hipMemcpy(f.data(), x.sg_P.p, 4 * NP, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < NP; k++) P[k] = f[k];
  if (mu) *mu = x.sg_mu;
  if (sigma) *sigma = x.sg_sigma;
  return TGX_OK;
}

// Host descriptors (n x dim, row-major, FP64) to the device sample matrix.
static int sg_upload(tgx_ctx* c, int n, const double* q, const double* y) {
  Extra& x = X(c);
  const int D = x.sg_dim;
  if (x.sg_q.ensure(sizeof(float) * (size_t)n * D) || x.mc_sy.ensure(8 * (size_t)n))
    return set_err(c, TGX_ERR_HIP, "allocation failed (surrogate)");
  std::vector<float> f((size_t)n * D);
  for (size_t k = 0; k < f.size(); k++) f[k] = (float)q[k];
  TGX_HIP(c, hipMemcpyAsync(x.sg_q.p, f.data(), f.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (y) TGX_HIP(c, hipMemcpyAsync(x.mc_sy.p, y, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  return TGX_OK;
}

int tgx_surrogate_predict(tgx_ctx* c, int n, const double* q, double* e) {
  if (!c) return TGX_ERR_ARG;
  int rc = sg_ready(c);
  if (rc) return rc;
  if (n < 1 || !q || !e) return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  Extra& x = X(c);
  rc = sg_upload(c, n, q, nullptr);
  if (!rc) rc = sg_work(c, n);
  if (rc) return rc;
  TGX_HIP(c, launch_mlp_forward<float>(n, x.sg_dim, x.sg_q.as<float>(), x.sg_dim, nullptr,
                                       x.sg_P.as<float>(), x.sg_W1t.as<unsigned short>(),
                                       x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                       x.sg_e.as<double>(), nullptr, nullptr, nullptr, c->stream));
  TGX_HIP(c, hipMemcpyAsync(e, x.sg_e.p, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  return TGX_OK;
}

int tgx_surrogate_gradient(tgx_ctx* c, int n, const double* q, const double* y, double* grad,
                           double* loss) {
  if (!c) return TGX_ERR_ARG;
  int rc = sg_ready(c);
  if (rc) return rc;
  if (n < 1 || !q || !y || !grad) return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  Extra& x = X(c);
  const int D = x.sg_dim;
  rc = sg_upload(c, n, q, y);
  if (!rc) rc = sg_work(c, n);
  if (rc) return rc;
  TGX_HIP(c, launch_mlp_forward<float>(n, D, x.sg_q.as<float>(), D, nullptr, x.sg_P.as<float>(),
                                       x.sg_W1t.as<unsigned short>(),
                                       x.sg_W2t.as<unsigned short>(), x.sg_norm.as<float>(),
                                       x.sg_e.as<double>(), x.sg_z1.as<float>(),
                                       x.sg_z2.as<float>(), nullptr, c->stream));
  TGX_HIP(c, launch_mlp_gradient<float>(n, D, x.sg_q.as<float>(), D, nullptr,
                                        x.mc_sy.as<double>(), x.sg_e.as<double>(),
                                        x.sg_z1.as<float>(), x.sg_z2.as<float>(),
                                        x.sg_P.as<float>(), x.sg_W2t.as<unsigned short>(),
                                        x.sg_norm.as<float>(), x.sg_part.as<float>(),
                                        x.sg_grad.as<float>(), c->stream));
  const size_t NP = mlp_nparams(D);
  std::vector<float> g(NP);
  std::vector<double> e(n);
  TGX_HIP(c, hipMemcpyAsync(g.data(), x.sg_grad.p, 4 * NP, hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipMemcpyAsync(e.data(), x.sg_e.p, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < NP; k++) grad[k] = g[k];
  if (loss) {
    double a = 0;
    for (int i = 0; i < n; i++) {
      const double r = (e[i] - y[i]) / x.sg_sigma;
      a += r * r;
    }
    *loss = a / n;
  }
  return TGX_OK;
}

int tgx_surrogate_fit(tgx_ctx* c, int n, const double* q, const double* y, int epochs,
                      int batch, double lr, uint64_t seed, double* rmse) {
  if (!c) return TGX_ERR_ARG;
  int rc = sg_ready(c);
  if (rc) return rc;
  if (n < 16 || !q || !y || epochs < 1 || batch < 16 || !(lr > 0))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  Extra& x = X(c);
  rc = sg_upload(c, n, q, y);
  if (rc) return rc;
  return sg_train(c, n, x.sg_q.as<float>(), x.mc_sy.as<double>(), epochs, batch, (float)lr, seed,
                  rmse);
}

// ---- molecular dynamics ---------------------------------------------------

static int md_step_launch(tgx_ctx* c) {
  Extra& x = X(c);
  MdState& m = x.md;
  int* ctl = m.ctl.as<int>();
  const int n = m.n;
  TGX_HIP(c, launch_md_vv(n, c->pos.as<double>(), c->md_vel.as<double>(), c->forces.as<double>(),
                          c->md_fprev.as<double>(), c->md_mass.as<double>(), m.dt, ctl + 2,
                          m.ke.as<double>(), m.pref.as<double>(), 0.25 * m.skin * m.skin,
                          ctl + 0, c->stream));
  TGX_HIP(c, launch_neighbors(c->pos.as<double>(), n, m.cell, m.rcut_list, c->cell_of.as<int>(),
                              x.cell_count.as<int>(), c->cell_start.as<int>(),
                              x.cell_fill.as<int>(), c->cell_atoms.as<int>(),
                              c->nb_count.as<int>(), c->maxnb, c->nb_list.as<int>(),
                              c->nb_rev.as<int>(), c->nb_flag.as<int>(), c->stream, ctl + 0,
                              m.pref.as<double>()));
  // the final energy/virial sum is done by the finish kernel
  int rcode = ef_launch(c, n, m.n_sites, Geom{m.cell}, 1, false, 0, true);
  if (rcode) return rcode;
  TGX_HIP(c, launch_md_finish(n, m.ke.as<double>(), x.out.as<double>(), ctl + 0, ctl + 1, ctl + 2,
                              m.ep.as<double>(), m.ek.as<double>(), m.cap, c->stream));
  return TGX_OK;
}

static void md_release_graph(MdState& m) {
  if (m.exec) (void)hipGraphExecDestroy(m.exec);
  if (m.graph) (void)hipGraphDestroy(m.graph);
  m.exec = nullptr;
  m.graph = nullptr;
}

// Atom order along a Z-order (Morton) curve of the wrapped positions, 10 bits
// per axis of the orthorhombic box (La, Lb, Lc); ties keep the input order.
static std::vector<int> zorder_atoms(int n, const double* pos, double La, double Lb, double Lc) {
  std::vector<std::pair<uint64_t, int>> key(n);
  const double L[3] = {La, Lb, Lc};
  for (int i = 0; i < n; i++) {
    uint64_t k = 0;
    for (int d = 0; d < 3; d++) {
      double f = pos[3 * (size_t)i + d] / L[d];
      f -= std::floor(f);
      const uint64_t q = std::min<uint64_t>(1023, (uint64_t)(f * 1024.0));
      for (int b = 0; b < 10; b++) k |= ((q >> b) & 1) << (3 * b + d);
    }
    key[i] = {k, i};
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<uint64_t, int>& x, const std::pair<uint64_t, int>& y) {
                     return x.first < y.first;
                   });
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = key[i].second;
  return perm;
}

int tgx_md_end(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  MdState& m = X(c).md;
  md_release_graph(m);
  m.active = false;
  return TGX_OK;
}

int tgx_md_begin(tgx_ctx* c, int n, const int* species, const double* lattice,
                 const double* mass, const double* pos, const double* vel, double dt,
                 double skin, int history) {
  if (!c) return TGX_ERR_ARG;
  if (n < 2 || !species || !lattice || !mass || !pos || !vel || !(dt > 0) || skin < 0 ||
      history < 1)
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  if (!has_model(c)) return set_err(c, TGX_ERR_STATE, "no model loaded");
  (void)hipSetDevice(c->device);
  Extra& x = X(c);
  MdState& m = x.md;
  if (m.active) tgx_md_end(c);
  for (int i = 0; i < n; i++)
    if (!(mass[i] > 0)) return set_err(c, TGX_ERR_ARG, "mass[%d] must be positive", i);
  m.n = n;
  m.dt = dt;
  m.skin = skin;
  m.cap = history;
  m.rcut_list = rcut_all(c) + skin;
  std::string err;
  int rcode = make_cell(lattice, m.rcut_list, &m.cell, &err);
  if (rcode) return set_err(c, rcode, "%s", err.c_str());
  if ((long long)n * m.cell.nrep >= (1LL << TGX_IMG_SHIFT))
    return set_err(c, TGX_ERR_ARG, "too many atoms");
  // the run works on the atoms in a Z-order of their wrapped positions:
  // the SOAP kernels, the force gather and the list build then read the
  // data of nearby atoms from nearby memory; tgx_md_read restores the order
  m.perm = zorder_atoms(n, pos, m.cell.La, m.cell.Lb, m.cell.Lc);
  std::vector<double> pos_p(3 * (size_t)n), vel_p(3 * (size_t)n), mass_p(n);
  std::vector<int> sp_p(n);
  for (int k = 0; k < n; k++) {
    const int a = m.perm[k];
    for (int d = 0; d < 3; d++) {
      pos_p[3 * (size_t)k + d] = pos[3 * (size_t)a + d];
      vel_p[3 * (size_t)k + d] = vel[3 * (size_t)a + d];
    }
    mass_p[k] = mass[a];
    sp_p[k] = species[a];
  }
  pos = pos_p.data();
  vel = vel_p.data();
  mass = mass_p.data();
  species = sp_p.data();
  rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  // size the list from a first build, with head room for density changes
  c->maxnb = 0;
  rcode = build_list(c, n, m.cell, m.rcut_list, true);
  if (rcode) return rcode;
  {
    std::vector<int> cnt(n);
    TGX_HIP(c, hipMemcpy(cnt.data(), c->nb_count.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    const int mx = *std::max_element(cnt.begin(), cnt.end());
    const int want = ((mx * 3) / 2 + 16 + 7) / 8 * 8;
    if (c->maxnb < want) {
      c->maxnb = want;
      rcode = build_list(c, n, m.cell, m.rcut_list, true);
      if (rcode) return rcode;
    }
  }
  rcode = setup_sites(c, n, species, &m.n_sites);
  if (rcode) return rcode;
  ENSURE(c, md_vel, 24 * (size_t)n);
  ENSURE(c, md_fprev, 24 * (size_t)n);
  ENSURE(c, md_mass, 8 * (size_t)n);
  if (m.pref.ensure(24 * (size_t)n) || m.ke.ensure(8 * (size_t)n) || m.ctl.ensure(64) ||
      m.ep.ensure(8 * (size_t)history) || m.ek.ensure(8 * (size_t)history) ||
      m.vout.ensure(24 * (size_t)n) || m.kesum.ensure(16))
    return set_err(c, TGX_ERR_HIP, "device allocation failed (md)");
  std::vector<double> me(n);
  for (int i = 0; i < n; i++) me[i] = mass[i] * 103.6426965268;  // amu -> eV fs^2/A^2
  TGX_HIP(c, hipMemcpyAsync(c->md_vel.p, vel, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemcpyAsync(c->md_mass.p, me.data(), 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  TGX_HIP(c, hipMemsetAsync(c->md_fprev.p, 0, 24 * (size_t)n, c->stream));
  TGX_HIP(c, hipMemcpyAsync(m.pref.p, c->pos.p, 24 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
  TGX_HIP(c, hipMemsetAsync(m.ctl.p, 0, 64, c->stream));
  TGX_HIP(c, hipMemsetAsync(m.ep.p, 0, 8 * (size_t)history, c->stream));
  TGX_HIP(c, hipMemsetAsync(m.ek.p, 0, 8 * (size_t)history, c->stream));
  // forces and energy at the initial positions
  rcode = ef_launch(c, n, m.n_sites, Geom{m.cell}, 1, false);
  if (rcode) return rcode;
  TGX_HIP(c, hipMemcpyAsync(m.ep.p, x.out.as<double>() + 9, 8, hipMemcpyDeviceToDevice, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  m.active = true;
  const char* ng = getenv("TGX_NO_GRAPH");
  m.use_graph = !(ng && ng[0] == '1');
  if (m.use_graph) {
    TGX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rcode = md_step_launch(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rcode) {
      if (g) (void)hipGraphDestroy(g);
      return rcode;
    }
    if (e != hipSuccess) return set_err(c, TGX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    m.graph = g;
    e = hipGraphInstantiate(&m.exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      md_release_graph(m);
      return set_err(c, TGX_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
  }
  return TGX_OK;
}

int tgx_md_steps(tgx_ctx* c, int n_steps) {
  if (!c) return TGX_ERR_ARG;
  MdState& m = X(c).md;
  if (!m.active) return set_err(c, TGX_ERR_STATE, "tgx_md_begin has not been called");
  if (n_steps < 0) return set_err(c, TGX_ERR_ARG, "n_steps < 0");
  (void)hipSetDevice(c->device);
  for (int s = 0; s < n_steps; s++) {
    if (m.exec) {
      TGX_HIP(c, hipGraphLaunch(m.exec, c->stream));
    } else {
      int rcode = md_step_launch(c);
      if (rcode) return rcode;
    }
  }
  return TGX_OK;
}

int tgx_md_read(tgx_ctx* c, double* pos, double* vel, double* forces, double* e_pot,
                double* e_kin, int* n_steps_done, int* n_rebuilds) {
  if (!c) return TGX_ERR_ARG;
  Extra& x = X(c);
  MdState& m = x.md;
  if (!m.active) return set_err(c, TGX_ERR_STATE, "tgx_md_begin has not been called");
  (void)hipSetDevice(c->device);
  const int n = m.n;
  TGX_HIP(c, launch_md_sync_vel(n, c->md_vel.as<double>(), c->forces.as<double>(),
                                c->md_fprev.as<double>(), c->md_mass.as<double>(), m.dt,
                                m.ctl.as<int>() + 2, m.vout.as<double>(), m.ke.as<double>(),
                                m.kesum.as<double>(), c->stream));
  int ctl[4], flag[4];
  double kes = 0.0;
  TGX_HIP(c, hipMemcpyAsync(ctl, m.ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipMemcpyAsync(flag, c->nb_flag.p, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipMemcpyAsync(&kes, m.kesum.p, 8, hipMemcpyDeviceToHost, c->stream));
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  if (flag[0] > c->maxnb)
    return set_err(c, TGX_ERR_CAPACITY, "neighbour capacity %d exceeded during MD (%d needed)",
                   c->maxnb, flag[0]);
  if (flag[1]) return set_err(c, TGX_ERR_STATE, "neighbour list lost symmetry during MD");
  const int steps = ctl[2];
  if (n_steps_done) *n_steps_done = steps;
  if (n_rebuilds) *n_rebuilds = ctl[1];
  // per-atom arrays back in the caller's atom order
  std::vector<double> tmp(3 * (size_t)n);
  auto out3 = [&](double* dst, const void* src) -> hipError_t {
    const hipError_t e = hipMemcpy(tmp.data(), src, 24 * (size_t)n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (int k = 0; k < n; k++)
      for (int d = 0; d < 3; d++) dst[3 * (size_t)m.perm[k] + d] = tmp[3 * (size_t)k + d];
    return hipSuccess;
  };
  if (pos) TGX_HIP(c, out3(pos, c->pos.p));
  if (vel) TGX_HIP(c, out3(vel, m.vout.p));
  if (forces) TGX_HIP(c, out3(forces, c->forces.p));
  if (e_pot || e_kin) {
    const int cap = m.cap;
    std::vector<double> ep(cap), ek(cap);
    TGX_HIP(c, hipMemcpy(ep.data(), m.ep.p, 8 * (size_t)cap, hipMemcpyDeviceToHost));
    TGX_HIP(c, hipMemcpy(ek.data(), m.ek.p, 8 * (size_t)cap, hipMemcpyDeviceToHost));
    const int first = steps + 1 > cap ? steps + 1 - cap : 0;
    for (int s = first; s <= steps; s++) {
      if (e_pot) e_pot[s - first] = ep[s % cap];
      if (e_kin) e_kin[s - first] = s == steps ? kes : ek[s % cap];
    }
  }
  return TGX_OK;
}

int tgx_md_run(tgx_ctx* c, int n, int n_steps, double dt, double skin, const int* species,
               const double* lattice, const double* mass, double* pos, double* vel,
               double* forces, double* e_pot, double* e_kin, int* n_rebuilds) {
  if (!c) return TGX_ERR_ARG;
  if (n_steps < 0 || !pos || !vel) return set_err(c, TGX_ERR_ARG, "bad arguments");
  int rcode = tgx_md_begin(c, n, species, lattice, mass, pos, vel, dt, skin, n_steps + 1);
  if (rcode) return rcode;
  rcode = tgx_md_steps(c, n_steps);
  if (!rcode) rcode = tgx_md_read(c, pos, vel, forces, e_pot, e_kin, nullptr, n_rebuilds);
  tgx_md_end(c);
  return rcode;
}

int tgx_sync(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  (void)hipSetDevice(c->device);
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  return TGX_OK;
}

int tgx_poly3_orthonormalization(int alpha_max, double* S, double* W) {
  if (alpha_max < 1 || alpha_max > TGX_MAX_NR || !W) return TGX_ERR_ARG;
  std::vector<double> w((size_t)TGX_MAX_NR * TGX_MAX_NR);
  poly3_W(alpha_max, w.data());
  for (int i = 0; i < alpha_max; i++)
    for (int j = 0; j < alpha_max; j++) {
      W[j * alpha_max + i] = w[(size_t)i * TGX_MAX_NR + j];
      if (S)
        S[j * alpha_max + i] =
            std::sqrt((2.0 * i + 7.0) * (2.0 * j + 7.0)) / (double)(i + j + 7);
    }
  return TGX_OK;
}

int tgx_poly3gauss_orthonormalization(int alpha_max, double atom_sigma, double rcut_hard,
                                      double* S, double* W) {
  if (alpha_max < 1 || alpha_max > TGX_MAX_NR || !W || !(atom_sigma > 0.0) ||
      !(rcut_hard > 0.0) || !std::isfinite(atom_sigma) || !std::isfinite(rcut_hard))
    return TGX_ERR_ARG;
  std::vector<double> w((size_t)TGX_MAX_NR * TGX_MAX_NR), sm((size_t)alpha_max * alpha_max);
  poly3gauss_W(alpha_max, rcut_hard, atom_sigma, sm.data(), w.data());
  for (int i = 0; i < alpha_max; i++)
    for (int j = 0; j < alpha_max; j++) {
      W[j * alpha_max + i] = w[(size_t)i * TGX_MAX_NR + j];
      if (S) S[j * alpha_max + i] = sm[(size_t)i * alpha_max + j];
    }
  return TGX_OK;
}

int tgx_stage_times(const tgx_ctx* c, double* ms7) {
  if (!c || !ms7) return TGX_ERR_ARG;
  for (int k = 0; k < 7; k++) ms7[k] = c->stage_ms[k];
  return TGX_OK;
}

int tgx_set_list_buffer(tgx_ctx* c, double buffer) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (!(buffer >= 0.0) || buffer > 10.0) return set_err(c, TGX_ERR_ARG, "buffer out of range");
  c->list_buffer = buffer;
  c->maxnb = 0;  // re-sized at the next build
  return TGX_OK;
}

int tgx_add_local_property(tgx_ctx* c, int block, int dim, int n_sparse, const double* Qs,
                           const double* alphas, double delta, double zeta, double V0,
                           int* index_out) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (!c->has_soap || block < 0 || block >= (int)c->blocks.size())
    return set_err(c, TGX_ERR_ARG, "no SOAP block %d", block);
  Extra& x = X(c);
  std::unique_ptr<SoapBlock> b(new (std::nothrow) SoapBlock());
  if (!b) return set_err(c, TGX_ERR_ARG, "out of host memory");
  SoapBlock* const keep = c->blk;
  c->blk = b.get();
  const int rcode = load_soap_block(c, &c->blocks[block]->hyp, dim, n_sparse, Qs, alphas, delta,
                                    zeta);
  c->blk = keep;
  if (rcode) {
    b->release();
    return rcode;
  }
  x.lp.push_back(std::move(b));
  x.lp_block.push_back(block);
  x.lp_v0.push_back(V0);
  if (index_out) *index_out = (int)x.lp.size() - 1;
  return TGX_OK;
}

int tgx_clear_local_properties(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  Extra& x = X(c);
  (void)hipSetDevice(c->device);
  for (auto& b : x.lp) b->release();
  x.lp.clear();
  x.lp_block.clear();
  x.lp_v0.clear();
  x.ts_lp = -1;  // the vdW term read one of them
  return TGX_OK;
}

// The property's adjoint on block P (its kernel step already run with the
// gradient): dV_i/dr_j of every slot of every site row into c->gpair (zero
// beyond the descriptor cutoff), rows of atoms that are no site untouched.
static int lp_adjoint(tgx_ctx* c, SoapBlock* P, const Geom& cell, int n, int ns) {
  const int prec = c->precision;
  ENSURE(c, gpair, tsize(prec) * 4 * (size_t)n * c->maxnb);
  ENSURE(c, blk->vsite, 48 * (size_t)ns);
  AdjIn ai;
  const int arc = adjoint_inputs(c, ns, &ai);
  if (arc) return arc;
  hipError_t e;
  if (prec == TGX_FP64)
    e = launch_soap_bwd<double>(
        P->sd, P->sd_dev.as<SoapDev>(), P->W_dev.as<double>(), P->rtab.as<float>(),
        c->pos.as<double>(), c->species.as<int>(), cell, c->maxnb, c->nb_count.as<int>(),
        c->nb_list.as<int>(), P->site_atom.as<int>(), ns, P->pairn.as<int>(),
        static_cast<const double*>(ai.q), ai.q_ld, P->cbuf.as<double>(),
        static_cast<const double*>(ai.pn), static_cast<const double*>(ai.G), ai.g_ld,
        c->gpair.as<double>(), P->vsite.as<double>(), nullptr, 0, c->eatom.as<double>(),
        c->stream, 0);
  else
    e = launch_soap_bwd<float>(
        P->sd, P->sd_dev.as<SoapDev>(), P->W_dev.as<double>(), P->rtab.as<float>(),
        c->pos.as<double>(), c->species.as<int>(), cell, c->maxnb, c->nb_count.as<int>(),
        c->nb_list.as<int>(), P->site_atom.as<int>(), ns, P->pairn.as<int>(),
        static_cast<const float*>(ai.q), ai.q_ld, P->cbuf.as<float>(),
        static_cast<const float*>(ai.pn), static_cast<const float*>(ai.G), ai.g_ld,
        c->gpair.as<float>(), P->vsite.as<double>(), nullptr, 0, c->eatom.as<double>(),
        c->stream, 0);
  TGX_HIP(c, e);
  return TGX_OK;
}

// local_property_predict (gap/local_properties.f90:46-146) on the device:
// the block's descriptors, the property's kernel step V = delta^2 sum_s
// alphas_s K^zeta + V0 clipped at 0, and with V_der the adjoint's per-slot
// gradients dV_i/dr_j (the self entry minus their sum), in the order of the
// list tgx_neighbors_build returns at rcut_list.
static int local_property_impl(tgx_ctx* c, int index, int n, const double* pos,
                               const int* species, const double* lattice, double rcut_list,
                               double* V, int64_t cap, double* V_der, int64_t* n_pairs) {
  Extra& x = X(c);
  SoapBlock* P = x.lp[index].get();
  const double rc = P->sd.rcut_max;
  if (V_der && !(rcut_list >= rc))
    return set_err(c, TGX_ERR_ARG, "rcut_list %g below the descriptor cutoff %g", rcut_list, rc);
  int rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  Cell cell;
  std::string err;
  int n_ext = n;
  const bool general = !is_orthorhombic(lattice);
  if (general && V_der)
    return set_err(c, TGX_ERR_UNSUPPORTED, "V_der needs an orthorhombic cell");
  const double rl = V_der ? rcut_list : rc;
  if (general) {
    rcode = prepare_image_cell(c, n, lattice, rl, false, &cell, &n_ext);
  } else {
    rcode = make_cell(lattice, rl, &cell, &err);
    if (rcode) return set_err(c, rcode, "%s", err.c_str());
    rcode = build_list(c, n, cell, rl, false);
  }
  if (rcode) return rcode;
  rcode = setup_block_sites(c, P, n, species, n_ext);
  if (rcode) return rcode;
  c->blk = P;
  const int ns = P->n_sites;
  const int prec = c->precision;
  if (ns > 0) {
    rcode = soap_forward(c, n, Geom{cell}, ns, false, true);
    if (!rcode) rcode = kernel_step(c, ns, V_der != nullptr);
    if (rcode) return rcode;
    if (V_der) rcode = lp_adjoint(c, P, Geom{cell}, n, ns);
    if (rcode) return rcode;
  }
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> es(ns);
  std::vector<int> sa(ns);
  if (ns > 0) {
    TGX_HIP(c, hipMemcpy(es.data(), P->esite.p, 8 * (size_t)ns, hipMemcpyDeviceToHost));
    TGX_HIP(c, hipMemcpy(sa.data(), P->site_atom.p, 4 * (size_t)ns, hipMemcpyDeviceToHost));
  }
  std::vector<double> vv(n, 0.0);
  std::vector<char> live(n, 0);
  for (int s = 0; s < ns; s++) {
    const double v = es[s] + x.lp_v0[index];
    vv[sa[s]] = v < 0.0 ? 0.0 : v;
    live[sa[s]] = vv[sa[s]] != 0.0;
  }
  std::memcpy(V, vv.data(), 8 * (size_t)n);
  if (!V_der) return TGX_OK;
  std::vector<int> cnt(n);
  TGX_HIP(c, hipMemcpy(cnt.data(), c->nb_count.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
  std::vector<long long> off(n + 1, 0);
  for (int i = 0; i < n; i++) off[i + 1] = off[i] + std::min(cnt[i], c->maxnb) + 1;
  *n_pairs = off[n];
  if (off[n] > cap)
    return set_err(c, TGX_ERR_CAPACITY, "need %lld pairs, cap %lld", off[n], (long long)cap);
  const size_t slots = (size_t)n * c->maxnb;
  std::vector<double> g(4 * slots, 0.0);
  if (ns > 0) {
    if (prec == TGX_FP64) {
      TGX_HIP(c, hipMemcpy(g.data(), c->gpair.p, 32 * slots, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> gf(4 * slots);
      TGX_HIP(c, hipMemcpy(gf.data(), c->gpair.p, 16 * slots, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < gf.size(); k++) g[k] = gf[k];
    }
  }
  for (int i = 0; i < n; i++) {
    double* row = V_der + 3 * off[i];
    const int m = (int)(off[i + 1] - off[i]) - 1;
    for (int k = 0; k < 3 * (m + 1); k++) row[k] = 0.0;
    if (!live[i]) continue;
    for (int s = 0; s < m; s++)
      for (int k = 0; k < 3; k++) {
        const double v = g[4 * ((size_t)i * c->maxnb + s) + k];
        row[3 * (s + 1) + k] = v;
        row[k] -= v;
      }
  }
  return TGX_OK;
}

int tgx_local_property(tgx_ctx* c, int index, int n, const double* pos, const int* species,
                       const double* lattice, double rcut_list, double* V, int64_t cap,
                       double* V_der, int64_t* n_pairs) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  Extra& x = X(c);
  if (index < 0 || index >= (int)x.lp.size())
    return set_err(c, TGX_ERR_ARG, "no local property %d", index);
  if (n < 1 || !pos || !species || !lattice || !V || (V_der && !n_pairs))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  (void)hipSetDevice(c->device);
  const int saved = c->maxnb;
  c->maxnb = 0;
  const int rcode = local_property_impl(c, index, n, pos, species, lattice, rcut_list, V, cap,
                                        V_der, n_pairs);
  c->maxnb = saved;
  c->blk = c->blocks[0].get();
  return rcode;
}

// get_vdw_energy_and_forces / get_ts_energy_and_forces
// (vdw/vdw_interface.f90:40-385) on the device: one list at max(rcut, the
// property's descriptor cutoff), the property's descriptors, kernel step
// and (with forces) adjoint, then the volume, pair and Hirshfeld-gather
// kernels.  Energies per atom, forces and the virial in FP64.
static int vdw_ts_impl(tgx_ctx* c, int n, const double* pos, const int* species,
                       const double* lattice, int do_forces, double* energies, double* forces,
                       double* virial) {
  Extra& x = X(c);
  SoapBlock* P = x.lp[x.ts_lp].get();
  const double rl = std::max(x.ts.rcut, P->sd.rcut_max);
  int rcode = upload_structure(c, n, pos, species);
  if (rcode) return rcode;
  Cell cell;
  int n_ext = n;
  if (!is_orthorhombic(lattice)) {
    rcode = prepare_image_cell(c, n, lattice, rl, do_forces != 0, &cell, &n_ext);
  } else {
    std::string err;
    rcode = make_cell(lattice, rl, &cell, &err);
    if (rcode) return set_err(c, rcode, "%s", err.c_str());
    if ((long long)n * cell.nrep >= (1LL << TGX_IMG_SHIFT))
      return set_err(c, TGX_ERR_ARG, "too many atoms");
    if (cell.nrep >= (1 << (31 - TGX_IMG_SHIFT)))
      return set_err(c, TGX_ERR_UNSUPPORTED, "box too small for the cutoff (%d images)",
                     cell.nrep);
    rcode = build_list(c, n, cell, rl, do_forces != 0);
  }
  if (rcode) return rcode;
  rcode = setup_block_sites(c, P, n, species, n_ext);
  if (rcode) return rcode;
  c->blk = P;
  const int ns = P->n_sites;
  const Geom geo{cell};
  if (ns > 0) {
    rcode = soap_forward(c, n, geo, ns, false, true);
    if (!rcode) rcode = kernel_step(c, ns, do_forces != 0);
    if (!rcode && do_forces) rcode = lp_adjoint(c, P, geo, n, ns);
    if (rcode) return rcode;
  }
  if (grow(c, x.ts_V, 8 * (size_t)n) || grow(c, x.ts_pref, 8 * (size_t)n) ||
      grow(c, x.ts_vatom, 48 * (size_t)n))
    return grow_err(c, 64 * (size_t)n, "vdW buffers");
  ENSURE(c, eatom, 8 * (size_t)n);
  ENSURE(c, forces, 24 * (size_t)n);
  if (ns > 0)
    TGX_HIP(c, launch_ts_volume(n, P->site_of_atom.as<int>(), P->esite.as<double>(),
                                x.lp_v0[x.ts_lp], x.ts_V.as<double>(), c->stream));
  else
    TGX_HIP(c, hipMemsetAsync(x.ts_V.p, 0, 8 * (size_t)n, c->stream));
  TGX_HIP(c, launch_ts_pairs(x.ts, c->pos.as<double>(), c->species.as<int>(), n, geo, c->maxnb,
                             c->nb_count.as<int>(), c->nb_list.as<int>(), x.ts_V.as<double>(),
                             c->eatom.as<double>(), c->forces.as<double>(),
                             x.ts_vatom.as<double>(), x.ts_pref.as<double>(), do_forces,
                             c->stream));
  if (do_forces && ns > 0) {
    if (c->precision == TGX_FP64)
      TGX_HIP(c, launch_ts_gather<double>(x.ts, c->pos.as<double>(), n, geo, c->maxnb,
                                          c->nb_count.as<int>(), c->nb_list.as<int>(),
                                          c->nb_rev.as<int>(), c->gpair.as<double>(),
                                          x.ts_pref.as<double>(), c->forces.as<double>(),
                                          x.ts_vatom.as<double>(), c->stream));
    else
      TGX_HIP(c, launch_ts_gather<float>(x.ts, c->pos.as<double>(), n, geo, c->maxnb,
                                         c->nb_count.as<int>(), c->nb_list.as<int>(),
                                         c->nb_rev.as<int>(), c->gpair.as<float>(),
                                         x.ts_pref.as<double>(), c->forces.as<double>(),
                                         x.ts_vatom.as<double>(), c->stream));
  }
  TGX_HIP(c, hipStreamSynchronize(c->stream));
  TGX_HIP(c, hipMemcpy(energies, c->eatom.p, 8 * (size_t)n, hipMemcpyDeviceToHost));
  if (!do_forces) return TGX_OK;
  TGX_HIP(c, hipMemcpy(forces, c->forces.p, 24 * (size_t)n, hipMemcpyDeviceToHost));
  std::vector<double> va(6 * (size_t)n);
  TGX_HIP(c, hipMemcpy(va.data(), x.ts_vatom.p, 48 * (size_t)n, hipMemcpyDeviceToHost));
  double sv[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 6; k++) sv[k] += va[6 * (size_t)i + k];
  virial[0] = sv[0];
  virial[4] = sv[1];
  virial[8] = sv[2];
  virial[1] = virial[3] = sv[3];
  virial[2] = virial[6] = sv[4];
  virial[5] = virial[7] = sv[5];
  return TGX_OK;
}

int tgx_set_vdw_ts(tgx_ctx* c, int local_property, int n_species, const double* c6_ref,
                   const double* r0_ref, const double* alpha0_ref, double rcut, double buffer,
                   double rcut_inner, double buffer_inner, double sR, double d) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  Extra& x = X(c);
  if (local_property < 0 || local_property >= (int)x.lp.size())
    return set_err(c, TGX_ERR_ARG, "no local property %d", local_property);
  if (n_species < 1 || n_species > TGX_MAX_SPECIES || !c6_ref || !r0_ref || !alpha0_ref)
    return set_err(c, TGX_ERR_ARG, "bad species references");
  TsDev t{};
  for (int s = 0; s < n_species; s++) {
    if (!std::isfinite(c6_ref[s]) || !std::isfinite(r0_ref[s]) || !std::isfinite(alpha0_ref[s]))
      return set_err(c, TGX_ERR_ARG, "reference of species %d is not finite", s + 1);
    t.c6[s] = c6_ref[s];
    t.r0[s] = r0_ref[s];
    t.a0[s] = alpha0_ref[s];
  }
  if (!(rcut > 0.0) || !std::isfinite(rcut) || !(buffer >= 0.0) || !(buffer <= rcut) ||
      !(rcut_inner >= 0.0) || !(buffer_inner >= 0.0) || !std::isfinite(buffer_inner) ||
      !(rcut_inner < rcut) || !(sR > 0.0) || !std::isfinite(sR) || !std::isfinite(d))
    return set_err(c, TGX_ERR_ARG, "bad vdW cutoffs or damping parameters");
  t.n_species = n_species;
  t.rcut = rcut;
  t.buffer = buffer;
  t.rcut_inner = rcut_inner;
  t.buffer_inner = buffer_inner;
  t.sR = sR;
  t.d = d;
  x.ts = t;
  x.ts_lp = local_property;
  return TGX_OK;
}

int tgx_clear_vdw(tgx_ctx* c) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  X(c).ts_lp = -1;
  return TGX_OK;
}

int tgx_vdw_energy_forces(tgx_ctx* c, int n, const double* pos, const int* species,
                          const double* lattice, int do_forces, double* energies,
                          double* forces, double* virial) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  Extra& x = X(c);
  if (x.ts_lp < 0) return set_err(c, TGX_ERR_STATE, "no vdW model set");
  if (n < 1 || !pos || !species || !lattice || !energies || (do_forces && (!forces || !virial)))
    return set_err(c, TGX_ERR_ARG, "bad arguments");
  for (int i = 0; i < n; i++)
    if (species[i] < 1 || species[i] > x.ts.n_species)
      return set_err(c, TGX_ERR_ARG, "species[%d] = %d has no vdW reference", i, species[i]);
  (void)hipSetDevice(c->device);
  const int saved = c->maxnb;
  c->maxnb = 0;
  const int rcode = vdw_ts_impl(c, n, pos, species, lattice, do_forces, energies, forces, virial);
  c->maxnb = saved;
  c->blk = c->blocks[0].get();
  return rcode;
}

int tgx_set_mc_batch(tgx_ctx* c, int trials_per_batch) {
  if (!c) return TGX_ERR_ARG;
  if (md_busy(c)) return TGX_ERR_STATE;
  if (trials_per_batch < 0 || trials_per_batch > 64)
    return set_err(c, TGX_ERR_ARG, "trials per batch must be 0..64");
  X(c).mc_batch = trials_per_batch;
  return TGX_OK;
}

int tgx_set_timing(tgx_ctx* c, int on) {
  if (!c) return TGX_ERR_ARG;
  X(c).timing = on != 0;
  return TGX_OK;
}

int tgx_kernel_times(const tgx_ctx* c, double* ms4) {
  if (!c || !ms4) return TGX_ERR_ARG;
  for (int k = 0; k < 4; k++) ms4[k] = c->kernel_ms[k];
  return TGX_OK;
}

}  // extern "C"
