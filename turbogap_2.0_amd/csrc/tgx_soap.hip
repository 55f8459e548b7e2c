// SOAP launchers: the entry that picks the pair geometry (one periodic cell,
// or a cell per structure for batched structures).  The kernels and their
// launchers are instantiated in tgx_soap_inst.hip (one object per direction,
// precision and geometry).
#include "tgx_device.h"
#include "tgx_internal.h"

namespace tgx {

typedef unsigned short u16;

template <typename T, typename GEO>
hipError_t launch_soap_fwd_g(const SoapDev& sdh, const SoapDev* sd, const double* W,
                             const float* rtab, const double* pos, const int* species,
                             const GEO& cell, int maxnb, const int* count, const int* list,
                             const int* site_atom, int n_sites, const int* dmap, T* q, int q_ld,
                             u16* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st, int split,
                             int wpb);
template <typename T, typename GEO>
hipError_t launch_soap_bwd_g(const SoapDev& sdh, const SoapDev* sd, const double* W,
                             const float* rtab, const double* pos, const int* species,
                             const GEO& cell, int maxnb, const int* count, const int* list,
                             const int* site_atom, int n_sites, const int* dmap, const T* q,
                             int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                             T* gpair, double* vsite, const TwoBodyDev* twob, int n2b,
                             double* eatom, hipStream_t st, int wpb);

template <typename T>
hipError_t launch_soap_fwd(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, T* q, int q_ld,
                           u16* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st, int split,
                           int wpb) {
  if (g.many)
    return launch_soap_fwd_g<T, CellMany>(sdh, sd, W, rtab, pos, species,
                                          CellMany{g.many, g.sid}, maxnb, count, list, site_atom,
                                          n_sites, dmap, q, q_ld, qb, qb_ld, cbuf, pnorm, st, 0,
                                          wpb);
  return launch_soap_fwd_g<T, CellOne>(sdh, sd, W, rtab, pos, species, CellOne{g.one}, maxnb,
                                       count, list, site_atom, n_sites, dmap, q, q_ld, qb, qb_ld,
                                       cbuf, pnorm, st, split, wpb);
}

template <typename T>
hipError_t launch_soap_bwd(const SoapDev& sdh, const SoapDev* sd, const double* W,
                           const float* rtab, const double* pos, const int* species,
                           const Geom& g, int maxnb, const int* count, const int* list,
                           const int* site_atom, int n_sites, const int* dmap, const T* q,
                           int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                           T* gpair, double* vsite, const TwoBodyDev* twob, int n2b,
                           double* eatom, hipStream_t st, int wpb) {
  if (g.many)
    return launch_soap_bwd_g<T, CellMany>(sdh, sd, W, rtab, pos, species,
                                          CellMany{g.many, g.sid}, maxnb, count, list, site_atom,
                                          n_sites, dmap, q, q_ld, cbuf, pnorm, G, g_ld, gpair,
                                          vsite, twob, n2b, eatom, st, wpb);
  return launch_soap_bwd_g<T, CellOne>(sdh, sd, W, rtab, pos, species, CellOne{g.one}, maxnb,
                                       count, list, site_atom, n_sites, dmap, q, q_ld, cbuf,
                                       pnorm, G, g_ld, gpair, vsite, twob, n2b, eatom, st, wpb);
}

template hipError_t launch_soap_fwd<float>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Geom&, int, const int*, const int*, const int*, int, const int*, float*, int, u16*, int, float*, float*, hipStream_t, int, int);
template hipError_t launch_soap_fwd<double>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Geom&, int, const int*, const int*, const int*, int, const int*, double*, int, u16*, int, double*, double*, hipStream_t, int, int);
template hipError_t launch_soap_bwd<float>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Geom&, int, const int*, const int*, const int*, int, const int*, const float*, int, const float*, const float*, const float*, int, float*, double*, const TwoBodyDev*, int, double*, hipStream_t, int);
template hipError_t launch_soap_bwd<double>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const Geom&, int, const int*, const int*, const int*, int, const int*, const double*, int, const double*, const double*, const double*, int, double*, double*, const TwoBodyDev*, int, double*, hipStream_t, int);

}  // namespace tgx
