// SOAP launchers for batched structures (a periodic cell per structure).
#include "tgx_soap_impl.h"

namespace tgx {

template <typename T>
hipError_t launch_soap_fwd_many(const SoapDev& sdh, const SoapDev* sd, const double* W,
                                const float* rtab, const double* pos, const int* species,
                                const CellMany& g, int maxnb, const int* count, const int* list,
                                const int* site_atom, int n_sites, const int* dmap, T* q,
                                int q_ld, u16* qb, int qb_ld, T* cbuf, T* pnorm, hipStream_t st,
                                int wpb) {
  return launch_soap_fwd_g<T, CellMany>(sdh, sd, W, rtab, pos, species, g, maxnb, count, list,
                                        site_atom, n_sites, dmap, q, q_ld, qb, qb_ld, cbuf,
                                        pnorm, st, 0, wpb);
}

template <typename T>
hipError_t launch_soap_bwd_many(const SoapDev& sdh, const SoapDev* sd, const double* W,
                                const float* rtab, const double* pos, const int* species,
                                const CellMany& g, int maxnb, const int* count, const int* list,
                                const int* site_atom, int n_sites, const int* dmap, const T* q,
                                int q_ld, const T* cbuf, const T* pnorm, const T* G, int g_ld,
                                T* gpair, double* vsite, const TwoBodyDev* twob, int n2b,
                                double* eatom, hipStream_t st, int wpb) {
  return launch_soap_bwd_g<T, CellMany>(sdh, sd, W, rtab, pos, species, g, maxnb, count, list,
                                        site_atom, n_sites, dmap, q, q_ld, cbuf, pnorm, G, g_ld,
                                        gpair, vsite, twob, n2b, eatom, st, wpb);
}

template hipError_t launch_soap_fwd_many<float>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const CellMany&, int, const int*, const int*, const int*, int, const int*, float*, int, u16*, int, float*, float*, hipStream_t, int);
template hipError_t launch_soap_fwd_many<double>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const CellMany&, int, const int*, const int*, const int*, int, const int*, double*, int, u16*, int, double*, double*, hipStream_t, int);
template hipError_t launch_soap_bwd_many<float>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const CellMany&, int, const int*, const int*, const int*, int, const int*, const float*, int, const float*, const float*, const float*, int, float*, double*, const TwoBodyDev*, int, double*, hipStream_t, int);
template hipError_t launch_soap_bwd_many<double>(const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*, const CellMany&, int, const int*, const int*, const int*, int, const int*, const double*, int, const double*, const double*, const double*, int, double*, double*, const TwoBodyDev*, int, double*, hipStream_t, int);

}  // namespace tgx
