// One instantiation of the SOAP kernel launchers: the Makefile compiles this
// file once per (direction, precision, pair geometry), chosen by
//   TGX_INST_BWD  0 forward, 1 adjoint
//   TGX_INST_F64  0 float,   1 double
//   TGX_INST_MANY 0 one periodic cell, 1 a cell per structure
// so that the eight instantiation sets build in parallel.
#include "tgx_soap_impl.h"

#if !defined(TGX_INST_BWD) || !defined(TGX_INST_F64) || !defined(TGX_INST_MANY)
#error "TGX_INST_BWD, TGX_INST_F64 and TGX_INST_MANY must be defined"
#endif

namespace tgx {

#if TGX_INST_F64
typedef double InstT;
#else
typedef float InstT;
#endif
#if TGX_INST_MANY
typedef CellMany InstGeo;
#else
typedef CellOne InstGeo;
#endif

#if TGX_INST_BWD
template hipError_t launch_soap_bwd_g<InstT, InstGeo>(
    const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*,
    const InstGeo&, int, const int*, const int*, const int*, int, const int*, const InstT*, int,
    const InstT*, const InstT*, const InstT*, int, InstT*, double*, const TwoBodyDev*, int,
    double*, hipStream_t, int);
#else
template hipError_t launch_soap_fwd_g<InstT, InstGeo>(
    const SoapDev&, const SoapDev*, const double*, const float*, const double*, const int*,
    const InstGeo&, int, const int*, const int*, const int*, int, const int*, InstT*, int, u16*,
    int, InstT*, InstT*, hipStream_t, int, int);
#endif

}  // namespace tgx
