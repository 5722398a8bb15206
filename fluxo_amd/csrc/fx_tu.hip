// One compilation unit of the engine's per-degree kernels: degree FX_TU_N (2..5) and kernel
// group FX_TU_GROUP, so the template kernels of different degrees and groups compile in
// parallel (fluxo_amd/build.py builds this file once per pair and links the objects with
// fx_engine.hip into libfluxo_gpu.so).
//   group 0  split-form volume kernel (every VolumeFlux, mode, precision and tile)
//   group 1  surface kernel (every Riemann solver), SurfInt alone, NFVSE shock capturing,
//            the mortar interpolation and projection, the face-point table
//   group 2  lifting, CALCTIMESTEP, CalcEntropy, analysis
//   group 3  GLM-MHD volume, surface and CALCTIMESTEP kernels
#include <hip/hip_runtime.h>

#include "fx_launch_n.hpp"

#ifndef FX_TU_N
#error "FX_TU_N (degree) not defined"
#endif
#ifndef FX_TU_GROUP
#error "FX_TU_GROUP not defined"
#endif

namespace fx {

#if FX_TU_GROUP == 0
template void launch_volume_n<FX_TU_N>(const Batch &, int, RKArgs);
#elif FX_TU_GROUP == 1
template void launch_surface_n<FX_TU_N>(const Batch &);
template void launch_face_table_n<FX_TU_N>(const Batch &, int, int);
template void launch_nfvse_alpha_n<FX_TU_N>(const Batch &);
template void launch_nfvse_fv_n<FX_TU_N>(const Batch &, const double *);
template void launch_surfint_n<FX_TU_N>(const Batch &, const double *, const double *);
template void launch_mortar_u_n<FX_TU_N>(const Batch &);
template void launch_flux_mortar_n<FX_TU_N>(const Batch &);
#elif FX_TU_GROUP == 2
template void launch_lifting_n<FX_TU_N>(const Batch &, int, const DtFuse *);
template void launch_dt_n<FX_TU_N>(const Batch &, double, double);
template void launch_entropy_n<FX_TU_N>(const Batch &, double *, double *);
template void launch_analyze_n<FX_TU_N>(const Batch &, const AnaArgs &, const double *, const double *, double *,
                                        double *);
#elif FX_TU_GROUP == 3
template void launch_mhd_volume_n<FX_TU_N>(const Batch &, int, RKArgs);
template void launch_mhd_surface_n<FX_TU_N>(const Batch &);
template void launch_mhd_dt_n<FX_TU_N>(const Batch &, double, double);
#else
#error "unknown FX_TU_GROUP"
#endif

namespace {
// this unit's copy of the operators (its kernels read c_ops[FX_TU_N])
int upload_ops(const Ops &o, int N)
{
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_ops), &o, sizeof(Ops), sizeof(Ops) * N, hipMemcpyHostToDevice));
  return FX_OK;
}
const int registered = register_ops_uploader(upload_ops);
}  // namespace

}  // namespace fx
