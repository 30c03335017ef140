// Temporary entry points for parts of the export set not yet built here.
#include "solver_common.hpp"
#include "../../include/ddscat_matvec_hip.h"

using namespace ddhip;

#define DDH_STUB_SOLVER(name, code)                                                                   \
  extern "C" DDSCAT_HIP_API int name(const void*, void*, const void*, const void*, int, int, float, int*, \
                                     float*, int*) {                                                   \
    set_error(#name, "not available in this build");                                                  \
    return code;                                                                                       \
  }

DDH_STUB_SOLVER(ddscat_cuda_solve_gpbicg_f32, 30)
DDH_STUB_SOLVER(ddscat_cuda_solve_qmrccg_f32, 50)
DDH_STUB_SOLVER(ddscat_cuda_solve_pbcgst_f32, 70)
DDH_STUB_SOLVER(ddscat_cuda_solve_pbcgs2_f32, 90)
DDH_STUB_SOLVER(ddscat_cuda_solve_petrkp_f32, 110)
DDH_STUB_SOLVER(ddscat_cuda_solve_bicgstab4_f32, 130)
DDH_STUB_SOLVER(ddscat_cuda_solve_gpbicgstab2_f32, 150)
DDH_STUB_SOLVER(ddscat_cuda_solve_gpbicgstab4_f32, 180)

extern "C" DDSCAT_HIP_API int ddscat_cuda_evalq_f32(int, const void*, const void*, const void*, const void*, float,
                                                    float, float, float, double*, double*, double*) {
  set_error("GPU EVALQ", "not available in this build");
  return 300;
}

extern "C" DDSCAT_HIP_API int ddscat_cuda_scat_f32(int, int, const float*, const float*, const float*, const float*,
                                                   const float*, float, float, const void*, const float*, double*,
                                                   double*, double*, double*, int*, void*, void*) {
  set_error("GPU SCAT", "not available in this build");
  return 320;
}
