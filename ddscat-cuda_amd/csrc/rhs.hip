// Per-RHS operator and incident-field setup on the device, for batched
// solves: for each right-hand side r (incidence k_r, polarization e0_r)
//   b_r[c][q]    = e0_r[c] exp(i k_r . x_q)          (EVALE: incident field)
//   adia_r[c][q] = 1 / alpha_LDR(eps_c(q), kd, khat_r, e0_r)   (ALPHADIAG)
//   aoff_r       = 0                                  (isotropic / diagonal eps)
// on the compact site list, with x_q = ((ix+1) + x0) dx as in the far field
// (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:800-802).  In DDSCAT these
// are Fortran-side inputs of the ABI (EVALE, ALPHADIAG / LATTDR); here they
// are built where the batched solver needs them.  Phases are FP64.
//
// Lattice dispersion relation, DDSCAT's two LDR prescriptions (CALPHA,
// ddscat_cuda_unified/fortran/reapar_cuda.f90:497-513):
//   LATTDR (Draine & Goodman 1993, ApJ 405, 685), isotropic:
//     1/alpha = 1/alpha_CM + (b1 + b2 eps + b3 S eps) (kd)^2 - (2/3) i (kd)^3,
//     S = sum_j (khat_j ehat_j)^2;
//   GKDLDR (Gutkowicz-Krusin & Draine 2004, arXiv:astro-ph/0403082), a
//     diagonal tensor whose jj element takes S_j = khat_j^2 (direction of
//     propagation only) in place of S;
//   alpha_CM = (3/4pi)(eps - 1)/(eps + 2),
//   b1 = -1.8915316, b2 = 0.1648469, b3 = -1.7700004.
// For a real unit polarization ehat, sum_j ehat_j^2 (1/alpha_jj) of GKDLDR
// equals the LATTDR 1/alpha, since sum_j ehat_j^2 khat_j^2 = S.
#include "solver_common.hpp"
#include "../../include/ddscat_matvec_hip.h"
#include <cmath>
#include <vector>

namespace ddhip {
namespace {

constexpr double LDR_B1 = -1.8915316, LDR_B2 = 0.1648469, LDR_B3 = -1.7700004;

struct RhsParams {
  double k[3];      // k d
  double e0[6];     // complex e0 (re, im) x 3
  double S[3];      // S of component j: LATTDR sum (khat_i ehat_i)^2 for all j, GKDLDR khat_j^2
  double kd;
};

int g_prescription = DDSCAT_HIP_LATTDR;

__global__ __launch_bounds__(256) void k_setup_rhs(const RhsParams* __restrict__ prm, const int* __restrict__ occ,
                                                   int nx, int ny, int nat0, double x00, double x01, double x02,
                                                   double dx0, double dx1, double dx2,
                                                   const float2* __restrict__ eps, int eps_per_comp,
                                                   float2* __restrict__ b, float2* __restrict__ adia,
                                                   float2* __restrict__ aoff) {
  const int r = blockIdx.y;
  const RhsParams P = prm[r];
  const size_t n = (size_t)3 * nat0, off = (size_t)r * n;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nat0; q += gridDim.x * blockDim.x) {
    const int j = occ[q];
    const int ix = j % nx, iy = (j / nx) % ny, iz = j / (nx * ny);
    const double ph = P.k[0] * ((ix + 1 + x00) * dx0) + P.k[1] * ((iy + 1 + x01) * dx1) +
                      P.k[2] * ((iz + 1 + x02) * dx2);
    double sn, cs;
    sincos(ph, &sn, &cs);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double er = P.e0[2 * c], ei = P.e0[2 * c + 1];
      b[off + (size_t)c * nat0 + q] = make_float2((float)(er * cs - ei * sn), (float)(er * sn + ei * cs));
      const float2 ef = eps[(eps_per_comp ? (size_t)c * nat0 : 0) + q];
      const double2 e = make_double2(ef.x, ef.y);
      // 1/alpha_CM = (4 pi/3)(eps + 2)/(eps - 1)
      const double2 inv_cm = zscale(zdiv(make_double2(e.x + 2.0, e.y), make_double2(e.x - 1.0, e.y)),
                                    4.0 * M_PI / 3.0);
      const double kd2 = P.kd * P.kd;
      const double f = LDR_B2 + LDR_B3 * P.S[c];
      const double2 corr = make_double2((LDR_B1 + f * e.x) * kd2, f * e.y * kd2 - (2.0 / 3.0) * kd2 * P.kd);
      const double2 ai = zadd(inv_cm, corr);
      adia[off + (size_t)c * nat0 + q] = make_float2((float)ai.x, (float)ai.y);
      aoff[off + (size_t)c * nat0 + q] = make_float2(0.f, 0.f);
    }
  }
}

}  // namespace
}  // namespace ddhip

using namespace ddhip;

extern "C" DDSCAT_HIP_API int ddscat_hip_set_polarizability(int prescription) {
  if (prescription != DDSCAT_HIP_LATTDR && prescription != DDSCAT_HIP_GKDLDR) {
    set_error("set_polarizability", "prescription must be DDSCAT_HIP_LATTDR or DDSCAT_HIP_GKDLDR");
    return 1;
  }
  g_prescription = prescription;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_setup_rhs(int nrhs, const float* ak, const float* e0, const float* x0,
                                                   const float* dx, const void* eps, int eps_per_comp, void* b,
                                                   void* adia, void* aoff) {
  Context& c = ctx();
  clear_error();
  if (!c.ready || nrhs <= 0 || !ak || !e0 || !x0 || !dx || !eps || !b || !adia || !aoff ||
      (eps_per_comp != 0 && eps_per_comp != 1)) {
    set_error("setup_rhs", "invalid argument");
    return 1;
  }
  std::vector<RhsParams> h(nrhs);
  for (int r = 0; r < nrhs; ++r) {
    RhsParams& P = h[r];
    double kk = 0.0;
    for (int i = 0; i < 3; ++i) {
      P.k[i] = ak[3 * r + i];
      kk += P.k[i] * P.k[i];
      P.e0[2 * i] = e0[6 * r + 2 * i];
      P.e0[2 * i + 1] = e0[6 * r + 2 * i + 1];
    }
    P.kd = std::sqrt(kk);
    if (!(P.kd > 0.0)) { set_error("setup_rhs", "zero wave vector"); return 1; }
    double ea[3], en = 0.0;
    for (int i = 0; i < 3; ++i) {
      ea[i] = std::hypot(P.e0[2 * i], P.e0[2 * i + 1]);
      en += ea[i] * ea[i];
    }
    en = std::sqrt(en);
    if (!(en > 0.0)) { set_error("setup_rhs", "zero polarization vector"); return 1; }
    double S = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double t = P.k[i] / P.kd * ea[i] / en;
      S += t * t;
    }
    for (int i = 0; i < 3; ++i) {
      const double a = P.k[i] / P.kd;
      P.S[i] = g_prescription == DDSCAT_HIP_GKDLDR ? a * a : S;
    }
  }
  hipStream_t st = c.stream;
  RhsParams* dp = nullptr;
  DDH_TRY(hipMalloc((void**)&dp, nrhs * sizeof(RhsParams)), "setup_rhs", 29);
  if (hipMemcpyAsync(dp, h.data(), nrhs * sizeof(RhsParams), hipMemcpyHostToDevice, st) != hipSuccess) {
    hipFree(dp);
    set_error("setup_rhs", "parameter upload failed");
    return 2;
  }
  const unsigned nb = (unsigned)std::min(1024, (c.nat0 + 255) / 256);
  hipLaunchKernelGGL(k_setup_rhs, dim3(nb, nrhs), dim3(256), 0, st, dp, c.occ, c.nx, c.ny, c.nat0, (double)x0[0],
                     (double)x0[1], (double)x0[2], (double)dx[0], (double)dx[1], (double)dx[2], (const float2*)eps,
                     eps_per_comp, (float2*)b, (float2*)adia, (float2*)aoff);
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(st);
  hipFree(dp);
  if (le != hipSuccess || se != hipSuccess) {
    set_hip_error("setup_rhs", le != hipSuccess ? le : se);
    return 3;
  }
  return 0;
}
