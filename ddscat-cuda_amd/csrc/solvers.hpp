// Shared pieces of the solver entry points.
#pragma once
#include "solver_common.hpp"

namespace ddhip {

// reliable-residual policy (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:221-222)
constexpr int RELIABLE_PERIOD = 20;
constexpr double RELIABLE_GAP = 1.e-3;

// device scalar block layout (doubles)
enum {
  SL_STATUS = 0,  // first breakdown code raised by a scalar kernel (0 = none)
  SL_RESID = 1,   // relative residual computed on the device
  SL_FLAGS = 2,   // PBCGS2 reliable-update flags
  SL_D = 8,       // dot-product results, complex i at SL_D + 2 i (up to 24 doubles)
  SL_X = 64,      // solver complex scalars
  SL_R = 160,     // solver real scalars
};

struct SolverArgs {
  const void* b;
  void* x;
  const void* adia;
  const void* aoff;
  int ndim, maxit;
  float tol;
  int* itno;
  float* tole;
  int* ncompte;
  // 0: reference boundary (full [3][NAT] host vectors, ndim = 3 NAT);
  // 1: compact [3][NAT0] device vectors on the library stream, ndim = 3 NAT0
  int device = 0;
};

struct Reliable {
  double true_res = INFINITY, gap = INFINITY;
  bool converged = false, restart = false;
  bool replace = false;  // QMRCCG under RELIABLE_REPLACE: take the true residual, keep the recurrence
};

// argument checks, boundary copies (full host -> compact device), clears the
// scalar block; initial guess gathered into xs
int solver_begin(Context& c, const SolverArgs& a, int code_arg, int code_io, const char* name, float2* xs);
// publishes xs as the resident solution, copies it to the host full vector
int solver_end(Context& c, const SolverArgs& a, const float2* xs, int it, double resid, int nc, int code);
// replaceable: the solver's recursive residual only monitors convergence
// (QMRCCG), so under c.reliable_mode == RELIABLE_REPLACE a failed check asks
// for the true residual to replace it (rr.replace) instead of a restart
int reliable_check(Context& c, const float2* rhs, const float2* xcur, const float2* rrec, double rhsnorm,
                   double rec_res, double tol, int& nc, Reliable& rr, bool replaceable = false);

// Single-RHS QMRCCG vector steps (cu:437, cu:438, cu:443) on the fused
// kernels of the batched solver with one RHS; coefficients from c.scal.
//   pq:   p = v~/gamma - mu p ; q = (A w~ - gamma mu q)/ksi
//   vw:   v~ = Ap - (tau/gamma) v~ ; w~ = q - (tau/ksi) w~
//   dsxr: d = theta d + kappa p ; s = theta s + kappa Ap ; x += d ; r -= s,
//         with |r|^2 into c.scal[dst] (imaginary part 0 at dst + 1)
int qmr1_pq(Context& c, float2* p, float2* q, const float2* vt, const float2* atw, size_t n);
int qmr1_vw(Context& c, float2* vt, float2* wt, const float2* ap, const float2* q, size_t n);
int qmr1_dsxr(Context& c, float2* d, float2* sv, float2* x, float2* r, const float2* p, const float2* ap, size_t n,
              int dst);
// Single-RHS PBCGS2 vector steps (cu:466-473) on the batched solver's fused
// kernels, coefficients from c.scal:
//   xr (K = 1, 2): x += alpha u0 ; r_j -= alpha u_{j+1}, j < K
//   u2:            u_j = r_j - beta u_j, j = 0, 1
//   comb:          u0 -= y1 u1 + y2 u2 ; x += y1 r0 + y2 r1 ; r0 -= y1 r1 + y2 r2
int p21_xr(Context& c, int K, float2* x, float2* r0, float2* r1, const float2* u0, const float2* u1, const float2* u2,
           size_t n);
int p21_u2(Context& c, float2* u0, float2* u1, const float2* r0, const float2* r1, size_t n);
int p21_comb(Context& c, float2* u0, float2* x, float2* r0, const float2* u1, const float2* u2, const float2* r1,
             const float2* r2, size_t n);

int solve_pbcgs2(Context& c, const SolverArgs& a);
int solve_qmrccg(Context& c, const SolverArgs& a);
int solve_petrkp(Context& c, const SolverArgs& a);
int solve_gpbicg(Context& c, const SolverArgs& a);
int solve_pbcgst(Context& c, const SolverArgs& a);
int solve_bicgstab4(Context& c, const SolverArgs& a);
int solve_gpbicgstab2(Context& c, const SolverArgs& a);
int solve_gpbicgstab4(Context& c, const SolverArgs& a);

}  // namespace ddhip
