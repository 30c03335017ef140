// Batched Krylov solves: many right-hand sides of one prepared target (for
// instance incident orientations x polarizations, each with its own LDR
// operator) advanced together.  Every kernel of an iteration covers all RHS
// (grid.y / grid.z = RHS), so one iteration is a fixed sequence of ~16
// launches whatever the batch size: it is captured once in a HIP graph and
// replayed.  A device flag per RHS (active[r]) masks finished RHS: their
// workgroups exit on entry, their vectors stay frozen.
//
// The per-RHS arithmetic is the single-RHS solver's (solvers.hip, shared FP64
// recurrences in krylov_scalars.hpp): QMRCCG
// (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:434-443, 1071-1103) and
// PBCGS2 (cu:459-480, 1168-1278), with the reliable-residual policy of
// cu:543-556 applied per RHS.  BLAS-1 updates of one step are fused into one
// kernel; the per-step dots and the scalar recurrence share the reduction's
// finishing kernel.
#include "solver_common.hpp"
#include "dots.hpp"
#include "solvers.hpp"
#include "krylov_scalars.hpp"
#include "../../include/ddscat_matvec_hip.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>
#include <thread>

namespace ddhip {
namespace {

constexpr int BT = 256;             // threads per workgroup
constexpr int BMAX_BLOCKS = 2048;   // reduction partial blocks per RHS
constexpr int SLOT_BNORM = SL_R + 12;  // |b| of the RHS
constexpr int SLOT_R2 = SL_D + 20;     // |r|^2 dot slot of the fused update kernels

struct BatchState {
  int B = 0, nvec = 0;
  size_t n = 0;
  size_t cap_vec = 0, cap_a1 = 0, cap_a2 = 0;
  float2* W = nullptr;
  float2* a1 = nullptr;
  float2* a2 = nullptr;
  double* scal = nullptr;   // [B][SCAL_COUNT]
  double* part = nullptr;   // [B][2 DOT_MAX][BMAX_BLOCKS]
  int* active = nullptr;    // [B]
  int* sel = nullptr;       // [B]
  int* m1 = nullptr;        // [B] device-computed masks (PBCGS2 reliable update)
  int* m2 = nullptr;
  double* packed = nullptr; // [B][3] status, resid, flags
  double* hpacked = nullptr;
  int* hflags = nullptr;    // pinned [B]
  // iteration bookkeeping on the device (kb_step_end): [4][B] ints =
  // iterations, check request, active flags saved at a pause, reliable-update
  // MATVECs; then the pause flag.  hist: recursive residual of iteration i
  // at row (i - 1) mod HIST_ROWS.  hctl / hhist: pinned copies.
  int* ctl = nullptr;
  double* hist = nullptr;
  int* hctl = nullptr;
  double* hhist = nullptr;
  int capB = 0;
};
constexpr int HIST_ROWS = 32;  // > RELIABLE_PERIOD: a chunk never wraps
BatchState g_bs[MAX_TARGETS + 1];  // one per target slot
BatchState& bstate(const Context& c) { return g_bs[c.slot]; }

void bs_free(BatchState& S) {
  for (void** p : {(void**)&S.W, (void**)&S.a1, (void**)&S.a2, (void**)&S.scal, (void**)&S.part,
                   (void**)&S.active, (void**)&S.sel, (void**)&S.m1, (void**)&S.m2, (void**)&S.packed,
                   (void**)&S.ctl, (void**)&S.hist}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
  for (void** p : {(void**)&S.hpacked, (void**)&S.hflags, (void**)&S.hctl, (void**)&S.hhist}) {
    if (*p) hipHostFree(*p);
    *p = nullptr;
  }
  S.cap_vec = S.cap_a1 = S.cap_a2 = 0;
  S.capB = 0;
}

int bs_ensure(Context& c, BatchState& S, int B, int nvec) {
  const size_t n = (size_t)3 * c.nat0;
  const size_t need_vec = (size_t)nvec * B * n;
  // FFT work per RHS on FFT3D; SLICES runs the RHS through the context's slice buffers
  const bool fft3d = c.backend == BACKEND_FFT3D;
  const size_t need_a1 = fft3d ? (size_t)B * matvec_a1_count(c) : 0;
  const size_t need_a2 = fft3d ? (size_t)B * matvec_a2_count(c) : 0;
  if (need_vec > S.cap_vec) {
    if (S.W) hipFree(S.W);
    S.W = nullptr;
    S.cap_vec = 0;
    DDH_TRY(hipMalloc((void**)&S.W, need_vec * sizeof(float2)), "batched Krylov vectors", 29);
    S.cap_vec = need_vec;
  }
  if (need_a1 > S.cap_a1) {
    if (S.a1) hipFree(S.a1);
    S.a1 = nullptr;
    S.cap_a1 = 0;
    DDH_TRY(hipMalloc((void**)&S.a1, need_a1 * sizeof(float2)), "batched FFT work a1", 29);
    S.cap_a1 = need_a1;
  }
  if (need_a2 > S.cap_a2) {
    if (S.a2) hipFree(S.a2);
    S.a2 = nullptr;
    S.cap_a2 = 0;
    DDH_TRY(hipMalloc((void**)&S.a2, need_a2 * sizeof(float2)), "batched FFT work a2", 29);
    S.cap_a2 = need_a2;
  }
  if (B > S.capB) {
    for (void** p : {(void**)&S.scal, (void**)&S.part, (void**)&S.active, (void**)&S.sel, (void**)&S.m1,
                     (void**)&S.m2, (void**)&S.packed, (void**)&S.ctl, (void**)&S.hist}) {
      if (*p) hipFree(*p);
      *p = nullptr;
    }
    for (void** p : {(void**)&S.hpacked, (void**)&S.hflags, (void**)&S.hctl, (void**)&S.hhist}) {
      if (*p) hipHostFree(*p);
      *p = nullptr;
    }
    S.capB = 0;
    DDH_TRY(hipMalloc((void**)&S.scal, (size_t)B * SCAL_COUNT * sizeof(double)), "batched scalars", 29);
    DDH_TRY(hipMalloc((void**)&S.part, (size_t)B * 2 * DOT_MAX * BMAX_BLOCKS * sizeof(double)), "batched partials",
            29);
    DDH_TRY(hipMalloc((void**)&S.active, (size_t)B * sizeof(int)), "batched flags", 29);
    DDH_TRY(hipMalloc((void**)&S.sel, (size_t)B * sizeof(int)), "batched flags", 29);
    DDH_TRY(hipMalloc((void**)&S.m1, (size_t)B * sizeof(int)), "batched flags", 29);
    DDH_TRY(hipMalloc((void**)&S.m2, (size_t)B * sizeof(int)), "batched flags", 29);
    DDH_TRY(hipMalloc((void**)&S.packed, (size_t)B * 3 * sizeof(double)), "batched status", 29);
    DDH_TRY(hipHostMalloc((void**)&S.hpacked, (size_t)B * 3 * sizeof(double)), "pinned status", 29);
    DDH_TRY(hipHostMalloc((void**)&S.hflags, (size_t)B * sizeof(int)), "pinned flags", 29);
    DDH_TRY(hipMalloc((void**)&S.ctl, ((size_t)4 * B + 1) * sizeof(int)), "batched iteration state", 29);
    DDH_TRY(hipMalloc((void**)&S.hist, (size_t)HIST_ROWS * B * sizeof(double)), "batched residual history", 29);
    DDH_TRY(hipHostMalloc((void**)&S.hctl, ((size_t)4 * B + 1) * sizeof(int)), "pinned iteration state", 29);
    DDH_TRY(hipHostMalloc((void**)&S.hhist, (size_t)HIST_ROWS * B * sizeof(double)), "pinned residual history", 29);
    S.capB = B;
  }
  S.B = B;
  S.n = n;
  S.nvec = nvec;
  return 0;
}

// ------------------------------------------------------------ reductions --
template <typename Acc>
__device__ __forceinline__ Acc wsum(Acc v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Writes this workgroup's partial sums re[t], im[t] (t < nt) of RHS r.
template <typename Acc, int NT>
__device__ __forceinline__ void put_partials(const Acc* re, const Acc* im, int nt, double* __restrict__ part, int r,
                                             int nbk) {
  __shared__ Acc sred[2 * NT][BT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < nt) {
      const Acc a = wsum(re[t]);
      const Acc b = wsum(im[t]);
      if (lane == 0) { sred[2 * t][wave] = a; sred[2 * t + 1][wave] = b; }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * nt) {
    Acc s = 0;
    for (int w = 0; w < BT / 64; ++w) s += sred[threadIdx.x][w];
    part[((size_t)r * 2 * DOT_MAX + threadIdx.x) * nbk + blockIdx.x] = (double)s;
  }
}

template <typename Acc, int NT, int NV>
__global__ __launch_bounds__(BT) void kb_dots(DotOperands L, size_t n, const int* __restrict__ active,
                                              double* __restrict__ part, int nbk) {
  const int r = blockIdx.y;
  if (active && !active[r]) return;
  Acc re[NT], im[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) { re[t] = 0; im[t] = 0; }
  dot_accumulate<Acc, NT, NV>(L, (size_t)r * n, (size_t)blockIdx.x * BT + threadIdx.x, (size_t)nbk * BT, n, re, im);
  put_partials<Acc, NT>(re, im, L.nt, part, r, nbk);
}

// what the finishing lane does with the reduced dots
enum FinOp {
  FIN_NONE = 0,
  FIN_BNORM,        // s[BNORM] = sqrt(dot0); zero RHS -> status 52
  FIN_RESID,        // s[SL_RESID] = sqrt(dot at SLOT_R2) / s[BNORM]
  FIN_QMR_INIT,     // qmr::init_sc
  FIN_QMR_RECUR,    // qmr::recur
  FIN_P2_INIT,      // p2::init_reals + p2::init
  FIN_P2_BETA1,     // p2::outer, then p2::beta(has_norm = 0) (first half-step)
  FIN_P2_BETA2,     // p2::beta(has_norm = 1)
  FIN_P2_ALPHA,     // p2::alpha
  FIN_P2_CONVEX,    // p2::convex
};

// One workgroup per RHS, one wave per sum: lane l loads partials l, l + 64,
// ... < nbk (all in flight), adds them in order, then a fixed butterfly;
// the first lane then stores the sums and applies `op`.
template <typename Acc>
__global__ __launch_bounds__(BT) void kb_finish(const double* __restrict__ part, int nbk, int nsums,
                                                double* __restrict__ scal, int dst, const int* __restrict__ active,
                                                int op) {
  const int r = blockIdx.x;
  if (active && !active[r]) return;
  double* s = scal + (size_t)r * SCAL_COUNT;
  __shared__ double sums[2 * DOT_MAX];
  constexpr int PER_LANE = BMAX_BLOCKS / 64;
  const int lane = threadIdx.x & 63;
  for (int k = threadIdx.x >> 6; k < nsums; k += BT / 64) {
    const double* pp = part + ((size_t)r * 2 * DOT_MAX + k) * nbk;
    double p[PER_LANE];
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) p[j] = pp[min(lane + 64 * j, nbk - 1)];
    Acc v = 0;
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) v += lane + 64 * j < nbk ? (Acc)p[j] : (Acc)0;
    v = wsum(v);
    if (lane == 0) sums[k] = (double)v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 0; k < nsums; ++k) s[dst + k] = sums[k];
  switch (op) {
    case FIN_BNORM: {
      const double b = sqrt(fmax(s[SL_D], 0.0));
      s[SLOT_BNORM] = b;
      if (b == 0.0) fail(s, 52);
      break;
    }
    case FIN_RESID: {
      const double b = s[SLOT_BNORM];
      s[SL_RESID] = b > 0.0 ? sqrt(fmax(s[SLOT_R2], 0.0)) / b : 0.0;
      break;
    }
    case FIN_QMR_INIT: qmr::init_sc(s); break;
    case FIN_QMR_RECUR: qmr::recur(s); break;
    case FIN_P2_INIT: p2::init_reals(s); p2::init(s); break;
    case FIN_P2_BETA1: p2::outer(s); p2::beta(s, 0); break;
    case FIN_P2_BETA2: p2::beta(s, 1); break;
    case FIN_P2_ALPHA: p2::alpha(s); break;
    case FIN_P2_CONVEX: p2::convex(s); break;
    default: break;
  }
}

// ---------------------------------------------------- vector arithmetic --
template <typename R>
struct Cx {
  R x, y;
};
template <typename R>
__device__ __forceinline__ Cx<R> coef(const double* s, int i, R sign = (R)1) {
  return Cx<R>{sign * (R)s[i], sign * (R)s[i + 1]};
}
template <typename R>
__device__ __forceinline__ Cx<R> cv(float2 v) {
  return Cx<R>{(R)v.x, (R)v.y};
}
template <typename R>
__device__ __forceinline__ Cx<R> cm(Cx<R> a, Cx<R> b) {
  return Cx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ Cx<R> ca(Cx<R> a, Cx<R> b) {
  return Cx<R>{a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ Cx<R> cs(Cx<R> a, Cx<R> b) {
  return Cx<R>{a.x - b.x, a.y - b.y};
}
template <typename R>
__device__ __forceinline__ float2 f2(Cx<R> a) {
  return make_float2((float)a.x, (float)a.y);
}

#define BATCH_PROLOGUE                                    \
  const int r = blockIdx.y;                               \
  if (active && !active[r]) return;                       \
  const double* s = scal + (size_t)r * SCAL_COUNT;        \
  const size_t off = (size_t)r * n;                       \
  const size_t gstride = (size_t)gridDim.x * blockDim.x;  \
  (void)s;

// generic batched linear combination d = sum_t coef_t v_t (LinComb terms)
struct BLin {
  const float2* v[LC_MAX];
  double lre[LC_MAX], lim[LC_MAX];
  int sidx[LC_MAX], neg[LC_MAX];
  int n;
};
template <typename R>
__global__ __launch_bounds__(BT) void kb_lin(float2* __restrict__ d, BLin L, const double* __restrict__ scal,
                                             size_t n, const int* __restrict__ active) {
  BATCH_PROLOGUE
  Cx<R> cf[LC_MAX];
#pragma unroll
  for (int t = 0; t < LC_MAX; ++t) {
    if (t < L.n) {
      if (L.sidx[t] >= 0) cf[t] = coef<R>(s, L.sidx[t], L.neg[t] ? (R)-1 : (R)1);
      else cf[t] = Cx<R>{(R)L.lre[t], (R)L.lim[t]};
    } else {
      cf[t] = Cx<R>{0, 0};
    }
  }
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    Cx<R> acc{0, 0};
#pragma unroll
    for (int t = 0; t < LC_MAX; ++t)
      if (t < L.n) acc = ca(acc, cm(cf[t], cv<R>(L.v[t][off + i])));
    d[off + i] = f2(acc);
  }
}

// QMRCCG step A (cu:437): p = v~/gamma - mu p ; q = (A w~ - gamma mu q) / ksi
template <typename R>
__global__ __launch_bounds__(BT) void kb_qmr_pq(float2* __restrict__ p, float2* __restrict__ q,
                                                const float2* __restrict__ vt, const float2* __restrict__ atw,
                                                const double* __restrict__ scal, size_t n,
                                                const int* __restrict__ active) {
  BATCH_PROLOGUE
  const Cx<R> ginv = coef<R>(s, qmr::GINV), mu = coef<R>(s, qmr::MU), gmu = coef<R>(s, qmr::GMU),
              kinv = coef<R>(s, qmr::KINV);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    const size_t k = off + i;
    p[k] = f2(cs(cm(ginv, cv<R>(vt[k])), cm(mu, cv<R>(p[k]))));
    q[k] = f2(cm(kinv, cs(cv<R>(atw[k]), cm(gmu, cv<R>(q[k])))));
  }
}

// QMRCCG step B (cu:438): v~ = Ap - (tau/gamma) v~ ; w~ = q - (tau/ksi) w~
template <typename R>
__global__ __launch_bounds__(BT) void kb_qmr_vw(float2* __restrict__ vt, float2* __restrict__ wt,
                                                const float2* __restrict__ ap, const float2* __restrict__ q,
                                                const double* __restrict__ scal, size_t n,
                                                const int* __restrict__ active) {
  BATCH_PROLOGUE
  const Cx<R> tg = coef<R>(s, qmr::TG), tk = coef<R>(s, qmr::TK);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    const size_t k = off + i;
    vt[k] = f2(cs(cv<R>(ap[k]), cm(tg, cv<R>(vt[k]))));
    wt[k] = f2(cs(cv<R>(q[k]), cm(tk, cv<R>(wt[k]))));
  }
}

// QMRCCG step C (cu:443): d = theta d + kappa p ; s = theta s + kappa Ap ;
// x += d ; r -= s ; partial |r|^2
template <typename R>
__global__ __launch_bounds__(BT) void kb_qmr_dsxr(float2* __restrict__ d, float2* __restrict__ sv,
                                                  float2* __restrict__ x, float2* __restrict__ rv,
                                                  const float2* __restrict__ p, const float2* __restrict__ ap,
                                                  const double* __restrict__ scal, size_t n,
                                                  const int* __restrict__ active, double* __restrict__ part,
                                                  int nbk) {
  BATCH_PROLOGUE
  const Cx<R> th = coef<R>(s, qmr::THETA), ka = coef<R>(s, qmr::KAPPA);
  R re[1] = {0}, im[1] = {0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    const size_t k = off + i;
    const float2 dn = f2(ca(cm(th, cv<R>(d[k])), cm(ka, cv<R>(p[k]))));
    const float2 sn = f2(ca(cm(th, cv<R>(sv[k])), cm(ka, cv<R>(ap[k]))));
    d[k] = dn;
    sv[k] = sn;
    const float2 xo = x[k], ro = rv[k];
    x[k] = make_float2(xo.x + dn.x, xo.y + dn.y);
    const float2 rn = make_float2(ro.x - sn.x, ro.y - sn.y);
    rv[k] = rn;
    re[0] += (R)rn.x * (R)rn.x + (R)rn.y * (R)rn.y;
  }
  put_partials<R, 1>(re, im, 1, part, r, nbk);
}

// PBCGS2 first half-step update (cu:466-469): x += alpha u0 ; r_j -= alpha u_{j+1}, j < K
template <typename R, int K>
__global__ __launch_bounds__(BT) void kb_p2_xr(float2* __restrict__ x, float2* __restrict__ r0, float2* __restrict__ r1,
                                               const float2* __restrict__ u0, const float2* __restrict__ u1,
                                               const float2* __restrict__ u2, const double* __restrict__ scal,
                                               size_t n, const int* __restrict__ active) {
  BATCH_PROLOGUE
  const Cx<R> al = coef<R>(s, p2::ALPHA);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    const size_t k = off + i;
    x[k] = f2(ca(cv<R>(x[k]), cm(al, cv<R>(u0[k]))));
    r0[k] = f2(cs(cv<R>(r0[k]), cm(al, cv<R>(u1[k]))));
    if (K == 2) r1[k] = f2(cs(cv<R>(r1[k]), cm(al, cv<R>(u2[k]))));
  }
}

// PBCGS2 second half-step direction update: u_j = r_j - beta u_j, j = 0, 1
template <typename R>
__global__ __launch_bounds__(BT) void kb_p2_u2(float2* __restrict__ u0, float2* __restrict__ u1,
                                               const float2* __restrict__ r0, const float2* __restrict__ r1,
                                               const double* __restrict__ scal, size_t n,
                                               const int* __restrict__ active) {
  BATCH_PROLOGUE
  const Cx<R> be = coef<R>(s, p2::BETA);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    const size_t k = off + i;
    u0[k] = f2(cs(cv<R>(r0[k]), cm(be, cv<R>(u0[k]))));
    u1[k] = f2(cs(cv<R>(r1[k]), cm(be, cv<R>(u1[k]))));
  }
}

// PBCGS2 minimal-residual combination (cu:472-473):
// u0 -= y1 u1 + y2 u2 ; x += y1 r0 + y2 r1 ; r0 -= y1 r1 + y2 r2
template <typename R>
__global__ __launch_bounds__(BT) void kb_p2_comb(float2* __restrict__ u0, float2* __restrict__ x,
                                                 float2* __restrict__ r0, const float2* __restrict__ u1,
                                                 const float2* __restrict__ u2, const float2* __restrict__ r1,
                                                 const float2* __restrict__ r2, const double* __restrict__ scal,
                                                 size_t n, const int* __restrict__ active) {
  BATCH_PROLOGUE
  const Cx<R> y1 = coef<R>(s, p2::Y1), y2 = coef<R>(s, p2::Y2);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gstride) {
    const size_t k = off + i;
    const Cx<R> vr0 = cv<R>(r0[k]), vr1 = cv<R>(r1[k]);
    u0[k] = f2(cs(cs(cv<R>(u0[k]), cm(y1, cv<R>(u1[k]))), cm(y2, cv<R>(u2[k]))));
    x[k] = f2(ca(ca(cv<R>(x[k]), cm(y1, vr0)), cm(y2, vr1)));
    r0[k] = f2(cs(cs(vr0, cm(y1, vr1)), cm(y2, cv<R>(r2[k]))));
  }
}

// PBCGS2 reliable-update masks from the convex step's flags (cu:474):
// m1 = active and rcmp, m2 = active and xpdt
__global__ void kb_p2_masks(const double* __restrict__ scal, const int* __restrict__ active, int* __restrict__ m1,
                            int* __restrict__ m2, int B) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const double* s = scal + (size_t)r * SCAL_COUNT;
  const int on = active[r] && s[SL_STATUS] == 0.0;
  const int f = (int)s[SL_FLAGS];
  m1[r] = on && (f & 1);
  m2[r] = on && (f & 2);
}

// PBCGS2 reliable-update bookkeeping for the RHS in m1 (xpdt from m2)
__global__ void kb_p2_reliable(double* __restrict__ scal, const int* __restrict__ m1, const int* __restrict__ m2,
                               int B) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B || !m1[r]) return;
  p2::reliable_reals(scal + (size_t)r * SCAL_COUNT, m2[r]);
}

// PBCGS2 restart bookkeeping for the RHS in sel: |r| = sqrt(dot slot 0)
__global__ void kb_p2_restart(double* __restrict__ scal, const int* __restrict__ sel, int B) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B || !sel[r]) return;
  double* s = scal + (size_t)r * SCAL_COUNT;
  p2::restart_reals(s, sqrt(fmax(s[SL_D], 0.0)));
}

// EVALQ local term per RHS (cu:770-783): q = (adia + i (2/3) k^3) P + offdiag(aoff) P
__global__ __launch_bounds__(BT) void kb_evalq_q(const float2* __restrict__ p, float2* __restrict__ q,
                                                 const float2* __restrict__ ad, const float2* __restrict__ ao,
                                                 size_t os, const float* __restrict__ rabs, int nat0) {
  const int r = blockIdx.y;
  const size_t n = (size_t)3 * nat0, off = (size_t)r * n, oo = (size_t)r * os;
  const float2 ri = make_float2(0.f, rabs[r]);
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < (size_t)nat0; j += (size_t)gridDim.x * blockDim.x) {
    const float2 p1 = p[off + j], p2 = p[off + j + nat0], p3 = p[off + j + 2 * nat0];
    const float2 d1 = ad[oo + j], d2 = ad[oo + j + nat0], d3 = ad[oo + j + 2 * nat0];
    const float2 a23 = ao[oo + j], a31 = ao[oo + j + nat0], a12 = ao[oo + j + 2 * nat0];
    q[off + j] = cadd(cadd(cmul(cadd(d1, ri), p1), cmul(a31, p3)), cmul(a12, p2));
    q[off + j + nat0] = cadd(cadd(cmul(cadd(d2, ri), p2), cmul(a12, p1)), cmul(a23, p3));
    q[off + j + 2 * nat0] = cadd(cadd(cmul(cadd(d3, ri), p3), cmul(a23, p2)), cmul(a31, p1));
  }
}

// Adjoint of Cabs = P^H W P, W = -(4 pi k / E0^2) Im(M) per site, M the EVALQ
// site matrix (diag adia + i (2/3) k^3, symmetric off-diagonal aoff).  A is
// complex symmetric, so the adjoint A^H lam = W P is solved as A mu = W conj(P)
// with mu = conj(lam): the same MATVEC and solver as the forward problem.
// adjoint right-hand side W conj(P); cfac[r] = 4 pi k / E0^2, kk3[r] = (2/3) k^3
__global__ __launch_bounds__(BT) void kb_adj_rhs(const float2* __restrict__ p, const float2* __restrict__ ad,
                                                 const float2* __restrict__ ao, size_t os,
                                                 const float* __restrict__ cfac, const float* __restrict__ kk3,
                                                 float2* __restrict__ out, int nat0) {
  const int r = blockIdx.y;
  const size_t n = (size_t)3 * nat0, off = (size_t)r * n, oo = (size_t)r * os;
  const float cf = -cfac[r], k3 = kk3[r];
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < (size_t)nat0; j += (size_t)gridDim.x * blockDim.x) {
    const float2 p1 = p[off + j], p2 = p[off + j + nat0], p3 = p[off + j + 2 * nat0];
    const float2 c1 = make_float2(p1.x, -p1.y), c2 = make_float2(p2.x, -p2.y), c3 = make_float2(p3.x, -p3.y);
    const float w1 = ad[oo + j].y + k3, w2 = ad[oo + j + nat0].y + k3, w3 = ad[oo + j + 2 * nat0].y + k3;
    const float o23 = ao[oo + j].y, o31 = ao[oo + j + nat0].y, o12 = ao[oo + j + 2 * nat0].y;
    out[off + j] = cscale(cadd(cadd(cscale(c1, w1), cscale(c3, o31)), cscale(c2, o12)), cf);
    out[off + j + nat0] = cscale(cadd(cadd(cscale(c2, w2), cscale(c1, o12)), cscale(c3, o23)), cf);
    out[off + j + 2 * nat0] = cscale(cadd(cadd(cscale(c3, w3), cscale(c2, o23)), cscale(c1, o31)), cf);
  }
}

// dCabs = Re(g_adia . d adia) + Re(g_aoff . d aoff):
//   g_adia_c = i cfac |P_c|^2 - 2 mu_c P_c
//   g_a23 = i cfac 2 Re(conj(P2) P3) - 2 (mu2 P3 + mu3 P2), a31 (1, 3), a12 (1, 2) alike
__global__ __launch_bounds__(BT) void kb_adj_grad(const float2* __restrict__ p, const float2* __restrict__ mu,
                                                  const float* __restrict__ cfac, float2* __restrict__ gad,
                                                  float2* __restrict__ gao, int nat0) {
  const int r = blockIdx.y;
  const size_t n = (size_t)3 * nat0, off = (size_t)r * n;
  const float cf = cfac[r];
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < (size_t)nat0; j += (size_t)gridDim.x * blockDim.x) {
    const float2 P[3] = {p[off + j], p[off + j + nat0], p[off + j + 2 * nat0]};
    const float2 U[3] = {mu[off + j], mu[off + j + nat0], mu[off + j + 2 * nat0]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float2 up = cmul(U[c], P[c]);
      gad[off + j + c * (size_t)nat0] = make_float2(-2.f * up.x, cf * (P[c].x * P[c].x + P[c].y * P[c].y) - 2.f * up.y);
    }
    if (gao) {
      // slots a23 (2,3), a31 (3,1), a12 (1,2)
      const int pa[3] = {1, 2, 0}, pb[3] = {2, 0, 1};
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const float2 A = P[pa[s]], Bv = P[pb[s]];
        const float re2 = 2.f * (A.x * Bv.x + A.y * Bv.y);
        const float2 m = cadd(cmul(U[pa[s]], Bv), cmul(U[pb[s]], A));
        gao[off + j + s * (size_t)nat0] = make_float2(-2.f * m.x, cf * re2 - 2.f * m.y);
      }
    }
  }
}

// packed (status, relative residual, flags) per RHS for the host read-back
__global__ void kb_pack(const double* __restrict__ scal, double* __restrict__ out, int B) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  out[3 * r] = scal[(size_t)r * SCAL_COUNT + SL_STATUS];
  out[3 * r + 1] = scal[(size_t)r * SCAL_COUNT + SL_RESID];
  out[3 * r + 2] = scal[(size_t)r * SCAL_COUNT + SL_FLAGS];
}

// End of a batched iteration, on the device, so that several iterations run
// between host synchronisations: for every RHS that was active, count the
// iteration (and, P2, a reliable-update MATVEC when flag bit 0 is set), keep
// its recursive residual in the history, retire it on a breakdown status,
// and request a host check when the reliable-residual period is reached,
// the residual is at or below tol, or maxit is reached.  A request pauses
// the whole batch: the active flags are saved and cleared, so the rest of
// the launched iterations exit on entry until the host has run the checks.
// One workgroup.
template <bool P2>
__global__ void kb_step_end(const double* __restrict__ packed, int* __restrict__ active, int* __restrict__ ctl,
                            double* __restrict__ hist, int B, double tol, int maxit, int period) {
  __shared__ int pause;
  if (threadIdx.x == 0) pause = 0;
  __syncthreads();
  int* it = ctl;
  int* sel = ctl + B;
  int* sav = ctl + 2 * B;
  int* nxt = ctl + 3 * B;
  for (int r = threadIdx.x; r < B; r += blockDim.x) {
    if (!active[r]) continue;
    const double st = packed[3 * r], rs = packed[3 * r + 1];
    const int k = ++it[r];
    if (P2) nxt[r] += ((int)packed[3 * r + 2]) & 1;
    hist[(size_t)((k - 1) % HIST_ROWS) * B + r] = rs;
    if (st != 0.0) {
      active[r] = 0;
      continue;
    }
    if (k % period == 0 || rs <= tol || k >= maxit) {
      sel[r] = 1;
      pause = 1;
    }
  }
  __syncthreads();
  if (pause) {
    for (int r = threadIdx.x; r < B; r += blockDim.x) {
      sav[r] = active[r];
      active[r] = 0;
    }
    if (threadIdx.x == 0) ctl[4 * B] = 1;
  }
}

}  // namespace

// ------------------------------------- single-RHS QMRCCG and PBCGS2 steps --
// One RHS: grid.y = 1, no activity mask, scalars at c.scal.  The |r|^2
// partials of dsxr go to c.red with the stride finalize_partials reads.
static_assert(BMAX_BLOCKS == RED_MAX_BLOCKS, "single-RHS partials share the reduction buffer layout");
static dim3 single_grid(size_t n) {
  return dim3((unsigned)std::min<size_t>(RED_MAX_BLOCKS, std::max<size_t>(1, (n + (size_t)BT * 4 - 1) / ((size_t)BT * 4))));
}

int qmr1_pq(Context& c, float2* p, float2* q, const float2* vt, const float2* atw, size_t n) {
  if (c.precision != PREC_FP32)
    hipLaunchKernelGGL(kb_qmr_pq<double>, single_grid(n), dim3(BT), 0, c.stream, p, q, vt, atw, c.scal, n, nullptr);
  else
    hipLaunchKernelGGL(kb_qmr_pq<float>, single_grid(n), dim3(BT), 0, c.stream, p, q, vt, atw, c.scal, n, nullptr);
  DDH_LAUNCH_OK("QMRCCG p, q update", 1);
  return 0;
}

int qmr1_vw(Context& c, float2* vt, float2* wt, const float2* ap, const float2* q, size_t n) {
  if (c.precision != PREC_FP32)
    hipLaunchKernelGGL(kb_qmr_vw<double>, single_grid(n), dim3(BT), 0, c.stream, vt, wt, ap, q, c.scal, n, nullptr);
  else
    hipLaunchKernelGGL(kb_qmr_vw<float>, single_grid(n), dim3(BT), 0, c.stream, vt, wt, ap, q, c.scal, n, nullptr);
  DDH_LAUNCH_OK("QMRCCG v~, w~ update", 1);
  return 0;
}

int qmr1_dsxr(Context& c, float2* d, float2* sv, float2* x, float2* r, const float2* p, const float2* ap, size_t n,
              int dst) {
  const dim3 g = single_grid(n);
  if (c.precision != PREC_FP32)
    hipLaunchKernelGGL(kb_qmr_dsxr<double>, g, dim3(BT), 0, c.stream, d, sv, x, r, p, ap, c.scal, n, nullptr, c.red,
                       RED_MAX_BLOCKS);
  else
    hipLaunchKernelGGL(kb_qmr_dsxr<float>, g, dim3(BT), 0, c.stream, d, sv, x, r, p, ap, c.scal, n, nullptr, c.red,
                       RED_MAX_BLOCKS);
  DDH_LAUNCH_OK("QMRCCG d, s, x, r update", 1);
  return finalize_partials(c, (int)g.x, 2, dst, c.stream);
}

int p21_xr(Context& c, int K, float2* x, float2* r0, float2* r1, const float2* u0, const float2* u1, const float2* u2,
           size_t n) {
  const dim3 g = single_grid(n);
  const bool f64 = c.precision != PREC_FP32;
  if (K == 1 && f64) hipLaunchKernelGGL((kb_p2_xr<double, 1>), g, dim3(BT), 0, c.stream, x, r0, r1, u0, u1, u2, c.scal, n, nullptr);
  else if (K == 1) hipLaunchKernelGGL((kb_p2_xr<float, 1>), g, dim3(BT), 0, c.stream, x, r0, r1, u0, u1, u2, c.scal, n, nullptr);
  else if (f64) hipLaunchKernelGGL((kb_p2_xr<double, 2>), g, dim3(BT), 0, c.stream, x, r0, r1, u0, u1, u2, c.scal, n, nullptr);
  else hipLaunchKernelGGL((kb_p2_xr<float, 2>), g, dim3(BT), 0, c.stream, x, r0, r1, u0, u1, u2, c.scal, n, nullptr);
  DDH_LAUNCH_OK("PBCGS2 x, r update", 1);
  return 0;
}

int p21_u2(Context& c, float2* u0, float2* u1, const float2* r0, const float2* r1, size_t n) {
  if (c.precision != PREC_FP32)
    hipLaunchKernelGGL(kb_p2_u2<double>, single_grid(n), dim3(BT), 0, c.stream, u0, u1, r0, r1, c.scal, n, nullptr);
  else
    hipLaunchKernelGGL(kb_p2_u2<float>, single_grid(n), dim3(BT), 0, c.stream, u0, u1, r0, r1, c.scal, n, nullptr);
  DDH_LAUNCH_OK("PBCGS2 u update", 1);
  return 0;
}

int p21_comb(Context& c, float2* u0, float2* x, float2* r0, const float2* u1, const float2* u2, const float2* r1,
             const float2* r2, size_t n) {
  if (c.precision != PREC_FP32)
    hipLaunchKernelGGL(kb_p2_comb<double>, single_grid(n), dim3(BT), 0, c.stream, u0, x, r0, u1, u2, r1, r2, c.scal,
                       n, nullptr);
  else
    hipLaunchKernelGGL(kb_p2_comb<float>, single_grid(n), dim3(BT), 0, c.stream, u0, x, r0, u1, u2, r1, r2, c.scal,
                       n, nullptr);
  DDH_LAUNCH_OK("PBCGS2 convex combination", 1);
  return 0;
}

int batch_matvec(Context& c, int cwhat, const float2* x, float2* y, const float2* adia, const float2* aoff,
                 const MvBatch& bt, hipStream_t st) {
  if (c.backend != BACKEND_SLICES) return matvec_launch(c, cwhat, x, y, adia, aoff, 1, st, nullptr, &bt);
  for (int r = 0; r < bt.nrhs; ++r) {
    const int* act = bt.k.active ? bt.k.active + r : nullptr;
    if (int rc = slices_matvec(c, cwhat, x + r * bt.k.xs, y + r * bt.k.xs, adia + r * bt.k.os, aoff + r * bt.k.os, 1,
                               st, act))
      return rc;
  }
  return 0;
}

namespace {

// ------------------------------------------------------------- host ops --
struct Ops {
  Context& c;
  BatchState& S;
  hipStream_t st;
  bool f64;
  const float2* adia;
  const float2* aoff;
  size_t os;
  int nbk_red = 1, nbk_vec = 1;

  Ops(Context& cc, BatchState& ss, const float2* ad, const float2* ao, size_t ostride)
      : c(cc), S(ss), st(cc.stream), f64(cc.precision != PREC_FP32), adia(ad), aoff(ao), os(ostride) {
    nbk_red = (int)std::min<size_t>(BMAX_BLOCKS, std::max<size_t>(1, (S.n + (size_t)BT * 4 - 1) / ((size_t)BT * 4)));
    nbk_vec = (int)std::min<size_t>(2048, std::max<size_t>(1, (S.n + (size_t)BT * 4 - 1) / ((size_t)BT * 4)));
  }
  float2* V(int k) const { return S.W + (size_t)k * S.B * S.n; }
  dim3 vgrid() const { return dim3((unsigned)nbk_vec, (unsigned)S.B); }

  int mv(const float2* x, float2* y, const int* mask) {
    MvBatch bt;
    bt.nrhs = S.B;
    bt.k.xs = S.n;
    bt.k.os = os;
    bt.k.a1s = matvec_a1_count(c);
    bt.k.a2s = matvec_a2_count(c);
    bt.k.active = mask;
    bt.a1 = S.a1;
    bt.a2 = S.a2;
    return batch_matvec(c, 'N', x, y, adia, aoff, bt, st);
  }
  template <typename Acc>
  struct DotsRun {
    const DotOperands& L;
    const Ops& o;
    const int* mask;
    template <int NT, int NV>
    void run() const {
      hipLaunchKernelGGL((kb_dots<Acc, NT, NV>), dim3((unsigned)o.nbk_red, (unsigned)o.S.B), dim3(BT), 0, o.st, L,
                         o.S.n, mask, o.S.part, o.nbk_red);
    }
  };
  int dots(const DotSet& d, int dst, const int* mask, int op) {
    DotOperands L;
    const int nv = dot_operands(d, 0, d.n, L);
    if (nv < 0) {
      set_error("batched dots", "too many distinct operand vectors");
      return 1;
    }
    if (f64) {
      dot_dispatch(d.n, nv, DotsRun<double>{L, *this, mask});
      hipLaunchKernelGGL(kb_finish<double>, dim3(S.B), dim3(BT), 0, st, S.part, nbk_red, 2 * d.n, S.scal, dst,
                         mask, op);
    } else {
      dot_dispatch(d.n, nv, DotsRun<float>{L, *this, mask});
      hipLaunchKernelGGL(kb_finish<float>, dim3(S.B), dim3(BT), 0, st, S.part, nbk_red, 2 * d.n, S.scal, dst,
                         mask, op);
    }
    DDH_LAUNCH_OK("batched dots", 1);
    return 0;
  }
  int lin(float2* d, const LinComb& lc, const int* mask) {
    BLin L{};
    L.n = lc.n;
    for (int t = 0; t < lc.n; ++t) {
      L.v[t] = lc.t[t].v;
      L.sidx[t] = lc.t[t].sidx;
      L.lre[t] = lc.t[t].re;
      L.lim[t] = lc.t[t].im;
      L.neg[t] = lc.t[t].negate;
    }
    for (int t = lc.n; t < LC_MAX; ++t) { L.v[t] = d; L.sidx[t] = -1; }
    if (f64) hipLaunchKernelGGL(kb_lin<double>, vgrid(), dim3(BT), 0, st, d, L, S.scal, S.n, mask);
    else hipLaunchKernelGGL(kb_lin<float>, vgrid(), dim3(BT), 0, st, d, L, S.scal, S.n, mask);
    DDH_LAUNCH_OK("batched linear combination", 1);
    return 0;
  }
  int fin(int dst, int nsums, const int* mask, int op) {
    if (f64)
      hipLaunchKernelGGL(kb_finish<double>, dim3(S.B), dim3(BT), 0, st, S.part, nbk_vec, nsums, S.scal, dst, mask,
                         op);
    else
      hipLaunchKernelGGL(kb_finish<float>, dim3(S.B), dim3(BT), 0, st, S.part, nbk_vec, nsums, S.scal, dst, mask,
                         op);
    DDH_LAUNCH_OK("batched finish", 1);
    return 0;
  }
  int pack() {
    hipLaunchKernelGGL(kb_pack, dim3((S.B + 63) / 64), dim3(64), 0, st, S.scal, S.packed, S.B);
    DDH_TRY(hipMemcpyAsync(S.hpacked, S.packed, (size_t)S.B * 3 * sizeof(double), hipMemcpyDeviceToHost, st),
            "batched status D2H", 1);
    return 0;
  }
  // device iteration bookkeeping after the pack of an iteration
  int step_end(bool p2, double tol, int maxit) {
    if (p2) hipLaunchKernelGGL(kb_step_end<true>, dim3(1), dim3(256), 0, st, S.packed, S.active, S.ctl, S.hist, S.B,
                               tol, maxit, RELIABLE_PERIOD);
    else hipLaunchKernelGGL(kb_step_end<false>, dim3(1), dim3(256), 0, st, S.packed, S.active, S.ctl, S.hist, S.B,
                            tol, maxit, RELIABLE_PERIOD);
    DDH_LAUNCH_OK("batched step end", 1);
    return 0;
  }
  // check requests and the pause flag cleared before a chunk of iterations
  int chunk_reset() {
    DDH_TRY(hipMemsetAsync(S.ctl + S.B, 0, (size_t)S.B * sizeof(int), st), "batched iteration state", 1);
    DDH_TRY(hipMemsetAsync(S.ctl + 4 * S.B, 0, sizeof(int), st), "batched iteration state", 1);
    return 0;
  }
  // iteration state and residual history to the host, then a sync
  int chunk_fetch() {
    DDH_TRY(hipMemcpyAsync(S.hctl, S.ctl, ((size_t)4 * S.B + 1) * sizeof(int), hipMemcpyDeviceToHost, st),
            "batched iteration state D2H", 1);
    DDH_TRY(hipMemcpyAsync(S.hhist, S.hist, (size_t)HIST_ROWS * S.B * sizeof(double), hipMemcpyDeviceToHost, st),
            "batched residual history D2H", 1);
    return sync();
  }
  int upload(int* dev, const std::vector<int>& h) {
    for (int r = 0; r < S.B; ++r) S.hflags[r] = h[r];
    DDH_TRY(hipMemcpyAsync(dev, S.hflags, (size_t)S.B * sizeof(int), hipMemcpyHostToDevice, st), "flags H2D", 1);
    // the pinned staging buffer is reused: wait for the copy
    DDH_TRY(hipStreamSynchronize(st), "flags H2D", 1);
    return 0;
  }
  int sync() {
    DDH_TRY(hipStreamSynchronize(st), "batched sync", 1);
    return 0;
  }
  // fetch (status, resid) of every RHS
  int fetch() {
    if (pack()) return 1;
    return sync();
  }
  // full scalar slot k of RHS r (after a sync): read via a small copy
  int fetch_slots(int first, int count, std::vector<double>& out) {
    out.assign((size_t)S.B * count, 0.0);
    for (int r = 0; r < S.B; ++r)
      DDH_TRY(hipMemcpyAsync(out.data() + (size_t)r * count, S.scal + (size_t)r * SCAL_COUNT + first,
                             (size_t)count * sizeof(double), hipMemcpyDeviceToHost, st),
              "batched scalars D2H", 1);
    return sync();
  }
};

bool any_set(const std::vector<int>& v) {
  return std::any_of(v.begin(), v.end(), [](int e) { return e != 0; });
}

struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  ~Graph() {
    if (e) hipGraphExecDestroy(e);
    if (g) hipGraphDestroy(g);
  }
};

// Iterations run eagerly before the iteration graph is captured: a solve
// that ends within them (a few iterations on a small target) does not pay the
// capture and instantiation (about 0.3 ms); graph replay and eager launches
// give the same results bit for bit.
constexpr int GRAPH_FROM_ITER = 4;

// runs `body` once, either directly or through a captured graph replay
template <typename F>
int run_step(Ops& o, Graph& gr, bool use_graph, F&& body) {
  if (!use_graph) return body();
  if (!gr.e) {
    DDH_TRY(hipStreamBeginCapture(o.st, hipStreamCaptureModeThreadLocal), "graph capture", 1);
    const int rc = body();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(o.st, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    DDH_TRY(ec, "graph capture", 1);
    gr.g = g;
    DDH_TRY(hipGraphInstantiate(&gr.e, gr.g, nullptr, nullptr, 0), "graph instantiate", 1);
  }
  DDH_TRY(hipGraphLaunch(gr.e, o.st), "graph launch", 1);
  return 0;
}

// Iterations the next chunk runs before the host looks again: up to the
// next reliable-residual check (every active RHS has the same iteration
// count) and maxit, and no further than the first active RHS is expected to
// reach tol at the residual decay of the last chunk (each iteration launched
// after a pause costs only its early-exit launches, each host round trip a
// synchronisation).  Grows from 2 while no estimate is available.
int next_chunk(int itc, int maxit, int prev_k, double pred) {
  int k = RELIABLE_PERIOD - itc % RELIABLE_PERIOD;
  k = std::min(k, std::max(1, maxit - itc));
  const int grow = std::min(RELIABLE_PERIOD, std::max(2, 2 * prev_k));
  const int est = pred < 1e9 ? std::max(1, (int)std::floor(0.8 * pred)) : grow;
  return std::max(1, std::min(k, std::min(grow, est)));
}
// iterations until RHS k reaches tol at the decay from r0 to r1 over n iterations
double iters_to_tol(double r0, double r1, int n, double tol) {
  if (n <= 0 || !(r0 > 0.0) || !(r1 > 0.0) || !(r1 < r0) || !(r1 > tol)) return 1e9;
  return std::log(tol / r1) / (std::log(r1 / r0) / n);
}

// ----------------------------------------------------------- QMRCCG -------
int bsolve_qmrccg(Context& c, BatchState& S, const float2* b, float2* x, const float2* adia, const float2* aoff,
                  size_t os, int maxit, double tol, int* itno, float* tole, int* ncompte, bool use_graph) {
  const char* nm = "QMRCCG batch";
  Ops o(c, S, adia, aoff, os);
  const int B = S.B;
  c.log.reset(B);
  enum { R_ = 0, P_, AP_, Q_, D_, SV_, VT_, WT_, ATW_, X_, SY_ };
  float2 *r = o.V(R_), *p = o.V(P_), *ap = o.V(AP_), *q = o.V(Q_), *d = o.V(D_), *sv = o.V(SV_), *vt = o.V(VT_),
         *wt = o.V(WT_), *atw = o.V(ATW_), *xs = o.V(X_), *sy = o.V(SY_);
  const size_t vb = (size_t)B * S.n * sizeof(float2);
  DDH_TRY(hipMemcpyAsync(xs, x, vb, hipMemcpyDeviceToDevice, o.st), nm, 51);
  DDH_TRY(hipMemsetAsync(S.scal, 0, (size_t)B * SCAL_COUNT * sizeof(double), o.st), nm, 51);
  std::vector<int> act(B, 1), sel(B, 0), nc(B, 0), it(B, 0), rcode(B, 0);
  std::vector<double> resid(B, 1.0);
  if (o.upload(S.active, act)) return 51;
  const int* A = S.active;
  const int* Sm = S.sel;
  LinComb z;  // no terms: exact zeros, whatever the target held
  // |b|
  { DotSet ds; ds.add(b, b, DOT_N2); if (o.dots(ds, SL_D, A, FIN_BNORM)) return 52; }
  auto restart_from = [&](const float2* rs, const int* m) -> int {
    LinComb cp; cp.lit(rs, 1.0);
    if (o.lin(r, cp, m) || o.lin(vt, cp, m) || o.lin(wt, cp, m)) return 53;
    for (float2* v : {p, ap, q, d, sv, atw})
      if (o.lin(v, z, m)) return 53;
    return 0;
  };
  auto init_scalars = [&](const int* m) -> int {
    if (int rc = o.mv(wt, atw, m)) return rc;
    DotSet ds;
    ds.add(vt, vt, DOT_N2).add(wt, wt, DOT_N2).add(vt, wt, DOT_U).add(vt, atw, DOT_U);
    if (o.dots(ds, SL_D, m, FIN_QMR_INIT)) return 54;
    return 0;
  };
  auto residual = [&](const int* m) -> int {
    DotSet ds; ds.add(r, r, DOT_N2);
    return o.dots(ds, SLOT_R2, m, FIN_RESID) ? 55 : 0;
  };
  // r = b - A x0
  if (int rc = o.mv(xs, sy, A)) return rc;
  { LinComb l; l.lit(b, 1.0).lit(sy, -1.0); if (o.lin(sy, l, A)) return 53; }
  if (int rc = restart_from(sy, A)) return rc;
  if (int rc = init_scalars(A)) return rc;
  if (int rc = residual(A)) return rc;
  if (o.fetch()) return 55;
  for (int k = 0; k < B; ++k) {
    nc[k] = 2;
    resid[k] = S.hpacked[3 * k + 1];
    const int status = (int)S.hpacked[3 * k];
    if (status) { rcode[k] = status == 52 ? 52 : 55; act[k] = 0; }
    else if (!(resid[k] > tol)) act[k] = 0;  // converged at the initial guess (true residual)
  }
  if (o.upload(S.active, act)) return 55;
  Graph gr;
  // true residual / gap for the RHS in `sel`; returns per-RHS outcomes
  std::vector<double> slots;
  std::vector<int> conv(B), rst(B), rplc(B, 0);
  std::vector<double> tres(B, 0.0);

  auto reliable = [&](std::vector<int>& conv, std::vector<int>& rst, std::vector<double>& tres) -> int {
    if (o.upload(S.sel, sel)) return 59;
    if (int rc = o.mv(xs, sy, Sm)) return rc;
    DotSet ds; ds.add(b, sy, DOT_D2).add(b, sy, DOT_G2, r);
    if (o.dots(ds, SL_D, Sm, FIN_NONE)) return 59;
    if (o.sync() || o.fetch_slots(SL_D, 4, slots)) return 59;
    std::vector<double> bn;
    if (o.fetch_slots(SLOT_BNORM, 1, bn)) return 59;
    for (int k = 0; k < B; ++k) {
      conv[k] = rst[k] = 0;
      if (!sel[k]) continue;
      ++nc[k];
      const double tr = std::sqrt(std::max(0.0, slots[(size_t)k * 4]));
      const double gp = std::sqrt(std::max(0.0, slots[(size_t)k * 4 + 2]));
      tres[k] = bn[k] > 0.0 ? tr / bn[k] : 0.0;
      const double gap = tr > 0.0 ? gp / tr : (gp == 0.0 ? 0.0 : INFINITY);
      conv[k] = tres[k] <= tol;
      rst[k] = !conv[k] && (gap >= RELIABLE_GAP || resid[k] <= tol);
      int action = 0;
      if (rst[k] && it[k] < maxit)
        action = c.reliable_mode == RELIABLE_REPLACE ? SolveLog::REPLACE : SolveLog::RESTART;
      rplc[k] = action == SolveLog::REPLACE;
      c.log.check(k, tres[k], gap, action);
    }
    return 0;
  };
  auto body = [&]() -> int {
    if (o.f64) hipLaunchKernelGGL(kb_qmr_pq<double>, o.vgrid(), dim3(BT), 0, o.st, p, q, vt, atw, S.scal, S.n, A);
    else hipLaunchKernelGGL(kb_qmr_pq<float>, o.vgrid(), dim3(BT), 0, o.st, p, q, vt, atw, S.scal, S.n, A);
    if (int rc = o.mv(p, ap, A)) return rc;
    if (o.f64) hipLaunchKernelGGL(kb_qmr_vw<double>, o.vgrid(), dim3(BT), 0, o.st, vt, wt, ap, q, S.scal, S.n, A);
    else hipLaunchKernelGGL(kb_qmr_vw<float>, o.vgrid(), dim3(BT), 0, o.st, vt, wt, ap, q, S.scal, S.n, A);
    if (int rc = o.mv(wt, atw, A)) return rc;
    DotSet ds;
    ds.add(vt, vt, DOT_N2).add(wt, wt, DOT_N2).add(vt, wt, DOT_U).add(vt, atw, DOT_U);
    if (o.dots(ds, SL_D, A, FIN_QMR_RECUR)) return 57;
    if (o.f64)
      hipLaunchKernelGGL(kb_qmr_dsxr<double>, o.vgrid(), dim3(BT), 0, o.st, d, sv, xs, r, p, ap, S.scal, S.n, A,
                         S.part, o.nbk_vec);
    else
      hipLaunchKernelGGL(kb_qmr_dsxr<float>, o.vgrid(), dim3(BT), 0, o.st, d, sv, xs, r, p, ap, S.scal, S.n, A,
                         S.part, o.nbk_vec);
    if (o.fin(SLOT_R2, 2, A, FIN_RESID)) return 59;
    DDH_LAUNCH_OK("QMRCCG batch step", 56);
    if (o.pack()) return 56;
    return o.step_end(false, tol, maxit);
  };
  // iterations run in chunks: kb_step_end counts them on the device and
  // pauses the batch when a RHS needs the host (check, convergence, maxit)
  DDH_TRY(hipMemsetAsync(S.ctl, 0, ((size_t)4 * B + 1) * sizeof(int), o.st), nm, 51);
  int kchunk = 0;
  while (any_set(act)) {
    int itc = 0;
    for (int k = 0; k < B; ++k)
      if (act[k]) { itc = it[k]; break; }
    kchunk = next_chunk(itc, maxit, kchunk, 1e9);
    if (o.chunk_reset()) return 59;
    for (int i = 0; i < kchunk; ++i)
      if (int rc = run_step(o, gr, use_graph && itc >= GRAPH_FROM_ITER, body)) return rc;
    if (o.chunk_fetch()) return 59;
    std::fill(sel.begin(), sel.end(), 0);
    double pred = 1e9;
    for (int k = 0; k < B; ++k) {
      if (!act[k]) continue;
      const int n_it = S.hctl[k] - it[k];
      nc[k] += 2 * n_it;
      const int status = (int)S.hpacked[3 * k];
      for (int j = 0; j < (status ? n_it - 1 : n_it); ++j)
        c.log.iter(k, S.hhist[(size_t)((it[k] + j) % HIST_ROWS) * B + k]);
      it[k] = S.hctl[k];
      if (status) {
        rcode[k] = 58;
        act[k] = 0;
        continue;
      }
      const double r0 = resid[k];
      resid[k] = S.hpacked[3 * k + 1];
      sel[k] = S.hctl[B + k];
      pred = std::min(pred, iters_to_tol(r0, resid[k], n_it, tol));
    }
    kchunk = next_chunk(0, RELIABLE_PERIOD, kchunk, pred);
    if (any_set(sel)) {
      if (int rc = reliable(conv, rst, tres)) return rc;
      std::vector<int> restart(B, 0), replace(B, 0);
      for (int k = 0; k < B; ++k) {
        if (!sel[k]) continue;
        if (conv[k] || it[k] >= maxit) {
          resid[k] = tres[k];
          act[k] = 0;
        } else if (rst[k]) {
          (rplc[k] ? replace : restart)[k] = 1;
          resid[k] = tres[k];
        }
      }
      if (any_set(replace)) {
        // r = b - A x (A x in sy from the check); the Lanczos recurrence goes on
        if (o.upload(S.sel, replace)) return 59;
        LinComb l; l.lit(b, 1.0).lit(sy, -1.0);
        if (o.lin(r, l, Sm)) return 59;
      }
      if (any_set(restart)) {
        sel = restart;
        if (o.upload(S.sel, sel)) return 59;
        LinComb l; l.lit(b, 1.0).lit(sy, -1.0);
        if (o.lin(sy, l, Sm)) return 59;
        if (int rc = restart_from(sy, Sm)) return rc;
        if (int rc = init_scalars(Sm)) return rc;
        if (o.fetch()) return 59;
        for (int k = 0; k < B; ++k) {
          if (!restart[k]) continue;
          ++nc[k];
          if ((int)S.hpacked[3 * k]) { rcode[k] = 55; act[k] = 0; }
        }
      }
    }
    // a pause cleared the device flags
    if (o.upload(S.active, act)) return 59;
  }
  DDH_TRY(hipMemcpyAsync(x, xs, vb, hipMemcpyDeviceToDevice, o.st), nm, 61);
  if (o.sync()) return 61;
  int first = 0;
  for (int k = 0; k < B; ++k) {
    itno[k] = it[k];
    tole[k] = (float)resid[k];
    ncompte[k] = nc[k];
    if (rcode[k] && !first) first = rcode[k];
  }
  if (first) {
    char msg[96];
    std::snprintf(msg, sizeof(msg), "breakdown in at least one right-hand side (code %d)", first);
    set_error(nm, msg);
  }
  return first;
}

// ----------------------------------------------------------- PBCGS2 -------
int bsolve_pbcgs2(Context& c, BatchState& S, const float2* b, float2* x, const float2* adia, const float2* aoff,
                  size_t os, int maxit, double tol, int* itno, float* tole, int* ncompte, bool use_graph) {
  const char* nm = "PBCGS2 batch";
  Ops o(c, S, adia, aoff, os);
  const int B = S.B;
  c.log.reset(B);
  // work layout per RHS as the single-RHS solver (cu:1173-1174) plus A x scratch
  enum { RR_ = 0, R0_, R1_, R2_, U0_, U1_, U2_, XP_, BP_, X_, SY_ };
  float2 *rr = o.V(RR_), *r0 = o.V(R0_), *r1 = o.V(R1_), *r2 = o.V(R2_), *u0 = o.V(U0_), *u1 = o.V(U1_),
         *u2 = o.V(U2_), *xp = o.V(XP_), *bp = o.V(BP_), *xv = o.V(X_), *sy = o.V(SY_);
  const size_t vb = (size_t)B * S.n * sizeof(float2);
  DDH_TRY(hipMemcpyAsync(xp, x, vb, hipMemcpyDeviceToDevice, o.st), nm, 91);
  DDH_TRY(hipMemsetAsync(S.scal, 0, (size_t)B * SCAL_COUNT * sizeof(double), o.st), nm, 91);
  std::vector<int> act(B, 1), sel(B, 0), nc(B, 0), it(B, 0), rcode(B, 0);
  std::vector<double> resid(B, 0.0), rnrm0(B, 0.0), slots;
  if (o.upload(S.active, act)) return 92;
  const int* A = S.active;
  const int* Sm = S.sel;
  LinComb z;  // no terms: exact zeros, whatever the target held
  // r0 = b - A xp for nonzero initial guesses, b otherwise (ZBCG2 NONZERO_X)
  { DotSet ds; ds.add(xp, xp, DOT_N2); if (o.dots(ds, SL_D, A, FIN_NONE)) return 92; }
  if (o.fetch_slots(SL_D, 1, slots)) return 92;
  for (int k = 0; k < B; ++k) sel[k] = slots[k] > 0.0;
  { LinComb l; l.lit(b, 1.0); if (o.lin(r0, l, A)) return 92; }
  if (any_set(sel)) {
    if (o.upload(S.sel, sel)) return 92;
    if (int rc = o.mv(xp, sy, Sm)) return rc;
    LinComb l; l.lit(b, 1.0).lit(sy, -1.0);
    if (o.lin(r0, l, Sm)) return 92;
    for (int k = 0; k < B; ++k) nc[k] += sel[k];
  }
  { LinComb cp; cp.lit(r0, 1.0); if (o.lin(rr, cp, A) || o.lin(bp, cp, A)) return 92; }
  for (float2* v : {r1, r2, u0, u1, u2, xv})
    if (o.lin(v, z, A)) return 92;
  { DotSet ds; ds.add(r0, r0, DOT_N2); if (o.dots(ds, SL_D, A, FIN_P2_INIT)) return 92; }
  if (o.fetch_slots(p2::RNRM0, 1, slots)) return 92;
  for (int k = 0; k < B; ++k) {
    rnrm0[k] = slots[k];
    resid[k] = rnrm0[k] > 0.0 ? 1.0 : 0.0;
    if (!(resid[k] > tol)) act[k] = 0;
  }
  if (o.upload(S.active, act)) return 92;
  const dim3 sg((B + 63) / 64), sb(64);
  auto body = [&]() -> int {
    // first half-step (j = 0)
    { DotSet ds; ds.add(rr, r0, DOT_C); if (o.dots(ds, SL_D, A, FIN_P2_BETA1)) return 94; }
    { LinComb l; l.lit(r0, 1.0).dev(u0, p2::BETA, 1); if (o.lin(u0, l, A)) return 96; }
    if (int rc = o.mv(u0, u1, A)) return rc;
    { DotSet ds; ds.add(rr, u1, DOT_C); if (o.dots(ds, SL_D, A, FIN_P2_ALPHA)) return 97; }
    if (o.f64) hipLaunchKernelGGL((kb_p2_xr<double, 1>), o.vgrid(), dim3(BT), 0, o.st, xv, r0, r1, u0, u1, u2, S.scal, S.n, A);
    else hipLaunchKernelGGL((kb_p2_xr<float, 1>), o.vgrid(), dim3(BT), 0, o.st, xv, r0, r1, u0, u1, u2, S.scal, S.n, A);
    if (int rc = o.mv(r0, r1, A)) return rc;
    // second half-step (j = 1)
    { DotSet ds; ds.add(rr, r1, DOT_C).add(r0, r0, DOT_N2); if (o.dots(ds, SL_D, A, FIN_P2_BETA2)) return 94; }
    if (o.f64) hipLaunchKernelGGL(kb_p2_u2<double>, o.vgrid(), dim3(BT), 0, o.st, u0, u1, r0, r1, S.scal, S.n, A);
    else hipLaunchKernelGGL(kb_p2_u2<float>, o.vgrid(), dim3(BT), 0, o.st, u0, u1, r0, r1, S.scal, S.n, A);
    if (int rc = o.mv(u1, u2, A)) return rc;
    { DotSet ds; ds.add(rr, u2, DOT_C); if (o.dots(ds, SL_D, A, FIN_P2_ALPHA)) return 97; }
    if (o.f64) hipLaunchKernelGGL((kb_p2_xr<double, 2>), o.vgrid(), dim3(BT), 0, o.st, xv, r0, r1, u0, u1, u2, S.scal, S.n, A);
    else hipLaunchKernelGGL((kb_p2_xr<float, 2>), o.vgrid(), dim3(BT), 0, o.st, xv, r0, r1, u0, u1, u2, S.scal, S.n, A);
    if (int rc = o.mv(r1, r2, A)) return rc;
    // minimal-residual polynomial with the convex correction
    {
      DotSet g;
      g.add(r0, r0, DOT_C).add(r1, r0, DOT_C).add(r1, r1, DOT_C).add(r2, r0, DOT_C).add(r2, r1, DOT_C).add(r2, r2, DOT_C);
      if (o.dots(g, SL_D, A, FIN_P2_CONVEX)) return 101;
    }
    if (o.f64) hipLaunchKernelGGL(kb_p2_comb<double>, o.vgrid(), dim3(BT), 0, o.st, u0, xv, r0, u1, u2, r1, r2, S.scal, S.n, A);
    else hipLaunchKernelGGL(kb_p2_comb<float>, o.vgrid(), dim3(BT), 0, o.st, u0, xv, r0, u1, u2, r1, r2, S.scal, S.n, A);
    // reliable update of r (and x into xp) where the convex step asked for it
    hipLaunchKernelGGL(kb_p2_masks, sg, sb, 0, o.st, S.scal, A, S.m1, S.m2, B);
    if (int rc = o.mv(xv, sy, S.m1)) return rc;
    { LinComb l; l.lit(bp, 1.0).lit(sy, -1.0); if (o.lin(r0, l, S.m1)) return 104; }
    { LinComb l; l.lit(xv, 1.0).lit(xp, 1.0); if (o.lin(xp, l, S.m2)) return 104; }
    if (o.lin(xv, z, S.m2)) return 104;
    { LinComb cp; cp.lit(r0, 1.0); if (o.lin(bp, cp, S.m2)) return 104; }
    hipLaunchKernelGGL(kb_p2_reliable, sg, sb, 0, o.st, S.scal, S.m1, S.m2, B);
    DDH_LAUNCH_OK("PBCGS2 batch step", 103);
    if (o.pack()) return 103;
    return o.step_end(true, tol, maxit);
  };
  Graph gr;
  std::vector<double> bn;
  // iterations in chunks between host checks, as in the batched QMRCCG
  DDH_TRY(hipMemsetAsync(S.ctl, 0, ((size_t)4 * B + 1) * sizeof(int), o.st), nm, 92);
  std::vector<int> nxt(B, 0);
  int kchunk = 0;
  while (any_set(act)) {
    int itc = 0;
    for (int k = 0; k < B; ++k)
      if (act[k]) { itc = it[k]; break; }
    kchunk = next_chunk(itc, maxit, kchunk, 1e9);
    if (o.chunk_reset()) return 103;
    for (int i = 0; i < kchunk; ++i)
      if (int rc = run_step(o, gr, use_graph && itc >= GRAPH_FROM_ITER, body)) return rc;
    if (o.chunk_fetch()) return 103;
    std::fill(sel.begin(), sel.end(), 0);
    double pred = 1e9;
    for (int k = 0; k < B; ++k) {
      if (!act[k]) continue;
      const int n_it = S.hctl[k] - it[k];
      const int nx_now = S.hctl[3 * B + k];
      nc[k] += 4 * n_it + (nx_now - nxt[k]);
      nxt[k] = nx_now;
      const int status = (int)S.hpacked[3 * k];
      for (int j = 0; j < (status ? n_it - 1 : n_it); ++j)
        c.log.iter(k, S.hhist[(size_t)((it[k] + j) % HIST_ROWS) * B + k]);
      it[k] = S.hctl[k];
      if (status) {
        rcode[k] = status;
        act[k] = 0;
        continue;
      }
      const double r0 = resid[k];
      resid[k] = S.hpacked[3 * k + 1];
      sel[k] = S.hctl[B + k];
      pred = std::min(pred, iters_to_tol(r0, resid[k], n_it, tol));
    }
    kchunk = next_chunk(0, RELIABLE_PERIOD, kchunk, pred);
    if (!any_set(sel)) {
      if (o.upload(S.active, act)) return 104;
      continue;
    }
    // true residual |bp - A x| and gap |bp - A x - r0| (cu:543-556)
    if (o.upload(S.sel, sel)) return 104;
    if (int rc = o.mv(xv, sy, Sm)) return rc;
    { DotSet ds; ds.add(bp, sy, DOT_D2).add(bp, sy, DOT_G2, r0); if (o.dots(ds, SL_D, Sm, FIN_NONE)) return 104; }
    if (o.fetch_slots(SL_D, 4, slots)) return 104;
    std::vector<int> restart(B, 0);
    for (int k = 0; k < B; ++k) {
      if (!sel[k]) continue;
      ++nc[k];
      const double tr = std::sqrt(std::max(0.0, slots[(size_t)k * 4]));
      const double gp = std::sqrt(std::max(0.0, slots[(size_t)k * 4 + 2]));
      const double tres = rnrm0[k] > 0.0 ? tr / rnrm0[k] : 0.0;
      const double gap = tr > 0.0 ? gp / tr : (gp == 0.0 ? 0.0 : INFINITY);
      const bool conv = tres <= tol;
      if (conv || it[k] >= maxit) {
        act[k] = 0;
      } else if (gap >= RELIABLE_GAP || resid[k] <= tol) {
        restart[k] = 1;
      }
      c.log.check(k, tres, gap, restart[k] ? SolveLog::RESTART : 0);
      resid[k] = tres;
    }
    if (any_set(restart)) {
      sel = restart;
      if (o.upload(S.sel, sel)) return 104;
      LinComb lr; lr.lit(bp, 1.0).lit(sy, -1.0);
      LinComb lx; lx.lit(xv, 1.0).lit(xp, 1.0);
      LinComb cp; cp.lit(r0, 1.0);
      if (o.lin(r0, lr, Sm) || o.lin(xp, lx, Sm) || o.lin(xv, z, Sm) || o.lin(bp, cp, Sm) || o.lin(rr, cp, Sm))
        return 104;
      for (float2* v : {r1, r2, u0, u1, u2})
        if (o.lin(v, z, Sm)) return 104;
      hipLaunchKernelGGL(kb_p2_restart, sg, sb, 0, o.st, S.scal, Sm, B);
      DDH_LAUNCH_OK("PBCGS2 batch restart", 104);
    }
    // a pause cleared the device flags
    if (o.upload(S.active, act)) return 104;
  }
  // ZBCG2 returns X + XP with one final true residual (cu:1258-1264)
  std::fill(act.begin(), act.end(), 1);
  if (o.upload(S.active, act)) return 105;
  { LinComb lx; lx.lit(xv, 1.0).lit(xp, 1.0); if (o.lin(xv, lx, A)) return 105; }
  if (int rc = o.mv(xv, sy, A)) return rc;
  { DotSet ds; ds.add(b, sy, DOT_D2); if (o.dots(ds, SL_D, A, FIN_NONE)) return 106; }
  if (o.fetch_slots(SL_D, 1, slots)) return 106;
  DDH_TRY(hipMemcpyAsync(x, xv, vb, hipMemcpyDeviceToDevice, o.st), nm, 107);
  if (o.sync()) return 107;
  int first = 0;
  for (int k = 0; k < B; ++k) {
    ++nc[k];
    itno[k] = it[k];
    tole[k] = (float)(rnrm0[k] > 0.0 ? std::sqrt(std::max(0.0, slots[k])) / rnrm0[k] : 0.0);
    ncompte[k] = nc[k];
    if (rcode[k] && !first) first = rcode[k];
  }
  if (first) {
    char msg[96];
    std::snprintf(msg, sizeof(msg), "ZBCG2 breakdown in at least one right-hand side (code %d)", first);
    set_error(nm, msg);
  }
  return first;
}

}  // namespace

void batch_release(Context& c) { bs_free(bstate(c)); }

int batch_poison(Context& c) {
  BatchState& S = bstate(c);
  if (S.W && S.cap_vec)
    DDH_TRY(hipMemsetAsync(S.W, 0xFF, S.cap_vec * sizeof(float2), c.stream), "poison_workspace", 12);
  return 0;
}

}  // namespace ddhip

using namespace ddhip;

extern "C" DDSCAT_HIP_API int ddscat_hip_apply_batch(int cwhat, int nrhs, const void* x, void* y, const void* adia,
                                                     const void* aoff, int op_per_rhs, void* stream) {
  Context& c = ctx();
  if (!c.ready) { set_error("apply_batch", "not prepared"); return 10; }
  if ((cwhat != 'N' && cwhat != 'C') || nrhs <= 0 || !x || !y || !adia || !aoff ||
      (op_per_rhs != 0 && op_per_rhs != 1)) {
    set_error("apply_batch", "invalid argument");
    return 11;
  }
  BatchState& S = bstate(c);
  if (int rc = bs_ensure(c, S, nrhs, 0)) return rc;
  MvBatch bt;
  bt.nrhs = nrhs;
  bt.k.xs = S.n;
  bt.k.os = op_per_rhs ? S.n : 0;
  bt.k.a1s = matvec_a1_count(c);
  bt.k.a2s = matvec_a2_count(c);
  bt.a1 = S.a1;
  bt.a2 = S.a2;
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  return batch_matvec(c, cwhat, (const float2*)x, (float2*)y, (const float2*)adia, (const float2*)aoff, bt, st);
}

extern "C" DDSCAT_HIP_API int ddscat_hip_evalq_batch(int nrhs, const void* e, const void* p, const void* adia,
                                                     const void* aoff, int op_per_rhs, const float* ak, float e02,
                                                     double* out) {
  Context& c = ctx();
  clear_error();
  if (!c.ready || nrhs <= 0 || !e || !p || !adia || !aoff || !ak || !out || !(e02 > 0.f) ||
      (op_per_rhs != 0 && op_per_rhs != 1)) {
    set_error("evalq_batch", "invalid argument");
    return 300;
  }
  BatchState& S = bstate(c);
  if (int rc = bs_ensure(c, S, nrhs, 1)) return rc;
  std::vector<float> rabs(nrhs);
  std::vector<double> akk(nrhs);
  for (int r = 0; r < nrhs; ++r) {
    const double ak2 = (double)ak[3 * r] * ak[3 * r] + (double)ak[3 * r + 1] * ak[3 * r + 1] +
                       (double)ak[3 * r + 2] * ak[3 * r + 2];
    akk[r] = std::sqrt(ak2);
    rabs[r] = (float)(akk[r] * ak2 / 1.5);
  }
  hipStream_t st = c.stream;
  // the per-RHS k^3 terms travel in the partials buffer's head (overwritten by the dots below)
  float* drabs = (float*)S.part;
  DDH_TRY(hipMemcpyAsync(drabs, rabs.data(), nrhs * sizeof(float), hipMemcpyHostToDevice, st), "evalq_batch", 303);
  float2* q = S.W;
  const size_t os = op_per_rhs ? S.n : 0;
  const unsigned nb = (unsigned)std::min<size_t>(1024, (c.nat0 + BT - 1) / BT);
  hipLaunchKernelGGL(kb_evalq_q, dim3(nb, nrhs), dim3(BT), 0, st, (const float2*)p, q, (const float2*)adia,
                     (const float2*)aoff, os, drabs, c.nat0);
  DDH_LAUNCH_OK("batched EVALQ local term", 306);
  // EVALQ sums are FP64 whatever the solver precision mode
  const int prec = c.precision;
  c.precision = PREC_MIXED;
  Ops o(c, S, nullptr, nullptr, 0);
  DotSet d;
  d.add((const float2*)p, q, DOT_C).add((const float2*)e, (const float2*)p, DOT_C);
  const int drc = o.dots(d, SL_D, nullptr, FIN_NONE);
  c.precision = prec;
  std::vector<double> sl;
  if (drc || o.fetch_slots(SL_D, 4, sl)) return 307;
  const double pi = 3.14159265358979323846;
  for (int r = 0; r < nrhs; ++r) {
    const double* v = sl.data() + (size_t)r * 4;
    out[3 * r] = -4.0 * pi * akk[r] * v[1] / (double)e02;
    out[3 * r + 1] = 4.0 * pi * akk[r] * v[3] / (double)e02;
    out[3 * r + 2] = 2.0 * pi * akk[r] * v[2] / (double)e02;
  }
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_cabs_adjoint(int method, int nrhs, const void* p, const void* adia,
                                                      const void* aoff, int op_per_rhs, const float* ak, float e02,
                                                      int maxit, float tol, void* g_adia, void* g_aoff, int* itno,
                                                      float* tole) {
  Context& c = ctx();
  clear_error();
  if (!c.ready || nrhs <= 0 || !p || !adia || !aoff || !ak || !g_adia || !itno || !tole || !(e02 > 0.f) ||
      maxit <= 0 || !(tol > 0.f) || (op_per_rhs != 0 && op_per_rhs != 1)) {
    set_error("cabs_adjoint", "invalid argument");
    return 1;
  }
  if (method != DDSCAT_HIP_QMRCCG && method != DDSCAT_HIP_PBCGS2) {
    set_error("cabs_adjoint", "adjoint solves run batched QMRCCG or PBCGS2");
    return 2;
  }
  BatchState& S = bstate(c);
  // solver vectors 0..10, adjoint right-hand side 11, adjoint solution 12
  if (int rc = bs_ensure(c, S, nrhs, 13)) return rc;
  float2* rhs = S.W + (size_t)11 * nrhs * S.n;
  float2* mu = S.W + (size_t)12 * nrhs * S.n;
  std::vector<float> fac(2 * (size_t)nrhs);
  const double pi = 3.14159265358979323846;
  for (int r = 0; r < nrhs; ++r) {
    const double ak2 = (double)ak[3 * r] * ak[3 * r] + (double)ak[3 * r + 1] * ak[3 * r + 1] +
                       (double)ak[3 * r + 2] * ak[3 * r + 2];
    const double k = std::sqrt(ak2);
    fac[r] = (float)(4.0 * pi * k / (double)e02);
    fac[nrhs + r] = (float)(k * ak2 / 1.5);
  }
  hipStream_t st = c.stream;
  float* dfac = (float*)S.packed;  // 2 nrhs floats fit the 3 nrhs doubles of the status buffer
  DDH_TRY(hipMemcpyAsync(dfac, fac.data(), fac.size() * sizeof(float), hipMemcpyHostToDevice, st), "cabs_adjoint", 3);
  const size_t os = op_per_rhs ? S.n : 0;
  const unsigned nb = (unsigned)std::min<size_t>(1024, (c.nat0 + BT - 1) / BT);
  hipLaunchKernelGGL(kb_adj_rhs, dim3(nb, nrhs), dim3(BT), 0, st, (const float2*)p, (const float2*)adia,
                     (const float2*)aoff, os, dfac, dfac + nrhs, rhs, c.nat0);
  DDH_TRY(hipMemsetAsync(mu, 0, (size_t)nrhs * S.n * sizeof(float2), st), "cabs_adjoint", 3);
  DDH_LAUNCH_OK("adjoint right-hand side", 3);
  std::vector<int> nc(nrhs);
  auto fn = method == DDSCAT_HIP_QMRCCG ? bsolve_qmrccg : bsolve_pbcgs2;
  const int src = fn(c, S, rhs, mu, (const float2*)adia, (const float2*)aoff, os, maxit, (double)tol, itno, tole,
                     nc.data(), true);
  if (src) return src;
  // the solve used the status buffer: upload the factors again
  DDH_TRY(hipMemcpyAsync(dfac, fac.data(), nrhs * sizeof(float), hipMemcpyHostToDevice, st), "cabs_adjoint", 3);
  hipLaunchKernelGGL(kb_adj_grad, dim3(nb, nrhs), dim3(BT), 0, st, (const float2*)p, (const float2*)mu, dfac,
                     (float2*)g_adia, (float2*)g_aoff, c.nat0);
  DDH_LAUNCH_OK("adjoint gradient", 3);
  DDH_TRY(hipStreamSynchronize(st), "cabs_adjoint", 3);
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_solve_batch(int method, int nrhs, const void* b, void* x, const void* adia,
                                                     const void* aoff, int op_per_rhs, int maxit, float tol,
                                                     int* itno, float* tole, int* ncompte, int use_graph) {
  Context& c = ctx();
  clear_error();
  if (!c.ready || nrhs <= 0 || !b || !x || !adia || !aoff || !itno || !tole || !ncompte || maxit <= 0 ||
      !(tol > 0.f) || (op_per_rhs != 0 && op_per_rhs != 1)) {
    set_error("solve_batch", "invalid argument");
    return 1;
  }
  if (method != DDSCAT_HIP_QMRCCG && method != DDSCAT_HIP_PBCGS2) {
    set_error("solve_batch", "method not available in batched form");
    return 2;
  }
  BatchState& S = bstate(c);
  if (int rc = bs_ensure(c, S, nrhs, 11)) return rc;
  c.solution_valid = false;
  const size_t os = op_per_rhs ? S.n : 0;
  auto fn = method == DDSCAT_HIP_QMRCCG ? bsolve_qmrccg : bsolve_pbcgs2;
  return fn(c, S, (const float2*)b, (float2*)x, (const float2*)adia, (const float2*)aoff, os, maxit, (double)tol,
            itno, tole, ncompte, use_graph != 0);
}

extern "C" DDSCAT_HIP_API int ddscat_hip_select_target(int slot) {
  if (slot < 0 || slot > MAX_TARGETS) {
    set_error("select_target", "slot out of range");
    return 1;
  }
  select_target(slot);
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_solve_targets(int method, int ntargets, const int* slots, const int* nrhs,
                                                       void* const* b, void* const* x, void* const* adia,
                                                       void* const* aoff, int op_per_rhs, int maxit, float tol,
                                                       int* itno, float* tole, int* ncompte, int use_graph) {
  clear_error();
  if (ntargets <= 0 || !slots || !nrhs || !b || !x || !adia || !aoff || !itno || !tole || !ncompte || maxit <= 0 ||
      !(tol > 0.f) || (op_per_rhs != 0 && op_per_rhs != 1)) {
    set_error("solve_targets", "invalid argument");
    return 1;
  }
  if (method != DDSCAT_HIP_QMRCCG && method != DDSCAT_HIP_PBCGS2) {
    set_error("solve_targets", "method not available in batched form");
    return 2;
  }
  std::vector<size_t> first(ntargets + 1, 0);
  for (int t = 0; t < ntargets; ++t) {
    const int s = slots[t];
    if (s < 0 || s > MAX_TARGETS || nrhs[t] <= 0 || !b[t] || !x[t] || !adia[t] || !aoff[t]) {
      set_error("solve_targets", "invalid target entry");
      return 1;
    }
    for (int u = 0; u < t; ++u)
      if (slots[u] == s) { set_error("solve_targets", "a target slot appears twice"); return 1; }
    Context& c = ctx_at(s);
    if (!c.ready) {
      set_error("solve_targets", "target not prepared");
      return 2;
    }
    first[t + 1] = first[t] + (size_t)nrhs[t];
  }
  // workspaces first, from this thread: no allocation runs while another
  // target's iteration is being captured
  for (int t = 0; t < ntargets; ++t) {
    Context& c = ctx_at(slots[t]);
    if (int rc = bs_ensure(c, bstate(c), nrhs[t], 11)) return rc;
    c.solution_valid = false;
  }
  int dev = 0;
  DDH_TRY(hipGetDevice(&dev), "solve_targets", 1);
  auto fn = method == DDSCAT_HIP_QMRCCG ? bsolve_qmrccg : bsolve_pbcgs2;
  std::vector<int> rcs(ntargets, 0);
  auto run = [&](int t) {
    if (hipSetDevice(dev) != hipSuccess) { rcs[t] = 1; return; }
    Context& c = ctx_at(slots[t]);
    const size_t os = op_per_rhs ? (size_t)3 * c.nat0 : 0;
    const size_t f = first[t];
    rcs[t] = fn(c, bstate(c), (const float2*)b[t], (float2*)x[t], (const float2*)adia[t], (const float2*)aoff[t], os,
                maxit, (double)tol, itno + f, tole + f, ncompte + f, use_graph != 0);
  };
  // every target on its own stream, driven from its own host thread, so
  // their kernels share the GPU
  std::vector<std::thread> pool;
  pool.reserve(ntargets);
  for (int t = 1; t < ntargets; ++t) pool.emplace_back(run, t);
  run(0);
  for (std::thread& th : pool) th.join();
  for (int t = 0; t < ntargets; ++t)
    if (rcs[t]) return rcs[t];
  return 0;
}
