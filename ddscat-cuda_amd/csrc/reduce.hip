// FP64 (or FP32) multi-dot reductions, linear combinations and layout maps.
#include "solver_common.hpp"
#include "dots.hpp"

namespace ddhip {

// NT: compile-time term slots (>= L.nt); NV: compile-time operand slots
// (>= distinct vectors)
template <typename Acc, int NT, int NV>
__global__ __launch_bounds__(RED_THREADS) void k_dots(DotOperands L, size_t n, double* __restrict__ part) {
  Acc re[NT], im[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) { re[t] = 0; im[t] = 0; }
  dot_accumulate<Acc, NT, NV>(L, 0, (size_t)blockIdx.x * RED_THREADS + threadIdx.x, (size_t)gridDim.x * RED_THREADS,
                              n, re, im);
  __shared__ Acc sred[2 * NT][RED_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < L.nt) {
      const Acc r = wave_sum(re[t]);
      const Acc m = wave_sum(im[t]);
      if (lane == 0) { sred[2 * t][wave] = r; sred[2 * t + 1][wave] = m; }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * L.nt) {
    Acc s = 0;
    for (int w = 0; w < RED_THREADS / 64; ++w) s += sred[threadIdx.x][w];
    part[(size_t)threadIdx.x * RED_MAX_BLOCKS + blockIdx.x] = (double)s;
  }
}

// one wave per sum: lane l loads partials l, l + 64, ... (all in flight at
// once; they were written on other XCDs and come from beyond this XCD's L2),
// adds them in order, then a fixed butterfly; deterministic for a given
// block count
constexpr int FIN_THREADS = 1024;
template <typename Acc>
__global__ __launch_bounds__(FIN_THREADS) void k_dots_finalize(const double* __restrict__ part, int nblocks,
                                                              int nsums, double* __restrict__ out) {
  constexpr int PER_LANE = RED_MAX_BLOCKS / 64;
  const int lane = threadIdx.x & 63;
  for (int s = threadIdx.x >> 6; s < nsums; s += FIN_THREADS / 64) {
    const double* ps = part + (size_t)s * RED_MAX_BLOCKS + lane;
    double p[PER_LANE];  // the buffer holds RED_MAX_BLOCKS partials per sum
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) p[j] = ps[64 * j];
    Acc v = 0;
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) v += lane + 64 * j < nblocks ? (Acc)p[j] : (Acc)0;
    v = wave_sum(v);
    if (lane == 0) out[s] = (double)v;
  }
}

int finalize_partials(Context& c, int nblocks, int nsums, int dst, hipStream_t st) {
  if (c.precision == PREC_FP32)
    hipLaunchKernelGGL(k_dots_finalize<float>, dim3(1), dim3(FIN_THREADS), 0, st, c.red, nblocks, nsums, c.scal + dst);
  else
    hipLaunchKernelGGL(k_dots_finalize<double>, dim3(1), dim3(FIN_THREADS), 0, st, c.red, nblocks, nsums,
                       c.scal + dst);
  DDH_LAUNCH_OK("reduction finalize", 1);
  return 0;
}

template <typename Acc>
struct DotsRun {
  const DotOperands& L;
  size_t n;
  int nb;
  double* part;
  hipStream_t st;
  template <int NT, int NV>
  void run() const {
    hipLaunchKernelGGL((k_dots<Acc, NT, NV>), dim3(nb), dim3(RED_THREADS), 0, st, L, n, part);
  }
};

// launches the terms [t0, t1) of s as one reduction (operand list fits DOT_VMAX)
static int dots_chunk(Context& c, const DotSet& s, int t0, int t1, size_t n, int dst, hipStream_t st) {
  DotOperands L;
  const int nv = dot_operands(s, t0, t1, L);
  if (nv < 0) {
    set_error("dot reduction", "too many distinct operand vectors");
    return 1;
  }
  size_t want = (n + (size_t)RED_THREADS * 4 - 1) / ((size_t)RED_THREADS * 4);
  int nb = (int)std::min<size_t>(RED_MAX_BLOCKS, std::max<size_t>(1, want));
  if (c.precision == PREC_FP32) dot_dispatch(L.nt, nv, DotsRun<float>{L, n, nb, c.red, st});
  else dot_dispatch(L.nt, nv, DotsRun<double>{L, n, nb, c.red, st});
  DDH_LAUNCH_OK("FP64 dot reduction", 1);
  return finalize_partials(c, nb, 2 * L.nt, dst, st);
}

int dots(Context& c, const DotSet& s, size_t n, int dst, hipStream_t st) {
  // split where a term would bring the distinct operands past DOT_VMAX
  int t0 = 0;
  while (t0 < s.n) {
    const float2* seen[3 * DOT_MAX];
    int nv = 0, t = t0;
    for (; t < s.n; ++t) {
      const float2* ops[3] = {s.t[t].a, s.t[t].b ? s.t[t].b : s.t[t].a,
                              s.t[t].c ? s.t[t].c : (s.t[t].b ? s.t[t].b : s.t[t].a)};
      int add = 0;
      const float2* fresh[3];
      for (const float2* p : ops) {
        bool have = false;
        for (int v = 0; v < nv && !have; ++v) have = seen[v] == p;
        for (int v = 0; v < add && !have; ++v) have = fresh[v] == p;
        if (!have) fresh[add++] = p;
      }
      if (nv + add > DOT_VMAX) break;
      for (int v = 0; v < add; ++v) seen[nv++] = fresh[v];
    }
    if (int rc = dots_chunk(c, s, t0, t, n, dst + 2 * t0, st)) return rc;
    t0 = t;
  }
  return 0;
}

struct LinLaunch {
  const float2* v[LC_MAX];
  double lre[LC_MAX], lim[LC_MAX];
  int sidx[LC_MAX];
  int neg[LC_MAX];
  int n;
};

template <bool F64>
__global__ __launch_bounds__(256) void k_lincomb(float2* __restrict__ d, LinLaunch L, const double* __restrict__ scal,
                                                 size_t n) {
  double cr[LC_MAX], ci[LC_MAX];
#pragma unroll
  for (int t = 0; t < LC_MAX; ++t) {
    if (t < L.n) {
      if (L.sidx[t] >= 0) {
        const double s = L.neg[t] ? -1.0 : 1.0;
        cr[t] = s * scal[L.sidx[t]];
        ci[t] = s * scal[L.sidx[t] + 1];
      } else {
        cr[t] = L.lre[t];
        ci[t] = L.lim[t];
      }
    } else {
      cr[t] = ci[t] = 0.0;
    }
  }
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (F64) {
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int t = 0; t < LC_MAX; ++t) {
        if (t < L.n) {
          const float2 v = L.v[t][i];
          ar += cr[t] * (double)v.x - ci[t] * (double)v.y;
          ai += cr[t] * (double)v.y + ci[t] * (double)v.x;
        }
      }
      d[i] = make_float2((float)ar, (float)ai);
    } else {
      float ar = 0.f, ai = 0.f;
#pragma unroll
      for (int t = 0; t < LC_MAX; ++t) {
        if (t < L.n) {
          const float2 v = L.v[t][i];
          ar += (float)cr[t] * v.x - (float)ci[t] * v.y;
          ai += (float)cr[t] * v.y + (float)ci[t] * v.x;
        }
      }
      d[i] = make_float2(ar, ai);
    }
  }
}

int lincomb(Context& c, float2* d, const LinComb& lc, size_t n, hipStream_t st) {
  LinLaunch L{};
  L.n = lc.n;
  for (int t = 0; t < lc.n; ++t) {
    L.v[t] = lc.t[t].v;
    L.sidx[t] = lc.t[t].sidx;
    L.lre[t] = lc.t[t].re;
    L.lim[t] = lc.t[t].im;
    L.neg[t] = lc.t[t].negate;
  }
  for (int t = lc.n; t < LC_MAX; ++t) { L.v[t] = d; L.sidx[t] = -1; }
  const int th = 256;
  size_t want = (n + th * 4 - 1) / (th * 4);
  int nb = (int)std::min<size_t>(8192, std::max<size_t>(1, want));
  if (c.precision == PREC_FP32)
    hipLaunchKernelGGL(k_lincomb<false>, dim3(nb), dim3(th), 0, st, d, L, c.scal, n);
  else
    hipLaunchKernelGGL(k_lincomb<true>, dim3(nb), dim3(th), 0, st, d, L, c.scal, n);
  DDH_LAUNCH_OK("vector linear combination", 1);
  return 0;
}

int fetch_scalars(Context& c, int first, int count, hipStream_t st) {
  DDH_TRY(hipMemcpyAsync(c.hscal + first, c.scal + first, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, st),
          "scalar D2H", 1);
  DDH_TRY(hipStreamSynchronize(st), "scalar sync", 1);
  return 0;
}

__global__ void k_gather(const float2* __restrict__ full, float2* __restrict__ comp, const int* __restrict__ occ,
                         int nat, int nat0) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)3 * nat0) return;
  const int cc = (int)(q / nat0);
  const int k = (int)(q - (size_t)cc * nat0);
  comp[q] = full[(size_t)cc * nat + occ[k]];
}

__global__ void k_zero(float2* __restrict__ d, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) d[t] = make_float2(0.f, 0.f);
}

__global__ void k_scatter(const float2* __restrict__ comp, float2* __restrict__ full, const int* __restrict__ occ,
                          int nat, int nat0) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)3 * nat0) return;
  const int cc = (int)(q / nat0);
  const int k = (int)(q - (size_t)cc * nat0);
  full[(size_t)cc * nat + occ[k]] = comp[q];
}

int gather_compact(Context& c, const float2* full, float2* compact, hipStream_t st) {
  const size_t n = (size_t)3 * c.nat0;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, full, compact, c.occ, c.nat, c.nat0);
  DDH_LAUNCH_OK("gather full->compact", 1);
  return 0;
}

int scatter_full(Context& c, const float2* compact, float2* full, hipStream_t st) {
  const size_t n = (size_t)3 * c.nat, cn = (size_t)3 * c.nat0;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, full, n);
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((cn + 255) / 256)), dim3(256), 0, st, compact, full, c.occ, c.nat,
                     c.nat0);
  DDH_LAUNCH_OK("scatter compact->full", 1);
  return 0;
}

int run_mv(Context& c, int cwhat, const float2* x, float2* y, int& nc) {
  int rc = (c.backend == BACKEND_SLICES) ? slices_matvec(c, cwhat, x, y, c.adia_c, c.aoff_c, 1, c.stream)
                                         : matvec_launch(c, cwhat, x, y, c.adia_c, c.aoff_c, 1, c.stream);
  if (rc) return rc;
  ++nc;
  return 0;
}

}  // namespace ddhip
