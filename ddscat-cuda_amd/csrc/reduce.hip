// FP64 (or FP32) multi-dot reductions, linear combinations and layout maps.
#include "solver_common.hpp"

namespace ddhip {

struct DotLaunch {
  const float2* a[DOT_MAX];
  const float2* b[DOT_MAX];
  const float2* c[DOT_MAX];
  int kind[DOT_MAX];
  int nt;
};

template <typename Acc>
__device__ __forceinline__ Acc wave_sum(Acc v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// NT: compile-time number of term slots (>= L.nt); accumulators for NT terms only
template <typename Acc, int NT>
__global__ __launch_bounds__(RED_THREADS) void k_dots(DotLaunch L, size_t n, double* __restrict__ part) {
  Acc re[NT], im[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) { re[t] = 0; im[t] = 0; }
  const size_t stride = (size_t)gridDim.x * RED_THREADS;
  for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < L.nt) {
        const float2 a = L.a[t][i];
        const int kd = L.kind[t];
        if (kd == DOT_N2) {
          re[t] += (Acc)a.x * (Acc)a.x + (Acc)a.y * (Acc)a.y;
        } else if (kd >= DOT_D2) {
          const float2 b = L.b[t][i];
          Acc dx = (Acc)a.x - (Acc)b.x, dy = (Acc)a.y - (Acc)b.y;
          if (kd == DOT_G2) {
            const float2 cc = L.c[t][i];
            dx -= (Acc)cc.x;
            dy -= (Acc)cc.y;
          }
          re[t] += dx * dx + dy * dy;
        } else {
          const float2 b = L.b[t][i];
          const Acc ay = L.kind[t] == 0 ? -(Acc)a.y : (Acc)a.y;
          re[t] += (Acc)a.x * (Acc)b.x - ay * (Acc)b.y;
          im[t] += (Acc)a.x * (Acc)b.y + ay * (Acc)b.x;
        }
      }
    }
  }
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

template <typename Acc>
__global__ __launch_bounds__(RED_THREADS) void k_dots_finalize(const double* __restrict__ part, int nblocks,
                                                              int nsums, double* __restrict__ out) {
  __shared__ Acc sh[RED_THREADS];
  for (int s = 0; s < nsums; ++s) {
    Acc v = 0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS) v += (Acc)part[(size_t)s * RED_MAX_BLOCKS + b];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = (double)sh[0];
    __syncthreads();
  }
}

int dots(Context& c, const DotSet& s, size_t n, int dst, hipStream_t st) {
  DotLaunch L{};
  L.nt = s.n;
  for (int t = 0; t < s.n; ++t) {
    L.a[t] = s.t[t].a;
    L.b[t] = s.t[t].b ? s.t[t].b : s.t[t].a;
    L.c[t] = s.t[t].c ? s.t[t].c : L.b[t];
    L.kind[t] = s.t[t].kind;
  }
  for (int t = s.n; t < DOT_MAX; ++t) { L.a[t] = L.a[0]; L.b[t] = L.b[0]; L.c[t] = L.c[0]; L.kind[t] = DOT_N2; }
  size_t want = (n + (size_t)RED_THREADS * 4 - 1) / ((size_t)RED_THREADS * 4);
  int nb = (int)std::min<size_t>(RED_MAX_BLOCKS, std::max<size_t>(1, want));
  const bool f32 = c.precision == PREC_FP32;
#define DDH_DOTS(NT_)                                                                                    \
  if (f32) hipLaunchKernelGGL((k_dots<float, NT_>), dim3(nb), dim3(RED_THREADS), 0, st, L, n, c.red);  \
  else hipLaunchKernelGGL((k_dots<double, NT_>), dim3(nb), dim3(RED_THREADS), 0, st, L, n, c.red);
  if (s.n <= 1) { DDH_DOTS(1) }
  else if (s.n <= 2) { DDH_DOTS(2) }
  else if (s.n <= 4) { DDH_DOTS(4) }
  else if (s.n <= 6) { DDH_DOTS(6) }
  else { DDH_DOTS(DOT_MAX) }
#undef DDH_DOTS
  if (f32)
    hipLaunchKernelGGL(k_dots_finalize<float>, dim3(1), dim3(RED_THREADS), 0, st, c.red, nb, 2 * s.n, c.scal + dst);
  else
    hipLaunchKernelGGL(k_dots_finalize<double>, dim3(1), dim3(RED_THREADS), 0, st, c.red, nb, 2 * s.n, c.scal + dst);
  DDH_LAUNCH_OK("FP64 dot reduction", 1);
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
