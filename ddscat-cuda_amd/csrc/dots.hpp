// Multi-dot accumulation over a deduplicated operand list, shared by the
// single-RHS reductions (reduce.hip) and the batched ones (bsolve.hip).
//
// Term t of a DotSet reads v[ia[t]], v[ib[t]], v[ic[t]], so each distinct
// vector is loaded once per element however many terms share it (the
// QMRCCG set reads v~, w~ and A w~ for its four dots).  Each lane takes two
// elements per step and issues all their loads before any arithmetic.
#pragma once
#include "solver_common.hpp"

namespace ddhip {

constexpr int DOT_VMAX = 12;  // distinct operand vectors per launch
struct DotOperands {
  const float2* v[DOT_VMAX];
  unsigned char ia[DOT_MAX], ib[DOT_MAX], ic[DOT_MAX];
  int kind[DOT_MAX];
  int nt;
};

// Terms [t0, t1) of s as operand slots; returns the number of distinct
// vectors, or -1 when they exceed DOT_VMAX.  Unused slots repeat v[0].
inline int dot_operands(const DotSet& s, int t0, int t1, DotOperands& L) {
  L = DotOperands{};
  L.nt = t1 - t0;
  int nv = 0;
  auto slot = [&](const float2* p) -> int {
    for (int v = 0; v < nv; ++v)
      if (L.v[v] == p) return v;
    if (nv == DOT_VMAX) return -1;
    L.v[nv] = p;
    return nv++;
  };
  for (int t = t0; t < t1; ++t) {
    const float2* a = s.t[t].a;
    const float2* b = s.t[t].b ? s.t[t].b : a;
    const float2* c = s.t[t].c ? s.t[t].c : b;
    const int ia = slot(a), ib = slot(b), ic = slot(c);
    if (ia < 0 || ib < 0 || ic < 0) return -1;
    L.ia[t - t0] = (unsigned char)ia;
    L.ib[t - t0] = (unsigned char)ib;
    L.ic[t - t0] = (unsigned char)ic;
    L.kind[t - t0] = s.t[t].kind;
  }
  for (int v = nv; v < DOT_VMAX; ++v) L.v[v] = L.v[0];
  return nv;
}

// Calls f.template run<NT, NV>() for the smallest compiled slot counts
// covering nt terms over nv distinct vectors.
template <typename F>
inline void dot_dispatch(int nt, int nv, const F& f) {
  if (nt <= 1 && nv <= 1) f.template run<1, 1>();
  else if (nt <= 2 && nv <= 2) f.template run<2, 2>();
  else if (nt <= 2 && nv <= 3) f.template run<2, 3>();
  else if (nt <= 4 && nv <= 3) f.template run<4, 3>();
  else if (nt <= 6 && nv <= 4) f.template run<6, 4>();
  else if (nv <= 6) f.template run<DOT_MAX, 6>();
  else f.template run<DOT_MAX, DOT_VMAX>();
}

template <typename Acc>
__device__ __forceinline__ Acc wave_sum(Acc v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// x[k] for a wave-uniform k < NV as a chain of selects; the empty asm keeps
// the compiler from folding the chain back into an indexed array access,
// which it would place in scratch memory
template <int NV>
__device__ __forceinline__ float2 pick(const float2* x, int k) {
  float rx = x[0].x, ry = x[0].y;
#pragma unroll
  for (int v = 1; v < NV; ++v) {
    rx = k == v ? x[v].x : rx;
    ry = k == v ? x[v].y : ry;
    asm volatile("" : "+v"(rx), "+v"(ry));
  }
  return make_float2(rx, ry);
}

// kind: DOT_C sum conj(a) b, DOT_U sum a b, DOT_N2 |a|^2, DOT_D2 |a - b|^2,
// DOT_G2 |a - b - c|^2
template <typename Acc, int NT, int NV>
__device__ __forceinline__ void dot_terms(const DotOperands& L, const float2* x, Acc* re, Acc* im) {
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < L.nt) {
      const int kd = L.kind[t];
      const float2 a = pick<NV>(x, L.ia[t]);
      if (kd == DOT_N2) {
        re[t] += (Acc)a.x * (Acc)a.x + (Acc)a.y * (Acc)a.y;
      } else {
        const float2 b = pick<NV>(x, L.ib[t]);
        if (kd >= DOT_D2) {
          Acc dx = (Acc)a.x - (Acc)b.x, dy = (Acc)a.y - (Acc)b.y;
          if (kd == DOT_G2) {
            const float2 cc = pick<NV>(x, L.ic[t]);
            dx -= (Acc)cc.x;
            dy -= (Acc)cc.y;
          }
          re[t] += dx * dx + dy * dy;
        } else {
          const Acc ay = kd == DOT_C ? -(Acc)a.y : (Acc)a.y;
          re[t] += (Acc)a.x * (Acc)b.x - ay * (Acc)b.y;
          im[t] += (Acc)a.x * (Acc)b.y + ay * (Acc)b.x;
        }
      }
    }
  }
}

// Accumulates elements i0, i0 + stride, ... < n of the vectors at offset
// `off` into re[t], im[t].
template <typename Acc, int NT, int NV>
__device__ __forceinline__ void dot_accumulate(const DotOperands& L, size_t off, size_t i0, size_t stride, size_t n,
                                               Acc* re, Acc* im) {
  size_t i = i0;
  for (; i + stride < n; i += 2 * stride) {
    float2 x0[NV], x1[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      x0[v] = L.v[v][off + i];
      x1[v] = L.v[v][off + i + stride];
    }
    dot_terms<Acc, NT, NV>(L, x0, re, im);
    dot_terms<Acc, NT, NV>(L, x1, re, im);
  }
  if (i < n) {
    float2 x0[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) x0[v] = L.v[v][off + i];
    dot_terms<Acc, NT, NV>(L, x0, re, im);
  }
}

}  // namespace ddhip
