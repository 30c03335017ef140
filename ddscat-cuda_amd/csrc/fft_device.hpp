// LDS-resident batched 1-D FFTs for gfx950 (wave64).
//
// A workgroup transforms C independent sequences of length L that sit in LDS
// at address e*se + col*sc (e = element, col = sequence).  Each radix-R stage
// is a Stockham autosort step: butterfly j of a sequence reads elements
// j + r*L/R (r < R), multiplies by the twiddle w^{r*k} (k = j mod Ns, Ns the
// product of the radices already applied), does an R-point DFT in registers
// and writes element (j/Ns)*Ns*R + k + q*Ns (q < R).  Stages run in place:
// every thread loads all its butterflies into registers, the workgroup
// synchronises, then all threads store.  Twiddles come from an LDS copy of
// the table e^{+2 pi i m/L} (FP64-computed, rounded once to FP32); the
// inverse direction uses its conjugate.  DIR=+1 is the e^{+i} transform.
#pragma once
#include "common.hpp"

namespace ddhip {

template <int DIR>
__device__ __forceinline__ void dft2(float2* a) {
  float2 t = a[1];
  a[1] = csub(a[0], t);
  a[0] = cadd(a[0], t);
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* a) {
  const float h = 0.86602540378443864676f;  // sqrt(3)/2
  float2 t = cadd(a[1], a[2]);
  float2 d = csub(a[1], a[2]);
  float2 m = make_float2(a[0].x - 0.5f * t.x, a[0].y - 0.5f * t.y);
  float2 n = cmul_si<DIR>(cscale(d, h));
  a[0] = cadd(a[0], t);
  a[1] = cadd(m, n);
  a[2] = csub(m, n);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* a) {
  float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
  float2 s13 = cadd(a[1], a[3]), d13 = cmul_si<DIR>(csub(a[1], a[3]));
  a[0] = cadd(s02, s13);
  a[2] = csub(s02, s13);
  a[1] = cadd(d02, d13);
  a[3] = csub(d02, d13);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2* a) {
  const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
  const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
  const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
  const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
  float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
  float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
  float2 m1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
  float2 m2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
  float2 n1 = cmul_si<DIR>(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
  float2 n2 = cmul_si<DIR>(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
  a[0] = cadd(a[0], cadd(t1, t2));
  a[1] = cadd(m1, n1);
  a[4] = csub(m1, n1);
  a[2] = cadd(m2, n2);
  a[3] = csub(m2, n2);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* a) {
  const float r = 0.70710678118654752440f;  // 1/sqrt(2)
  float2 u[4], v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    u[i] = cadd(a[i], a[i + 4]);
    v[i] = csub(a[i], a[i + 4]);
  }
  // v_i *= w8^i, w8 = e^{DIR 2 pi i/8}
  v[1] = make_float2(r * (v[1].x - DIR * v[1].y), r * (v[1].y + DIR * v[1].x));
  v[2] = cmul_si<DIR>(v[2]);
  v[3] = make_float2(r * (-v[3].x - DIR * v[3].y), r * (-v[3].y + DIR * v[3].x));
  dft4<DIR>(u);
  dft4<DIR>(v);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a[2 * q] = u[q];
    a[2 * q + 1] = v[q];
  }
}

template <int DIR>
__device__ __forceinline__ void dft7(float2* a) {
  const float c1 = 0.62348980185873353053f, c2 = -0.22252093395631440429f, c3 = -0.90096886790241912624f;
  const float s1 = 0.78183148246802980871f, s2 = 0.97492791218182360702f, s3 = 0.43388373911755812048f;
  const float2 t1 = cadd(a[1], a[6]), t2 = cadd(a[2], a[5]), t3 = cadd(a[3], a[4]);
  const float2 d1 = csub(a[1], a[6]), d2 = csub(a[2], a[5]), d3 = csub(a[3], a[4]);
  const float2 a0 = a[0];
  auto comb = [&](float ca, float cb, float cc, float sa, float sb, float sc, float2& lo, float2& hi) {
    const float2 C = make_float2(a0.x + ca * t1.x + cb * t2.x + cc * t3.x, a0.y + ca * t1.y + cb * t2.y + cc * t3.y);
    const float2 S = cmul_si<DIR>(make_float2(sa * d1.x + sb * d2.x + sc * d3.x, sa * d1.y + sb * d2.y + sc * d3.y));
    lo = cadd(C, S);
    hi = csub(C, S);
  };
  float2 b1, b6, b2, b5, b3, b4;
  comb(c1, c2, c3, s1, s2, s3, b1, b6);
  comb(c2, c3, c1, s2, -s3, -s1, b2, b5);
  comb(c3, c1, c2, s3, -s1, s2, b3, b4);
  a[0] = cadd(a0, cadd(t1, cadd(t2, t3)));
  a[1] = b1; a[2] = b2; a[3] = b3; a[4] = b4; a[5] = b5; a[6] = b6;
}

template <int R, int DIR>
__device__ __forceinline__ void dftR(float2* a) {
  if constexpr (R == 2) dft2<DIR>(a);
  else if constexpr (R == 3) dft3<DIR>(a);
  else if constexpr (R == 4) dft4<DIR>(a);
  else if constexpr (R == 5) dft5<DIR>(a);
  else if constexpr (R == 7) dft7<DIR>(a);
  else dft8<DIR>(a);
}

// One in-place Stockham stage over C sequences.  PER bounds the number of
// elements a thread handles per stage (C*L <= PER*blockDim.x).
// COLFAST: butterfly b -> (col = b % C, j = b / C), so neighbouring lanes
// touch neighbouring sequences (layout e*C + col); otherwise
// (j = b % (L/R), col = b / (L/R)) for row layout col*sc + e.
template <int R, int DIR, bool COLFAST, int PER>
__device__ __forceinline__ void fft_stage(float2* __restrict__ buf, const float2* __restrict__ tw,
                                          int L, int C, int Ns, int se, int sc) {
  constexpr int NB = (PER + R - 1) / R;
  const int Lr = L / R;
  const int total = C * Lr;
  const int T = blockDim.x;
  const int tstep = L / (Ns * R);
  float2 v[NB][R];
  int ob[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + i * T;
    ob[i] = -1;
    if (b < total) {
      int col, j;
      if (COLFAST) { col = b % C; j = b / C; }
      else { j = b % Lr; col = b / Lr; }
      const int k = j % Ns;
      const int base = col * sc;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[base + (j + r * Lr) * se];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * tstep];
          if (DIR < 0) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      dftR<R, DIR>(v[i]);
      ob[i] = base + ((j / Ns) * Ns * R + k) * se;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (ob[i] >= 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) buf[ob[i] + q * Ns * se] = v[i][q];
    }
  }
  __syncthreads();
}

// Full transform: all stages of the plan.  The caller has synchronised after
// filling buf and tw.
template <int DIR, bool COLFAST, int PER>
__device__ __forceinline__ void fft_run(float2* __restrict__ buf, const float2* __restrict__ tw,
                                        const FftPlan& p, int C, int se, int sc) {
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.rad[s];
    switch (R) {
      case 2: fft_stage<2, DIR, COLFAST, PER>(buf, tw, p.L, C, Ns, se, sc); break;
      case 3: fft_stage<3, DIR, COLFAST, PER>(buf, tw, p.L, C, Ns, se, sc); break;
      case 4: fft_stage<4, DIR, COLFAST, PER>(buf, tw, p.L, C, Ns, se, sc); break;
      case 5: fft_stage<5, DIR, COLFAST, PER>(buf, tw, p.L, C, Ns, se, sc); break;
      case 7: fft_stage<7, DIR, COLFAST, PER>(buf, tw, p.L, C, Ns, se, sc); break;
      default: fft_stage<8, DIR, COLFAST, PER>(buf, tw, p.L, C, Ns, se, sc); break;
    }
    Ns *= R;
  }
}

}  // namespace ddhip
