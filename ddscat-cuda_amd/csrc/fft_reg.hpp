// Register-resident two-level FFTs for gfx950.
//
// A length L = R1 * R2 transform of one sequence is shared by R1 threads.
// With n = n1 + R1 n2 and k = k2 + R2 k1,
//   X[k2 + R2 k1] = sum_n1 W_R1^{n1 k1} W_L^{n1 k2} sum_n2 x[n1 + R1 n2] W_R2^{n2 k2}.
// Forward: thread n1 holds x[n1 + R1 n2] (n2 < R2) in registers, does the
// R2-point DFT, multiplies by W_L^{n1 k2}, writes to LDS; after one barrier
// thread t holds, for each owned k2 = t + R1 i (i < R2/R1), the R1 values
// n1 and does the R1-point DFT, ending with X[k2 + R2 k1] in registers.
// Inverse runs the same two steps backwards from that distribution and ends
// with y[n1 + R1 n2] on thread n1.  One LDS exchange per transform; the
// R-point DFTs are fully unrolled with compile-time twiddles.
#pragma once
#include "common.hpp"
#include "fft_device.hpp"

namespace ddhip {

// cos(2 pi m / 32), m any integer
__host__ __device__ constexpr float cos32(int m) {
  constexpr float C[9] = {1.0f,
                          0.98078528040323044913f,
                          0.92387953251128675613f,
                          0.83146961230254523708f,
                          0.70710678118654752440f,
                          0.55557023301960222474f,
                          0.38268343236508977173f,
                          0.19509032201612826785f,
                          0.0f};
  m &= 31;
  return m <= 8 ? C[m] : m <= 16 ? -C[16 - m] : m <= 24 ? -C[m - 16] : C[32 - m];
}
__host__ __device__ constexpr float sin32(int m) { return cos32(m - 8); }

// multiply by e^{DIR 2 pi i m / N} for N dividing 32, m compile-time after unrolling
template <int N, int DIR>
__device__ __forceinline__ float2 twc(float2 a, int m) {
  const int q = m * (32 / N);
  const float c = cos32(q), s = DIR * sin32(q);
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

template <int R, int DIR>
__device__ __forceinline__ void dft_reg(float2* a);

// N = P * Q composition: n = n1 + P n2, k = k2 + Q k1
template <int P, int Q, int DIR>
__device__ __forceinline__ void dft_comp(float2* a) {
  constexpr int N = P * Q;
  float2 b[N];
#pragma unroll
  for (int n1 = 0; n1 < P; ++n1) {
    float2 t[Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) t[n2] = a[n1 + P * n2];
    dft_reg<Q, DIR>(t);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) b[n1 * Q + k2] = (n1 * k2 == 0) ? t[k2] : twc<N, DIR>(t[k2], n1 * k2);
  }
#pragma unroll
  for (int k2 = 0; k2 < Q; ++k2) {
    float2 t[P];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) t[n1] = b[n1 * Q + k2];
    dft_reg<P, DIR>(t);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) a[k2 + Q * k1] = t[k1];
  }
}

template <int R, int DIR>
__device__ __forceinline__ void dft_reg(float2* a) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<DIR>(a);
  } else if constexpr (R == 4) {
    dft4<DIR>(a);
  } else if constexpr (R == 8) {
    dft8<DIR>(a);
  } else if constexpr (R == 16) {
    dft_comp<4, 4, DIR>(a);
  } else if constexpr (R == 32) {
    dft_comp<4, 8, DIR>(a);
  } else {
    static_assert(R == 1, "unsupported register DFT size");
  }
}

// LDS exchange layout for C sequences: element (row, s) at row * P + s,
// row = k2 * R1 + n1, P = C + 1 (pad against bank conflicts)
template <int R1, int R2>
struct Reg2 {
  static constexpr int L = R1 * R2;
  static constexpr int G = R2 / R1;  // k2 groups per thread in the second step
  static_assert(R2 % R1 == 0, "R2 must be a multiple of R1");
};

// forward: v[n2] = x[t + R1 n2] in, v[i * R1 + k1] = X[(t + R1 i) + R2 k1] out
template <int R1, int R2, int DIR>
__device__ __forceinline__ void fft2_fwd(float2* v, float2* __restrict__ lds, const float2* __restrict__ twL, int t,
                                         int s, int P) {
  constexpr int L = R1 * R2;
  dft_reg<R2, DIR>(v);
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    float2 w = twL[t * k2];
    if (DIR < 0) w.y = -w.y;
    lds[(k2 * R1 + t) * P + s] = cmul(v[k2], w);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R2 / R1; ++i) {
    const int k2 = t + R1 * i;
    float2 u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[(k2 * R1 + n1) * P + s];
    dft_reg<R1, DIR>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[i * R1 + k1] = u[k1];
  }
  (void)L;
}

// inverse of the forward distribution: v[i * R1 + k1] = Y[(t + R1 i) + R2 k1]
// in, v[n2] = y[t + R1 n2] out (direction DIR, normally -1)
template <int R1, int R2, int DIR>
__device__ __forceinline__ void fft2_inv(float2* v, float2* __restrict__ lds, const float2* __restrict__ twL, int t,
                                         int s, int P) {
#pragma unroll
  for (int i = 0; i < R2 / R1; ++i) {
    const int k2 = t + R1 * i;
    float2 u[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) u[k1] = v[i * R1 + k1];
    dft_reg<R1, DIR>(u);
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      float2 w = twL[n1 * k2];
      if (DIR < 0) w.y = -w.y;
      lds[(k2 * R1 + n1) * P + s] = cmul(u[n1], w);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) v[k2] = lds[(k2 * R1 + t) * P + s];
  dft_reg<R2, DIR>(v);
}

}  // namespace ddhip
