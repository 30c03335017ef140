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

// cos / sin of 2 pi m / N for any N, evaluated at compile time: the angle is
// reduced to [-pi/4, pi/4] by octant and expanded in Taylor series to double
// precision, so a table entry is exact to float rounding.
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int k = 1; k < 12; ++k) {
    t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 12; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}
// (cos, sin) of 2 pi m / N via the octant o = round(8 m / N) and r = 2 pi m / N - o pi / 4
constexpr double cos_frac(long m, long N, bool want_sin) {
  m %= N;
  if (m < 0) m += N;
  const long o = (8 * m + N / 2) / N;  // nearest eighth of a turn
  const double r = kTwoPi * ((double)m / (double)N) - (double)o * (kTwoPi / 8.0);
  const double c = taylor_cos(r), s = taylor_sin(r);
  const double h = 0.70710678118654752440;
  // rotate (c, s) by o * 45 degrees
  double cr = c, sr = s;
  switch (o & 7) {
    case 0: cr = c; sr = s; break;
    case 1: cr = h * (c - s); sr = h * (c + s); break;
    case 2: cr = -s; sr = c; break;
    case 3: cr = -h * (c + s); sr = h * (c - s); break;
    case 4: cr = -c; sr = -s; break;
    case 5: cr = h * (s - c); sr = -h * (c + s); break;
    case 6: cr = s; sr = -c; break;
    default: cr = h * (c + s); sr = h * (s - c); break;
  }
  return want_sin ? sr : cr;
}
template <int N>
struct TwTable {
  float c[N], s[N];
  constexpr TwTable() : c(), s() {
    for (int m = 0; m < N; ++m) {
      c[m] = (float)cos_frac(m, N, false);
      s[m] = (float)cos_frac(m, N, true);
    }
  }
};
template <int N>
struct Tw {
  static constexpr TwTable<N> t{};
};

// multiply by e^{DIR 2 pi i m / N}, m compile-time after unrolling
template <int N, int DIR>
__device__ __forceinline__ float2 twc(float2 a, int m) {
  float c, s;
  if constexpr (32 % N == 0) {
    const int q = m * (32 / N);
    c = cos32(q);
    s = DIR * sin32(q);
  } else {
    const int q = m % N;
    c = Tw<N>::t.c[q];
    s = DIR * Tw<N>::t.s[q];
  }
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
  } else if constexpr (R == 3) {
    dft3<DIR>(a);
  } else if constexpr (R == 5) {
    dft5<DIR>(a);
  } else if constexpr (R == 16) {
    dft_comp<4, 4, DIR>(a);
  } else if constexpr (R == 32) {
    dft_comp<4, 8, DIR>(a);
  } else if constexpr (R % 8 == 0 && R > 8) {
    dft_comp<R / 8, 8, DIR>(a);
  } else if constexpr (R % 4 == 0 && R > 4) {
    dft_comp<R / 4, 4, DIR>(a);
  } else if constexpr (R % 5 == 0 && R > 5) {
    dft_comp<R / 5, 5, DIR>(a);
  } else if constexpr (R % 3 == 0 && R > 3) {
    dft_comp<R / 3, 3, DIR>(a);
  } else if constexpr (R % 2 == 0 && R > 2) {
    dft_comp<R / 2, 2, DIR>(a);
  } else {
    static_assert(R == 1, "unsupported register DFT size");
  }
}

// Work split of a length L = R1 * R2 transform over R1 threads.  The first
// step is one R2-point DFT per thread; the second step's R2 DFTs of length R1
// go round-robin to the threads: thread t owns k2 = t + R1 i for i < G while
// k2 < R2 (all of them when R1 divides R2).  A thread's register array holds
// VN values.  LDS exchange layout for C sequences: element (row, s) at
// row * P + s, row = k2 * R1 + n1, P = C + 1 (pad against bank conflicts).
template <int R1, int R2>
struct Reg2 {
  static constexpr int L = R1 * R2;
  static constexpr int G = (R2 + R1 - 1) / R1;
  static constexpr int VN = R2 > G * R1 ? R2 : G * R1;
};
// Column tile of block bx (tiles of C complex64 columns, NT tiles) for
// kernels whose tiles are narrower than a 128-byte line (C < 16).  Blocks go
// round-robin over the 8 XCDs by linear id, and each XCD has its own L2, so
// the W = 16/C tiles sharing one line are given to blocks 8 ids apart (same
// XCD, back to back); with `mirror` the group also holds the kx mirror tiles
// NT-1-b, which read the same Green octant entries.  Falls back to the plain
// order (mirror: b and NT-1-b h = ceil(NT/2) apart) when NT is not a multiple
// of 8 times the group size or C >= 16.
template <int C>
__device__ __forceinline__ int xcd_tile(int bx, int NT, bool mirror) {
  constexpr int W = C >= 16 ? 1 : 16 / C;
  const int g = mirror ? 2 * W : W;
  if (W > 1 && NT % (8 * g) == 0) {
    const int m = bx >> 3, r = bx & 7;
    const int q = r + 8 * (m / g), slot = m % g;
    if (!mirror) return W * q + slot;
    return slot < W ? W * q + slot : NT - 1 - (W * q + (slot - W));
  }
  if (!mirror) return bx;
  const int h = (NT + 1) / 2;
  return bx < h ? bx : NT - 1 - (bx - h);
}

// The padded axes have length L = R1 R2 = 2N with data (forward input) and
// kept results (inverse output) in n < N.  Element n = t + R1 n2 of thread t
// (t < R1) lies below N for every t when 2 n2 + 1 < R2 and above it when
// 2 n2 + 1 > R2; only the middle element of an odd R2 depends on t.  With
// the test in this form the zero half of a forward input is a compile-time
// constant (its butterflies fold away) and the discarded half of an inverse
// output is never computed.
template <int R2>
__device__ __forceinline__ bool lower_half(int n2, int n, int N) {
  return (2 * n2 + 1 < R2) || (2 * n2 + 1 == R2 && n < N);
}
// lane-pair layout: element i of lane-parity h holds n = t + Q (h + 2 i), t < Q
template <int Q>
__device__ __forceinline__ bool lower_half_pair(int i, int n, int N) {
  return (2 * i + 1 < Q / 2) || (2 * i < Q / 2 && n < N);
}

template <int R1, int R2>
__device__ __forceinline__ bool reg2_owns(int t, int i) {
  return Reg2<R1, R2>::G * R1 == R2 || t + R1 * i < R2;
}

// LDS exchange index of element (k2, n1) of sequence s.
//  LAY 0 (column passes, lanes run over s): (k2 R1 + n1) P + s, P = C + 1.
//  LAY 1 (row passes, lanes run over n1 when writing and over k2 when
//  reading): s P + k2 (R1 + 1) + n1 with P >= R2 (R1 + 1): consecutive n1 and
//  k2 strides of R1 + 1 both hit distinct banks.
template <int LAY, int R1>
__device__ __forceinline__ int lidx(int k2, int n1, int s, int P) {
  return LAY == 0 ? (k2 * R1 + n1) * P + s : s * P + k2 * (R1 + 1) + n1;
}
// per-sequence stride for LAY 1: R2 (R1 + 1) rounded up so that two
// sequences of one 32-lane half-wave fall on disjoint 4-byte banks (odd multiple of 16)
template <int R1, int R2>
__host__ __device__ constexpr int row_stride() {
  const int base = R2 * (R1 + 1);
  int s = (base + 15) / 16 * 16;
  if ((s / 16) % 2 == 0) s += 16;
  return s;
}

// forward: v[n2] = x[t + R1 n2] (n2 < R2) in; v[i R1 + k1] = X[(t + R1 i) + R2 k1]
// out for the groups i the thread owns
template <int R1, int R2, int DIR, int LAY = 0>
__device__ __forceinline__ void fft2_fwd(float2* v, float2* __restrict__ lds, const float2* __restrict__ twL, int t,
                                         int s, int P) {
  dft_reg<R2, DIR>(v);
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    float2 w = twL[t * k2];
    if (DIR < 0) w.y = -w.y;
    lds[lidx<LAY, R1>(k2, t, s, P)] = k2 == 0 ? v[0] : cmul(v[k2], w);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < Reg2<R1, R2>::G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
    const int k2 = t + R1 * i;
    float2 u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[lidx<LAY, R1>(k2, n1, s, P)];
    dft_reg<R1, DIR>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[i * R1 + k1] = u[k1];
  }
}

// inverse of the forward distribution: v[i R1 + k1] = Y[(t + R1 i) + R2 k1] in
// (owned groups), v[n2] = y[t + R1 n2] out (direction DIR, normally -1)
template <int R1, int R2, int DIR, int LAY = 0>
__device__ __forceinline__ void fft2_inv(float2* v, float2* __restrict__ lds, const float2* __restrict__ twL, int t,
                                         int s, int P) {
#pragma unroll
  for (int i = 0; i < Reg2<R1, R2>::G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
    const int k2 = t + R1 * i;
    float2 u[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) u[k1] = v[i * R1 + k1];
    dft_reg<R1, DIR>(u);
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      float2 w = twL[n1 * k2];
      if (DIR < 0) w.y = -w.y;
      lds[lidx<LAY, R1>(k2, n1, s, P)] = n1 == 0 ? u[0] : cmul(u[n1], w);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) v[k2] = lds[lidx<LAY, R1>(k2, t, s, P)];
  dft_reg<R2, DIR>(v);
}

// Split-exchange form of fft2_fwd: the real and the imaginary parts cross
// LDS in turn, so the exchange buffer holds L*P floats instead of L*P
// float2.  Half the LDS per workgroup doubles the workgroups a CU can hold
// for kernels whose occupancy LDS limits (pass 1); two more barriers.
template <int R1, int R2, int DIR, int LAY = 0>
__device__ __forceinline__ void fft2_fwd_s(float2* v, float* __restrict__ lds, const float2* __restrict__ twL, int t,
                                           int s, int P) {
  constexpr int G = Reg2<R1, R2>::G;
  dft_reg<R2, DIR>(v);
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) {
    float2 w = twL[t * k2];
    if (DIR < 0) w.y = -w.y;
    v[k2] = cmul(v[k2], w);
  }
  float re[G * R1];
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) lds[lidx<LAY, R1>(k2, t, s, P)] = v[k2].x;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) re[i * R1 + n1] = lds[lidx<LAY, R1>(t + R1 * i, n1, s, P)];
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) lds[lidx<LAY, R1>(k2, t, s, P)] = v[k2].y;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
    float2 u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = make_float2(re[i * R1 + n1], lds[lidx<LAY, R1>(t + R1 * i, n1, s, P)]);
    dft_reg<R1, DIR>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[i * R1 + k1] = u[k1];
  }
}

// Split-exchange form of fft2_inv (real and imaginary parts cross LDS in turn).
template <int R1, int R2, int DIR, int LAY = 0>
__device__ __forceinline__ void fft2_inv_s(float2* v, float* __restrict__ lds, const float2* __restrict__ twL, int t,
                                           int s, int P) {
  constexpr int G = Reg2<R1, R2>::G;
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
    const int k2 = t + R1 * i;
    float2 u[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) u[k1] = v[i * R1 + k1];
    dft_reg<R1, DIR>(u);
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      float2 w = twL[n1 * k2];
      if (DIR < 0) w.y = -w.y;
      v[i * R1 + n1] = cmul(u[n1], w);
    }
  }
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) lds[lidx<LAY, R1>(t + R1 * i, n1, s, P)] = v[i * R1 + n1].x;
  }
  __syncthreads();
  float re[R2];
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) re[k2] = lds[lidx<LAY, R1>(k2, t, s, P)];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) lds[lidx<LAY, R1>(t + R1 * i, n1, s, P)] = v[i * R1 + n1].y;
  }
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < R2; ++k2) v[k2] = make_float2(re[k2], lds[lidx<LAY, R1>(k2, t, s, P)]);
  dft_reg<R2, DIR>(v);
}

// e^{DIR 2 pi i q / M} for M = 2, 3
template <int M, int DIR>
__device__ __forceinline__ float2 root_m(int q) {
  if (M == 2) return make_float2(q ? -1.f : 1.f, 0.f);
  constexpr float h = 0.8660254037844386f;  // sqrt(3) / 2
  return q == 0 ? make_float2(1.f, 0.f) : make_float2(-0.5f, ((q == 1) == (DIR > 0)) ? h : -h);
}
// output m of an M-point DFT (sign DIR) of z[0..M)
template <int M, int DIR>
__device__ __forceinline__ float2 dft_m_out(const float2* z, int m) {
  float2 r = z[0];
#pragma unroll
  for (int q = 1; q < M; ++q) r = cadd(r, cmul(root_m<M, DIR>((q * m) % M), z[q]));
  return r;
}

// ---- lane-pair split DFTs: a Q-point DFT (Q even) held by two adjacent
// lanes (h = lane & 1), Q/2 values each; one radix-2 step crosses the pair
// through a lane shuffle.
// DIT form: a[j] = x[h + 2j] in, a[j] = X[(Q/2) h + j] out.
template <int Q, int DIR>
__device__ __forceinline__ void dit_pair(float2* a, int h) {
  constexpr int H = Q / 2;
  dft_reg<H, DIR>(a);
  const float sg = h ? -1.f : 1.f;
#pragma unroll
  for (int m = 0; m < H; ++m) {
    const float2 tw = twc<Q, DIR>(a[m], m);
    const float2 z = h ? tw : a[m];  // lane 0: Y0[m], lane 1: w^m Y1[m]
    const float2 b = make_float2(__shfl_xor(z.x, 1), __shfl_xor(z.y, 1));
    a[m] = make_float2(b.x + sg * z.x, b.y + sg * z.y);  // Y0 + w Y1 | Y0 - w Y1
  }
}
// DIF form: a[j] = x[(Q/2) h + j] in, a[j] = X[h + 2j] out.
template <int Q, int DIR>
__device__ __forceinline__ void dif_pair(float2* a, int h) {
  constexpr int H = Q / 2;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const float2 b = make_float2(__shfl_xor(a[j].x, 1), __shfl_xor(a[j].y, 1));
    const float2 sum = make_float2(a[j].x + b.x, a[j].y + b.y);         // lane 0: x[j] + x[j+H]
    const float2 dif = twc<Q, DIR>(make_float2(b.x - a[j].x, b.y - a[j].y), j);  // lane 1: (x[j] - x[j+H]) w^j
    a[j] = h ? dif : sum;
  }
  dft_reg<H, DIR>(a);
}

// LDS index for the lane-pair transform of L = Q^2 points over C columns:
// element (k2, n1) of column s; the two halves of k2 sit side by side in a
// padded row of P2 slots so that a lane pair (h = 0, 1) hits disjoint banks.
template <int Q>
__device__ __forceinline__ int pidx(int k2, int n1, int s, int C, int P2) {
  constexpr int H = Q / 2;
  return ((k2 % H) * Q + n1) * P2 + (k2 / H) * C + s;
}
__host__ __device__ constexpr int pair_stride(int C) {
  int p = (2 * C + 15) / 16 * 16;
  if ((p / 16) % 2 == 0) p += 16;
  return p;
}

// forward two-level transform with lane pairs (virtual thread t of Q per
// column, lane half h): v[j] = x[t + Q (h + 2j)] in,
// v[j] = X[t + Q ((Q/2) h + j)] out
template <int Q, int DIR>
__device__ __forceinline__ void fft2p_fwd(float2* v, float2* __restrict__ lds, const float2* __restrict__ twL, int t,
                                          int h, int s, int C, int P2) {
  constexpr int H = Q / 2;
  dit_pair<Q, DIR>(v, h);
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const int k2 = H * h + j;
    float2 w = twL[t * k2];
    if (DIR < 0) w.y = -w.y;
    lds[pidx<Q>(k2, t, s, C, P2)] = cmul(v[j], w);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < H; ++j) v[j] = lds[pidx<Q>(t, h + 2 * j, s, C, P2)];
  dit_pair<Q, DIR>(v, h);
}

// inverse of fft2p_fwd's distribution: v[j] = Y[t + Q ((Q/2) h + j)] in,
// v[j] = y[t + Q (h + 2j)] out
template <int Q, int DIR>
__device__ __forceinline__ void fft2p_inv(float2* v, float2* __restrict__ lds, const float2* __restrict__ twL, int t,
                                          int h, int s, int C, int P2) {
  constexpr int H = Q / 2;
  dif_pair<Q, DIR>(v, h);
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const int n1 = h + 2 * j;
    float2 w = twL[n1 * t];
    if (DIR < 0) w.y = -w.y;
    lds[pidx<Q>(t, n1, s, C, P2)] = cmul(v[j], w);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < H; ++j) v[j] = lds[pidx<Q>(H * h + j, t, s, C, P2)];
  dif_pair<Q, DIR>(v, h);
}

// ---- register Stockham stages
// A length L = R0 R1 R2 R3 transform (R3 = 1: three stages) held by T = L / V
// threads per sequence, V values per thread.  Forward stage s of radix R and
// span P = R0 ... R_{s-1}: butterfly i = j + T b (b < V / R) of thread j
// gathers u_r = d[i + r L / R] (r < R), multiplies u_r by w_{PR}^{r (i mod P)},
// does the R-point DFT and scatters y_r to (i - i mod P) R + i mod P + r P.
// After the last stage d is the transform in natural order (self-sorting
// Stockham).  Stage 0 gathers from registers the caller loaded, the stages
// between exchange through LDS, and the last stage's outputs stay in
// registers: thread j ends with X[i + r L / R_last] at v[b R_last + r].  The
// inverse runs the transposed stages in reverse order (gather from the
// scatter positions, conjugate DFT, conjugate twiddle, scatter to the gather
// positions) and leaves y where stage 0 gathered from.  Half the registers
// of the two-level form for the same length, one more LDS exchange.
constexpr int sk_rad(int s, int r0, int r1, int r2, int r3) { return s == 0 ? r0 : s == 1 ? r1 : s == 2 ? r2 : r3; }
template <int V_, int R0, int R1, int R2, int R3 = 1>
struct SkPlan {
  static constexpr int V = V_, L = R0 * R1 * R2 * R3, T = L / V, S = R3 > 1 ? 4 : 3;
  static constexpr int rad(int s) { return sk_rad(s, R0, R1, R2, R3); }
  static constexpr int span(int s) { return s == 0 ? 1 : span(s - 1) * rad(s - 1); }
  // LDS swizzle shift: stage 0 scatters with stride R0
  static constexpr int SH = R0 == 8 ? 3 : 2;
  static_assert(V % R0 == 0 && V % R1 == 0 && V % R2 == 0 && V % R3 == 0, "every radix divides V");
  static_assert(R0 % 2 == 0, "even first radix: the zero half of a padded input is a compile-time set");
  static_assert(L % 32 == 0, "the LDS swizzle permutes within 32-point blocks");
};
// LDS slot of point idx of sequence col: C sequences interleaved, the point
// index XOR-swizzled within 32-point blocks so that the stride-R0 scatter of
// stage 0 and the unit-stride gathers of a wave both hit distinct banks.
template <int C, int SH>
__device__ __forceinline__ int sk_lds(int idx, int col) {
  constexpr int M = C >= 32 ? 0 : 32 / C - 1;
  return C * (idx ^ ((idx >> SH) & M)) + col;
}
template <class PL, int s>
__device__ __forceinline__ int sk_scatter(int i, int r) {
  constexpr int R = PL::rad(s), P = PL::span(s);
  const int k = i % P;
  return (i - k) * R + k + r * P;
}
template <class PL, int s>
__device__ __forceinline__ int sk_gather(int i, int r) {
  return i + r * (PL::L / PL::rad(s));
}
// stage-s twiddles and DFTs of thread j's butterflies (forward order: twiddle, DFT)
template <class PL, int s, int DIR>
__device__ __forceinline__ void sk_fwd_bfly(float2* v, const float2* __restrict__ twL, int j) {
  constexpr int R = PL::rad(s), P = PL::span(s), NB = PL::V / R, TS = PL::L / (P * R);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if constexpr (P > 1) {
      const int k = (j + PL::T * b) % P;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = twL[r * k * TS];
        if (DIR < 0) w.y = -w.y;
        v[b * R + r] = cmul(v[b * R + r], w);
      }
    }
    dft_reg<R, DIR>(v + b * R);
  }
}
// transposed stage s (DFT, then twiddle; DIR is the inverse's sign)
template <class PL, int s, int DIR>
__device__ __forceinline__ void sk_inv_bfly(float2* v, const float2* __restrict__ twL, int j) {
  constexpr int R = PL::rad(s), P = PL::span(s), NB = PL::V / R, TS = PL::L / (P * R);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    dft_reg<R, DIR>(v + b * R);
    if constexpr (P > 1) {
      const int k = (j + PL::T * b) % P;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = twL[r * k * TS];
        if (DIR < 0) w.y = -w.y;
        v[b * R + r] = cmul(v[b * R + r], w);
      }
    }
  }
}
// LDS exchange buffer of an exchange (K sequences of L points per column,
// K L C slots): with NBUF = 2 consecutive exchanges alternate between two
// buffers, so a write needs no barrier of its own (every thread passed the
// previous exchange's barrier after its reads of the buffer it now
// overwrites).
// RC: the stage functions recompute the exchange addresses from fresh
// copies of the thread indices at every exchange instead of letting the
// compiler keep them live across transforms (fewer registers, more VALU).
template <class PL, int C, int NBUF, int K, bool RC = false>
struct SkLds {
  float2* xb;
  int e = 0;
  __device__ __forceinline__ float2* next() {
    if (NBUF == 1) {
      __syncthreads();
      return xb;
    }
    float2* b = xb + (e & 1) * (K * PL::L * C);
    ++e;
    return b;
  }
};
// K sequences advance stage by stage together (v[q V .. q V + V) is
// sequence q), so one barrier serves all K and a stage's twiddles are
// read once for the K.
// forward stages s .. S-1 (stage s-1's butterflies done, its outputs in v)
template <class PL, int s, int C, int NBUF, int K, int DIR, bool RC>
__device__ __forceinline__ void sk_fwd_from(float2* v, SkLds<PL, C, NBUF, K, RC>& lx, const float2* __restrict__ twL,
                                            int j, int col) {
  if constexpr (s < PL::S) {
    constexpr int V = PL::V, LC = PL::L * C;
    constexpr int Rp = PL::rad(s - 1), NBp = V / Rp, R = PL::rad(s), NB = V / R;
    float2* xb = lx.next();
    // fresh copies of the thread indices: the exchange addresses are then
    // recomputed here instead of kept live from an earlier transform
    if constexpr (RC) asm volatile("" : "+v"(j), "+v"(col));
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int b = 0; b < NBp; ++b)
#pragma unroll
        for (int r = 0; r < Rp; ++r)
          xb[q * LC + sk_lds<C, PL::SH>(sk_scatter<PL, s - 1>(j + PL::T * b, r), col)] = v[q * V + b * Rp + r];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[q * V + b * R + r] = xb[q * LC + sk_lds<C, PL::SH>(sk_gather<PL, s>(j + PL::T * b, r), col)];
#pragma unroll
    for (int q = 0; q < K; ++q) sk_fwd_bfly<PL, s, DIR>(v + q * V, twL, j);
    sk_fwd_from<PL, s + 1, C, NBUF, K, DIR>(v, lx, twL, j, col);
  }
}
// whole forward transform of K sequences: stage-0 gathers in, last stage's outputs out
template <class PL, int C, int NBUF, int K, int DIR, bool RC>
__device__ __forceinline__ void sk_forward(float2* v, SkLds<PL, C, NBUF, K, RC>& lx, const float2* __restrict__ twL, int j,
                                           int col) {
#pragma unroll
  for (int q = 0; q < K; ++q) sk_fwd_bfly<PL, 0, DIR>(v + q * PL::V, twL, j);
  sk_fwd_from<PL, 1, C, NBUF, K, DIR>(v, lx, twL, j, col);
}
// transposed stages s .. 0 (v holds stage s's scatter positions)
template <class PL, int s, int C, int NBUF, int K, int DIR, bool RC>
__device__ __forceinline__ void sk_inv_from(float2* v, SkLds<PL, C, NBUF, K, RC>& lx, const float2* __restrict__ twL,
                                            int j, int col) {
  constexpr int V = PL::V, LC = PL::L * C;
#pragma unroll
  for (int q = 0; q < K; ++q) sk_inv_bfly<PL, s, DIR>(v + q * V, twL, j);
  if constexpr (s > 0) {
    constexpr int R = PL::rad(s), NB = V / R, Rn = PL::rad(s - 1), NBn = V / Rn;
    float2* xb = lx.next();
    if constexpr (RC) asm volatile("" : "+v"(j), "+v"(col));
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
          xb[q * LC + sk_lds<C, PL::SH>(sk_gather<PL, s>(j + PL::T * b, r), col)] = v[q * V + b * R + r];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
      for (int b = 0; b < NBn; ++b)
#pragma unroll
        for (int r = 0; r < Rn; ++r)
          v[q * V + b * Rn + r] = xb[q * LC + sk_lds<C, PL::SH>(sk_scatter<PL, s - 1>(j + PL::T * b, r), col)];
    sk_inv_from<PL, s - 1, C, NBUF, K, DIR>(v, lx, twL, j, col);
  }
}
// whole inverse transform of K sequences: the forward's output distribution
// in, stage-0 gather positions out
template <class PL, int C, int NBUF, int K, int DIR, bool RC>
__device__ __forceinline__ void sk_inverse(float2* v, SkLds<PL, C, NBUF, K, RC>& lx, const float2* __restrict__ twL, int j,
                                           int col) {
  sk_inv_from<PL, PL::S - 1, C, NBUF, K, DIR>(v, lx, twL, j, col);
}

}  // namespace ddhip
