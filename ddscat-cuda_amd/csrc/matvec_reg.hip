// FFT3D MATVEC passes built on the register-resident two-level FFT
// (fft_reg.hpp) for the padded axis lengths 2N listed in kShapes below.
// Same five passes, buffers and operator as matvec.hip; the difference is
// that a transform keeps its data in registers, crosses LDS once, and pass 3
// does the 3x3 Green multiply in registers between its forward and inverse
// z-transforms.
//
// Compact vectors: when every lattice row's occupied sites form one interval
// [rowlo, rowlo + count) (true for convex targets), the compact index of
// (row g, ix) is rowptr[g] + ix - rowlo[g] and no per-site index is read;
// otherwise a full-site -> compact map (qmap, -1 for void sites) is used.
#include "state.hpp"
#include "fft_reg.hpp"
#include "solver_common.hpp"
#include <type_traits>

namespace ddhip {


struct SiteMap {
  const int* rowptr;
  const int* rowlo;
  const int* qmap;
  int contiguous;
};

// compact index of site (row g, ix), or -1 when void
__device__ __forceinline__ int compact_index(const SiteMap& m, int g, int ix, int nx) {
  if (m.contiguous) {
    const int lo = m.rowlo[g], q0 = m.rowptr[g], cnt = m.rowptr[g + 1] - q0;
    const int d = ix - lo;
    return (d >= 0 && d < cnt) ? q0 + d : -1;
  }
  return m.qmap[(size_t)g * nx + ix];
}

// ------------------------------------------------------------------ pass 1
template <int R1, int R2, int C>
__global__ __launch_bounds__(R1 * C) void kr_pass1(const float2* __restrict__ x, float2* __restrict__ a1, SiteMap sm,
                                               const float2* __restrict__ twg, int nx, int nrows, int nat, int nat0,
                                               int compact, int cj, BatchArgs ba) {
  constexpr int L = R1 * R2, RT = R1 * C, P = row_stride<R1, R2>(), G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float buf[C * P];  // split exchange, row layout
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  x += rb * ba.xs;
  a1 += rb * ba.a1s;
  const int t = threadIdx.x % R1, s = threadIdx.x / R1;
  const int g = blockIdx.x * C + s, c = blockIdx.y;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  float2 v[VN];
  const bool live = g < nrows;
  // site of each unpadded input n = t + R1 n2 (n2 < H), resolved with the
  // layout test outside the element loop; the loads then run unconditionally
  // (an empty site or a row past the end reads entry 0 and is zeroed), so
  // they are in flight together instead of waiting on per-element branches
  constexpr int H = (R2 + 1) / 2;
  size_t o[H];
  bool ok[H];
  if (!compact) {
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) {
      const int n = t + R1 * n2;
      ok[n2] = live && lower_half<R2>(n2, n, nx);
      o[n2] = ok[n2] ? (size_t)g * nx + n : 0;
    }
  } else if (sm.contiguous) {
    int lo = 0, q0 = 0, cnt = 0;
    if (live) { lo = sm.rowlo[g]; q0 = sm.rowptr[g]; cnt = sm.rowptr[g + 1] - q0; }
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) {
      const int n = t + R1 * n2, d = n - lo;
      ok[n2] = live && lower_half<R2>(n2, n, nx) && d >= 0 && d < cnt;
      o[n2] = ok[n2] ? (size_t)(q0 + d) : 0;
    }
  } else {
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) {
      const int n = t + R1 * n2;
      const int q = live && lower_half<R2>(n2, n, nx) ? sm.qmap[(size_t)g * nx + n] : -1;
      ok[n2] = q >= 0;
      o[n2] = ok[n2] ? (size_t)q : 0;
    }
  }
  const float2* xc = x + (size_t)c * (compact ? nat0 : nat);
#pragma unroll
  for (int n2 = 0; n2 < H; ++n2) {
    const float2 val = xc[o[n2]];
    v[n2] = ok[n2] ? val : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int n2 = H; n2 < R2; ++n2) v[n2] = make_float2(0.f, 0.f);
  if (cj) {
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) v[n2].y = -v[n2].y;
  }
  __syncthreads();
  fft2_fwd_s<R1, R2, +1, 1>(v, buf, twL, t, s, P);
  if (live) {
    float2* out = a1 + ((size_t)c * nrows + g) * L;
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1)
        if (reg2_owns<R1, R2>(t, i)) out[(t + R1 * i) + R2 * k1] = v[i * R1 + k1];
  }
}

// ------------------------------------------------------------ passes 2, 4
// element j of component c of a MATVEC input as BatchArgs' gather fields
// name it, or 0 when !take (the SLICES z-forward pass reads x itself, compact
// or conjugated).  MAP: compact input through the site map.  Branch-free: a
// lane that takes nothing loads site 0 and discards it, so a thread's
// site-map loads, then its x loads, issue back to back.
template <bool MAP>
__device__ __forceinline__ float2 gathered(const BatchArgs& ba, int c, size_t j, bool take) {
  j = take ? j : 0;
  float2 v;
  if constexpr (MAP) {
    const int q = ba.gmap[j];
    take = take && q >= 0;
    v = ba.gsrc[(size_t)c * ba.gcs + (q >= 0 ? q : 0)];
  } else {
    v = ba.gsrc[(size_t)c * ba.gcs + j];
  }
  if (!take) v = make_float2(0.f, 0.f);
  if (ba.gconj) v.y = -v.y;
  return v;
}

// in [3][nz][nin][gx] -> out [3][nz][nout][gx] along the middle axis.
// GA (DIR = +1, nz = 1: the SLICES z pass): input element n of column kx is
// gathered(ba, c, n gx + kx) instead of in[c ins + n gx + kx].
template <int R1, int R2, int C, int DIR, bool GA = false>
__global__ __launch_bounds__(R1 * C) void kr_passy(const float2* __restrict__ in, float2* __restrict__ out,
                                               const float2* __restrict__ twg, int gx, int nin, int nout, int nz,
                                               size_t ins, size_t outs, BatchArgs ba) {
  constexpr int L = R1 * R2, RT = R1 * C, P = C + 1, G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
  const int rb = blockIdx.z / 3;
  if (ba.active && !ba.active[rb]) return;
  in += rb * (DIR > 0 ? ba.a1s : ba.a2s);
  out += rb * (DIR > 0 ? ba.a2s : ba.a1s);
  const int s = threadIdx.x % C, t = threadIdx.x / C;
  const int kx = xcd_tile<C>((int)blockIdx.x, (int)gridDim.x, false) * C + s, iz = blockIdx.y, c = blockIdx.z % 3;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  const float2* src = in + ((size_t)c * nz + iz) * ins + kx;
  float2* dst = out + ((size_t)c * nz + iz) * outs + kx;
  float2 v[VN];
  if (DIR > 0) {
    auto load = [&](auto mapc) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int n = t + R1 * n2;
        const bool take = live && lower_half<R2>(n2, n, nin);
        if constexpr (GA)
          v[n2] = gathered<decltype(mapc)::value>(ba, c, (size_t)n * gx + kx, take);
        else
          v[n2] = take ? src[(size_t)n * gx] : make_float2(0.f, 0.f);
      }
    };
    if (GA && ba.gmap)
      load(std::true_type{});
    else
      load(std::false_type{});
    __syncthreads();
    fft2_fwd<R1, R2, +1>(v, buf, twL, t, s, P);
    if (live) {
#pragma unroll
      for (int i = 0; i < G; ++i)
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1)
          if (reg2_owns<R1, R2>(t, i)) dst[(size_t)((t + R1 * i) + R2 * k1) * gx] = v[i * R1 + k1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1)
        v[i * R1 + k1] = (live && reg2_owns<R1, R2>(t, i)) ? src[(size_t)((t + R1 * i) + R2 * k1) * gx]
                                                           : make_float2(0.f, 0.f);
    __syncthreads();
    fft2_inv<R1, R2, -1>(v, buf, twL, t, s, P);
    if (live) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int n = t + R1 * n2;
        if (lower_half<R2>(n2, n, nout)) dst[(size_t)n * gx] = v[n2];
      }
    }
  }
}

// ------------------------------------------------ passes 2, 4, L = M * 256
// Column pass of kr_passy for L = M * 256 (M = 2, 3) by the decimation of
// kr_pass3m: one component per thread, 16 values each, exchanges of real and
// imaginary parts in turn (256-point FFTs and the M-point combine), so a
// 768-point column tile of C = 8 needs about 36 KB of LDS instead of the
// 61 KB of the two-level 32 x 24 shape.  DIR = +1: nin (<= L/2, pruned)
// inputs -> L spectral outputs; DIR = -1: L spectral inputs -> nout (<= L/2).
// GA as in kr_passy.
template <int M, int C, int DIR, bool GA = false>
__global__ __launch_bounds__(M * 16 * C) void kr_passym(const float2* __restrict__ in, float2* __restrict__ out,
                                                    const float2* __restrict__ twg, int gx, int nin, int nout, int nz,
                                                    size_t ins, size_t outs, BatchArgs ba) {
  constexpr int R = 16, S = R * R, L = M * S, GT = R * C, RT = M * GT;
  constexpr int PS = M * C + 1;  // 256-point exchange: M C sequences per workgroup (floats)
  constexpr int PC = C + 1;      // M-point exchange layout (u S + k1) PC + s, u < M (floats)
  constexpr int XB = S * PS > M * S * PC ? S * PS : M * S * PC;
  __shared__ float2 twL[L];
  __shared__ float2 twS[S];
  __shared__ float xf[XB];
  const int rb = blockIdx.z / 3;
  if (ba.active && !ba.active[rb]) return;
  in += rb * (DIR > 0 ? ba.a1s : ba.a2s);
  out += rb * (DIR > 0 ? ba.a2s : ba.a1s);
  const int m = threadIdx.x / GT, lt = threadIdx.x % GT;
  const int s = lt % C, t = lt / C, sv = m * C + s;
  const int kx = xcd_tile<C>((int)blockIdx.x, (int)gridDim.x, false) * C + s, iz = blockIdx.y, c = blockIdx.z % 3;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  for (int i = threadIdx.x; i < S; i += RT) twS[i] = twg[M * i];
  const bool live = kx < gx;
  const float2* src = in + ((size_t)c * nz + iz) * ins + kx;
  float2* dst = out + ((size_t)c * nz + iz) * outs + kx;
  float2 v[R];
  // M-point DFT across the groups, through LDS in two halves: the real
  // parts cross first and are held (M R floats), then each output reads the
  // imaginary parts at its k1 and combines (DIR: sign of the M-point DFT;
  // the inverse then applies the conjugate twiddle w_L^{-m k1})
  auto combine = [&](auto dirc) {
    constexpr int D = decltype(dirc)::value;
    float zr[M][R];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) xf[(m * S + t + R * i) * PC + s] = v[i].x;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int u = 0; u < M; ++u) zr[u][i] = xf[(u * S + t + R * i) * PC + s];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) xf[(m * S + t + R * i) * PC + s] = v[i].y;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) {
      float2 z[M];
#pragma unroll
      for (int u = 0; u < M; ++u) z[u] = make_float2(zr[u][i], xf[(u * S + t + R * i) * PC + s]);
      v[i] = dft_m_out<M, D>(z, m);
      if (D < 0) {
        float2 w = twL[m * (t + R * i)];
        w.y = -w.y;
        v[i] = cmul(v[i], w);
      }
    }
  };
  if (DIR > 0) {
    auto load = [&](auto mapc) {
#pragma unroll
      for (int n2 = 0; n2 < R; ++n2) {
        const int n = M * (t + R * n2) + m;
        const bool take = live && n2 < R / 2 && n < nin;
        if constexpr (GA)
          v[n2] = gathered<decltype(mapc)::value>(ba, c, (size_t)n * gx + kx, take);
        else
          v[n2] = take ? src[(size_t)n * gx] : make_float2(0.f, 0.f);
      }
    };
    if (GA && ba.gmap)
      load(std::true_type{});
    else
      load(std::false_type{});
    __syncthreads();
    fft2_fwd_s<R, R, +1>(v, xf, twS, t, sv, PS);  // v[i] = Y_m[t + 16 i]
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = cmul(v[i], twL[m * (t + R * i)]);
    combine(std::integral_constant<int, +1>{});  // v[i] = X[t + 16 i + 256 m]
    if (live) {
#pragma unroll
      for (int i = 0; i < R; ++i) dst[(size_t)(t + R * i + S * m) * gx] = v[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = live ? src[(size_t)(t + R * i + S * m) * gx] : make_float2(0.f, 0.f);
    combine(std::integral_constant<int, -1>{});  // v[i] = W_m[t + 16 i]
    __syncthreads();
    fft2_inv_s<R, R, -1>(v, xf, twS, t, sv, PS);  // v[n2] = x[M (t + 16 n2) + m]
    if (live) {
#pragma unroll
      for (int n2 = 0; n2 < R / 2; ++n2) {
        const int n = M * (t + R * n2) + m;
        if (n < nout) dst[(size_t)n * gx] = v[n2];
      }
    }
  }
}

// ------------------------------------------------------------------ pass 3
// One column of the Green convolution pass.  FFT3D: column (kx, ky) of a2,
// transform along z; SLICES: column (kx, kz = k0 + plane) of a slice batch
// [planes][3][NY][2NX], transform along y.  Element n of component c is at
// col[n es + c cs]; the transform has nf nonzero inputs and kept outputs.
// The Green entry of point k along the transform is at pb + step min(k, mf - k);
// the off-diagonal signs are c12, c13, c23 times the parity of k
// (k <= tf) for the components that involve the transform axis
// (octant mirror rules for IPBC=0, cu:109-115 and cu:166-186; IPBC=1 has
// no mirrors and all signs +1).
struct P3Col {
  float2* col;
  size_t es, cs;
  int nf;
  size_t pb, step;
  int mf, tf;
  float c12, c13, c23;
};
template <bool SLY>
__device__ __forceinline__ P3Col p3_col(float2* a2, int kx, int row, int nx, int ny, int nz, int gx, size_t plane,
                                        const GreenGeo& gg) {
  P3Col q;
  const int ir = kx < gx ? min(kx, gg.mx - kx) : 0;
  const float sx = kx <= gg.tx ? 1.f : -1.f;
  if (!SLY) {  // row = ky
    const int jr = min(row, gg.my - row);
    const float sy = row <= gg.ty ? 1.f : -1.f;
    q.col = a2 + (size_t)row * gx + kx;
    q.es = plane;
    q.cs = (size_t)nz * plane;
    q.nf = nz;
    q.pb = (size_t)ir + (size_t)gg.ox * jr;
    q.step = (size_t)gg.ox * gg.oy;
    q.mf = gg.mz; q.tf = gg.tz;
    q.c12 = sx * sy; q.c13 = sx; q.c23 = sy;
  } else {  // row = kz; a2 points at plane 0 of the batch, plane = index of the batch's first kz
    const int b = row - (int)plane, kr = min(row, gg.mz - row);
    const float sz = row <= gg.tz ? 1.f : -1.f;
    q.col = a2 + (size_t)b * 3 * ny * gx + kx;
    q.es = gx;
    q.cs = (size_t)ny * gx;
    q.nf = ny;
    q.pb = (size_t)ir + (size_t)gg.ox * gg.oy * kr;
    q.step = gg.ox;
    q.mf = gg.my; q.tf = gg.ty;
    q.c12 = sx; q.c13 = sx * sz; q.c23 = sz;
  }
  return q;
}
// Green multiply of point k along the transform: p0..p2 in, G p out
template <bool SLY>
__device__ __forceinline__ void p3_green(const P3Col& q, const float2* __restrict__ green, int k, float2& p0,
                                         float2& p1, float2& p2) {
  const int kr = min(k, q.mf - k);
  const float sf = k <= q.tf ? 1.f : -1.f;
  const Green6 G = green_at(green, q.pb + q.step * kr);
  const float2 g12 = cscale(G.xy, SLY ? q.c12 * sf : q.c12);
  const float2 g13 = cscale(G.xz, SLY ? q.c13 : q.c13 * sf);
  const float2 g23 = cscale(G.yz, q.c23 * sf);
  const float2 a = p0, b = p1, c = p2;
  p0 = cadd(cadd(cmul(G.xx, a), cmul(g12, b)), cmul(g13, c));
  p1 = cadd(cadd(cmul(g12, a), cmul(G.yy, b)), cmul(g23, c));
  p2 = cadd(cadd(cmul(g13, a), cmul(g23, b)), cmul(G.zz, c));
}
// g0 p0 + g1 p1 + g2 p2 as two FMA chains (6 instructions per part)
__device__ __forceinline__ float2 cdot3(float2 g0, float2 p0, float2 g1, float2 p1, float2 g2, float2 p2) {
  float re = g0.x * p0.x, im = g0.x * p0.y;
  re = fmaf(-g0.y, p0.y, re);
  im = fmaf(g0.y, p0.x, im);
  re = fmaf(g1.x, p1.x, re);
  im = fmaf(g1.x, p1.y, im);
  re = fmaf(-g1.y, p1.y, re);
  im = fmaf(g1.y, p1.x, im);
  re = fmaf(g2.x, p2.x, re);
  im = fmaf(g2.x, p2.y, im);
  re = fmaf(-g2.y, p2.y, re);
  im = fmaf(g2.y, p2.x, im);
  return make_float2(re, im);
}
#ifndef P3S_GREEN_GROUP
#define P3S_GREEN_GROUP 4
#endif
__device__ __forceinline__ Green6 green_at_b(const char* __restrict__ g, unsigned off) {
  const float4 a = *reinterpret_cast<const float4*>(g + off);
  const float4 b = *reinterpret_cast<const float4*>(g + off + 16u);
  const float4 c = *reinterpret_cast<const float4*>(g + off + 32u);
  return Green6{make_float2(a.x, a.y), make_float2(a.z, a.w), make_float2(b.x, b.y),
                make_float2(b.z, b.w), make_float2(c.x, c.y), make_float2(c.z, c.w)};
}
// p3_green with the rows as FMA chains
template <bool SLY>
__device__ __forceinline__ void p3_green_fma(const P3Col& q, const float2* __restrict__ green, int k, float2& p0,
                                             float2& p1, float2& p2) {
  const int kr = min(k, q.mf - k);
  const float sf = k <= q.tf ? 1.f : -1.f;
  const Green6 G = green_at(green, q.pb + q.step * kr);
  const float2 g12 = cscale(G.xy, SLY ? q.c12 * sf : q.c12);
  const float2 g13 = cscale(G.xz, SLY ? q.c13 : q.c13 * sf);
  const float2 g23 = cscale(G.yz, q.c23 * sf);
  const float2 a = p0, b = p1, c = p2;
  p0 = cdot3(G.xx, a, g12, b, g13, c);
  p1 = cdot3(g12, a, G.yy, b, g23, c);
  p2 = cdot3(g13, a, g23, b, G.zz, c);
}
// row of block row index j: FFT3D orders ky as 0, NY, 1, 2NY-1, 2, ... so
// mirror rows (same Green octant entries) run back to back on one XCD
template <bool SLY>
__device__ __forceinline__ int p3_row(int j, int ny, int gy, int k0) {
  if (SLY) return k0 + j;
  return j == 0 ? 0 : j == 1 ? ny : ((j & 1) ? gy - ((j - 2) / 2 + 1) : (j - 2) / 2 + 1);
}

// Pass-3 grid: one dimension of ceil(U/8)*8*B workgroups over the U = NT x GY
// column units (NT column tiles, GY rows) and B right-hand sides.  Workgroups
// go round-robin over the 8 XCDs by linear id, so unit l = 8 s' + x of RHS r
// is workgroup (s' B + r) 8 + x: it runs on XCD x, and the B right-hand sides
// of a unit run back to back there, so the unit's Green octant entries come
// from HBM once and from that XCD's L2 for the other B - 1.  Units keep the
// single-RHS order l = bx + NT by (the mirror locality below).  False for the
// padding workgroups of the last group of eight.
struct P3Grid {
  int nt, units, nrhs;
};
struct P3Unit {
  int bx, by, rb;
};
__device__ __forceinline__ bool p3_unit(const P3Grid& g3, P3Unit& u) {
  const int L = (int)blockIdx.x, x = L & 7, s = L >> 3;
  const int sp = s / g3.nrhs, l = sp * 8 + x;
  if (l >= g3.units) return false;
  u.rb = s - sp * g3.nrhs;
  u.by = l / g3.nt;
  u.bx = l - u.by * g3.nt;
  return true;
}
static P3Grid p3_grid(int gx, int gy, int C, int nrhs) {
  P3Grid g3;
  g3.nt = (gx + C - 1) / C;
  g3.units = g3.nt * gy;
  g3.nrhs = nrhs;
  return g3;
}
static unsigned p3_blocks(const P3Grid& g3) { return (unsigned)(((g3.units + 7) / 8) * 8 * g3.nrhs); }

// PRE > 0: the Green loads of a thread's first PRE points are issued before
// the third component's forward transform, so they arrive while it runs.
template <int R1, int R2, int C, int PRE = 0, int WPE = 1>
__global__ __launch_bounds__(R1 * C) __attribute__((amdgpu_waves_per_eu(WPE))) void kr_pass3(float2* __restrict__ a2, const float2* __restrict__ green,
                                               const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                               int gy, size_t plane, BatchArgs ba, GreenGeo gg, P3Grid g3) {
  constexpr int L = R1 * R2, RT = R1 * C, P = C + 1, G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
  P3Unit u;
  if (!p3_unit(g3, u)) return;
  const int rb = u.rb;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int s = threadIdx.x % C, t = threadIdx.x / C;
  // Unit order puts Green-sharing blocks on one XCD: ky runs 0, ny, 1, gy-1,
  // 2, gy-2, ... so mirror rows are one unit row apart, and column tile b and
  // its mirror NT-1-b are h = ceil(NT/2) unit columns apart.
  const int j = u.by;
  const int ky = j == 0 ? 0 : j == 1 ? ny : ((j & 1) ? gy - ((j - 2) / 2 + 1) : (j - 2) / 2 + 1);
  const int NT = g3.nt, h = (NT + 1) / 2;
  const int tile = u.bx < h ? u.bx : NT - 1 - (u.bx - h);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  // workgroup-uniform base of the tile's first column per component, 32-bit
  // byte offsets per lane (the launcher checks a component spans < 4 GiB)
  const char* const c0 = reinterpret_cast<const char*>(a2 + (size_t)ky * gx + (size_t)tile * C);
  const size_t csb = (size_t)nz * plane * sizeof(float2);
  const unsigned esb = (unsigned)(plane * sizeof(float2)), lane0 = (unsigned)(s * sizeof(float2));
  auto at = [&](int c, unsigned o) -> float2* { return reinterpret_cast<float2*>(const_cast<char*>(c0) + c * csb + o); };
  // element n = t + R1 n2 at ob + n2 (R1 esb)
  const unsigned ob = lane0 + (unsigned)t * esb, sb = (unsigned)R1 * esb;
  float2 v0[VN], v1[VN], v2[VN];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int n = t + R1 * n2;
    const bool ok = live && lower_half<R2>(n2, n, nz);
    const unsigned o = ob + (unsigned)n2 * sb;
    v0[n2] = ok ? *at(0, o) : make_float2(0.f, 0.f);
    v1[n2] = ok ? *at(1, o) : make_float2(0.f, 0.f);
    v2[n2] = ok ? *at(2, o) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v1, buf, twL, t, s, P);
  __syncthreads();
  // 3x3 symmetric Green multiply; octant parity reconstruction for IPBC=0 (cu:109-115), whole grid for IPBC=1.
  // Each octant entry serves up to eight (kx, ky, kz) mirrors; kz mirrors
  // are in this block and the kx / ky mirror blocks run on the same XCD at
  // about the same time (grid mapping above), so repeats hit in L2.
  const int ox = gg.ox, oy = gg.oy;
  const int ir = live ? min(kx, gg.mx - kx) : 0;
  const int jr = min(ky, gg.my - ky);
  const float sxy = ((kx <= gg.tx) == (ky <= gg.ty)) ? 1.f : -1.f;
  const float sx = kx <= gg.tx ? 1.f : -1.f, sy = ky <= gg.ty ? 1.f : -1.f;
  const char* const gb = reinterpret_cast<const char*>(green);
  const unsigned pb48 = (unsigned)(ir + ox * jr) * 48u, st48 = (unsigned)(ox * oy) * 48u;
  // multiply of point e = i R1 + k1 (kz = t + R1 i + R2 k1) by its entry
  auto gmul = [&](const Green6& Gm, int i, int k1) {
    const int kz = (t + R1 * i) + R2 * k1;
    const float sz = kz <= gg.tz ? 1.f : -1.f;
    const float2 g12 = cscale(Gm.xy, sxy);
    const float2 g13 = cscale(Gm.xz, sx * sz);
    const float2 g23 = cscale(Gm.yz, sy * sz);
    const int e = i * R1 + k1;
    const float2 p0 = v0[e], p1 = v1[e], p2 = v2[e];
    v0[e] = cdot3(Gm.xx, p0, g12, p1, g13, p2);
    v1[e] = cdot3(g12, p0, Gm.yy, p1, g23, p2);
    v2[e] = cdot3(g13, p0, g23, p1, Gm.zz, p2);
  };
  auto gload = [&](int i, int k1) {
    const int kz = (t + R1 * i) + R2 * k1;
    const int kr = min(kz, gg.mz - kz);
    return green_at_b(gb, pb48 + (unsigned)kr * st48);
  };
  Green6 gpre[PRE > 0 ? PRE : 1];
  if constexpr (PRE > 0) {
    static_assert(G == 1 && PRE <= R1, "early Green loads: one k2 group per thread");
#pragma unroll
    for (int p = 0; p < PRE; ++p) gpre[p] = gload(0, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  fft2_fwd<R1, R2, +1>(v2, buf, twL, t, s, P);
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) gmul((PRE > 0 && k1 < PRE) ? gpre[k1 < PRE ? k1 : 0] : gload(i, k1), i, k1);
  }
  __syncthreads();
  fft2_inv<R1, R2, -1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v2, buf, twL, t, s, P);
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = t + R1 * n2;
      if (lower_half<R2>(n2, n, nz)) {
        const unsigned o = ob + (unsigned)n2 * sb;
        *at(0, o) = v0[n2];
        *at(1, o) = v1[n2];
        *at(2, o) = v2[n2];
      }
    }
  }
}

// SLICES y pass with the register FFT of kr_pass3: y-FFT, Green multiply at
// kz, inverse y-FFT over a batch of kz planes starting at k0 (a2 = slice
// buffer [planes][3][NY][2NX], plane = k0, grid rows = planes of the batch).
// A kernel of its own so that kr_pass3 keeps its register count.
template <int R1, int R2, int C>
__global__ __launch_bounds__(R1 * C) void kr_sly(float2* __restrict__ a2, const float2* __restrict__ green,
                                               const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                               int gy, size_t plane, BatchArgs ba, GreenGeo gg) {
  constexpr int L = R1 * R2, RT = R1 * C, P = C + 1, G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int s = threadIdx.x % C, t = threadIdx.x / C;
  // Grid order puts Green-sharing blocks on one XCD (workgroups go round-robin
  // over the 8 XCDs by linear id): mirror rows one grid row apart (p3_row),
  // and column tile b and its mirror NT-1-b h = ceil(NT/2) grid columns
  // apart (NT = column tiles).
  const int NT = gridDim.x, h = (NT + 1) / 2;
  const int tile = (int)blockIdx.x < h ? (int)blockIdx.x : NT - 1 - ((int)blockIdx.x - h);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  const P3Col q = p3_col<true>(a2, kx, p3_row<true>((int)blockIdx.y, ny, gy, (int)plane), nx, ny, nz, gx, plane, gg);
  float2 v0[VN], v1[VN], v2[VN];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int n = t + R1 * n2;
    const bool ok = live && lower_half<R2>(n2, n, q.nf);
    const size_t o = (size_t)n * q.es;
    v0[n2] = ok ? q.col[o] : make_float2(0.f, 0.f);
    v1[n2] = ok ? q.col[o + q.cs] : make_float2(0.f, 0.f);
    v2[n2] = ok ? q.col[o + 2 * q.cs] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v2, buf, twL, t, s, P);
  // 3x3 symmetric Green multiply.  Each octant entry serves up to eight
  // mirror points; the mirrors along the transform are in this block and the
  // others run on the same XCD at about the same time (grid order above), so
  // repeats hit in L2.
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int e = i * R1 + k1;
      p3_green<true>(q, green, (t + R1 * i) + R2 * k1, v0[e], v1[e], v2[e]);
    }
  }
  __syncthreads();
  fft2_inv<R1, R2, -1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v2, buf, twL, t, s, P);
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = t + R1 * n2;
      if (lower_half<R2>(n2, n, q.nf)) {
        const size_t o = (size_t)n * q.es;
        q.col[o] = v0[n2];
        q.col[o + q.cs] = v1[n2];
        q.col[o + 2 * q.cs] = v2[n2];
      }
    }
  }
}

// Pass 3 for long axes (32 values per thread and component): the three
// components go to three thread groups of one workgroup (R1 C threads each),
// so no thread holds more than one component.  Each group runs its own
// register FFT (split real / imaginary exchange); the spectra then meet in
// LDS, each group forms its own row of the symmetric 3x3 Green product, and
// runs its own inverse transform.  LDS: the three spectra, 3 L C complex,
// which the exchange buffers alias.
template <int R1, int R2, int C>
__global__ __launch_bounds__(3 * R1 * C) void kr_pass3c(float2* __restrict__ a2, const float2* __restrict__ green,
                                                     const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                                     int gy, size_t plane, BatchArgs ba, GreenGeo gg) {
  constexpr int L = R1 * R2, GT = R1 * C, RT = 3 * GT, P = C + 1, G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  constexpr int XB = L * P;  // floats of one group's exchange buffer
  static_assert(3 * XB <= 2 * 3 * L * C, "exchange buffers must fit inside the spectrum area");
  __shared__ float2 twL[L];
  __shared__ float2 spec[3 * L * C];  // [component][kz][column]; aliased by the exchange buffers
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int grp = threadIdx.x / GT, lt = threadIdx.x % GT;
  const int s = lt % C, t = lt / C;
  float* xb = reinterpret_cast<float*>(spec) + grp * XB;
  // XCD-aware grid order as in kr_pass3
  const int j = blockIdx.y;
  const int ky = j == 0 ? 0 : j == 1 ? ny : ((j & 1) ? gy - ((j - 2) / 2 + 1) : (j - 2) / 2 + 1);
  const int tile = xcd_tile<C>((int)blockIdx.x, (int)gridDim.x, true);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  float2* col = a2 + (size_t)grp * nz * plane + (size_t)ky * gx + kx;
  float2 v[VN];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int n = t + R1 * n2;
    v[n2] = (live && lower_half<R2>(n2, n, nz)) ? col[(size_t)n * plane] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_fwd_s<R1, R2, +1>(v, xb, twL, t, s, P);
  __syncthreads();
  // own spectrum out: v[i R1 + k1] = X[(t + R1 i) + R2 k1]
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) spec[(grp * L + (t + R1 * i) + R2 * k1) * C + s] = v[i * R1 + k1];
  }
  __syncthreads();
  // row grp of the Green product; octant mirror rules for IPBC=0 (cu:109-115), whole grid for IPBC=1
  const int ox = gg.ox, oy = gg.oy;
  const int ir = live ? min(kx, gg.mx - kx) : 0;
  const int jr = min(ky, gg.my - ky);
  const float sxy = ((kx <= gg.tx) == (ky <= gg.ty)) ? 1.f : -1.f;
  const float sx = kx <= gg.tx ? 1.f : -1.f, sy = ky <= gg.ty ? 1.f : -1.f;
  const size_t pb = (size_t)ir + (size_t)ox * jr;
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int kz = (t + R1 * i) + R2 * k1;
      const int kr = min(kz, gg.mz - kz);
      const float sz = kz <= gg.tz ? 1.f : -1.f;
      const Green6 Gm = green_at(green, pb + (size_t)ox * oy * kr);
      const float2 p0 = spec[(0 * L + kz) * C + s], p1 = spec[(1 * L + kz) * C + s], p2 = spec[(2 * L + kz) * C + s];
      float2 r;
      if (grp == 0)
        r = cadd(cadd(cmul(Gm.xx, p0), cmul(cscale(Gm.xy, sxy), p1)), cmul(cscale(Gm.xz, sx * sz), p2));
      else if (grp == 1)
        r = cadd(cadd(cmul(cscale(Gm.xy, sxy), p0), cmul(Gm.yy, p1)), cmul(cscale(Gm.yz, sy * sz), p2));
      else
        r = cadd(cadd(cmul(cscale(Gm.xz, sx * sz), p0), cmul(cscale(Gm.yz, sy * sz), p1)), cmul(Gm.zz, p2));
      v[i * R1 + k1] = r;
    }
  }
  __syncthreads();
  fft2_inv_s<R1, R2, -1>(v, xb, twL, t, s, P);
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = t + R1 * n2;
      if (lower_half<R2>(n2, n, nz)) col[(size_t)n * plane] = v[n2];
    }
  }
}

// ------------------------------------------------ pass 3, L = M * 256
// Pass 3 for L = M * 256 (M = 2: 512, M = 3: 768 points) by decimation in
// time over 256-point register FFTs (16 x 16).  Thread group m of M
// transforms x[M n1 + m] (zero for n1 >= 128, pruned at compile time as in
// the 256-point kernels); the twiddle w_L^{m k1} and an M-point DFT through
// LDS leave group m holding X[k1 + 256 m] for its k1 = t + 16 j.  The
// inverse runs the same steps backwards.  A thread keeps 16 values per
// component, so the three components and the Green multiply stay in
// registers; two-level shapes of 512 or 768 points hold 32 values per
// component and have to split the components over thread groups.
template <int M, int C, bool SLY = false>
__global__ __launch_bounds__(M * 16 * C) void kr_pass3m(float2* __restrict__ a2, const float2* __restrict__ green,
                                                    const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                                    int gy, size_t plane, BatchArgs ba, GreenGeo gg, P3Grid g3) {
  constexpr int R = 16, S = R * R, L = M * S, GT = R * C, RT = M * GT;
  constexpr int PS = M * C + 1;  // 256-point exchange: M C sequences per workgroup
  constexpr int PC = C + 1;      // M-point exchange layout (u S + k1) PC + s, u < M
  constexpr int XB = S * PS > M * S * PC ? S * PS : M * S * PC;
  __shared__ float2 twL[L];
  __shared__ float2 twS[S];
  __shared__ float2 xb[XB];
  // FFT3D: the unit grid of kr_pass3; SLICES: the batch's 3-D grid
  P3Unit u{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
  int NT = (int)gridDim.x;
  if (!SLY) {
    if (!p3_unit(g3, u)) return;
    NT = g3.nt;
  }
  const int rb = u.rb;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int m = threadIdx.x / GT, lt = threadIdx.x % GT;
  const int s = lt % C, t = lt / C, sv = m * C + s;
  // XCD-aware unit order as in kr_pass3
  const int tile = xcd_tile<C>(u.bx, NT, true);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  for (int i = threadIdx.x; i < S; i += RT) twS[i] = twg[M * i];
  const bool live = kx < gx;
  const P3Col q = p3_col<SLY>(a2, kx, p3_row<SLY>(u.by, ny, gy, (int)plane), nx, ny, nz, gx, plane, gg);
  float2 v[3][R];
#pragma unroll
  for (int n2 = 0; n2 < R; ++n2) {
    const bool ok = live && n2 < R / 2;
    const size_t o = (size_t)(M * (t + R * n2) + m) * q.es;
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c][n2] = ok ? q.col[o + c * q.cs] : make_float2(0.f, 0.f);
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c) __syncthreads();
    fft2_fwd<R, R, +1>(v[c], xb, twS, t, sv, PS);  // v[c][i] = Y_m[t + 16 i]
#pragma unroll
    for (int i = 0; i < R; ++i) v[c][i] = cmul(v[c][i], twL[m * (t + R * i)]);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) xb[(m * S + t + R * i) * PC + s] = v[c][i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) {
      float2 z[M];
#pragma unroll
      for (int u = 0; u < M; ++u) z[u] = xb[(u * S + t + R * i) * PC + s];
      v[c][i] = dft_m_out<M, +1>(z, m);  // X[t + 16 i + 256 m]
    }
  }
  // 3x3 symmetric Green multiply at k = t + 16 i + 256 m along the transform
#pragma unroll
  for (int i = 0; i < R; ++i) p3_green<SLY>(q, green, t + R * i + S * m, v[0][i], v[1][i], v[2][i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) xb[(m * S + t + R * i) * PC + s] = v[c][i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i) {
      float2 z[M];
#pragma unroll
      for (int u = 0; u < M; ++u) z[u] = xb[(u * S + t + R * i) * PC + s];
      float2 w = twL[m * (t + R * i)];
      w.y = -w.y;
      v[c][i] = cmul(dft_m_out<M, -1>(z, m), w);  // W_m[t + 16 i]
    }
    __syncthreads();
    fft2_inv<R, R, -1>(v[c], xb, twS, t, sv, PS);  // v[c][n2] = x[M (t + 16 n2) + m]
  }
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < R / 2; ++n2) {
      const size_t o = (size_t)(M * (t + R * n2) + m) * q.es;
#pragma unroll
      for (int c = 0; c < 3; ++c) q.col[o + c * q.cs] = v[c][n2];
    }
  }
}

// Pass 3 with register Stockham transforms (fft_reg.hpp, SkPlan): a column
// of L = 2 NZ points is held by T = L / V threads with V values per
// component, so the three components take 6 V registers (48 at V = 8, where
// kr_pass3m keeps 96) and a CU holds two workgroups: one computes while the
// other waits on its loads.  C columns per workgroup, thread (col, j) at
// threadIdx.x = j C + col so that a wave's loads run along kx.  Unit grid
// and Green reconstruction as in kr_pass3m.  Addresses are a workgroup-
// uniform base per component plus a 32-bit byte offset per lane (one
// register per address instead of two); the launcher checks that a
// component of a2 and the Green array span less than 4 GiB.
template <class PL, int C, int NBUF, int K, int WPE, bool SLY = false, int GS = P3S_GREEN_GROUP, bool RC = false,
          int PRE = 0>
__global__ __launch_bounds__(PL::T * C) __attribute__((amdgpu_waves_per_eu(WPE))) void kr_pass3s(
    float2* __restrict__ a2, const float2* __restrict__ green, const float2* __restrict__ twg, int nx, int ny, int nz,
    int gx, int gy, size_t plane, BatchArgs ba, GreenGeo gg, P3Grid g3) {
  constexpr int L = PL::L, T = PL::T, V = PL::V, NTH = T * C, R0 = PL::rad(0), RL = PL::rad(PL::S - 1);
  __shared__ float2 twL[L];
  __shared__ float2 xb[NBUF * K * L * C];
  P3Unit u{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
  int NT = (int)gridDim.x;
  if (!SLY) {
    if (!p3_unit(g3, u)) return;
    NT = g3.nt;
  }
  const int rb = u.rb;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int col = threadIdx.x % C, j = threadIdx.x / C;
  const int tile = xcd_tile<C>(u.bx, NT, true);
  const int kx = tile * C + col;
  for (int i = threadIdx.x; i < L; i += NTH) twL[i] = twg[i];
  const bool live = kx < gx;
  const int row = p3_row<SLY>(u.by, ny, gy, (int)plane);
  const P3Col q = p3_col<SLY>(a2, kx, row, nx, ny, nz, gx, plane, gg);
  // uniform base of the tile's first column; lane offset (col + n es) * 8 bytes
  char* const cb = reinterpret_cast<char*>(
      a2 + (SLY ? (size_t)(row - (int)plane) * 3 * ny * gx : (size_t)row * gx) + (size_t)tile * C);
  const size_t csb = q.cs * sizeof(float2);
  const unsigned esb = (unsigned)(q.es * sizeof(float2));
  const unsigned lane0 = (unsigned)(col * sizeof(float2));
  // stage-0 gathers: v[c][b R0 + r] = x_c[j + T b + r L / R0]; the upper
  // half (r >= R0 / 2) is the zero padding
  float2 v[3][V];
#pragma unroll
  for (int b = 0; b < V / R0; ++b)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const unsigned o = lane0 + (unsigned)sk_gather<PL, 0>(j + T * b, r) * esb;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        v[c][b * R0 + r] = (live && 2 * r < R0) ? *reinterpret_cast<const float2*>(cb + c * csb + o)
                                                : make_float2(0.f, 0.f);
    }
  __syncthreads();
  // K = 3: the three components advance stage by stage together; K = 1: one after another
  SkLds<PL, C, NBUF, K, RC> lx{xb};
  // 3x3 symmetric Green multiply at k = i + r L / R_last (octant point
  // pb + step min(k, mf - k), 48 bytes each); with PRE > 0 (components
  // one after another) the first PRE points' entries are loaded before the
  // last component's forward transform
  const char* const gb = reinterpret_cast<const char*>(green);
  const unsigned pb48 = (unsigned)q.pb * 48u, st48 = (unsigned)q.step * 48u;
  auto gload = [&](int e) {
    const int b = e / RL, r = e % RL, k = j + T * b + r * (L / RL);
    return green_at_b(gb, pb48 + (unsigned)min(k, q.mf - k) * st48);
  };
  Green6 gpre[PRE > 0 ? PRE : 1];
  static_assert(PRE == 0 || K == 1, "early Green loads with the components one after another");
#pragma unroll
  for (int c = 0; c < 3; c += K) {
    if constexpr (PRE > 0) {
      if (c == 2) {
#pragma unroll
        for (int p = 0; p < PRE; ++p) gpre[p] = gload(p);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    sk_forward<PL, C, NBUF, K, +1>(v[c], lx, twL, j, col);
  }
#pragma unroll
  for (int b = 0; b < V / RL; ++b)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int e = b * RL + r, k = j + T * b + r * (L / RL);
      const float sf = k <= q.tf ? 1.f : -1.f;
      const Green6 G = e < PRE ? gpre[e < PRE ? e : 0] : gload(e);
      const float2 g12 = cscale(G.xy, SLY ? q.c12 * sf : q.c12);
      const float2 g13 = cscale(G.xz, SLY ? q.c13 : q.c13 * sf);
      const float2 g23 = cscale(G.yz, q.c23 * sf);
      const float2 p0 = v[0][e], p1 = v[1][e], p2 = v[2][e];
      v[0][e] = cdot3(G.xx, p0, g12, p1, g13, p2);
      v[1][e] = cdot3(g12, p0, G.yy, p1, g23, p2);
      v[2][e] = cdot3(g13, p0, g23, p1, G.zz, p2);
      // at most GS points' Green loads (12 registers each) in flight
      if ((e + 1) % GS == 0) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
  for (int c = 0; c < 3; c += K) sk_inverse<PL, C, NBUF, K, -1>(v[c], lx, twL, j, col);
  if (live) {
#pragma unroll
    for (int b = 0; b < V / R0; ++b)
#pragma unroll
      for (int r = 0; 2 * r < R0; ++r) {
        const unsigned o = lane0 + (unsigned)sk_gather<PL, 0>(j + T * b, r) * esb;
#pragma unroll
        for (int c = 0; c < 3; ++c) *reinterpret_cast<float2*>(cb + c * csb + o) = v[c][b * R0 + r];
      }
  }
}

// Pass 3 with the mirror rows ky and gy - ky in the two halves of one
// workgroup: both halves read the same Green octant entries at the same
// time, so one of each pair of loads is served by the CU's L1.
template <int R1, int R2, int C>
__global__ __launch_bounds__(2 * R1 * C) void kr_pass3y(float2* __restrict__ a2, const float2* __restrict__ green,
                                               const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                               int gy, size_t plane, BatchArgs ba, GreenGeo gg) {
  constexpr int L = R1 * R2, RT = 2 * R1 * C, P = C + 1, G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float2 bufs[2 * L * P];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int hp = threadIdx.x / (R1 * C), lt = threadIdx.x % (R1 * C);
  const int s = lt % C, t = lt / C;
  float2* buf = bufs + hp * L * P;
  // Grid order puts Green-sharing blocks on one XCD (workgroups go round-robin
  // over the 8 XCDs by linear id): ky runs 0, ny, 1, gy-1, 2, gy-2, ... so
  // mirror rows are one grid row apart, and column tile b and its mirror
  // NT-1-b are h = ceil(NT/2) grid columns apart (NT = column tiles).
  // grid row j holds ky = j and gy - j (j = 0: the self-mirror rows 0 and ny)
  const int j = blockIdx.y;
  const int ky = j == 0 ? (hp ? ny : 0) : (hp ? gy - j : j);
  const int NT = gridDim.x, h = (NT + 1) / 2;
  const int tile = (int)blockIdx.x < h ? (int)blockIdx.x : NT - 1 - ((int)blockIdx.x - h);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  // workgroup-uniform base of the tile's first column per component, 32-bit
  // byte offsets per lane (the launcher checks a component spans < 4 GiB)
  const char* const c0 = reinterpret_cast<const char*>(a2 + (size_t)ky * gx + (size_t)tile * C);
  const size_t csb = (size_t)nz * plane * sizeof(float2);
  const unsigned esb = (unsigned)(plane * sizeof(float2)), lane0 = (unsigned)(s * sizeof(float2));
  auto at = [&](int c, unsigned o) -> float2* { return reinterpret_cast<float2*>(const_cast<char*>(c0) + c * csb + o); };
  // element n = t + R1 n2 at ob + n2 (R1 esb)
  const unsigned ob = lane0 + (unsigned)t * esb, sb = (unsigned)R1 * esb;
  float2 v0[VN], v1[VN], v2[VN];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int n = t + R1 * n2;
    const bool ok = live && lower_half<R2>(n2, n, nz);
    const unsigned o = ob + (unsigned)n2 * sb;
    v0[n2] = ok ? *at(0, o) : make_float2(0.f, 0.f);
    v1[n2] = ok ? *at(1, o) : make_float2(0.f, 0.f);
    v2[n2] = ok ? *at(2, o) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v2, buf, twL, t, s, P);
  // 3x3 symmetric Green multiply; octant parity reconstruction for IPBC=0 (cu:109-115), whole grid for IPBC=1.
  // Each octant entry serves up to eight (kx, ky, kz) mirrors; kz mirrors
  // are in this block and the kx / ky mirror blocks run on the same XCD at
  // about the same time (grid mapping above), so repeats hit in L2.
  const int ox = gg.ox, oy = gg.oy;
  const int ir = live ? min(kx, gg.mx - kx) : 0;
  const int jr = min(ky, gg.my - ky);
  const float sxy = ((kx <= gg.tx) == (ky <= gg.ty)) ? 1.f : -1.f;
  const float sx = kx <= gg.tx ? 1.f : -1.f, sy = ky <= gg.ty ? 1.f : -1.f;
  const char* const gb = reinterpret_cast<const char*>(green);
  const unsigned pb48 = (unsigned)(ir + ox * jr) * 48u, st48 = (unsigned)(ox * oy) * 48u;
#pragma unroll
  for (int i = 0; i < G; ++i) {
    if (!reg2_owns<R1, R2>(t, i)) continue;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int kz = (t + R1 * i) + R2 * k1;
      const int kr = min(kz, gg.mz - kz);
      const float sz = kz <= gg.tz ? 1.f : -1.f;
      const Green6 G = green_at_b(gb, pb48 + (unsigned)kr * st48);
      const float2 g12 = cscale(G.xy, sxy);
      const float2 g13 = cscale(G.xz, sx * sz);
      const float2 g23 = cscale(G.yz, sy * sz);
      const int e = i * R1 + k1;
      const float2 p0 = v0[e], p1 = v1[e], p2 = v2[e];
      v0[e] = cdot3(G.xx, p0, g12, p1, g13, p2);
      v1[e] = cdot3(g12, p0, G.yy, p1, g23, p2);
      v2[e] = cdot3(g13, p0, g23, p1, G.zz, p2);
    }
  }
  __syncthreads();
  fft2_inv<R1, R2, -1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v2, buf, twL, t, s, P);
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = t + R1 * n2;
      if (lower_half<R2>(n2, n, nz)) {
        const unsigned o = ob + (unsigned)n2 * sb;
        *at(0, o) = v0[n2];
        *at(1, o) = v1[n2];
        *at(2, o) = v2[n2];
      }
    }
  }
}

// Pass 3 with lane pairs (L = Q^2): each column's Q virtual threads are lane
// pairs holding Q/2 values per component, half the registers of kr_pass3, so
// more waves fit a SIMD to hide the load / Green-read latency.
template <int Q, int C>
__global__ __launch_bounds__(2 * Q * C) void kr_pass3p(float2* __restrict__ a2, const float2* __restrict__ green,
                                                    const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                                    int gy, size_t plane, BatchArgs ba, GreenGeo gg) {
  constexpr int L = Q * Q, H = Q / 2, RT = 2 * Q * C, P2 = pair_stride(C);
  __shared__ float2 twL[L];
  __shared__ float2 buf[H * Q * P2];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int h = threadIdx.x & 1, s = (threadIdx.x >> 1) % C, t = (threadIdx.x >> 1) / C;
  // XCD-aware grid order as in kr_pass3
  const int j = blockIdx.y;
  const int ky = j == 0 ? 0 : j == 1 ? ny : ((j & 1) ? gy - ((j - 2) / 2 + 1) : (j - 2) / 2 + 1);
  const int NT = gridDim.x, hx = (NT + 1) / 2;
  const int tile = (int)blockIdx.x < hx ? (int)blockIdx.x : NT - 1 - ((int)blockIdx.x - hx);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  const size_t cs = (size_t)nz * plane;
  float2* col = a2 + (size_t)ky * gx + kx;
  float2 v0[H], v1[H], v2[H];
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const int n = t + Q * (h + 2 * i);
    const bool ok = live && lower_half_pair<Q>(i, n, nz);
    const size_t o = (size_t)n * plane;
    v0[i] = ok ? col[o] : make_float2(0.f, 0.f);
    v1[i] = ok ? col[o + cs] : make_float2(0.f, 0.f);
    v2[i] = ok ? col[o + 2 * cs] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2p_fwd<Q, +1>(v0, buf, twL, t, h, s, C, P2);
  __syncthreads();
  fft2p_fwd<Q, +1>(v1, buf, twL, t, h, s, C, P2);
  __syncthreads();
  fft2p_fwd<Q, +1>(v2, buf, twL, t, h, s, C, P2);
  // 3x3 symmetric Green multiply; octant parity reconstruction for IPBC=0 (cu:109-115), whole grid for IPBC=1
  const int ox = gg.ox, oy = gg.oy;
  const int ir = live ? min(kx, gg.mx - kx) : 0;
  const int jr = min(ky, gg.my - ky);
  const float sxy = ((kx <= gg.tx) == (ky <= gg.ty)) ? 1.f : -1.f;
  const float sx = kx <= gg.tx ? 1.f : -1.f, sy = ky <= gg.ty ? 1.f : -1.f;
  const size_t pb = (size_t)ir + (size_t)ox * jr;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const int kz = t + Q * (H * h + i);
    const int kr = min(kz, gg.mz - kz);
    const float sz = kz <= gg.tz ? 1.f : -1.f;
    const Green6 G = green_at(green, pb + (size_t)ox * oy * kr);
    const float2 g12 = cscale(G.xy, sxy);
    const float2 g13 = cscale(G.xz, sx * sz);
    const float2 g23 = cscale(G.yz, sy * sz);
    const float2 p0 = v0[i], p1 = v1[i], p2 = v2[i];
    v0[i] = cadd(cadd(cmul(G.xx, p0), cmul(g12, p1)), cmul(g13, p2));
    v1[i] = cadd(cadd(cmul(g12, p0), cmul(G.yy, p1)), cmul(g23, p2));
    v2[i] = cadd(cadd(cmul(g13, p0), cmul(g23, p1)), cmul(G.zz, p2));
  }
  __syncthreads();
  fft2p_inv<Q, -1>(v0, buf, twL, t, h, s, C, P2);
  __syncthreads();
  fft2p_inv<Q, -1>(v1, buf, twL, t, h, s, C, P2);
  __syncthreads();
  fft2p_inv<Q, -1>(v2, buf, twL, t, h, s, C, P2);
  if (live) {
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const int n = t + Q * (h + 2 * i);
      if (lower_half_pair<Q>(i, n, nz)) {
        const size_t o = (size_t)n * plane;
        col[o] = v0[i];
        col[o + cs] = v1[i];
        col[o + 2 * cs] = v2[i];
      }
    }
  }
}

// ------------------------------------------------------------------ pass 5
// finish of row g (cu:116-118): y = adia x - w/G + offdiag(aoff) x on the
// unpadded outputs; w0..w2[n2] hold the inverse-transformed components at
// row index ix(n2) for n2 < H, kept(n2, ix) says whether ix < nx (compact
// or full layout, conj for 'C').  Site indices are resolved first with the
// layout test outside the element loop; the loads then run unconditionally
// in groups of FG outputs (an unoccupied site reads entry 0 and stores
// nothing), so each group's 9 FG loads are in flight together instead of
// waiting on per-output branches.
template <int H, int FG, typename IX, typename KEPT>
__device__ __forceinline__ void p5_finish_at(const float2* w0, const float2* w1, const float2* w2, IX ixof, KEPT kept,
                                             int g, const float2* __restrict__ x, float2* __restrict__ y,
                                             const float2* __restrict__ adia, const float2* __restrict__ aoff,
                                             const int16_t* __restrict__ iocc, const SiteMap& sm, int nx, int nat,
                                             int nat0, int compact, int cj, float inv) {
  size_t o[H];
  bool st[H], zr[H];
  if (!compact) {
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) {
      const int ix = ixof(n2);
      st[n2] = kept(n2, ix);
      o[n2] = st[n2] ? (size_t)g * nx + ix : 0;
      zr[n2] = iocc && iocc[o[n2]] == 0;
    }
  } else if (sm.contiguous) {
    const int lo = sm.rowlo[g], q0 = sm.rowptr[g], cnt = sm.rowptr[g + 1] - q0;
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) {
      const int ix = ixof(n2), d = ix - lo;
      st[n2] = kept(n2, ix) && d >= 0 && d < cnt;
      o[n2] = st[n2] ? (size_t)(q0 + d) : 0;
      zr[n2] = false;
    }
  } else {
#pragma unroll
    for (int n2 = 0; n2 < H; ++n2) {
      const int ix = ixof(n2);
      const int q = kept(n2, ix) ? sm.qmap[(size_t)g * nx + ix] : -1;
      st[n2] = q >= 0;
      o[n2] = st[n2] ? (size_t)q : 0;
      zr[n2] = false;
    }
  }
  const size_t sd = compact ? (size_t)nat0 : (size_t)nat;
#pragma unroll
  for (int g0 = 0; g0 < H; g0 += FG) {
    float2 xv[FG][3], dv[FG][3], av[FG][3];
#pragma unroll
    for (int j = 0; j < FG && g0 + j < H; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const size_t e = o[g0 + j] + k * sd;
        xv[j][k] = x[e];
        dv[j][k] = adia[e];
        av[j][k] = aoff[e];
      }
#pragma unroll
    for (int j = 0; j < FG && g0 + j < H; ++j) {
      const int n2 = g0 + j;
      float2 x1 = xv[j][0], x2 = xv[j][1], x3 = xv[j][2];
      const float2 a23 = av[j][0], a31 = av[j][1], a12 = av[j][2];
      if (cj) { x1.y = -x1.y; x2.y = -x2.y; x3.y = -x3.y; }
      float2 r1 = csub(cmul(dv[j][0], x1), cscale(w0[n2], inv));
      float2 r2 = csub(cmul(dv[j][1], x2), cscale(w1[n2], inv));
      float2 r3 = csub(cmul(dv[j][2], x3), cscale(w2[n2], inv));
      r1 = cadd(r1, cadd(cmul(a31, x3), cmul(a12, x2)));
      r2 = cadd(r2, cadd(cmul(a12, x1), cmul(a23, x3)));
      r3 = cadd(r3, cadd(cmul(a23, x2), cmul(a31, x1)));
      if (cj) { r1.y = -r1.y; r2.y = -r2.y; r3.y = -r3.y; }
      if (zr[n2]) r1 = r2 = r3 = make_float2(0.f, 0.f);
      if (st[n2]) {
        const size_t e = o[n2];
        y[e] = r1; y[e + sd] = r2; y[e + 2 * sd] = r3;
      }
    }
  }
}

// finish in the register order of fft2_inv: output n2 < (R2 + 1) / 2 of
// thread t is row index t + R1 n2
template <int R1, int R2, int FG>
__device__ __forceinline__ void p5_finish(const float2* w0, const float2* w1, const float2* w2, int t, int g,
                                          const float2* __restrict__ x, float2* __restrict__ y,
                                          const float2* __restrict__ adia, const float2* __restrict__ aoff,
                                          const int16_t* __restrict__ iocc, const SiteMap& sm, int nx, int nat,
                                          int nat0, int compact, int cj, float inv) {
  p5_finish_at<(R2 + 1) / 2, FG>(
      w0, w1, w2, [=](int n2) { return t + R1 * n2; },
      [=](int n2, int ix) { return lower_half<R2>(n2, ix, nx); }, g, x, y, adia, aoff, iocc, sm, nx, nat, nat0,
      compact, cj, inv);
}

// pass-4 output of one component of row g in the register order of fft2_inv;
// src of a row past the end is clamped to the last row by the caller (its
// results are never stored), so only the register-ownership test remains
template <int R1, int R2>
__device__ __forceinline__ void p5_load(float2* v, const float2* __restrict__ src, int t) {
#pragma unroll
  for (int i = 0; i < Reg2<R1, R2>::G; ++i)
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1)
      v[i * R1 + k1] = reg2_owns<R1, R2>(t, i) ? src[(t + R1 * i) + R2 * k1] : make_float2(0.f, 0.f);
}

// FG: outputs per finish load group
template <int R1, int R2, int C, int FG = 4>
__global__ __launch_bounds__(R1 * C) void kr_pass5(const float2* __restrict__ a1, const float2* __restrict__ twg,
                                               const float2* __restrict__ x, float2* __restrict__ y,
                                               const float2* __restrict__ adia, const float2* __restrict__ aoff,
                                               const int16_t* __restrict__ iocc, SiteMap sm, int nx, int nrows,
                                               int nat, int nat0, int compact, int cj, float inv, BatchArgs ba) {
  constexpr int L = R1 * R2, RT = R1 * C, P = R2 * (R1 + 1), VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float2 buf[C * P];  // row layout
  const int rb = blockIdx.y;
  if (ba.active && !ba.active[rb]) return;
  a1 += rb * ba.a1s;
  x += rb * ba.xs;
  y += rb * ba.xs;
  adia += rb * ba.os;
  aoff += rb * ba.os;
  const int t = threadIdx.x % R1, s = threadIdx.x / R1;
  const int g = blockIdx.x * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = g < nrows;
  const size_t cs = (size_t)nrows * L;
  const float2* src = a1 + (size_t)(live ? g : nrows - 1) * L;
  float2 v0[VN], v1[VN], v2[VN];
  p5_load<R1, R2>(v0, src, t);
  p5_load<R1, R2>(v1, src + cs, t);
  p5_load<R1, R2>(v2, src + 2 * cs, t);
  __syncthreads();
  fft2_inv<R1, R2, -1, 1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1, 1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1, 1>(v2, buf, twL, t, s, P);
  // the transforms' kept outputs are materialised here: otherwise their last
  // butterflies sink into the finish, interleave with its loads and hold
  // both live at once (248 instead of 140 VGPRs at 16 x 16, 2 waves per SIMD)
#pragma unroll
  for (int n2 = 0; n2 < (R2 + 1) / 2; ++n2)
    asm volatile("" : "+v"(v0[n2].x), "+v"(v0[n2].y), "+v"(v1[n2].x), "+v"(v1[n2].y), "+v"(v2[n2].x),
                 "+v"(v2[n2].y));
  if (!live) return;
  p5_finish<R1, R2, FG>(v0, v1, v2, t, g, x, y, adia, aoff, iocc, sm, nx, nat, nat0, compact, cj, inv);
}

// ------------------------------------------------ pass 5, L = M * 256
// Pass 5 for L = M * 256 (M = 2, 3) by the decimation of kr_passym along
// the contiguous x rows: thread t of group m reads spectral elements
// t + 16 i + 256 m of row g (16 lanes cover 128 contiguous bytes).  The
// three components are transformed in turn.  The kept outputs
// x[M (t + 16 n2) + m], n2 < 8 (n < L/2), cross LDS once more into row order,
// so that thread q = 16 m + t of the row finishes row indices q + 16 M j,
// j < 8, with the coalesced loads and stores of the register kernels.
// KEEPL parks components 0 and 1 in LDS in row order until the finish
// instead of in registers: 167 instead of 194 VGPRs at M = 3, C = 4 (3
// waves per SIMD; 48 KB of LDS per workgroup).
// M-point DFT across the M groups of a decimated L = M * 256 transform, as
// in kr_passym
// (thread t of group m holds v[i], element t + 16 i of its 256-point
// sequence s), through LDS in two halves: the real parts cross first and are
// held (M 16 floats), then each output reads the imaginary parts at its
// element and combines.  D: sign of the M-point DFT; the inverse (D < 0)
// then applies the conjugate twiddle w_L^{-m (t + 16 i)}.  Layout of xf:
// (u 256 + t + 16 i) (C + 1) + s for group u; the caller's exchange buffer
// holds at least M 256 (C + 1) floats.
template <int M, int C, int D>
__device__ __forceinline__ void dec_combine(float2* v, float* xf, const float2* twL, int m,
                                            int t, int s) {
  constexpr int R = 16, S = R * R, PC = C + 1;
  float zr[M][R];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R; ++i) xf[(m * S + t + R * i) * PC + s] = v[i].x;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int u = 0; u < M; ++u) zr[u][i] = xf[(u * S + t + R * i) * PC + s];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R; ++i) xf[(m * S + t + R * i) * PC + s] = v[i].y;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < R; ++i) {
    float2 z[M];
#pragma unroll
    for (int u = 0; u < M; ++u) z[u] = make_float2(zr[u][i], xf[(u * S + t + R * i) * PC + s]);
    v[i] = dft_m_out<M, D>(z, m);
    if (D < 0) {
      float2 w = twL[m * (t + R * i)];
      w.y = -w.y;
      v[i] = cmul(v[i], w);
    }
  }
}

template <int M, int C, bool KEEPL>
__global__ __launch_bounds__(M * 16 * C) void kr_pass5m(const float2* __restrict__ a1, const float2* __restrict__ twg,
                                                    const float2* __restrict__ x, float2* __restrict__ y,
                                                    const float2* __restrict__ adia, const float2* __restrict__ aoff,
                                                    const int16_t* __restrict__ iocc, SiteMap sm, int nx, int nrows,
                                                    int nat, int nat0, int compact, int cj, float inv, BatchArgs ba) {
  constexpr int R = 16, S = R * R, L = M * S, H = L / 2, GT = R * C, RT = M * GT;
  constexpr int PS = M * C + 1;       // 256-point exchange (floats)
  constexpr int PC = C + 1;           // M-point exchange (floats)
  constexpr int PR = H + 1;           // row-order exchange: real at s PR + n, imaginary at (C + s) PR + n
  constexpr int XA = S * PS > M * S * PC ? S * PS : M * S * PC;
  constexpr int XB = XA > 2 * C * PR ? XA : 2 * C * PR;
  constexpr int RB = 2 * C * PR;      // one component in row order
  __shared__ float2 twL[L];
  __shared__ float2 twS[S];
  __shared__ float xf[XB];
  __shared__ float rowk[KEEPL ? 2 * RB : 1];  // KEEPL: components 0, 1 in row order
  const int rb = blockIdx.y;
  if (ba.active && !ba.active[rb]) return;
  a1 += rb * ba.a1s;
  x += rb * ba.xs;
  y += rb * ba.xs;
  adia += rb * ba.os;
  aoff += rb * ba.os;
  const int m = threadIdx.x / GT, lt = threadIdx.x % GT;
  const int t = lt % R, s = lt / R, sv = m * C + s, q = R * m + t;
  const int g = blockIdx.x * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  for (int i = threadIdx.x; i < S; i += RT) twS[i] = twg[M * i];
  const bool live = g < nrows;
  const size_t cs = (size_t)nrows * L;
  const float2* src = a1 + (size_t)(live ? g : nrows - 1) * L + t + S * m;
  float2 w[3][R / 2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float2 v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = src[c * cs + R * i];
    dec_combine<M, C, -1>(v, xf, twL, m, t, s);  // v[i] = W_m[t + 16 i]
    __syncthreads();
    fft2_inv_s<R, R, -1>(v, xf, twS, t, sv, PS);  // v[n2] = x[M (t + 16 n2) + m]
    float* ro = KEEPL && c < 2 ? rowk + c * RB : xf;
    if (!KEEPL || c == 2) __syncthreads();
#pragma unroll
    for (int n2 = 0; n2 < R / 2; ++n2) {
      const int n = M * (t + R * n2) + m;
      ro[s * PR + n] = v[n2].x;
      ro[(C + s) * PR + n] = v[n2].y;
    }
    if (!KEEPL || c == 2) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < R / 2; ++j) {
        const int n = q + R * M * j;
        w[c][j] = make_float2(xf[s * PR + n], xf[(C + s) * PR + n]);
      }
    }
  }
  if (KEEPL) {  // written before the barriers of component 2
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < R / 2; ++j) {
        const int n = q + R * M * j;
        w[c][j] = make_float2(rowk[c * RB + s * PR + n], rowk[c * RB + (C + s) * PR + n]);
      }
  }
  if (!live) return;
  p5_finish_at<R / 2, 4>(
      w[0], w[1], w[2], [=](int j) { return q + R * M * j; }, [=](int, int ix) { return ix < nx; }, g, x, y, adia,
      aoff, iocc, sm, nx, nat, nat0, compact, cj, inv);
}

// ------------------------------------------------------------------ host --
template <int R1, int R2, int C>
static void launch_reg(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                       const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  const MatvecPlan& m = c.mv;
  constexpr int RT = R1 * C;
  const int nrows = c.ny * c.nz;
  const unsigned nb = (unsigned)bt.nrhs;
  SiteMap sm{c.rowptr, c.rowlo, c.qmap, c.rows_contiguous};
  switch (which) {
    case 1:
      hipLaunchKernelGGL((kr_pass1<R1, R2, C>), dim3((nrows + C - 1) / C, 3, nb), dim3(RT), 0, st, x, bt.a1, sm,
                         m.twx, c.nx, nrows, c.nat, c.nat0, compact, cj, bt.k);
      break;
    case 2:
      hipLaunchKernelGGL((kr_passy<R1, R2, C, +1>), dim3((m.gx + C - 1) / C, c.nz, 3 * nb), dim3(RT), 0, st,
                         bt.a1, bt.a2, m.twy, m.gx, c.ny, m.gy, c.nz, m.p1, m.p2, bt.k);
      break;
    case 3: {
      const P3Grid g3 = p3_grid(m.gx, m.gy, C, (int)nb);
      hipLaunchKernelGGL((kr_pass3<R1, R2, C>), dim3(p3_blocks(g3)), dim3(RT), 0, st, bt.a2, c.green, m.twz, c.nx,
                         c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c), g3);
      break;
    }
    case 4:
      hipLaunchKernelGGL((kr_passy<R1, R2, C, -1>), dim3((m.gx + C - 1) / C, c.nz, 3 * nb), dim3(RT), 0, st,
                         bt.a2, bt.a1, m.twy, m.gx, m.gy, c.ny, c.nz, m.p2, m.p1, bt.k);
      break;
    default:
      hipLaunchKernelGGL((kr_pass5<R1, R2, C>), dim3((nrows + C - 1) / C, nb), dim3(RT), 0, st, bt.a1, m.twx, x,
                         y, adia, aoff, compact ? nullptr : c.iocc, sm, c.nx, nrows, c.nat, c.nat0, compact, cj,
                         1.f / (float)m.ngrid, bt.k);
      break;
  }
}

// shape whose pass 3 uses the lane-pair kernel; other passes as launch_reg<Q, Q, CR>
template <int Q, int CR, int CP>
static void launch_pair(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                        const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  if (which != 3) {
    launch_reg<Q, Q, CR>(c, cj, compact, x, y, adia, aoff, st, bt, which);
    return;
  }
  const MatvecPlan& m = c.mv;
  hipLaunchKernelGGL((kr_pass3p<Q, CP>), dim3((m.gx + CP - 1) / CP, m.gy, (unsigned)bt.nrhs), dim3(2 * Q * CP), 0,
                     st, bt.a2, c.green, m.twz, c.nx, c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c));
}

// shape whose pass 3 pairs mirror rows in one workgroup; other passes as launch_reg<R1, R2, C>
template <int R1, int R2, int C>
static void launch_ypair(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                         const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  if (which != 3) {
    launch_reg<R1, R2, C>(c, cj, compact, x, y, adia, aoff, st, bt, which);
    return;
  }
  const MatvecPlan& m = c.mv;
  hipLaunchKernelGGL((kr_pass3y<R1, R2, C>), dim3((m.gx + C - 1) / C, c.ny, (unsigned)bt.nrhs), dim3(2 * R1 * C), 0,
                     st, bt.a2, c.green, m.twz, c.nx, c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c));
}

// pass 5 by kr_pass5 with finish load groups of FG outputs; serves pass 5 only
template <int R1, int R2, int C, int FG>
static void launch_p5f(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                       const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  if (which != 5) return;
  const MatvecPlan& m = c.mv;
  const int nrows = c.ny * c.nz;
  SiteMap sm{c.rowptr, c.rowlo, c.qmap, c.rows_contiguous};
  hipLaunchKernelGGL((kr_pass5<R1, R2, C, FG>), dim3((nrows + C - 1) / C, (unsigned)bt.nrhs), dim3(R1 * C), 0,
                     st, bt.a1, m.twx, x, y, adia, aoff, compact ? nullptr : c.iocc, sm, c.nx, nrows, c.nat, c.nat0,
                     compact, cj, 1.f / (float)m.ngrid, bt.k);
}

// single-component passes only (1, 2, 4): shapes whose per-thread arrays are
// too large to hold the three components of passes 3 and 5 in registers
template <int R1, int R2, int C>
static void launch_reg1(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                        const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  const MatvecPlan& m = c.mv;
  constexpr int RT = R1 * C;
  const int nrows = c.ny * c.nz;
  const unsigned nb = (unsigned)bt.nrhs;
  SiteMap sm{c.rowptr, c.rowlo, c.qmap, c.rows_contiguous};
  if (which == 1)
    hipLaunchKernelGGL((kr_pass1<R1, R2, C>), dim3((nrows + C - 1) / C, 3, nb), dim3(RT), 0, st, x, bt.a1, sm,
                       m.twx, c.nx, nrows, c.nat, c.nat0, compact, cj, bt.k);
  else if (which == 2)
    hipLaunchKernelGGL((kr_passy<R1, R2, C, +1>), dim3((m.gx + C - 1) / C, c.nz, 3 * nb), dim3(RT), 0, st, bt.a1,
                       bt.a2, m.twy, m.gx, c.ny, m.gy, c.nz, m.p1, m.p2, bt.k);
  else if (which == 4)
    hipLaunchKernelGGL((kr_passy<R1, R2, C, -1>), dim3((m.gx + C - 1) / C, c.nz, 3 * nb), dim3(RT), 0, st, bt.a2,
                       bt.a1, m.twy, m.gx, m.gy, c.ny, c.nz, m.p2, m.p1, bt.k);
}

// passes 2 and 4 by kr_passym; listed for L = gy = M * 256
template <int M, int C>
static void launch_ym(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                      const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  (void)cj; (void)compact; (void)x; (void)y; (void)adia; (void)aoff;
  const MatvecPlan& m = c.mv;
  const dim3 grid((m.gx + C - 1) / C, c.nz, 3 * (unsigned)bt.nrhs), block(M * 16 * C);
  if (which == 2)
    hipLaunchKernelGGL((kr_passym<M, C, +1>), grid, block, 0, st, bt.a1, bt.a2, m.twy, m.gx, c.ny, m.gy, c.nz, m.p1,
                       m.p2, bt.k);
  else if (which == 4)
    hipLaunchKernelGGL((kr_passym<M, C, -1>), grid, block, 0, st, bt.a2, bt.a1, m.twy, m.gx, m.gy, c.ny, c.nz, m.p2,
                       m.p1, bt.k);
}

// pass 5 by kr_pass5m; serves pass 5 only, listed for L = gx = M * 256
template <int M, int C, bool KEEPL>
static void launch_p5m(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                       const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  const MatvecPlan& m = c.mv;
  const int nrows = c.ny * c.nz;
  SiteMap sm{c.rowptr, c.rowlo, c.qmap, c.rows_contiguous};
  hipLaunchKernelGGL((kr_pass5m<M, C, KEEPL>), dim3((nrows + C - 1) / C, (unsigned)bt.nrhs), dim3(M * 16 * C), 0, st,
                     bt.a1, m.twx, x, y, adia, aoff, compact ? nullptr : c.iocc, sm, c.nx, nrows, c.nat, c.nat0,
                     compact, cj, 1.f / (float)m.ngrid, bt.k);
}

// pass 3 by kr_pass3m; serves pass 3 only, listed for L = gz = 2 NZ = M * 256
template <int M, int C>
static void launch_p3m(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                       const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  (void)cj; (void)compact; (void)x; (void)y; (void)adia; (void)aoff;
  if (which != 3) return;
  const MatvecPlan& m = c.mv;
  const P3Grid g3 = p3_grid(m.gx, m.gy, C, bt.nrhs);
  hipLaunchKernelGGL((kr_pass3m<M, C>), dim3(p3_blocks(g3)), dim3(M * 16 * C), 0, st, bt.a2, c.green, m.twz, c.nx,
                     c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c), g3);
}

// kr_pass3s addresses a component of a2 and the Green array with 32-bit
// byte offsets
static bool p3s_fits(const Context& c) {
  const size_t comp = (size_t)c.nz * c.mv.p2 * sizeof(float2);
  const size_t gbytes = green_points(c.nx, c.ny, c.nz, c.ipbc) * 6 * sizeof(float2);
  return comp < ((size_t)1 << 32) && gbytes < ((size_t)1 << 32);
}
// pass 3 by kr_pass3 with early Green loads; serves pass 3 only
template <int R1, int R2, int C, int PRE, int WPE = 1>
static void launch_p3pre(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                         const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  (void)cj; (void)compact; (void)x; (void)y; (void)adia; (void)aoff;
  if (which != 3) return;
  const MatvecPlan& m = c.mv;
  const P3Grid g3 = p3_grid(m.gx, m.gy, C, bt.nrhs);
  hipLaunchKernelGGL((kr_pass3<R1, R2, C, PRE, WPE>), dim3(p3_blocks(g3)), dim3(R1 * C), 0, st, bt.a2, c.green, m.twz,
                     c.nx, c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c), g3);
}

// pass 3 by kr_pass3s (register Stockham); serves pass 3 only
template <class PL, int C, int NBUF, int K, int WPE, int GS = P3S_GREEN_GROUP, bool RC = false, int PRE = 0>
static void launch_p3s(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                       const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  (void)cj; (void)compact; (void)x; (void)y; (void)adia; (void)aoff;
  if (which != 3) return;
  const MatvecPlan& m = c.mv;
  const P3Grid g3 = p3_grid(m.gx, m.gy, C, bt.nrhs);
  hipLaunchKernelGGL((kr_pass3s<PL, C, NBUF, K, WPE, false, GS, RC, PRE>), dim3(p3_blocks(g3)), dim3(PL::T * C), 0, st, bt.a2, c.green, m.twz,
                     c.nx, c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c), g3);
}

// pass 3 by component-split thread groups (kr_pass3c); serves pass 3 only
template <int R1, int R2, int C>
static void launch_split3(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                          const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  (void)cj; (void)compact; (void)x; (void)y; (void)adia; (void)aoff;
  if (which != 3) return;
  const MatvecPlan& m = c.mv;
  hipLaunchKernelGGL((kr_pass3c<R1, R2, C>), dim3((m.gx + C - 1) / C, m.gy, (unsigned)bt.nrhs), dim3(3 * R1 * C), 0,
                     st, bt.a2, c.green, m.twz, c.nx, c.ny, c.nz, m.gx, m.gy, m.p2, bt.k, green_geo(c));
}

using RegLaunch = void (*)(Context&, int, int, const float2*, float2*, const float2*, const float2*, hipStream_t,
                           const MvBatch&, int);
constexpr int ALL_PASSES = 0x3e, ONE_COMPONENT_PASSES = (1 << 1) | (1 << 2) | (1 << 4);
struct RegShape {
  int L, r1, r2, c;
  RegLaunch fn;
  int passes = ALL_PASSES;  // bit p set: serves pass p
};
#define DDH_SHAPE(L_, R1_, R2_, C_) RegShape{L_, R1_, R2_, C_, &launch_reg<R1_, R2_, C_>}
#define DDH_SHAPE1(L_, R1_, R2_, C_) \
  RegShape{L_, R1_, R2_, C_, &launch_reg1<R1_, R2_, C_>, ONE_COMPONENT_PASSES}
// One register shape per padded axis length 2N (L = R1 R2, R1 threads per
// sequence, C sequences per workgroup of R1 C threads).  Shapes keep a
// thread's register array at <= 16 complex values so the three-component
// passes (3 and 5) stay free of spills; 200 = 20 x 10 and 216 = 18 x 12 hold
// 20 and 18 values per thread (153-217 VGPRs in passes 3 and 5, no spills;
// the 10 x 20 pass-3 tile at 200, 191 VGPRs).
// Lengths without such a split (162, primes of 7 ...) use the LDS-staged
// kernels.
// The first shape of a length is its default; further ones are autotuner
// candidates.
static const RegShape kShapes[] = {
    DDH_SHAPE(32, 4, 8, 64),    DDH_SHAPE(48, 8, 6, 32),    DDH_SHAPE(60, 10, 6, 32),  DDH_SHAPE(64, 8, 8, 32),
    DDH_SHAPE(64, 4, 16, 64),   DDH_SHAPE(72, 12, 6, 16),   DDH_SHAPE(80, 10, 8, 32),  DDH_SHAPE(96, 12, 8, 16),
    DDH_SHAPE(96, 8, 12, 32),   DDH_SHAPE(100, 10, 10, 32), DDH_SHAPE(108, 12, 9, 16), DDH_SHAPE(120, 12, 10, 16),
    DDH_SHAPE(120, 8, 15, 32),  DDH_SHAPE(128, 8, 16, 32),  DDH_SHAPE(128, 16, 8, 16), DDH_SHAPE(144, 12, 12, 16),
    DDH_SHAPE(144, 16, 9, 16),  DDH_SHAPE(150, 15, 10, 16), DDH_SHAPE(160, 16, 10, 16), DDH_SHAPE(180, 15, 12, 16),
    DDH_SHAPE(192, 16, 12, 16), DDH_SHAPE(240, 16, 15, 16), DDH_SHAPE(256, 16, 16, 16),
    // 200 and 216: pass 3 defaults to the fastest tile measured on 200^3 and
    // 240 x 240 x 216 grids (10 x 20 by 16 columns: 121 us; 18 x 12 by 32
    // columns: 171 us); the 16-column tiles of 20 x 10 and 18 x 12 ran at
    // 207 and 282 us, behind the LDS-staged kernel's 175 us at 200
    RegShape{200, 20, 10, 16, &launch_reg<20, 10, 16>, ALL_PASSES & ~(1 << 3)},
    RegShape{200, 10, 20, 16, &launch_reg<10, 20, 16>, 1 << 3},
    RegShape{216, 18, 12, 16, &launch_reg<18, 12, 16>, ALL_PASSES & ~(1 << 3)},
    RegShape{216, 18, 12, 32, &launch_reg<18, 12, 32>, 1 << 3},
    // 512 and 768 (256^3 and 384^3 targets): 32 values per thread and component;
    // single-component passes in registers, pass 3 split over component groups
    DDH_SHAPE1(512, 32, 16, 8), DDH_SHAPE1(512, 32, 16, 4),
    RegShape{512, 16, 16, 8, &launch_p3m<2, 8>, 1 << 3}, RegShape{512, 16, 16, 16, &launch_p3m<2, 16>, 1 << 3},
    RegShape{512, 16, 16, 4, &launch_p3m<2, 4>, 1 << 3},
    RegShape{512, 32, 16, 8, &launch_split3<32, 16, 8>, 1 << 3},
    RegShape{512, 32, 16, 4, &launch_split3<32, 16, 4>, 1 << 3},
    DDH_SHAPE1(768, 32, 24, 8), DDH_SHAPE1(768, 32, 24, 4),
    RegShape{768, 16, 16, 8, &launch_p3m<3, 8>, 1 << 3}, RegShape{768, 16, 16, 4, &launch_p3m<3, 4>, 1 << 3},
    RegShape{768, 32, 24, 8, &launch_split3<32, 24, 8>, 1 << 3},
    RegShape{768, 32, 24, 4, &launch_split3<32, 24, 4>, 1 << 3},
};
#define DDH_PAIR(L_, Q_, CR_, CP_) RegShape{L_, Q_, Q_, CR_, &launch_pair<Q_, CR_, CP_>, 1 << 3}
// Autotuner candidates listed after the defaults of their lengths: lane-pair
// and mirror-row pass-3 kernels, a narrower pass-1 tile, and the 512- and
// 768-point y passes and pass 5 by decimation over 256-point FFTs.
static const RegShape kPairShapes[] = {
    DDH_PAIR(64, 8, 32, 32), DDH_PAIR(144, 12, 16, 16), DDH_PAIR(256, 16, 16, 16), DDH_PAIR(256, 16, 16, 8),
    DDH_SHAPE(256, 16, 16, 32),
    RegShape{256, 16, 16, 16, &launch_ypair<16, 16, 16>, 1 << 3}, RegShape{128, 8, 16, 32, &launch_ypair<8, 16, 32>, 1 << 3},
    RegShape{64, 8, 8, 32, &launch_ypair<8, 8, 32>, 1 << 3},
    RegShape{256, 16, 16, 8, &launch_reg1<16, 16, 8>, 1 << 1},
    RegShape{216, 18, 12, 8, &launch_reg<18, 12, 8>, 1 << 3}, RegShape{216, 12, 18, 16, &launch_reg<12, 18, 16>, 1 << 3},
    RegShape{200, 20, 10, 8, &launch_reg<20, 10, 8>, 1 << 3}, RegShape{200, 20, 10, 32, &launch_reg<20, 10, 32>, 1 << 3},
    RegShape{200, 20, 10, 16, &launch_reg<20, 10, 16>, 1 << 3},
    RegShape{512, 16, 16, 8, &launch_ym<2, 8>, (1 << 2) | (1 << 4)},
    RegShape{512, 16, 16, 16, &launch_ym<2, 16>, (1 << 2) | (1 << 4)},
    RegShape{768, 16, 16, 8, &launch_ym<3, 8>, (1 << 2) | (1 << 4)},
    RegShape{768, 16, 16, 16, &launch_ym<3, 16>, (1 << 2) | (1 << 4)},
    RegShape{512, 16, 16, 4, &launch_p5m<2, 4, false>, 1 << 5}, RegShape{512, 16, 16, 8, &launch_p5m<2, 8, false>, 1 << 5},
    RegShape{768, 16, 16, 4, &launch_p5m<3, 4, true>, 1 << 5}, RegShape{768, 16, 16, 4, &launch_p5m<3, 4, false>, 1 << 5},
    RegShape{768, 16, 16, 8, &launch_p5m<3, 8, false>, 1 << 5},
    // register Stockham Green passes: SkPlan<V, radices>, C columns, NBUF
    // exchange buffers, K components per exchange, waves per SIMD the
    // registers must allow, Green points per load group, exchange addresses
    // recomputed per exchange (r1 = values per thread, r2 = first radix)
    RegShape{512, 8, 8, 16, &launch_p3s<SkPlan<8, 8, 8, 8>, 16, 1, 1, 4>, 1 << 3},
    RegShape{512, 8, 8, 16, &launch_p3s<SkPlan<8, 8, 8, 8>, 16, 1, 1, 4, 4, true>, 1 << 3},
    RegShape{512, 8, 8, 8, &launch_p3s<SkPlan<8, 8, 8, 8>, 8, 2, 1, 3>, 1 << 3},
    RegShape{512, 8, 8, 8, &launch_p3s<SkPlan<8, 8, 8, 8>, 8, 2, 1, 4, 4, true>, 1 << 3},
    RegShape{512, 16, 8, 8, &launch_p3s<SkPlan<16, 8, 8, 8>, 8, 1, 1, 2>, 1 << 3},
    RegShape{768, 12, 12, 8, &launch_p3s<SkPlan<12, 12, 4, 4, 4>, 8, 2, 1, 2>, 1 << 3},
    RegShape{768, 12, 12, 8, &launch_p3s<SkPlan<12, 12, 4, 4, 4>, 8, 1, 1, 2>, 1 << 3},
    RegShape{768, 12, 12, 4, &launch_p3s<SkPlan<12, 12, 4, 4, 4>, 4, 1, 1, 3, 4, true>, 1 << 3},
    RegShape{256, 8, 8, 8, &launch_p3s<SkPlan<8, 8, 8, 4>, 8, 1, 1, 4, 4, true>, 1 << 3},
    RegShape{256, 8, 8, 16, &launch_p3s<SkPlan<8, 8, 8, 4>, 16, 1, 1, 4, 4, true>, 1 << 3},
    // 256-point Green pass with the first two points' Green loads issued
    // before the third forward transform (three or four points: 170-192
    // VGPRs and 2 waves per SIMD, or spills at 3)
    RegShape{256, 16, 2, 16, &launch_p3pre<16, 16, 16, 2>, 1 << 3},
    // and at 144-192 points (180 and 160 points: 6-7 % faster)
    RegShape{192, 16, 2, 16, &launch_p3pre<16, 12, 16, 2>, 1 << 3},
    RegShape{180, 15, 2, 16, &launch_p3pre<15, 12, 16, 2>, 1 << 3},
    RegShape{160, 16, 2, 16, &launch_p3pre<16, 10, 16, 2>, 1 << 3},
    RegShape{144, 12, 2, 16, &launch_p3pre<12, 12, 16, 2>, 1 << 3},
    // pass 5 with finish load groups of 2 and 3 outputs (123 and 125 VGPRs,
    // 4 waves per SIMD)
    RegShape{256, 16, 16, 16, &launch_p5f<16, 16, 16, 2>, 1 << 5},
    RegShape{192, 16, 12, 16, &launch_p5f<16, 12, 16, 3>, 1 << 5},
    // Stockham Green passes with the first two points' Green loads issued
    // before the last component's forward transform
    RegShape{512, 8, 8, 8, &launch_p3s<SkPlan<8, 8, 8, 8>, 8, 2, 1, 4, 4, true, 2>, 1 << 3},
    RegShape{768, 12, 12, 8, &launch_p3s<SkPlan<12, 12, 4, 4, 4>, 8, 2, 1, 2, 4, false, 2>, 1 << 3},
};
#undef DDH_PAIR
#undef DDH_SHAPE

// ------------------------------------------------ SLICES x passes (register)
// Rows of a batch of kz planes, row r = (b 3 + c) NY + iy for plane
// kz = k0 + b and component c.  DIR = +1: x-FFT of zwork row
// [c][kz][iy] (NX values, zero-padded to 2NX) into slice row r (2NX
// spectral values).  DIR = -1: inverse x-FFT of slice row r back into the
// zwork row (its NX unpadded outputs).  A row past the end is clamped to
// the last row for loading and never stored.
template <int R1, int R2, int C, int DIR>
__global__ __launch_bounds__(R1 * C) void kr_slx(float2* __restrict__ zw, float2* __restrict__ sl,
                                             const float2* __restrict__ twg, int nx, int ny, int gz, int k0,
                                             int nrows, const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  constexpr int L = R1 * R2, RT = R1 * C, P = row_stride<R1, R2>(), G = Reg2<R1, R2>::G, VN = Reg2<R1, R2>::VN;
  __shared__ float2 twL[L];
  __shared__ float buf[C * P];  // split exchange, row layout
  const int t = threadIdx.x % R1, s = threadIdx.x / R1;
  const int r = blockIdx.x * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = r < nrows;
  const int rr = live ? r : nrows - 1;
  const int iy = rr % ny, bc = rr / ny, c = bc % 3, b = bc / 3;
  float2* zrow = zw + (((size_t)c * gz + k0 + b) * ny + iy) * nx;
  float2* srow = sl + (size_t)rr * L;
  float2 v[VN];
  if (DIR > 0) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = t + R1 * n2;
      v[n2] = lower_half<R2>(n2, n, nx) ? zrow[n] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft2_fwd_s<R1, R2, +1, 1>(v, buf, twL, t, s, P);
    if (live) {
#pragma unroll
      for (int i = 0; i < G; ++i)
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1)
          if (reg2_owns<R1, R2>(t, i)) srow[(t + R1 * i) + R2 * k1] = v[i * R1 + k1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1)
        v[i * R1 + k1] = reg2_owns<R1, R2>(t, i) ? srow[(t + R1 * i) + R2 * k1] : make_float2(0.f, 0.f);
    __syncthreads();
    fft2_inv_s<R1, R2, -1, 1>(v, buf, twL, t, s, P);
    if (live) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int n = t + R1 * n2;
        if (lower_half<R2>(n2, n, nx)) zrow[n] = v[n2];
      }
    }
  }
}

template <int R1, int R2, int C>
static void launch_slx(Context& c, int dir, float2* zw, float2* sl, int k0, int nb, hipStream_t st,
                       const int* act) {
  const MatvecPlan& m = c.mv;
  const int nrows = nb * 3 * c.ny;
  const dim3 grid((nrows + C - 1) / C);
  if (dir > 0)
    hipLaunchKernelGGL((kr_slx<R1, R2, C, +1>), grid, dim3(R1 * C), 0, st, zw, sl, m.twx, c.nx, c.ny, m.gz, k0, nrows,
                       act);
  else
    hipLaunchKernelGGL((kr_slx<R1, R2, C, -1>), grid, dim3(R1 * C), 0, st, zw, sl, m.twx, c.nx, c.ny, m.gz, k0, nrows,
                       act);
}
using SlxLaunch = void (*)(Context&, int, float2*, float2*, int, int, hipStream_t, const int*);
struct SlxShape {
  int L;
  SlxLaunch fn;
};
// the default row shape of each register length (passes 1 and 5 of FFT3D)
static const SlxShape kSlxShapes[] = {
    {32, &launch_slx<4, 8, 64>},     {48, &launch_slx<8, 6, 32>},     {60, &launch_slx<10, 6, 32>},
    {64, &launch_slx<8, 8, 32>},     {72, &launch_slx<12, 6, 16>},    {80, &launch_slx<10, 8, 32>},
    {96, &launch_slx<12, 8, 16>},    {100, &launch_slx<10, 10, 32>},  {108, &launch_slx<12, 9, 16>},
    {120, &launch_slx<12, 10, 16>},  {128, &launch_slx<8, 16, 32>},   {144, &launch_slx<12, 12, 16>},
    {150, &launch_slx<15, 10, 16>},  {160, &launch_slx<16, 10, 16>},  {180, &launch_slx<15, 12, 16>},
    {192, &launch_slx<16, 12, 16>},  {240, &launch_slx<16, 15, 16>},  {256, &launch_slx<16, 16, 16>},
    {512, &launch_slx<32, 16, 8>},   {768, &launch_slx<32, 24, 8>},
};

int slices_x_reg(Context& c, int dir, float2* zw, float2* sl, int k0, int nb, hipStream_t st, const int* act) {
  if (c.fft_generic) return 0;
  for (const SlxShape& r : kSlxShapes)
    if (r.L == c.mv.gx) {
      r.fn(c, dir, zw, sl, k0, nb, st, act);
      return 1;
    }
  return 0;
}

// ------------------------------------------------ SLICES z passes (register)
// Column FFTs along z over the NX NY columns of a plane (kr_passy with one
// "row", columns at element stride NX NY).  dir = +1: in [3][NZ][NY][NX]
// (the full-layout vector) -> out = zwork [3][2NZ][NY][NX]; dir = -1: zwork ->
// out [3][NZ][NY][NX] (the NZ unpadded outputs).
// one right-hand side (grid z = the 3 components, so RHS index 0) whose
// active flag, when given, is *act
static BatchArgs one_rhs(const int* act) {
  BatchArgs ba;
  ba.active = act;
  return ba;
}
// ba: one_rhs(act), with the gather fields set when the forward pass reads
// the MATVEC's x itself (compact or conjugated)
template <int R1, int R2, int C>
static void launch_slz(Context& c, int dir, const float2* in, float2* out, hipStream_t st, const BatchArgs& ba) {
  const MatvecPlan& m = c.mv;
  const int plane = c.nx * c.ny;
  const size_t full = (size_t)m.gz * plane, half = (size_t)c.nz * plane;
  const dim3 grid((plane + C - 1) / C, 1, 3);
  if (dir > 0 && ba.gsrc)
    hipLaunchKernelGGL((kr_passy<R1, R2, C, +1, true>), grid, dim3(R1 * C), 0, st, in, out, m.twz, plane, c.nz, m.gz,
                       1, half, full, ba);
  else if (dir > 0)
    hipLaunchKernelGGL((kr_passy<R1, R2, C, +1>), grid, dim3(R1 * C), 0, st, in, out, m.twz, plane, c.nz, m.gz, 1, half,
                       full, ba);
  else
    hipLaunchKernelGGL((kr_passy<R1, R2, C, -1>), grid, dim3(R1 * C), 0, st, in, out, m.twz, plane, m.gz, c.nz, 1, full,
                       half, ba);
}
template <int M, int C>
static void launch_slz_m(Context& c, int dir, const float2* in, float2* out, hipStream_t st, const BatchArgs& ba) {
  const MatvecPlan& m = c.mv;
  const int plane = c.nx * c.ny;
  const size_t full = (size_t)m.gz * plane, half = (size_t)c.nz * plane;
  const dim3 grid((plane + C - 1) / C, 1, 3), block(M * 16 * C);
  if (dir > 0 && ba.gsrc)
    hipLaunchKernelGGL((kr_passym<M, C, +1, true>), grid, block, 0, st, in, out, m.twz, plane, c.nz, m.gz, 1, half,
                       full, ba);
  else if (dir > 0)
    hipLaunchKernelGGL((kr_passym<M, C, +1>), grid, block, 0, st, in, out, m.twz, plane, c.nz, m.gz, 1, half, full,
                       ba);
  else
    hipLaunchKernelGGL((kr_passym<M, C, -1>), grid, block, 0, st, in, out, m.twz, plane, m.gz, c.nz, 1, full, half,
                       ba);
}
using SlzLaunch = void (*)(Context&, int, const float2*, float2*, hipStream_t, const BatchArgs&);
struct SlzShape {
  int L;
  SlzLaunch fn;
};
// the default column shape of each register length (passes 2 and 4 of FFT3D);
// 512 and 768 points by decimation over 256-point FFTs (kr_passym, the faster
// y pass of FFT3D at those lengths)
static const SlzShape kSlzShapes[] = {
    {32, &launch_slz<4, 8, 64>},     {48, &launch_slz<8, 6, 32>},     {60, &launch_slz<10, 6, 32>},
    {64, &launch_slz<8, 8, 32>},     {72, &launch_slz<12, 6, 16>},    {80, &launch_slz<10, 8, 32>},
    {96, &launch_slz<12, 8, 16>},    {100, &launch_slz<10, 10, 32>},  {108, &launch_slz<12, 9, 16>},
    {120, &launch_slz<12, 10, 16>},  {128, &launch_slz<8, 16, 32>},   {144, &launch_slz<12, 12, 16>},
    {150, &launch_slz<15, 10, 16>},  {160, &launch_slz<16, 10, 16>},  {180, &launch_slz<15, 12, 16>},
    {192, &launch_slz<16, 12, 16>},  {240, &launch_slz<16, 15, 16>},  {256, &launch_slz<16, 16, 16>},
    {512, &launch_slz_m<2, 16>},     {768, &launch_slz_m<3, 16>},
};

int slices_z_reg_length(int L) {
  for (const SlzShape& r : kSlzShapes)
    if (r.L == L) return 1;
  return 0;
}

int slices_z_reg(Context& c, int dir, const float2* in, float2* out, hipStream_t st, const int* act, bool gather,
                 int compact, int cj) {
  if (c.fft_generic) return 0;
  BatchArgs ba = one_rhs(act);
  if (gather && dir > 0) {
    ba.gsrc = in;
    ba.gmap = compact ? c.qmap : nullptr;
    ba.gcs = compact ? (size_t)c.nat0 : (size_t)c.nat;
    ba.gconj = cj;
  }
  for (const SlzShape& r : kSlzShapes)
    if (r.L == c.mv.gz) {
      r.fn(c, dir, in, out, st, ba);
      return 1;
    }
  return 0;
}

// ------------------------------------------------- SLICES y pass (register)
// y-FFT, Green multiply at kz and inverse y-FFT of nb slice planes starting
// at kz = k0, with the register kernels of pass 3 (column setup p3_col<true>)
template <int R1, int R2, int C>
static void launch_sly(Context& c, float2* sl, int k0, int nb, hipStream_t st, const int* act) {
  const MatvecPlan& m = c.mv;
  hipLaunchKernelGGL((kr_sly<R1, R2, C>), dim3((m.gx + C - 1) / C, nb, 1), dim3(R1 * C), 0, st, sl, c.green, m.twy,
                     c.nx, c.ny, c.nz, m.gx, m.gy, (size_t)k0, one_rhs(act), green_geo(c));
}
template <int M, int C>
static void launch_sly_m(Context& c, float2* sl, int k0, int nb, hipStream_t st, const int* act) {
  const MatvecPlan& m = c.mv;
  hipLaunchKernelGGL((kr_pass3m<M, C, true>), dim3((m.gx + C - 1) / C, nb, 1), dim3(M * 16 * C), 0, st, sl, c.green,
                     m.twy, c.nx, c.ny, c.nz, m.gx, m.gy, (size_t)k0, one_rhs(act), green_geo(c), P3Grid{});
}
// SLICES y pass by the register Stockham kernel (kr_pass3s along y)
template <class PL, int C, int NBUF, int K, int WPE, int GS, bool RC>
static void launch_sly_s(Context& c, float2* sl, int k0, int nb, hipStream_t st, const int* act) {
  const MatvecPlan& m = c.mv;
  hipLaunchKernelGGL((kr_pass3s<PL, C, NBUF, K, WPE, true, GS, RC>), dim3((m.gx + C - 1) / C, nb, 1), dim3(PL::T * C),
                     0, st, sl, c.green, m.twy, c.nx, c.ny, c.nz, m.gx, m.gy, (size_t)k0, one_rhs(act), green_geo(c),
                     P3Grid{});
}
using SlyLaunch = void (*)(Context&, float2*, int, int, hipStream_t, const int*);
struct SlyShape {
  int L;
  SlyLaunch fn;
  bool off32 = false;  // addresses the Green array with 32-bit byte offsets
};
// the default pass-3 shape of each register length
static const SlyShape kSlyShapes[] = {
    {32, &launch_sly<4, 8, 64>},     {48, &launch_sly<8, 6, 32>},     {60, &launch_sly<10, 6, 32>},
    {64, &launch_sly<8, 8, 32>},     {72, &launch_sly<12, 6, 16>},    {80, &launch_sly<10, 8, 32>},
    {96, &launch_sly<12, 8, 16>},    {100, &launch_sly<10, 10, 32>},  {108, &launch_sly<12, 9, 16>},
    {120, &launch_sly<12, 10, 16>},  {128, &launch_sly<8, 16, 32>},   {144, &launch_sly<12, 12, 16>},
    {150, &launch_sly<15, 10, 16>},  {160, &launch_sly<16, 10, 16>},  {180, &launch_sly<15, 12, 16>},
    {192, &launch_sly<16, 12, 16>},  {240, &launch_sly<16, 15, 16>},  {256, &launch_sly<16, 16, 16>},
    {512, &launch_sly_s<SkPlan<8, 8, 8, 8>, 8, 2, 1, 4, 4, true>, true},
    {768, &launch_sly_s<SkPlan<12, 12, 4, 4, 4>, 8, 2, 1, 2, 4, false>, true},
};

int slices_y_green_reg(Context& c, float2* sl, int k0, int nb, hipStream_t st, const int* act) {
  if (c.fft_generic) return 0;
  for (const SlyShape& r : kSlyShapes)
    if (r.L == c.mv.gy) {
      // the slice buffers are small; a Green array of 4 GiB or more takes
      // the LDS-staged kernel
      if (r.off32 && green_points(c.nx, c.ny, c.nz, c.ipbc) * 6 * sizeof(float2) >= ((size_t)1 << 32)) return 0;
      r.fn(c, sl, k0, nb, st, act);
      return 1;
    }
  return 0;
}

// shape `cand` (0 = default) of length L serving pass `which`, or nullptr
static const RegShape* find_shape(int L, int which, int cand) {
  for (const RegShape& r : kShapes)
    if (r.L == L && (r.passes >> which & 1) && cand-- == 0) return &r;
  for (const RegShape& r : kPairShapes)
    if (r.L == L && (r.passes >> which & 1) && cand-- == 0) return &r;
  return nullptr;
}

int reg_candidates(int L, int which) {
  int n = 0;
  for (const RegShape& r : kShapes) n += r.L == L && (r.passes >> which & 1);
  for (const RegShape& r : kPairShapes) n += r.L == L && (r.passes >> which & 1);
  return n;
}

// launches register pass `which` (1..5) with shape `cand` of its axis length;
// returns 1 when launched, 0 when the caller should use the generic kernel
int reg_pass(Context& c, int which, int cand, int cj, int compact, const float2* x, float2* y, const float2* adia,
             const float2* aoff, hipStream_t st, const MvBatch& bt) {
  const MatvecPlan& m = c.mv;
  const int L = (which == 1 || which == 5) ? m.gx : (which == 3 ? m.gz : m.gy);
  const RegShape* r = cand >= 0 ? find_shape(L, which, cand) : nullptr;
  if (!r) return 0;
  // register Green passes address a2 and the Green array with 32-bit byte
  // offsets; larger grids take the LDS-staged kernel
  if (which == 3 && !p3s_fits(c)) return 0;
  r->fn(c, cj, compact, x, y, adia, aoff, st, bt, which);
  return 1;
}

}  // namespace ddhip
