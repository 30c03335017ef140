// FFT3D MATVEC passes built on the register-resident two-level FFT
// (fft_reg.hpp) for padded axis lengths 2N = R1 * R2 in {64, 128, 256}.
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

namespace ddhip {

static constexpr int RT = 256;  // threads per workgroup

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
template <int R1, int R2>
__global__ __launch_bounds__(RT) void kr_pass1(const float2* __restrict__ x, float2* __restrict__ a1, SiteMap sm,
                                               const float2* __restrict__ twg, int nx, int nrows, int nat, int nat0,
                                               int compact, int cj, BatchArgs ba) {
  constexpr int L = R1 * R2, C = RT / R1, P = C + 1;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  x += rb * ba.xs;
  a1 += rb * ba.a1s;
  const int t = threadIdx.x % R1, s = threadIdx.x / R1;
  const int g = blockIdx.x * C + s, c = blockIdx.y;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  float2 v[R2];
  const bool live = g < nrows;
  int lo = 0, q0 = 0, cnt = 0;
  if (compact && sm.contiguous && live) {
    lo = sm.rowlo[g]; q0 = sm.rowptr[g]; cnt = sm.rowptr[g + 1] - q0;
  }
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int n = t + R1 * n2;
    float2 val = make_float2(0.f, 0.f);
    if (live && n < nx) {
      if (!compact) {
        val = x[(size_t)c * nat + (size_t)g * nx + n];
      } else if (sm.contiguous) {
        const int d = n - lo;
        if (d >= 0 && d < cnt) val = x[(size_t)c * nat0 + q0 + d];
      } else {
        const int q = sm.qmap[(size_t)g * nx + n];
        if (q >= 0) val = x[(size_t)c * nat0 + q];
      }
      if (cj) val.y = -val.y;
    }
    v[n2] = val;
  }
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v, buf, twL, t, s, P);
  if (live) {
    float2* out = a1 + ((size_t)c * nrows + g) * L;
#pragma unroll
    for (int i = 0; i < R2 / R1; ++i)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) out[(t + R1 * i) + R2 * k1] = v[i * R1 + k1];
  }
}

// ------------------------------------------------------------ passes 2, 4
// in [3][nz][nin][gx] -> out [3][nz][nout][gx] along the middle axis
template <int R1, int R2, int DIR>
__global__ __launch_bounds__(RT) void kr_passy(const float2* __restrict__ in, float2* __restrict__ out,
                                               const float2* __restrict__ twg, int gx, int nin, int nout, int nz,
                                               BatchArgs ba) {
  constexpr int L = R1 * R2, C = RT / R1, P = C + 1;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
  const int rb = blockIdx.z / 3;
  if (ba.active && !ba.active[rb]) return;
  in += rb * (DIR > 0 ? ba.a1s : ba.a2s);
  out += rb * (DIR > 0 ? ba.a2s : ba.a1s);
  const int s = threadIdx.x % C, t = threadIdx.x / C;
  const int kx = blockIdx.x * C + s, iz = blockIdx.y, c = blockIdx.z % 3;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  const float2* src = in + ((size_t)c * nz + iz) * (size_t)nin * gx + kx;
  float2* dst = out + ((size_t)c * nz + iz) * (size_t)nout * gx + kx;
  float2 v[R2];
  if (DIR > 0) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = t + R1 * n2;
      v[n2] = (live && n < nin) ? src[(size_t)n * gx] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft2_fwd<R1, R2, +1>(v, buf, twL, t, s, P);
    if (live) {
#pragma unroll
      for (int i = 0; i < R2 / R1; ++i)
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) dst[(size_t)((t + R1 * i) + R2 * k1) * gx] = v[i * R1 + k1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < R2 / R1; ++i)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1)
        v[i * R1 + k1] = live ? src[(size_t)((t + R1 * i) + R2 * k1) * gx] : make_float2(0.f, 0.f);
    __syncthreads();
    fft2_inv<R1, R2, -1>(v, buf, twL, t, s, P);
    if (live) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int n = t + R1 * n2;
        if (n < nout) dst[(size_t)n * gx] = v[n2];
      }
    }
  }
}

// ------------------------------------------------------------------ pass 3
template <int R1, int R2>
__global__ __launch_bounds__(RT) void kr_pass3(float2* __restrict__ a2, const float2* __restrict__ green,
                                               const float2* __restrict__ twg, int nx, int ny, int nz, int gx,
                                               int gy, BatchArgs ba) {
  constexpr int L = R1 * R2, C = RT / R1, P = C + 1;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int s = threadIdx.x % C, t = threadIdx.x / C;
  // Grid order puts Green-sharing blocks on one XCD (workgroups go round-robin
  // over the 8 XCDs by linear id): ky runs 0, ny, 1, gy-1, 2, gy-2, ... so
  // mirror rows are one grid row apart, and column tile b and its mirror
  // G-1-b are h = ceil(G/2) grid columns apart.
  const int j = blockIdx.y;
  const int ky = j == 0 ? 0 : j == 1 ? ny : ((j & 1) ? gy - ((j - 2) / 2 + 1) : (j - 2) / 2 + 1);
  const int G = gridDim.x, h = (G + 1) / 2;
  const int tile = (int)blockIdx.x < h ? (int)blockIdx.x : G - 1 - ((int)blockIdx.x - h);
  const int kx = tile * C + s;
  for (int i = threadIdx.x; i < L; i += RT) twL[i] = twg[i];
  const bool live = kx < gx;
  const size_t plane = (size_t)gy * gx, cs = (size_t)nz * plane;
  float2* col = a2 + (size_t)ky * gx + kx;
  float2 v0[R2], v1[R2], v2[R2];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int n = t + R1 * n2;
    const bool ok = live && n < nz;
    const size_t o = (size_t)n * plane;
    v0[n2] = ok ? col[o] : make_float2(0.f, 0.f);
    v1[n2] = ok ? col[o + cs] : make_float2(0.f, 0.f);
    v2[n2] = ok ? col[o + 2 * cs] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_fwd<R1, R2, +1>(v2, buf, twL, t, s, P);
  // 3x3 symmetric Green multiply with octant parity reconstruction (cu:109-115).
  // Each octant entry serves up to eight (kx, ky, kz) mirrors; kz mirrors
  // are in this block and the kx / ky mirror blocks run on the same XCD at
  // about the same time (grid mapping above), so repeats hit in L2.
  const int ox = nx + 1, oy = ny + 1;
  const int ir = live ? min(kx, gx - kx) : 0;
  const int jr = min(ky, gy - ky);
  const float sxy = ((kx <= nx) == (ky <= ny)) ? 1.f : -1.f;
  const float sx = kx <= nx ? 1.f : -1.f, sy = ky <= ny ? 1.f : -1.f;
  const size_t pb = (size_t)ir + (size_t)ox * jr;
#pragma unroll
  for (int i = 0; i < R2 / R1; ++i) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int kz = (t + R1 * i) + R2 * k1;
      const int kr = min(kz, L - kz);
      const float sz = kz <= nz ? 1.f : -1.f;
      const Green6 G = green_at(green, pb + (size_t)ox * oy * kr);
      const float2 g11 = G.xx;
      const float2 g12 = cscale(G.xy, sxy);
      const float2 g13 = cscale(G.xz, sx * sz);
      const float2 g22 = G.yy;
      const float2 g23 = cscale(G.yz, sy * sz);
      const float2 g33 = G.zz;
      const int e = i * R1 + k1;
      const float2 p0 = v0[e], p1 = v1[e], p2 = v2[e];
      v0[e] = cadd(cadd(cmul(g11, p0), cmul(g12, p1)), cmul(g13, p2));
      v1[e] = cadd(cadd(cmul(g12, p0), cmul(g22, p1)), cmul(g23, p2));
      v2[e] = cadd(cadd(cmul(g13, p0), cmul(g23, p1)), cmul(g33, p2));
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
      if (n < nz) {
        const size_t o = (size_t)n * plane;
        col[o] = v0[n2];
        col[o + cs] = v1[n2];
        col[o + 2 * cs] = v2[n2];
      }
    }
  }
}

// ------------------------------------------------------------------ pass 5
template <int R1, int R2>
__global__ __launch_bounds__(RT) void kr_pass5(const float2* __restrict__ a1, const float2* __restrict__ twg,
                                               const float2* __restrict__ x, float2* __restrict__ y,
                                               const float2* __restrict__ adia, const float2* __restrict__ aoff,
                                               const int16_t* __restrict__ iocc, SiteMap sm, int nx, int nrows,
                                               int nat, int nat0, int compact, int cj, float inv, BatchArgs ba) {
  constexpr int L = R1 * R2, C = RT / R1, P = C + 1;
  __shared__ float2 twL[L];
  __shared__ float2 buf[L * P];
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
  const float2* src = a1 + (size_t)g * L;
  float2 v0[R2], v1[R2], v2[R2];
#pragma unroll
  for (int i = 0; i < R2 / R1; ++i)
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int k = (t + R1 * i) + R2 * k1;
      const int e = i * R1 + k1;
      v0[e] = live ? src[k] : make_float2(0.f, 0.f);
      v1[e] = live ? src[k + cs] : make_float2(0.f, 0.f);
      v2[e] = live ? src[k + 2 * cs] : make_float2(0.f, 0.f);
    }
  __syncthreads();
  fft2_inv<R1, R2, -1>(v0, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v1, buf, twL, t, s, P);
  __syncthreads();
  fft2_inv<R1, R2, -1>(v2, buf, twL, t, s, P);
  if (!live) return;
  int lo = 0, q0 = 0, cnt = 0;
  if (compact && sm.contiguous) {
    lo = sm.rowlo[g]; q0 = sm.rowptr[g]; cnt = sm.rowptr[g + 1] - q0;
  }
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const int ix = t + R1 * n2;
    if (ix >= nx) continue;
    size_t o, stride;
    if (!compact) {
      o = (size_t)g * nx + ix;
      stride = (size_t)nat;
      if (iocc && iocc[o] == 0) {
        y[o] = y[o + stride] = y[o + 2 * stride] = make_float2(0.f, 0.f);
        continue;
      }
    } else {
      int q;
      if (sm.contiguous) {
        const int d = ix - lo;
        q = (d >= 0 && d < cnt) ? q0 + d : -1;
      } else {
        q = sm.qmap[(size_t)g * nx + ix];
      }
      if (q < 0) continue;
      o = (size_t)q;
      stride = (size_t)nat0;
    }
    float2 x1 = x[o], x2 = x[o + stride], x3 = x[o + 2 * stride];
    if (cj) { x1.y = -x1.y; x2.y = -x2.y; x3.y = -x3.y; }
    const float2 a23 = aoff[o], a31 = aoff[o + stride], a12 = aoff[o + 2 * stride];
    float2 r1 = csub(cmul(adia[o], x1), cscale(v0[n2], inv));
    float2 r2 = csub(cmul(adia[o + stride], x2), cscale(v1[n2], inv));
    float2 r3 = csub(cmul(adia[o + 2 * stride], x3), cscale(v2[n2], inv));
    r1 = cadd(r1, cadd(cmul(a31, x3), cmul(a12, x2)));
    r2 = cadd(r2, cadd(cmul(a12, x1), cmul(a23, x3)));
    r3 = cadd(r3, cadd(cmul(a23, x2), cmul(a31, x1)));
    if (cj) { r1.y = -r1.y; r2.y = -r2.y; r3.y = -r3.y; }
    y[o] = r1; y[o + stride] = r2; y[o + 2 * stride] = r3;
  }
}

// ------------------------------------------------------------------ host --
template <int R1, int R2>
static int launch_reg(Context& c, int cj, int compact, const float2* x, float2* y, const float2* adia,
                      const float2* aoff, hipStream_t st, const MvBatch& bt, int which) {
  const MatvecPlan& m = c.mv;
  constexpr int C = RT / R1;
  const int nrows = c.ny * c.nz;
  const unsigned nb = (unsigned)bt.nrhs;
  SiteMap sm{c.rowptr, c.rowlo, c.qmap, c.rows_contiguous};
  switch (which) {
    case 1:
      hipLaunchKernelGGL((kr_pass1<R1, R2>), dim3((nrows + C - 1) / C, 3, nb), dim3(RT), 0, st, x, bt.a1, sm,
                         m.twx, c.nx, nrows, c.nat, c.nat0, compact, cj, bt.k);
      break;
    case 2:
      hipLaunchKernelGGL((kr_passy<R1, R2, +1>), dim3((m.gx + C - 1) / C, c.nz, 3 * nb), dim3(RT), 0, st, bt.a1,
                         bt.a2, m.twy, m.gx, c.ny, m.gy, c.nz, bt.k);
      break;
    case 3:
      hipLaunchKernelGGL((kr_pass3<R1, R2>), dim3((m.gx + C - 1) / C, m.gy, nb), dim3(RT), 0, st, bt.a2, c.green,
                         m.twz, c.nx, c.ny, c.nz, m.gx, m.gy, bt.k);
      break;
    case 4:
      hipLaunchKernelGGL((kr_passy<R1, R2, -1>), dim3((m.gx + C - 1) / C, c.nz, 3 * nb), dim3(RT), 0, st, bt.a2,
                         bt.a1, m.twy, m.gx, m.gy, c.ny, c.nz, bt.k);
      break;
    default:
      hipLaunchKernelGGL((kr_pass5<R1, R2>), dim3((nrows + C - 1) / C, nb), dim3(RT), 0, st, bt.a1, m.twx, x, y,
                         adia, aoff, compact ? nullptr : c.iocc, sm, c.nx, nrows, c.nat, c.nat0, compact, cj,
                         1.f / (float)m.ngrid, bt.k);
      break;
  }
  return 0;
}

// axis length -> register FFT shape; 0 if not covered
int reg_shape(int L) {
  switch (L) {
    case 64: return 1;
    case 128: return 2;
    case 256: return 3;
    default: return 0;
  }
}



// launches register pass `which` (1..5) if its axis length is covered; returns
// 1 when launched, 0 when the caller should use the generic kernel
int reg_pass(Context& c, int which, int cj, int compact, const float2* x, float2* y, const float2* adia,
             const float2* aoff, hipStream_t st, const MvBatch& bt) {
  const MatvecPlan& m = c.mv;
  const int L = (which == 1 || which == 5) ? m.gx : (which == 3 ? m.gz : m.gy);
  switch (reg_shape(L)) {
    case 1: launch_reg<8, 8>(c, cj, compact, x, y, adia, aoff, st, bt, which); return 1;
    case 2: launch_reg<8, 16>(c, cj, compact, x, y, adia, aoff, st, bt, which); return 1;
    case 3: launch_reg<16, 16>(c, cj, compact, x, y, adia, aoff, st, bt, which); return 1;
    default: return 0;
  }
}

}  // namespace ddhip
