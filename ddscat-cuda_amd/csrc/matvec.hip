// FFT3D Green-tensor MATVEC for gfx950.
//
// y = adia o x + offdiag(aoff) x - (1/G) F-[ Ghat . F+[pad(x)] ]
// (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:104-118, 406-415; CPU
// semantics ddscat_cuda_unified/fortran/cprod_cuda.f90:446-607), computed as
// five pruned, LDS-staged 1-D FFT passes over the zero-padded 2NX x 2NY x 2NZ
// grid.  Only the non-zero input half of each axis is read and only the
// needed output half is written:
//
//   pass 1  x-FFT (+i)   x[3][NAT] or compact x      -> a1[3][NZ][NY][2NX]
//   pass 2  y-FFT (+i)   a1                            -> a2[3][NZ][2NY][2NX]
//   pass 3  z-FFT (+i), 3x3 Green multiply, z-FFT (-i), keep iz < NZ (in place in a2)
//   pass 4  y-FFT (-i)   a2, keep iy < NY              -> a1
//   pass 5  x-FFT (-i), keep ix < NX, finish with x, adia, aoff -> y
//
// The x axis stays contiguous in every buffer, so passes 2-4 load tiles of
// consecutive kx columns (coalesced rows) and transform along the strided
// axis inside LDS.
#include "state.hpp"
#include "fft_device.hpp"
#include <vector>
#include <cmath>
#include <algorithm>
#include <array>
#include <map>

namespace ddhip {

static constexpr int MV_THREADS = 256;
static constexpr int PER_WIDE = 16;   // passes 1, 2, 4
static constexpr int PER_NARROW = 8;  // passes 3, 5 (three components in LDS)
// Wide tiles: long axes (a 256-thread tile would hold fewer than 8 columns,
// i.e. rows of < 64 contiguous bytes) run passes 2-4 with 1024-thread
// workgroups and up to the device's per-workgroup LDS, so each tile reads
// rows of 8-16 columns.
static constexpr int WIDE_THREADS = 1024;

// ---------------------------------------------------------------- pass 1 --
template <int PER>
__global__ __launch_bounds__(MV_THREADS) void k_pass1_xfwd(
    const float2* __restrict__ x, float2* __restrict__ a1, const int* __restrict__ rowptr,
    const int* __restrict__ occ, const float2* __restrict__ twg, FftPlan p, int nx, int nrows,
    int rx, int nat, int nat0, int compact, int conj_in, BatchArgs ba) {
  extern __shared__ float2 sm[];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  x += rb * ba.xs;
  a1 += rb * ba.a1s;
  const int L = p.L;
  float2* tw = sm;
  float2* buf = sm + L;
  const int c = blockIdx.y;
  const int g0 = blockIdx.x * rx;
  const int T = blockDim.x;
  const int nr = min(rx, nrows - g0);
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  if (!compact) {
    const float2* xs = x + (size_t)c * nat + (size_t)g0 * nx;
    for (int idx = threadIdx.x; idx < rx * L; idx += T) {
      const int r = idx / L, e = idx - r * L;
      float2 v = make_float2(0.f, 0.f);
      if (r < nr && e < nx) {
        v = xs[(size_t)r * nx + e];
        if (conj_in) v.y = -v.y;
      }
      buf[idx] = v;
    }
  } else {
    for (int idx = threadIdx.x; idx < rx * L; idx += T) buf[idx] = make_float2(0.f, 0.f);
    __syncthreads();
    const int q0 = rowptr[g0], q1 = rowptr[g0 + nr];
    const float2* xs = x + (size_t)c * nat0;
    for (int q = q0 + threadIdx.x; q < q1; q += T) {
      const int j = occ[q];
      const int g = j / nx;
      const int e = j - g * nx;
      float2 v = xs[q];
      if (conj_in) v.y = -v.y;
      buf[(g - g0) * L + e] = v;
    }
  }
  __syncthreads();
  fft_run<+1, false, PER>(buf, tw, p, rx, 1, L);
  float2* out = a1 + ((size_t)c * nrows + g0) * L;
  for (int idx = threadIdx.x; idx < nr * L; idx += T) out[idx] = buf[idx];
}

// ------------------------------------------------------------- pass 2, 4 --
// in [3][nz][nin][gx] -> out [3][nz][nout][gx], FFT along the middle axis.
template <int DIR, int PER, int NT = MV_THREADS>
__global__ __launch_bounds__(NT) void k_pass_y(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ twg,
    FftPlan p, int gx, int nin, int nout, int nz, int cy_log2, size_t ins, size_t outs, BatchArgs ba) {
  extern __shared__ float2 sm[];
  const int rb = blockIdx.z / 3;
  if (ba.active && !ba.active[rb]) return;
  in += rb * (DIR > 0 ? ba.a1s : ba.a2s);
  out += rb * (DIR > 0 ? ba.a2s : ba.a1s);
  const int L = p.L;
  const int cy = 1 << cy_log2;
  float2* tw = sm;
  float2* buf = sm + L;
  const int c = blockIdx.z % 3, iz = blockIdx.y, kx0 = blockIdx.x * cy;
  const int T = blockDim.x;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  const float2* src = in + ((size_t)c * nz + iz) * ins;
  for (int idx = threadIdx.x; idx < L * cy; idx += T) {
    const int e = idx >> cy_log2, col = idx & (cy - 1);
    const int kx = kx0 + col;
    float2 v = make_float2(0.f, 0.f);
    if (e < nin && kx < gx) v = src[(size_t)e * gx + kx];
    buf[idx] = v;
  }
  __syncthreads();
  fft_run<DIR, true, PER>(buf, tw, p, cy, cy, 1);
  float2* dst = out + ((size_t)c * nz + iz) * outs;
  for (int idx = threadIdx.x; idx < nout * cy; idx += T) {
    const int e = idx >> cy_log2, col = idx & (cy - 1);
    const int kx = kx0 + col;
    if (kx < gx) dst[(size_t)e * gx + kx] = buf[idx];
  }
}

// ---------------------------------------------------------------- pass 3 --
template <int PER, int NT = MV_THREADS>
__global__ __launch_bounds__(NT) void k_pass3_z_green(
    float2* __restrict__ a2, const float2* __restrict__ green, const float2* __restrict__ twg,
    FftPlan p, int nx, int ny, int nz, int gx, int gy, int cz_log2, size_t plane, BatchArgs ba, GreenGeo gg) {
  extern __shared__ float2 sm[];
  const int rb = blockIdx.z;
  if (ba.active && !ba.active[rb]) return;
  a2 += rb * ba.a2s;
  const int L = p.L;  // = 2 nz
  const int cz = 1 << cz_log2;
  float2* tw = sm;
  float2* b0 = sm + L;
  float2* b1 = b0 + L * cz;
  float2* b2 = b1 + L * cz;
  const int ky = blockIdx.y, kx0 = blockIdx.x * cz;
  const int T = blockDim.x;
  const size_t cstride = (size_t)nz * plane;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  for (int idx = threadIdx.x; idx < L * cz; idx += T) {
    const int e = idx >> cz_log2, col = idx & (cz - 1);
    const int kx = kx0 + col;
    float2 v0 = make_float2(0.f, 0.f), v1 = v0, v2 = v0;
    if (e < nz && kx < gx) {
      const size_t o = (size_t)e * plane + (size_t)ky * gx + kx;
      v0 = a2[o];
      v1 = a2[o + cstride];
      v2 = a2[o + 2 * cstride];
    }
    b0[idx] = v0; b1[idx] = v1; b2[idx] = v2;
  }
  __syncthreads();
  fft_run<+1, true, PER>(b0, tw, p, cz, cz, 1);
  fft_run<+1, true, PER>(b1, tw, p, cz, cz, 1);
  fft_run<+1, true, PER>(b2, tw, p, cz, cz, 1);
  // symmetric 3x3 multiply; trimmed-octant reconstruction for IPBC=0
  // (cu:109-115), whole grid for IPBC=1
  const int ox = gg.ox, oy = gg.oy;
  const int jr = min(ky, gg.my - ky);
  const float sy = ky <= gg.ty ? 1.f : -1.f;
  for (int idx = threadIdx.x; idx < L * cz; idx += T) {
    const int kz = idx >> cz_log2, col = idx & (cz - 1);
    const int kx = kx0 + col;
    if (kx >= gx) continue;
    const int ir = min(kx, gg.mx - kx), kr = min(kz, gg.mz - kz);
    const float sx = kx <= gg.tx ? 1.f : -1.f;
    const float sz = kz <= gg.tz ? 1.f : -1.f;
    const Green6 G = green_at(green, (size_t)ir + (size_t)ox * ((size_t)jr + (size_t)oy * kr));
    const float2 g11 = G.xx;
    const float2 g12 = cscale(G.xy, sx * sy);
    const float2 g13 = cscale(G.xz, sx * sz);
    const float2 g22 = G.yy;
    const float2 g23 = cscale(G.yz, sy * sz);
    const float2 g33 = G.zz;
    const float2 p0 = b0[idx], p1 = b1[idx], p2 = b2[idx];
    b0[idx] = cadd(cadd(cmul(g11, p0), cmul(g12, p1)), cmul(g13, p2));
    b1[idx] = cadd(cadd(cmul(g12, p0), cmul(g22, p1)), cmul(g23, p2));
    b2[idx] = cadd(cadd(cmul(g13, p0), cmul(g23, p1)), cmul(g33, p2));
  }
  __syncthreads();
  fft_run<-1, true, PER>(b0, tw, p, cz, cz, 1);
  fft_run<-1, true, PER>(b1, tw, p, cz, cz, 1);
  fft_run<-1, true, PER>(b2, tw, p, cz, cz, 1);
  for (int idx = threadIdx.x; idx < nz * cz; idx += T) {
    const int e = idx >> cz_log2, col = idx & (cz - 1);
    const int kx = kx0 + col;
    if (kx < gx) {
      const size_t o = (size_t)e * plane + (size_t)ky * gx + kx;
      a2[o] = b0[idx];
      a2[o + cstride] = b1[idx];
      a2[o + 2 * cstride] = b2[idx];
    }
  }
}

// ---------------------------------------------------------------- pass 5 --
__device__ __forceinline__ void finish_site(float2 x1, float2 x2, float2 x3, float2 d1, float2 d2,
                                            float2 d3, float2 a23, float2 a31, float2 a12,
                                            float2 w1, float2 w2, float2 w3, float inv, bool cj,
                                            float2& r1, float2& r2, float2& r3) {
  if (cj) { x1.y = -x1.y; x2.y = -x2.y; x3.y = -x3.y; }
  r1 = csub(cmul(d1, x1), cscale(w1, inv));
  r2 = csub(cmul(d2, x2), cscale(w2, inv));
  r3 = csub(cmul(d3, x3), cscale(w3, inv));
  r1 = cadd(r1, cadd(cmul(a31, x3), cmul(a12, x2)));
  r2 = cadd(r2, cadd(cmul(a12, x1), cmul(a23, x3)));
  r3 = cadd(r3, cadd(cmul(a23, x2), cmul(a31, x1)));
  if (cj) { r1.y = -r1.y; r2.y = -r2.y; r3.y = -r3.y; }
}

template <int PER>
__global__ __launch_bounds__(MV_THREADS) void k_pass5_xinv_finish(
    const float2* __restrict__ a1, const float2* __restrict__ twg, FftPlan p,
    const float2* __restrict__ x, float2* __restrict__ y, const float2* __restrict__ adia,
    const float2* __restrict__ aoff, const int16_t* __restrict__ iocc,
    const int* __restrict__ rowptr, const int* __restrict__ occ, int nx, int nrows, int rx,
    int nat, int nat0, int compact, int cj, float inv, BatchArgs ba) {
  extern __shared__ float2 sm[];
  const int rb = blockIdx.y;
  if (ba.active && !ba.active[rb]) return;
  a1 += rb * ba.a1s;
  x += rb * ba.xs;
  y += rb * ba.xs;
  adia += rb * ba.os;
  aoff += rb * ba.os;
  const int L = p.L;
  float2* tw = sm;
  float2* b0 = sm + L;
  float2* b1 = b0 + rx * L;
  float2* b2 = b1 + rx * L;
  const int g0 = blockIdx.x * rx;
  const int T = blockDim.x;
  const int nr = min(rx, nrows - g0);
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  const size_t cs = (size_t)nrows * L;
  const float2* src = a1 + (size_t)g0 * L;
  for (int idx = threadIdx.x; idx < rx * L; idx += T) {
    float2 v0 = make_float2(0.f, 0.f), v1 = v0, v2 = v0;
    if (idx < nr * L) { v0 = src[idx]; v1 = src[idx + cs]; v2 = src[idx + 2 * cs]; }
    b0[idx] = v0; b1[idx] = v1; b2[idx] = v2;
  }
  __syncthreads();
  fft_run<-1, false, PER>(b0, tw, p, rx, 1, L);
  fft_run<-1, false, PER>(b1, tw, p, rx, 1, L);
  fft_run<-1, false, PER>(b2, tw, p, rx, 1, L);
  if (!compact) {
    for (int idx = threadIdx.x; idx < nr * nx; idx += T) {
      const int r = idx / nx, ix = idx - r * nx;
      const size_t j = (size_t)(g0 + r) * nx + ix;
      float2 r1, r2, r3;
      if (iocc && iocc[j] == 0) {
        r1 = r2 = r3 = make_float2(0.f, 0.f);
      } else {
        const int li = r * L + ix;
        finish_site(x[j], x[j + nat], x[j + 2 * (size_t)nat], adia[j], adia[j + nat],
                    adia[j + 2 * (size_t)nat], aoff[j], aoff[j + nat], aoff[j + 2 * (size_t)nat],
                    b0[li], b1[li], b2[li], inv, cj != 0, r1, r2, r3);
      }
      y[j] = r1; y[j + nat] = r2; y[j + 2 * (size_t)nat] = r3;
    }
  } else {
    const int q0 = rowptr[g0], q1 = rowptr[g0 + nr];
    for (int q = q0 + threadIdx.x; q < q1; q += T) {
      const int j = occ[q];
      const int g = j / nx;
      const int ix = j - g * nx;
      const int li = (g - g0) * L + ix;
      float2 r1, r2, r3;
      finish_site(x[q], x[q + nat0], x[q + 2 * (size_t)nat0], adia[q], adia[q + nat0],
                  adia[q + 2 * (size_t)nat0], aoff[q], aoff[q + nat0], aoff[q + 2 * (size_t)nat0],
                  b0[li], b1[li], b2[li], inv, cj != 0, r1, r2, r3);
      y[q] = r1; y[q + nat0] = r2; y[q + 2 * (size_t)nat0] = r3;
    }
  }
}

// ------------------------------------------------------------------ host --
static int pow2floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
static int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

bool fft_make_plan(int L, FftPlan& p) {
  p = FftPlan{};
  p.L = L;
  int m = L;
  std::vector<int> r;
  while (m % 8 == 0) { r.push_back(8); m /= 8; }
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  while (m % 2 == 0) { r.push_back(2); m /= 2; }
  while (m % 3 == 0) { r.push_back(3); m /= 3; }
  while (m % 5 == 0) { r.push_back(5); m /= 5; }
  while (m % 7 == 0) { r.push_back(7); m /= 7; }
  if (m != 1 || (int)r.size() > FFT_MAX_STAGES || L < 2) return false;
  p.nst = (int)r.size();
  for (int i = 0; i < p.nst; ++i) p.rad[i] = r[i];
  return true;
}

void fft_twiddles_host(int L, float2* out) {
  const double twopi = 6.283185307179586476925286766559;
  for (int m = 0; m < L; ++m) {
    const double a = twopi * (double)m / (double)L;
    out[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
}

static int upload_twiddles(int L, float2** dst) {
  std::vector<float2> h(L);
  fft_twiddles_host(L, h.data());
  DDH_TRY(hipMalloc((void**)dst, (size_t)L * sizeof(float2)), "twiddle alloc", 1);
  DDH_TRY(hipMemcpy(*dst, h.data(), (size_t)L * sizeof(float2), hipMemcpyHostToDevice), "twiddle H2D", 1);
  return 0;
}

// per-workgroup LDS bytes of the current device (160 KB on gfx950)
static size_t lds_limit() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    return 64 * 1024;
  return (size_t)v;
}

// lets the wide-tile kernels take more than 64 KB of dynamic LDS
static int allow_big_lds(size_t bytes) {
  const int b = (int)bytes;
  const hipError_t e[] = {
      hipFuncSetAttribute((const void*)k_pass_y<+1, PER_WIDE, WIDE_THREADS>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, b),
      hipFuncSetAttribute((const void*)k_pass_y<-1, PER_WIDE, WIDE_THREADS>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, b),
      hipFuncSetAttribute((const void*)k_pass3_z_green<PER_NARROW, WIDE_THREADS>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, b)};
  for (hipError_t x : e)
    if (x != hipSuccess) {
      set_hip_error("MATVEC plan (wide LDS tiles)", x);
      return 1;
    }
  return 0;
}

int matvec_setup(Context& c) {
  MatvecPlan& m = c.mv;
  m = MatvecPlan{};
  m.nx = c.nx; m.ny = c.ny; m.nz = c.nz; m.nat = c.nat; m.nat0 = c.nat0;
  m.gx = 2 * c.nx; m.gy = 2 * c.ny; m.gz = 2 * c.nz;
  m.ngrid = (size_t)m.gx * m.gy * m.gz;
  if (!fft_make_plan(m.gx, m.px) || !fft_make_plan(m.gy, m.py) || !fft_make_plan(m.gz, m.pz)) {
    set_error("MATVEC plan", "2*NX, 2*NY, 2*NZ must factor into 2, 3, 5 and 7");
    return 206;
  }
  const int wide = PER_WIDE * MV_THREADS, narrow = PER_NARROW * MV_THREADS;
  if (m.gx > wide || m.gy > wide || m.gz > narrow / 1 || m.gx > narrow) {
    set_error("MATVEC plan", "padded grid axis exceeds the LDS FFT limit (2N <= 2048)");
    return 206;
  }
  const int nrows = c.ny * c.nz;
  // pass 1 rows per workgroup: fill the LDS tile but keep >= ~4 workgroups per CU
  m.rx_fwd = std::max(1, wide / m.gx);
  while (m.rx_fwd > 1 && (size_t)((nrows + m.rx_fwd - 1) / m.rx_fwd) * 3 < 2048) m.rx_fwd /= 2;
  m.rx_inv = std::max(1, narrow / m.gx);
  while (m.rx_inv > 1 && (size_t)((nrows + m.rx_inv - 1) / m.rx_inv) < 1024) m.rx_inv /= 2;
  m.cy = std::min(16, pow2floor(std::max(1, wide / m.gy)));
  m.cz = std::min(16, pow2floor(std::max(1, narrow / m.gz)));
  m.ty = m.tz = MV_THREADS;
  const size_t lds_max = lds_limit();
  if (m.cy < 8) {
    int cw = std::min(16, pow2floor(std::max(1, PER_WIDE * WIDE_THREADS / m.gy)));
    while (cw > 1 && (size_t)(m.gy + cw * m.gy) * sizeof(float2) > lds_max) cw /= 2;
    if (cw > m.cy) { m.cy = cw; m.ty = WIDE_THREADS; }
  }
  if (m.cz < 8) {
    int cw = std::min(16, pow2floor(std::max(1, PER_NARROW * WIDE_THREADS / m.gz)));
    while (cw > 1 && (size_t)(m.gz + 3 * cw * m.gz) * sizeof(float2) > lds_max) cw /= 2;
    if (cw > m.cz) { m.cz = cw; m.tz = WIDE_THREADS; }
  }
  if ((m.ty == WIDE_THREADS || m.tz == WIDE_THREADS) && allow_big_lds(lds_max)) return 206;
  if (upload_twiddles(m.gx, &m.twx) || upload_twiddles(m.gy, &m.twy) || upload_twiddles(m.gz, &m.twz)) return 206;
  m.p1 = (size_t)c.ny * m.gx;
  m.p2 = (size_t)m.gy * m.gx;
  const size_t n1 = (size_t)3 * c.nz * m.p1, n2 = (size_t)3 * c.nz * m.p2;
  DDH_TRY(hipMalloc((void**)&c.a1, n1 * sizeof(float2)), "malloc FFT work a1", 204);
  DDH_TRY(hipMalloc((void**)&c.a2, n2 * sizeof(float2)), "malloc FFT work a2", 204);
  return 0;
}

size_t matvec_a1_count(const Context& c) { return (size_t)3 * c.nz * c.ny * c.mv.gx; }
size_t matvec_a2_count(const Context& c) { return (size_t)3 * c.nz * c.mv.p2; }

void matvec_release(Context& c) {
  for (float2** p : {&c.mv.twx, &c.mv.twy, &c.mv.twz, &c.a1, &c.a2}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
}

// One MATVEC pass with kernel choice `kind` (-1: LDS-staged, >= 0: register shape).
static int launch_pass(Context& c, int which, int kind, int cj, int compact, const float2* x, float2* y,
                       const float2* adia, const float2* aoff, hipStream_t st, const MvBatch& bt) {
  const MatvecPlan& m = c.mv;
  const int nrows = c.ny * c.nz;
  const unsigned nb = (unsigned)bt.nrhs;
  float2 *a1 = bt.a1, *a2 = bt.a2;
  if (!c.fft_generic && kind >= 0 && reg_pass(c, which, kind, cj, compact, x, y, adia, aoff, st, bt)) return 0;
  switch (which) {
    case 1: {
      dim3 grid((nrows + m.rx_fwd - 1) / m.rx_fwd, 3, nb);
      size_t lds = (size_t)(m.gx + m.rx_fwd * m.gx) * sizeof(float2);
      hipLaunchKernelGGL(k_pass1_xfwd<PER_WIDE>, grid, dim3(MV_THREADS), lds, st, x, a1, c.rowptr, c.occ, m.twx,
                         m.px, c.nx, nrows, m.rx_fwd, c.nat, c.nat0, compact, cj, bt.k);
      DDH_LAUNCH_OK("MATVEC pass 1 (x FFT)", 21);
      break;
    }
    case 2: {
      dim3 grid((m.gx + m.cy - 1) / m.cy, c.nz, 3 * nb);
      size_t lds = (size_t)(m.gy + m.cy * m.gy) * sizeof(float2);
      if (m.ty == WIDE_THREADS)
        hipLaunchKernelGGL((k_pass_y<+1, PER_WIDE, WIDE_THREADS>), grid, dim3(WIDE_THREADS), lds, st, a1, a2, m.twy,
                           m.py, m.gx, c.ny, m.gy, c.nz, ilog2(m.cy), m.p1, m.p2, bt.k);
      else
        hipLaunchKernelGGL((k_pass_y<+1, PER_WIDE>), grid, dim3(MV_THREADS), lds, st, a1, a2, m.twy, m.py, m.gx,
                           c.ny, m.gy, c.nz, ilog2(m.cy), m.p1, m.p2, bt.k);
      DDH_LAUNCH_OK("MATVEC pass 2 (y FFT)", 22);
      break;
    }
    case 3: {
      dim3 grid((m.gx + m.cz - 1) / m.cz, m.gy, nb);
      size_t lds = (size_t)(m.gz + 3 * m.cz * m.gz) * sizeof(float2);
      if (m.tz == WIDE_THREADS)
        hipLaunchKernelGGL((k_pass3_z_green<PER_NARROW, WIDE_THREADS>), grid, dim3(WIDE_THREADS), lds, st, a2,
                           c.green, m.twz, m.pz, c.nx, c.ny, c.nz, m.gx, m.gy, ilog2(m.cz), m.p2, bt.k, green_geo(c));
      else
        hipLaunchKernelGGL(k_pass3_z_green<PER_NARROW>, grid, dim3(MV_THREADS), lds, st, a2, c.green, m.twz, m.pz,
                           c.nx, c.ny, c.nz, m.gx, m.gy, ilog2(m.cz), m.p2, bt.k, green_geo(c));
      DDH_LAUNCH_OK("MATVEC pass 3 (z FFT + Green)", 23);
      break;
    }
    case 4: {
      dim3 grid((m.gx + m.cy - 1) / m.cy, c.nz, 3 * nb);
      size_t lds = (size_t)(m.gy + m.cy * m.gy) * sizeof(float2);
      if (m.ty == WIDE_THREADS)
        hipLaunchKernelGGL((k_pass_y<-1, PER_WIDE, WIDE_THREADS>), grid, dim3(WIDE_THREADS), lds, st, a2, a1, m.twy,
                           m.py, m.gx, m.gy, c.ny, c.nz, ilog2(m.cy), m.p2, m.p1, bt.k);
      else
        hipLaunchKernelGGL((k_pass_y<-1, PER_WIDE>), grid, dim3(MV_THREADS), lds, st, a2, a1, m.twy, m.py, m.gx,
                           m.gy, c.ny, c.nz, ilog2(m.cy), m.p2, m.p1, bt.k);
      DDH_LAUNCH_OK("MATVEC pass 4 (y inverse FFT)", 24);
      break;
    }
    default: {
      dim3 grid((nrows + m.rx_inv - 1) / m.rx_inv, nb);
      size_t lds = (size_t)(m.gx + 3 * m.rx_inv * m.gx) * sizeof(float2);
      const float inv = 1.f / (float)m.ngrid;
      hipLaunchKernelGGL(k_pass5_xinv_finish<PER_NARROW>, grid, dim3(MV_THREADS), lds, st, a1, m.twx, m.px, x, y,
                         adia, aoff, compact ? nullptr : c.iocc, c.rowptr, c.occ, c.nx, nrows, m.rx_inv, c.nat,
                         c.nat0, compact, cj, inv, bt.k);
      DDH_LAUNCH_OK("MATVEC pass 5 (x inverse FFT + finish)", 25);
      break;
    }
  }
  return 0;
}

int matvec_launch(Context& c, int cwhat, const float2* x, float2* y, const float2* adia,
                  const float2* aoff, int layout, hipStream_t st, hipEvent_t* ev, const MvBatch* btp) {
  const int cj = (cwhat == 'C') ? 1 : 0;
  const int compact = layout ? 1 : 0;
  MvBatch bt = btp ? *btp : MvBatch{};
  if (!bt.a1) bt.a1 = c.a1;
  if (!bt.a2) bt.a2 = c.a2;
  for (int p = 1; p <= 5; ++p) {
    if (ev) hipEventRecord(ev[p - 1], st);
    if (int rc = launch_pass(c, p, c.mv.kind[p], cj, compact, x, y, adia, aoff, st, bt)) return rc;
  }
  DDH_LAUNCH_OK("MATVEC register passes", 20);
  if (ev) hipEventRecord(ev[5], st);
  c.matvec_count += (unsigned)bt.nrhs;
  return 0;
}

// Autotuner: for each pass, every register shape of its axis length and the
// LDS-staged kernel run a few times on the prepared grid (compact vectors,
// stream events); the fastest is kept in c.mv.kind.  Choices are cached per
// grid for the life of the process, so re-preparing the same grid (for
// instance every target orientation of a DDSCAT run) does not re-time.
namespace {
struct TuneKey {
  int nx, ny, nz, ipbc;  // IPBC=1 reads the whole-grid Green in pass 3, not the octant
  bool operator<(const TuneKey& o) const {
    return nx != o.nx ? nx < o.nx : ny != o.ny ? ny < o.ny : nz != o.nz ? nz < o.nz : ipbc < o.ipbc;
  }
};
std::map<TuneKey, std::array<int, 6>>& tune_cache() {
  static std::map<TuneKey, std::array<int, 6>> m;
  return m;
}
}  // namespace

int matvec_autotune(Context& c) {
  MatvecPlan& m = c.mv;
  const int lens[6] = {0, m.gx, m.gy, m.gz, m.gy, m.gx};
  for (int p = 1; p <= 5; ++p) m.kind[p] = reg_candidates(lens[p], p) > 0 ? 0 : -1;
  m.tuned = 0;
  if (!c.autotune || c.fft_generic || !c.sx || !c.sy || !c.adia_c || !c.aoff_c) return 0;
  const TuneKey key{c.nx, c.ny, c.nz, c.ipbc};
  auto it = tune_cache().find(key);
  if (it != tune_cache().end()) {
    for (int p = 1; p <= 5; ++p) m.kind[p] = it->second[p];
    m.tuned = 1;
    return 0;
  }
  hipStream_t st = c.stream;
  const size_t cb = (size_t)3 * c.nat0 * sizeof(float2);
  DDH_TRY(hipMemsetAsync(c.sx, 0, cb, st), "autotune", 20);
  DDH_TRY(hipMemsetAsync(c.adia_c, 0, cb, st), "autotune", 20);
  DDH_TRY(hipMemsetAsync(c.aoff_c, 0, cb, st), "autotune", 20);
  hipEvent_t e0, e1;
  DDH_TRY(hipEventCreate(&e0), "autotune", 20);
  if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); set_error("autotune", "event"); return 20; }
  MvBatch bt;
  bt.a1 = c.a1;
  bt.a2 = c.a2;
  int rc = 0;
  for (int p = 1; p <= 5 && !rc; ++p) {
    const int ncand = reg_candidates(lens[p], p);
    if (ncand == 0) continue;
    float best = 1e30f;
    int best_kind = m.kind[p];
    for (int k = -1; k < ncand && !rc; ++k) {
      for (int w = 0; w < 2 && !rc; ++w) rc = launch_pass(c, p, k, 0, 1, c.sx, c.sy, c.adia_c, c.aoff_c, st, bt);
      hipEventRecord(e0, st);
      for (int r = 0; r < 5 && !rc; ++r) rc = launch_pass(c, p, k, 0, 1, c.sx, c.sy, c.adia_c, c.aoff_c, st, bt);
      hipEventRecord(e1, st);
      if (rc || hipEventSynchronize(e1) != hipSuccess) { rc = rc ? rc : 20; break; }
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; best_kind = k; }
    }
    m.kind[p] = best_kind;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  std::array<int, 6> k6;
  for (int p = 0; p < 6; ++p) k6[p] = m.kind[p];
  tune_cache()[key] = k6;
  m.tuned = 1;
  return 0;
}

}  // namespace ddhip
