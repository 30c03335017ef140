// SLICES (low-memory) MATVEC backend for gfx950.
//
// Same operator as FFT3D (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:136-214,
// 361-404): the z-FFT runs first over every (ix,iy) column into zwork
// [3][2NZ][NY][NX] (6 NAT complex, instead of the 18 NAT of the FFT3D work
// buffers), then the kz planes go through the xy transforms in batches of B
// planes held in a small slice buffer [B][3][NY][2NX]:
//   x-FFT (+i)  ->  y-FFT (+i), 3x3 Green multiply at that kz, y-FFT (-i), keep iy < NY
//               ->  x-FFT (-i), keep ix < NX, back into the zwork plane
// Batches alternate between two HIP streams, each with its own slice buffer,
// so consecutive batches overlap; the final inverse z-FFT is fused with the
// finish step (diag + off-diagonal blocks - convolution / G, IOCC mask).
#include "solver_common.hpp"
#include "fft_device.hpp"
#include <vector>
#include <algorithm>

namespace ddhip {

static constexpr int SL_THREADS = 256;
static constexpr int SL_PER = 8;    // elements per thread per stage (3-component kernels)
static constexpr int SL_PERW = 16;  // single-component kernels

struct SliceState {
  float2* zwork = nullptr;
  float2* zin = nullptr;  // [3][NZ][NY][NX] (full layout): inverse z-FFT output (register z path)
  float2* slice[2] = {nullptr, nullptr};
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev_in = nullptr, ev_done[2] = {nullptr, nullptr};
  int batch = 4;
  int cz = 1, cy = 1, rx_f = 1, rx_i = 1;
};
static SliceState g_sl[MAX_TARGETS + 1];  // one per target slot

// --------------------------------------------------- z-FFT forward (x -> zwork)
__global__ __launch_bounds__(SL_THREADS) void k_sl_zfwd(const float2* __restrict__ x, const int* __restrict__ qmap,
                                                       float2* __restrict__ zw, const float2* __restrict__ twg,
                                                       FftPlan p, int nx, int ny, int nz, int nat, int nat0,
                                                       int compact, int cj, int cz_log2, const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  extern __shared__ float2 sm[];
  const int L = p.L;
  const int cz = 1 << cz_log2;
  float2* tw = sm;
  float2* buf = sm + L;
  const int iy = blockIdx.y, ix0 = blockIdx.x * cz, c = blockIdx.z;
  const int T = blockDim.x;
  const size_t plane = (size_t)nx * ny;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  for (int idx = threadIdx.x; idx < L * cz; idx += T) {
    const int e = idx >> cz_log2, col = idx & (cz - 1);
    const int ix = ix0 + col;
    float2 v = make_float2(0.f, 0.f);
    if (e < nz && ix < nx) {
      const size_t j = (size_t)e * plane + (size_t)iy * nx + ix;
      if (compact) {
        const int q = qmap[j];
        if (q >= 0) v = x[(size_t)c * nat0 + q];
      } else {
        v = x[(size_t)c * nat + j];
      }
      if (cj) v.y = -v.y;
    }
    buf[idx] = v;
  }
  __syncthreads();
  fft_run<+1, true, SL_PERW>(buf, tw, p, cz, cz, 1);
  float2* out = zw + (size_t)c * L * plane + (size_t)iy * nx;
  for (int idx = threadIdx.x; idx < L * cz; idx += T) {
    const int e = idx >> cz_log2, col = idx & (cz - 1);
    const int ix = ix0 + col;
    if (ix < nx) out[(size_t)e * plane + ix] = buf[idx];
  }
}

// -------------------------------------------- x-FFT forward (zwork plane -> slice)
// rows r = (b, c, iy) of the batch; slice[(b*3 + c)*ny + iy][0..2nx)
__global__ __launch_bounds__(SL_THREADS) void k_sl_xfwd(const float2* __restrict__ zw, float2* __restrict__ sl,
                                                       const float2* __restrict__ twg, FftPlan p, int nx, int ny,
                                                       int gz, int k0, int nb, int rx, const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  extern __shared__ float2 sm[];
  const int L = p.L;
  float2* tw = sm;
  float2* buf = sm + L;
  const int T = blockDim.x;
  const int nrows = nb * 3 * ny;
  const int r0 = blockIdx.x * rx;
  const int nr = min(rx, nrows - r0);
  const size_t plane = (size_t)nx * ny;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  for (int idx = threadIdx.x; idx < rx * L; idx += T) {
    const int r = idx / L, e = idx - r * L;
    float2 v = make_float2(0.f, 0.f);
    if (r < nr && e < nx) {
      const int row = r0 + r;
      const int iy = row % ny, bc = row / ny;
      const int c = bc % 3, b = bc / 3;
      v = zw[((size_t)c * gz + (k0 + b)) * plane + (size_t)iy * nx + e];
    }
    buf[idx] = v;
  }
  __syncthreads();
  fft_run<+1, false, SL_PERW>(buf, tw, p, rx, 1, L);
  float2* out = sl + (size_t)r0 * L;
  for (int idx = threadIdx.x; idx < nr * L; idx += T) out[idx] = buf[idx];
}

// ------------------------------ y-FFT, Green multiply at kz, inverse y-FFT (in slice)
__global__ __launch_bounds__(SL_THREADS) void k_sl_y_green(float2* __restrict__ sl, const float2* __restrict__ green,
                                                          const float2* __restrict__ twg, FftPlan p, int nx, int ny,
                                                          int nz, int k0, int cy_log2, GreenGeo gg,
                                                          const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  extern __shared__ float2 sm[];
  const int L = p.L;  // 2 ny
  const int gx = 2 * nx;
  const int cy = 1 << cy_log2;
  float2* tw = sm;
  float2* b0 = sm + L;
  float2* b1 = b0 + L * cy;
  float2* b2 = b1 + L * cy;
  const int b = blockIdx.y, kx0 = blockIdx.x * cy;
  const int kz = k0 + b;
  const int T = blockDim.x;
  const size_t comp = (size_t)ny * gx;
  float2* base = sl + (size_t)b * 3 * comp;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  for (int idx = threadIdx.x; idx < L * cy; idx += T) {
    const int e = idx >> cy_log2, col = idx & (cy - 1);
    const int kx = kx0 + col;
    float2 v0 = make_float2(0.f, 0.f), v1 = v0, v2 = v0;
    if (e < ny && kx < gx) {
      const size_t o = (size_t)e * gx + kx;
      v0 = base[o]; v1 = base[o + comp]; v2 = base[o + 2 * comp];
    }
    b0[idx] = v0; b1[idx] = v1; b2[idx] = v2;
  }
  __syncthreads();
  fft_run<+1, true, SL_PER>(b0, tw, p, cy, cy, 1);
  fft_run<+1, true, SL_PER>(b1, tw, p, cy, cy, 1);
  fft_run<+1, true, SL_PER>(b2, tw, p, cy, cy, 1);
  // octant reconstruction for IPBC=0 (cu:166-186), whole grid for IPBC=1
  const int ox = gg.ox, oy = gg.oy;
  const int kr = min(kz, gg.mz - kz);
  const float sz = kz <= gg.tz ? 1.f : -1.f;
  for (int idx = threadIdx.x; idx < L * cy; idx += T) {
    const int ky = idx >> cy_log2, col = idx & (cy - 1);
    const int kx = kx0 + col;
    if (kx >= gx) continue;
    const int ir = min(kx, gg.mx - kx), jr = min(ky, gg.my - ky);
    const float sx = kx <= gg.tx ? 1.f : -1.f;
    const float sy = ky <= gg.ty ? 1.f : -1.f;
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
  fft_run<-1, true, SL_PER>(b0, tw, p, cy, cy, 1);
  fft_run<-1, true, SL_PER>(b1, tw, p, cy, cy, 1);
  fft_run<-1, true, SL_PER>(b2, tw, p, cy, cy, 1);
  for (int idx = threadIdx.x; idx < ny * cy; idx += T) {
    const int e = idx >> cy_log2, col = idx & (cy - 1);
    const int kx = kx0 + col;
    if (kx < gx) {
      const size_t o = (size_t)e * gx + kx;
      base[o] = b0[idx]; base[o + comp] = b1[idx]; base[o + 2 * comp] = b2[idx];
    }
  }
}

// ------------------------------------------ x-FFT inverse (slice -> zwork plane)
__global__ __launch_bounds__(SL_THREADS) void k_sl_xinv(const float2* __restrict__ sl, float2* __restrict__ zw,
                                                       const float2* __restrict__ twg, FftPlan p, int nx, int ny,
                                                       int gz, int k0, int nb, int rx, const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  extern __shared__ float2 sm[];
  const int L = p.L;
  float2* tw = sm;
  float2* buf = sm + L;
  const int T = blockDim.x;
  const int nrows = nb * 3 * ny;
  const int r0 = blockIdx.x * rx;
  const int nr = min(rx, nrows - r0);
  const size_t plane = (size_t)nx * ny;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  const float2* src = sl + (size_t)r0 * L;
  for (int idx = threadIdx.x; idx < rx * L; idx += T) buf[idx] = idx < nr * L ? src[idx] : make_float2(0.f, 0.f);
  __syncthreads();
  fft_run<-1, false, SL_PERW>(buf, tw, p, rx, 1, L);
  for (int idx = threadIdx.x; idx < nr * nx; idx += T) {
    const int r = idx / nx, e = idx - r * nx;
    const int row = r0 + r;
    const int iy = row % ny, bc = row / ny;
    const int c = bc % 3, b = bc / 3;
    zw[((size_t)c * gz + (k0 + b)) * plane + (size_t)iy * nx + e] = buf[r * L + e];
  }
}

// ------------------------------------------ z-FFT inverse + finish (zwork -> y)
__global__ __launch_bounds__(SL_THREADS) void k_sl_zinv_finish(
    const float2* __restrict__ zw, const float2* __restrict__ twg, FftPlan p, const float2* __restrict__ x,
    float2* __restrict__ y, const float2* __restrict__ adia, const float2* __restrict__ aoff,
    const int16_t* __restrict__ iocc, const int* __restrict__ qmap, int nx, int ny, int nz, int nat, int nat0,
    int compact, int cj, float inv, int cz_log2, const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  extern __shared__ float2 sm[];
  const int L = p.L;  // 2 nz
  const int cz = 1 << cz_log2;
  float2* tw = sm;
  float2* b0 = sm + L;
  float2* b1 = b0 + L * cz;
  float2* b2 = b1 + L * cz;
  const int iy = blockIdx.y, ix0 = blockIdx.x * cz;
  const int T = blockDim.x;
  const size_t plane = (size_t)nx * ny;
  const size_t cs = (size_t)L * plane;
  for (int i = threadIdx.x; i < L; i += T) tw[i] = twg[i];
  for (int idx = threadIdx.x; idx < L * cz; idx += T) {
    const int e = idx >> cz_log2, col = idx & (cz - 1);
    const int ix = ix0 + col;
    float2 v0 = make_float2(0.f, 0.f), v1 = v0, v2 = v0;
    if (ix < nx) {
      const size_t o = (size_t)e * plane + (size_t)iy * nx + ix;
      v0 = zw[o]; v1 = zw[o + cs]; v2 = zw[o + 2 * cs];
    }
    b0[idx] = v0; b1[idx] = v1; b2[idx] = v2;
  }
  __syncthreads();
  fft_run<-1, true, SL_PER>(b0, tw, p, cz, cz, 1);
  fft_run<-1, true, SL_PER>(b1, tw, p, cz, cz, 1);
  fft_run<-1, true, SL_PER>(b2, tw, p, cz, cz, 1);
  for (int idx = threadIdx.x; idx < nz * cz; idx += T) {
    const int iz = idx >> cz_log2, col = idx & (cz - 1);
    const int ix = ix0 + col;
    if (ix >= nx) continue;
    const size_t j = (size_t)iz * plane + (size_t)iy * nx + ix;
    size_t o;
    size_t stride;
    if (compact) {
      const int q = qmap[j];
      if (q < 0) continue;
      o = (size_t)q;
      stride = (size_t)nat0;
    } else {
      o = j;
      stride = (size_t)nat;
      if (iocc && iocc[j] == 0) {
        y[o] = y[o + stride] = y[o + 2 * stride] = make_float2(0.f, 0.f);
        continue;
      }
    }
    float2 x1 = x[o], x2 = x[o + stride], x3 = x[o + 2 * stride];
    if (cj) { x1.y = -x1.y; x2.y = -x2.y; x3.y = -x3.y; }
    const float2 a23 = aoff[o], a31 = aoff[o + stride], a12 = aoff[o + 2 * stride];
    float2 r1 = csub(cmul(adia[o], x1), cscale(b0[idx], inv));
    float2 r2 = csub(cmul(adia[o + stride], x2), cscale(b1[idx], inv));
    float2 r3 = csub(cmul(adia[o + 2 * stride], x3), cscale(b2[idx], inv));
    r1 = cadd(r1, cadd(cmul(a31, x3), cmul(a12, x2)));
    r2 = cadd(r2, cadd(cmul(a12, x1), cmul(a23, x3)));
    r3 = cadd(r3, cadd(cmul(a23, x2), cmul(a31, x1)));
    if (cj) { r1.y = -r1.y; r2.y = -r2.y; r3.y = -r3.y; }
    y[o] = r1; y[o + stride] = r2; y[o + 2 * stride] = r3;
  }
}

// ------------------------------------------- register z path: finish
// finish on the inverse-transformed w [3][NAT] (cu:116-118): y = adia x - w/G
// + offdiag(aoff) x at every occupied site, IOCC mask, conj for 'C'
__global__ __launch_bounds__(SL_THREADS) void k_sl_finish(const float2* __restrict__ w, const float2* __restrict__ x,
                                                         float2* __restrict__ y, const float2* __restrict__ adia,
                                                         const float2* __restrict__ aoff,
                                                         const int16_t* __restrict__ iocc,
                                                         const int* __restrict__ qmap, int nat, int nat0, int compact,
                                                         int cj, float inv, const int* __restrict__ act) {
  if (act && !*act) return;  // inactive right-hand side of a batch
  const size_t gs = (size_t)gridDim.x * SL_THREADS;
  for (size_t j = (size_t)blockIdx.x * SL_THREADS + threadIdx.x; j < (size_t)nat; j += gs) {
    size_t o = j, sd = (size_t)nat;
    bool st = true, zr = false;
    if (compact) {
      const int q = qmap[j];
      st = q >= 0;
      o = st ? (size_t)q : 0;
      sd = (size_t)nat0;
    } else {
      zr = iocc && iocc[j] == 0;
    }
    const float2 w0 = w[j], w1 = w[j + nat], w2 = w[j + 2 * (size_t)nat];
    float2 x1 = x[o], x2 = x[o + sd], x3 = x[o + 2 * sd];
    const float2 d1 = adia[o], d2 = adia[o + sd], d3 = adia[o + 2 * sd];
    const float2 a23 = aoff[o], a31 = aoff[o + sd], a12 = aoff[o + 2 * sd];
    if (cj) { x1.y = -x1.y; x2.y = -x2.y; x3.y = -x3.y; }
    float2 r1 = csub(cmul(d1, x1), cscale(w0, inv));
    float2 r2 = csub(cmul(d2, x2), cscale(w1, inv));
    float2 r3 = csub(cmul(d3, x3), cscale(w2, inv));
    r1 = cadd(r1, cadd(cmul(a31, x3), cmul(a12, x2)));
    r2 = cadd(r2, cadd(cmul(a12, x1), cmul(a23, x3)));
    r3 = cadd(r3, cadd(cmul(a23, x2), cmul(a31, x1)));
    if (cj) { r1.y = -r1.y; r2.y = -r2.y; r3.y = -r3.y; }
    if (zr) r1 = r2 = r3 = make_float2(0.f, 0.f);
    if (st) { y[o] = r1; y[o + sd] = r2; y[o + 2 * sd] = r3; }
  }
}

// ------------------------------------------------------------------ host --
static int p2floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
static int lg2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

int upload_twiddles_sl(int L, float2** dst) {
  std::vector<float2> h(L);
  fft_twiddles_host(L, h.data());
  DDH_TRY(hipMalloc((void**)dst, (size_t)L * sizeof(float2)), "twiddle alloc", 1);
  DDH_TRY(hipMemcpy(*dst, h.data(), (size_t)L * sizeof(float2), hipMemcpyHostToDevice), "twiddle H2D", 1);
  return 0;
}

int slices_setup(Context& c) {
  MatvecPlan& m = c.mv;
  m = MatvecPlan{};
  m.nx = c.nx; m.ny = c.ny; m.nz = c.nz; m.nat = c.nat; m.nat0 = c.nat0;
  m.gx = 2 * c.nx; m.gy = 2 * c.ny; m.gz = 2 * c.nz;
  m.ngrid = (size_t)m.gx * m.gy * m.gz;
  if (!fft_make_plan(m.gx, m.px) || !fft_make_plan(m.gy, m.py) || !fft_make_plan(m.gz, m.pz)) {
    set_error("SLICES plan", "2*NX, 2*NY, 2*NZ must factor into 2, 3, 5 and 7");
    return 206;
  }
  const int wide = SL_PERW * SL_THREADS, narrow = SL_PER * SL_THREADS;
  if (m.gx > narrow || m.gy > narrow || m.gz > narrow) {
    set_error("SLICES plan", "padded grid axis exceeds the LDS FFT limit (2N <= 2048)");
    return 206;
  }
  SliceState& s = g_sl[c.slot];
  s = SliceState{};
  // z passes: column tiles of consecutive ix; the forward pass is single component
  s.cz = std::min(16, p2floor(std::max(1, narrow / m.gz)));
  s.cy = std::min(16, p2floor(std::max(1, narrow / m.gy)));
  s.rx_f = std::max(1, wide / m.gx);
  s.rx_i = std::max(1, wide / m.gx);
  // planes per batch: 1/16 of the kz planes, at least 4 (the reference batch)
  // planes per slice batch: 2NZ/8 (two buffers of 3 NAT complex together);
  // 2NZ/16 ran 10 % slower at 128^3 and the same at 768^3
  s.batch = std::min(m.gz, std::max(4, m.gz / 8));
  if (upload_twiddles_sl(m.gx, &m.twx) || upload_twiddles_sl(m.gy, &m.twy) || upload_twiddles_sl(m.gz, &m.twz)) return 206;
  const size_t zcount = (size_t)3 * m.gz * c.nx * c.ny;
  const size_t scount = (size_t)s.batch * 3 * c.ny * m.gx;
  DDH_TRY(hipMalloc((void**)&s.zwork, zcount * sizeof(float2)), "malloc SLICES z workspace", 204);
  if (slices_z_reg_length(m.gz))
    DDH_TRY(hipMalloc((void**)&s.zin, (size_t)3 * c.nat * sizeof(float2)), "malloc SLICES z input", 204);
  for (int i = 0; i < 2; ++i) {
    DDH_TRY(hipMalloc((void**)&s.slice[i], scount * sizeof(float2)), "malloc SLICES slice batch", 204);
    DDH_TRY(hipStreamCreateWithFlags(&s.side[i], hipStreamNonBlocking), "create SLICES stream", 206);
    DDH_TRY(hipEventCreateWithFlags(&s.ev_done[i], hipEventDisableTiming), "create SLICES event", 206);
  }
  DDH_TRY(hipEventCreateWithFlags(&s.ev_in, hipEventDisableTiming), "create SLICES event", 206);
  m.rx_fwd = s.rx_f; m.rx_inv = s.rx_i; m.cy = s.cy; m.cz = s.cz;
  return 0;
}

void slices_release(Context& c) {
  SliceState& s = g_sl[c.slot];
  if (s.zwork) {
    for (int i = 0; i < 2; ++i) {
      if (s.side[i]) hipStreamSynchronize(s.side[i]);
    }
  }
  if (s.zwork) hipFree(s.zwork);
  if (s.zin) hipFree(s.zin);
  for (int i = 0; i < 2; ++i) {
    if (s.slice[i]) hipFree(s.slice[i]);
    if (s.side[i]) hipStreamDestroy(s.side[i]);
    if (s.ev_done[i]) hipEventDestroy(s.ev_done[i]);
  }
  if (s.ev_in) hipEventDestroy(s.ev_in);
  s = SliceState{};
  if (c.backend == BACKEND_SLICES) {
    for (float2** p : {&c.mv.twx, &c.mv.twy, &c.mv.twz}) {
      if (*p) hipFree(*p);
      *p = nullptr;
    }
  }
}

int slices_matvec(Context& c, int cwhat, const float2* x, float2* y, const float2* adia, const float2* aoff,
                  int layout, hipStream_t st, const int* act) {
  SliceState& s = g_sl[c.slot];
  const MatvecPlan& m = c.mv;
  if (!s.zwork) { set_error("SLICES MATVEC", "backend not prepared"); return 20; }
  const int cj = (cwhat == 'C') ? 1 : 0;
  const int compact = layout ? 1 : 0;
  // register z path: a full-layout 'N' input is already [3][NZ][NY][NX]
  const bool regz = s.zin && !c.fft_generic;
  const int ewg = (int)std::min<size_t>(8192, ((size_t)c.nat + SL_THREADS * 4 - 1) / (SL_THREADS * 4));
  if (regz) {
    // a compact or conjugated x is gathered inside the z-FFT's loads
    slices_z_reg(c, +1, x, s.zwork, st, act, compact || cj, compact, cj);
  } else {
    dim3 grid((c.nx + s.cz - 1) / s.cz, c.ny, 3);
    const size_t lds = (size_t)(m.gz + s.cz * m.gz) * sizeof(float2);
    hipLaunchKernelGGL(k_sl_zfwd, grid, dim3(SL_THREADS), lds, st, x, c.qmap, s.zwork, m.twz, m.pz, c.nx, c.ny,
                       c.nz, c.nat, c.nat0, compact, cj, lg2(s.cz), act);
  }
  DDH_LAUNCH_OK("SLICES z FFT", 21);
  DDH_TRY(hipEventRecord(s.ev_in, st), "SLICES record", 25);
  for (int i = 0; i < 2; ++i) DDH_TRY(hipStreamWaitEvent(s.side[i], s.ev_in, 0), "SLICES wait", 25);
  int k = 0;
  for (int k0 = 0; k0 < m.gz; k0 += s.batch, ++k) {
    const int nb = std::min(s.batch, m.gz - k0);
    hipStream_t ss = s.side[k & 1];
    float2* sl = s.slice[k & 1];
    const int nrows = nb * 3 * c.ny;
    if (!slices_x_reg(c, +1, s.zwork, sl, k0, nb, ss, act)) {
      dim3 grid((nrows + s.rx_f - 1) / s.rx_f);
      const size_t lds = (size_t)(m.gx + s.rx_f * m.gx) * sizeof(float2);
      hipLaunchKernelGGL(k_sl_xfwd, grid, dim3(SL_THREADS), lds, ss, s.zwork, sl, m.twx, m.px, c.nx, c.ny, m.gz, k0,
                         nb, s.rx_f, act);
    }
    DDH_LAUNCH_OK("SLICES x FFT", 23);
    if (!slices_y_green_reg(c, sl, k0, nb, ss, act)) {
      dim3 grid((m.gx + s.cy - 1) / s.cy, nb);
      const size_t lds = (size_t)(m.gy + 3 * s.cy * m.gy) * sizeof(float2);
      hipLaunchKernelGGL(k_sl_y_green, grid, dim3(SL_THREADS), lds, ss, sl, c.green, m.twy, m.py, c.nx, c.ny, c.nz,
                         k0, lg2(s.cy), green_geo(c), act);
    }
    DDH_LAUNCH_OK("SLICES y FFT + Green", 24);
    if (!slices_x_reg(c, -1, s.zwork, sl, k0, nb, ss, act)) {
      dim3 grid((nrows + s.rx_i - 1) / s.rx_i);
      const size_t lds = (size_t)(m.gx + s.rx_i * m.gx) * sizeof(float2);
      hipLaunchKernelGGL(k_sl_xinv, grid, dim3(SL_THREADS), lds, ss, sl, s.zwork, m.twx, m.px, c.nx, c.ny, m.gz, k0,
                         nb, s.rx_i, act);
    }
    DDH_LAUNCH_OK("SLICES x inverse FFT", 26);
  }
  for (int i = 0; i < 2; ++i) {
    DDH_TRY(hipEventRecord(s.ev_done[i], s.side[i]), "SLICES record", 25);
    DDH_TRY(hipStreamWaitEvent(st, s.ev_done[i], 0), "SLICES wait", 25);
  }
  if (regz) {
    slices_z_reg(c, -1, s.zwork, s.zin, st, act);
    hipLaunchKernelGGL(k_sl_finish, dim3(ewg), dim3(SL_THREADS), 0, st, s.zin, x, y, adia, aoff,
                       compact ? nullptr : c.iocc, c.qmap, c.nat, c.nat0, compact, cj, 1.f / (float)m.ngrid, act);
  } else {
    dim3 grid((c.nx + s.cz - 1) / s.cz, c.ny);
    const size_t lds = (size_t)(m.gz + 3 * s.cz * m.gz) * sizeof(float2);
    hipLaunchKernelGGL(k_sl_zinv_finish, grid, dim3(SL_THREADS), lds, st, s.zwork, m.twz, m.pz, x, y, adia, aoff,
                       compact ? nullptr : c.iocc, c.qmap, c.nx, c.ny, c.nz, c.nat, c.nat0, compact, cj,
                       1.f / (float)m.ngrid, lg2(s.cz), act);
  }
  DDH_LAUNCH_OK("SLICES z inverse FFT + finish", 29);
  c.matvec_count++;
  return 0;
}

}  // namespace ddhip
