// EVALQ and SCAT on the device (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:763-1002).
//
// EVALQ: q = (alpha^-1 + i (2/3) k^3) P + offdiag P on the occupied sites,
// Cabs = -4 pi k Im<P,q>/E0^2, Cext = 4 pi k Im<E,P>/E0^2,
// Cpha = 2 pi k Re<E,P>/E0^2 with FP64 dot products.
//
// SCAT: S(k_s) = sum_j P_j exp(-i k_s . r_j), r_j = ((i+1) + x0) dx per axis.
// The lattice makes the phase separable, exp(-i k.r) = Ex(ix) Ey(iy) Ez(iz),
// so per direction and lattice row (iy,iz) the kernel forms
// T = sum_ix Ex(ix) P(ix,iy,iz) over the row's occupied sites (a contiguous
// compact range) and adds Ey Ez T: three complex FP64 multiply-adds per site
// and direction, no transcendental per site.  Per-axis phase tables are built
// in FP64 on the device; row-chunk partials are reduced in a fixed order.
// Angular quadrature (directions, weights) is restated from cu:849-875 on
// the host.
#include "solvers.hpp"
#include "../../include/ddscat_matvec_hip.h"
#include <vector>
#include <cmath>
#include <complex>
#include <cstdio>

namespace ddhip {

struct FarfieldRegion {
  unsigned char* base = nullptr;
  size_t bytes = 0;
};

__global__ void k_evalq_q(const float2* __restrict__ p, float2* __restrict__ q, const float2* __restrict__ ad,
                          const float2* __restrict__ ao, int nat0, float rabs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nat0) return;
  const float2 p1 = p[j], p2 = p[j + nat0], p3 = p[j + 2 * nat0];
  const float2 d1 = ad[j], d2 = ad[j + nat0], d3 = ad[j + 2 * nat0];
  const float2 a23 = ao[j], a31 = ao[j + nat0], a12 = ao[j + 2 * nat0];
  const float2 ri = make_float2(0.f, rabs);
  q[j] = cadd(cadd(cmul(cadd(d1, ri), p1), cmul(a31, p3)), cmul(a12, p2));
  q[j + nat0] = cadd(cadd(cmul(cadd(d2, ri), p2), cmul(a12, p1)), cmul(a23, p3));
  q[j + 2 * nat0] = cadd(cadd(cmul(cadd(d3, ri), p3), cmul(a23, p2)), cmul(a31, p1));
}

int evalq(Context& c, int use_resident, const void* e_c, const void* p_c, const void* ad_c, const void* ao_c,
          float akx, float aky, float akz, float e02, double* cabs, double* cext, double* cpha) {
  clear_error();
  if (!c.ready || !cabs || !cext || !cpha || !(e02 > 0.f)) { set_error("GPU EVALQ", "invalid argument"); return 300; }
  const size_t cb = (size_t)3 * c.nat0 * sizeof(float2);
  hipStream_t st = c.stream;
  if (use_resident) {
    if (!c.solution_valid) { set_error("GPU EVALQ", "no resident final polarization published by solver"); return 301; }
  } else {
    if (!e_c || !p_c || !ad_c || !ao_c) { set_error("GPU EVALQ", "missing compact host arrays"); return 302; }
    DDH_TRY(hipMemcpyAsync(c.sb, e_c, cb, hipMemcpyHostToDevice, st), "GPU EVALQ E H2D", 303);
    DDH_TRY(hipMemcpyAsync(c.sx, p_c, cb, hipMemcpyHostToDevice, st), "GPU EVALQ P H2D", 303);
    DDH_TRY(hipMemcpyAsync(c.adia_c, ad_c, cb, hipMemcpyHostToDevice, st), "GPU EVALQ ADIA H2D", 305);
    DDH_TRY(hipMemcpyAsync(c.aoff_c, ao_c, cb, hipMemcpyHostToDevice, st), "GPU EVALQ AOFF H2D", 305);
    c.solution_valid = true;
  }
  const double ak2 = (double)akx * akx + (double)aky * aky + (double)akz * akz, akk = std::sqrt(ak2),
               ak3 = akk * ak2;
  hipLaunchKernelGGL(k_evalq_q, dim3((c.nat0 + 255) / 256), dim3(256), 0, st, c.sx, c.sy, c.adia_c, c.aoff_c,
                     c.nat0, (float)(ak3 / 1.5));
  DDH_LAUNCH_OK("GPU EVALQ local alpha^-1 P", 306);
  DotSet d;
  d.add(c.sx, c.sy, DOT_C).add(c.sb, c.sx, DOT_C);
  // EVALQ sums are always FP64, whatever the solver precision mode
  const int prec = c.precision;
  c.precision = PREC_MIXED;
  const int drc = dots(c, d, (size_t)3 * c.nat0, SL_D, st);
  c.precision = prec;
  if (drc || fetch_scalars(c, SL_D, 4, st)) return 307;
  const double pi = 3.14159265358979323846;
  const double zabs_im = c.hscal[SL_D + 1], zep_re = c.hscal[SL_D + 2], zep_im = c.hscal[SL_D + 3];
  *cabs = -4.0 * pi * akk * zabs_im / (double)e02;
  *cext = 4.0 * pi * akk * zep_im / (double)e02;
  *cpha = 2.0 * pi * akk * zep_re / (double)e02;
  return 0;
}

// ------------------------------------------------------------------ SCAT --
struct ScatDir {
  float k[3];
  double w[5];
};

static long long round_half_away(double v) { return v >= 0 ? (long long)std::floor(v + 0.5) : -(long long)std::floor(-v + 0.5); }

static int scat_integration_dirs(const float ak[3], const float2 e01[3], float etasca, int nat0,
                                 std::vector<ScatDir>& dirs) {
  const double pi = 3.14159265358979323846;
  const double ak2 = (double)ak[0] * ak[0] + (double)ak[1] * ak[1] + (double)ak[2] * ak[2], akk = std::sqrt(ak2);
  double a1[3] = {(double)e01[0].x, (double)e01[1].x, (double)e01[2].x};
  const double einc = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  if (!(einc > 0.0)) { set_error("GPU SCAT", "CXE01_TF is pure imaginary"); return 1; }
  for (double& v : a1) v *= akk / einc;
  const double a2[3] = {(ak[1] * a1[2] - ak[2] * a1[1]) / akk, (ak[2] * a1[0] - ak[0] * a1[2]) / akk,
                        (ak[0] * a1[1] - ak[1] * a1[0]) / akk};
  const double x = akk * std::cbrt(3.0 * (double)nat0 / (4.0 * pi)), theta0 = 2.0 * pi / (1.0 + x), afac = 1.0;
  const int ntheta = 1 + (int)round_half_away((2.0 * (3.0 + x) / (double)etasca) * (1.0 + afac / (pi + theta0)) /
                                              (1.0 + afac * theta0 / ((pi + theta0) * (pi + theta0))));
  double theta = 0.0, thetau = 0.0;
  dirs.clear();
  for (int ic = 1; ic <= ntheta; ++ic) {
    const double thetal = theta;
    theta = thetau;
    if (ic < ntheta) {
      const double si = (double)ic * (pi + afac * pi / (theta0 + pi)) / (double)(ntheta - 1), term = si - afac - theta0;
      thetau = 0.5 * (term + std::sqrt(term * term + 4.0 * theta0 * si));
    } else {
      thetau = pi;
    }
    const double domega = 2.0 * pi * (std::cos(0.5 * (thetal + theta)) - std::cos(0.5 * (theta + thetau)));
    const double ct = std::cos(theta), stt = std::sin(theta);
    int nphi = (ic == 1 || ic == ntheta) ? 1 : (int)round_half_away(4.0 * pi * stt / (thetau - thetal));
    if (nphi < 3 && ic != 1 && ic != ntheta) nphi = 3;
    const double w = domega / (double)nphi, dphi = 2.0 * pi / (double)nphi;
    for (int ip = 1; ip <= nphi; ++ip) {
      const double phi = dphi * ((double)ip - 0.5 * (double)(ic % 2)), sp = std::sin(phi), cp = std::cos(phi);
      ScatDir d{};
      for (int k = 0; k < 3; ++k) d.k[k] = (float)(ct * (double)ak[k] + stt * (cp * a1[k] + sp * a2[k]));
      d.w[0] = w; d.w[1] = w * ct; d.w[2] = w * stt * cp; d.w[3] = w * stt * sp; d.w[4] = w * ct * ct;
      dirs.push_back(d);
    }
  }
  return 0;
}

// per-axis phase tables tab[(axis offset + i) * ndir + d] = exp(-i k_a ((i+1)+x0_a) dx_a), FP64
// (direction fastest, so the lanes of a workgroup read consecutive entries)
__global__ void k_phase_tables(const float* __restrict__ kd, int ndir, int nx, int ny, int nz, double dx1, double dx2,
                               double dx3, double x01, double x02, double x03, double2* __restrict__ tab) {
  const int L = nx + ny + nz;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)ndir * L) return;
  const int d = (int)(id / L), i = (int)(id - (size_t)d * L);
  double r, k;
  if (i < nx) { r = ((double)(i + 1) + x01) * dx1; k = kd[3 * d]; }
  else if (i < nx + ny) { r = ((double)(i - nx + 1) + x02) * dx2; k = kd[3 * d + 1]; }
  else { r = ((double)(i - nx - ny + 1) + x03) * dx3; k = kd[3 * d + 2]; }
  double s, co;
  sincos(k * r, &s, &co);
  tab[(size_t)i * ndir + d] = make_double2(co, -s);
}

constexpr int FF_DIRS = 64;    // directions per workgroup (one per lane)
constexpr int FF_ROWS = 32;    // lattice rows per workgroup

// partial[chunk][d][3] = sum over the chunk's rows of Ey Ez sum_row Ex P
__global__ __launch_bounds__(FF_DIRS) void k_farfield_rows(
    const float2* __restrict__ p, const int* __restrict__ occ, const int* __restrict__ rowptr,
    const double2* __restrict__ tab, int ndir, int nx, int ny, int nz, int nat0, double2* __restrict__ part) {
  const int chunk = blockIdx.x;
  const int d = blockIdx.y * FF_DIRS + threadIdx.x;
  const int nrows = ny * nz;
  const int g0 = chunk * FF_ROWS, g1 = min(nrows, g0 + FF_ROWS);
  double2 s0 = make_double2(0, 0), s1 = s0, s2 = s0;
  if (d < ndir) {
    const double2* t = tab + d;
    const size_t sd = (size_t)ndir;
    for (int g = g0; g < g1; ++g) {
      const int q0 = rowptr[g], q1 = rowptr[g + 1];
      if (q0 == q1) continue;
      double2 r0 = make_double2(0, 0), r1 = r0, r2 = r0;
      const int base = g * nx;
      for (int q = q0; q < q1; ++q) {
        const double2 e = t[(size_t)(occ[q] - base) * sd];
        const float2 a = p[q], b = p[q + nat0], cc = p[q + 2 * nat0];
        r0.x += e.x * a.x - e.y * a.y; r0.y += e.x * a.y + e.y * a.x;
        r1.x += e.x * b.x - e.y * b.y; r1.y += e.x * b.y + e.y * b.x;
        r2.x += e.x * cc.x - e.y * cc.y; r2.y += e.x * cc.y + e.y * cc.x;
      }
      const int iy = g % ny, iz = g / ny;
      const double2 eyz = zmul(t[(size_t)(nx + iy) * sd], t[(size_t)(nx + ny + iz) * sd]);
      s0 = zadd(s0, zmul(eyz, r0));
      s1 = zadd(s1, zmul(eyz, r1));
      s2 = zadd(s2, zmul(eyz, r2));
    }
    double2* o = part + ((size_t)chunk * ndir + d) * 3;
    o[0] = s0; o[1] = s1; o[2] = s2;
  }
}

// Position-weighted far-field sums for the radiative torque:
// partial[chunk][d][12] = S_b (b = x, y, z) then R_ab = sum r_a P_b e^{-i k.r}
// (a = x, y, z, b fastest) over the chunk's rows.  Per row the x weight varies
// along the row (one extra multiply-add set); y and z weights are the row's.
__global__ __launch_bounds__(FF_DIRS) void k_farfield_rows_pos(
    const float2* __restrict__ p, const int* __restrict__ occ, const int* __restrict__ rowptr,
    const double2* __restrict__ tab, int ndir, int nx, int ny, int nz, int nat0, double dx1, double dx2, double dx3,
    double x01, double x02, double x03, double2* __restrict__ part) {
  const int chunk = blockIdx.x;
  const int d = blockIdx.y * FF_DIRS + threadIdx.x;
  const int nrows = ny * nz;
  const int g0 = chunk * FF_ROWS, g1 = min(nrows, g0 + FF_ROWS);
  if (d >= ndir) return;
  double2 acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = make_double2(0, 0);
  const double2* t = tab + d;
  const size_t sd = (size_t)ndir;
  for (int g = g0; g < g1; ++g) {
    const int q0 = rowptr[g], q1 = rowptr[g + 1];
    if (q0 == q1) continue;
    double2 r[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    double2 w[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    const int base = g * nx;
    for (int q = q0; q < q1; ++q) {
      const int ix = occ[q] - base;
      const double2 e = t[(size_t)ix * sd];
      const double xw = ((double)(ix + 1) + x01) * dx1;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const float2 a = p[q + (size_t)b * nat0];
        const double2 ea = make_double2(e.x * a.x - e.y * a.y, e.x * a.y + e.y * a.x);
        r[b] = zadd(r[b], ea);
        w[b].x += xw * ea.x;
        w[b].y += xw * ea.y;
      }
    }
    const int iy = g % ny, iz = g / ny;
    const double2 eyz = zmul(t[(size_t)(nx + iy) * sd], t[(size_t)(nx + ny + iz) * sd]);
    const double yw = ((double)(iy + 1) + x02) * dx2, zw = ((double)(iz + 1) + x03) * dx3;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double2 sb = zmul(eyz, r[b]);
      acc[b] = zadd(acc[b], sb);
      acc[3 + b] = zadd(acc[3 + b], zmul(eyz, w[b]));
      acc[6 + b] = zadd(acc[6 + b], zscale(sb, yw));
      acc[9 + b] = zadd(acc[9 + b], zscale(sb, zw));
    }
  }
  double2* o = part + ((size_t)chunk * ndir + d) * 12;
#pragma unroll
  for (int i = 0; i < 12; ++i) o[i] = acc[i];
}

// S[d][width] = sum over chunks of partial[chunk][d][width], fixed order
__global__ void k_farfield_reduce(const double2* __restrict__ part, int nchunks, int ndir, int width,
                                  double2* __restrict__ S) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)ndir * width) return;
  double2 acc = make_double2(0, 0);
  for (int ch = 0; ch < nchunks; ++ch) acc = zadd(acc, part[(size_t)ch * ndir * width + id]);
  S[id] = acc;
}

// Incident-beam torque terms, per workgroup: out[blk][0..2] = sum Im(P^* . E) r,
// out[blk][3..5] = sum Re(P x E^*), in FP64 over the compact sites.
constexpr int TQ_THREADS = 256;
__global__ __launch_bounds__(TQ_THREADS) void k_torque_local(const float2* __restrict__ p,
                                                             const float2* __restrict__ e, const int* __restrict__ occ,
                                                             int nx, int ny, int nat0, double dx1, double dx2,
                                                             double dx3, double x01, double x02, double x03,
                                                             double* __restrict__ out) {
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nat0; q += gridDim.x * blockDim.x) {
    const int j = occ[q];
    const int ix = j % nx, iy = (j / nx) % ny, iz = j / (nx * ny);
    const double r[3] = {((double)(ix + 1) + x01) * dx1, ((double)(iy + 1) + x02) * dx2,
                         ((double)(iz + 1) + x03) * dx3};
    double2 P[3], E[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float2 pb = p[q + (size_t)b * nat0], eb = e[q + (size_t)b * nat0];
      P[b] = make_double2(pb.x, pb.y);
      E[b] = make_double2(eb.x, eb.y);
    }
    // Im(conj(P) . E)
    double im = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) im += P[b].x * E[b].y - P[b].y * E[b].x;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] += im * r[a];
    // Re(P x conj(E))_a = Re(P_b conj(E_c) - P_c conj(E_b)), (a, b, c) cyclic
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int b = (a + 1) % 3, c = (a + 2) % 3;
      v[3 + a] += (P[b].x * E[c].x + P[b].y * E[c].y) - (P[c].x * E[b].x + P[c].y * E[b].y);
    }
  }
  __shared__ double red[TQ_THREADS];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    red[threadIdx.x] = v[k];
    __syncthreads();
    for (int h = TQ_THREADS / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * 6 + k] = red[0];
    __syncthreads();
  }
}

// S[d][3] for all directions (host result) of the compact polarization p.
// Device workspace (directions, phase tables, row-chunk partials, sums):
// allocated per call, or carved out of `region` (the allocation-free SCAT
// entry), where the direction slab shrinks to fit.  Returns 1 on failure, 2
// when the region cannot hold even one direction's workspace.
static int farfield_sums(Context& c, const float2* p, const std::vector<float>& kdirs, const float dx[3],
                         const float x0[3], std::vector<std::complex<double>>& S, const FarfieldRegion* region,
                         bool pos = false) {
  const int ndir = (int)(kdirs.size() / 3);
  const int width = pos ? 12 : 3;  // complex sums per direction: S, or S and the position-weighted R
  S.assign((size_t)ndir * width, 0.0);
  if (ndir == 0) return 0;
  hipStream_t st = c.stream;
  const int L = c.nx + c.ny + c.nz;
  const int nrows = c.ny * c.nz;
  const int nchunks = (nrows + FF_ROWS - 1) / FF_ROWS;
  float* dk = nullptr;
  double2 *tab = nullptr, *part = nullptr, *dS = nullptr;
  int rc = 0;
  int slab = 0;
  auto cleanup = [&]() {
    if (region) return;
    hipFree(dk); hipFree(tab); hipFree(part); hipFree(dS);
  };
  if (region) {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t fixed = up(kdirs.size() * sizeof(float)) + up((size_t)ndir * width * sizeof(double2));
    const size_t per_dir = (size_t)L * sizeof(double2) + (size_t)nchunks * width * sizeof(double2);
    if (region->bytes < fixed + 512 + per_dir) {
      char msg[200];
      std::snprintf(msg, sizeof(msg), "far-field workspace of %d directions on %d lattice rows needs more than the "
                    "%zu-byte scratch region", ndir, nrows, region->bytes);
      set_error("GPU SCAT", msg);
      return 2;
    }
    slab = (int)std::min<size_t>((size_t)ndir, (region->bytes - fixed - 512) / per_dir);
    unsigned char* b = region->base;
    dk = (float*)b; b += up(kdirs.size() * sizeof(float));
    dS = (double2*)b; b += up((size_t)ndir * width * sizeof(double2));
    tab = (double2*)b; b += up((size_t)slab * L * sizeof(double2));
    part = (double2*)b;
  } else {
    // bound the partial buffer: process directions in slabs
    slab = std::max(FF_DIRS, std::min(ndir, (int)std::max<size_t>(FF_DIRS, (size_t)(256u << 20) /
                                                                    ((size_t)nchunks * width * sizeof(double2)))));
    if (hipMalloc((void**)&dk, kdirs.size() * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&tab, (size_t)slab * L * sizeof(double2)) != hipSuccess ||
        hipMalloc((void**)&part, (size_t)nchunks * slab * width * sizeof(double2)) != hipSuccess ||
        hipMalloc((void**)&dS, (size_t)ndir * width * sizeof(double2)) != hipSuccess) {
      set_error("GPU SCAT", "far-field workspace allocation failed");
      cleanup();
      return 1;
    }
  }
  if (hipMemcpyAsync(dk, kdirs.data(), kdirs.size() * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error("GPU SCAT", "directions H2D failed");
    cleanup();
    return 1;
  }
  for (int d0 = 0; d0 < ndir && !rc; d0 += slab) {
    const int nd = std::min(slab, ndir - d0);
    const size_t nt = (size_t)nd * L;
    hipLaunchKernelGGL(k_phase_tables, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, dk + 3 * d0, nd, c.nx,
                       c.ny, c.nz, (double)dx[0], (double)dx[1], (double)dx[2], (double)x0[0], (double)x0[1],
                       (double)x0[2], tab);
    if (pos)
      hipLaunchKernelGGL(k_farfield_rows_pos, dim3(nchunks, (nd + FF_DIRS - 1) / FF_DIRS), dim3(FF_DIRS), 0, st, p,
                         c.occ, c.rowptr, tab, nd, c.nx, c.ny, c.nz, c.nat0, (double)dx[0], (double)dx[1],
                         (double)dx[2], (double)x0[0], (double)x0[1], (double)x0[2], part);
    else
      hipLaunchKernelGGL(k_farfield_rows, dim3(nchunks, (nd + FF_DIRS - 1) / FF_DIRS), dim3(FF_DIRS), 0, st, p,
                         c.occ, c.rowptr, tab, nd, c.nx, c.ny, c.nz, c.nat0, part);
    hipLaunchKernelGGL(k_farfield_reduce, dim3((unsigned)((nd * width + 255) / 256)), dim3(256), 0, st, part,
                       nchunks, nd, width, dS + (size_t)width * d0);
    if (hipGetLastError() != hipSuccess) { set_error("GPU SCAT", "far-field kernel launch failed"); rc = 1; }
  }
  if (!rc && (hipMemcpyAsync(S.data(), dS, (size_t)ndir * width * sizeof(double2), hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess)) {
    set_error("GPU SCAT", "far-field sums D2H failed");
    rc = 1;
  }
  cleanup();
  return rc;
}

int scat(Context& c, int jpbc, int ndir, const float* ak, const float* aks, const float* dx, const float* em1,
         const float* em2, float e02, float etasca, const void* e01v, const float* x0, double* cbksca, double* csca,
         double* cscag, double* cscag2, int* navg, void* f1v, void* f2v, const FarfieldRegion* region) {
  clear_error();
  if (!c.ready || !c.solution_valid || !ak || !dx || !x0 || !cbksca || !csca || !cscag || !cscag2 || !navg ||
      !(e02 > 0.f)) {
    set_error("GPU SCAT", "invalid state/argument");
    return 320;
  }
  if (ndir > 0 && (!aks || !em1 || !em2 || !f1v || !f2v)) { set_error("GPU SCAT", "missing selected-direction arrays"); return 321; }
  const double ak2 = (double)ak[0] * ak[0] + (double)ak[1] * ak[1] + (double)ak[2] * ak[2], akk = std::sqrt(ak2),
               ak3 = akk * ak2;
  *cbksca = 0.0; *csca = 0.0; cscag[0] = cscag[1] = cscag[2] = 0.0; *cscag2 = 0.0; *navg = 0;
  if (jpbc == 0) {
    if (!e01v || !(etasca > 0.f)) { set_error("GPU SCAT", "finite-target integration needs CXE01 and ETASCA>0"); return 323; }
    std::vector<ScatDir> dirs;
    if (scat_integration_dirs(ak, (const float2*)e01v, etasca, c.nat0, dirs)) return 324;
    *navg = (int)dirs.size();
    std::vector<float> kd(3 * dirs.size());
    for (size_t i = 0; i < dirs.size(); ++i)
      for (int k = 0; k < 3; ++k) kd[3 * i + k] = dirs[i].k[k];
    std::vector<std::complex<double>> S;
    if (int frc = farfield_sums(c, c.sx, kd, dx, x0, S, region)) return frc == 2 ? 334 : 325;
    double acc[5] = {0, 0, 0, 0, 0}, back = 0.0;
    for (size_t i = 0; i < dirs.size(); ++i) {
      const double nxv = (double)kd[3 * i] / akk, nyv = (double)kd[3 * i + 1] / akk, nzv = (double)kd[3 * i + 2] / akk;
      const std::complex<double> sx = S[3 * i], sy = S[3 * i + 1], sz = S[3 * i + 2];
      const std::complex<double> nd = nxv * sx + nyv * sy + nzv * sz;
      const double es2 = std::norm(sx - nxv * nd) + std::norm(sy - nyv * nd) + std::norm(sz - nzv * nd);
      for (int k = 0; k < 5; ++k) acc[k] += dirs[i].w[k] * es2;
      if (i + 1 == dirs.size()) back = es2;
    }
    const double fac = ak2 * ak2 / (double)e02;
    *csca = fac * acc[0];
    cscag[0] = fac * acc[1]; cscag[1] = fac * acc[2]; cscag[2] = fac * acc[3];
    *cscag2 = fac * acc[4];
    *cbksca = fac * back;
  }
  if (ndir > 0) {
    std::vector<float> kd(aks, aks + 3 * (size_t)ndir);
    std::vector<std::complex<double>> S;
    if (int frc = farfield_sums(c, c.sx, kd, dx, x0, S, region)) return frc == 2 ? 334 : 330;
    const double term = ak3 / std::sqrt((double)e02);
    float2* f1 = (float2*)f1v;
    float2* f2 = (float2*)f2v;
    for (int i = 0; i < ndir; ++i) {
      std::complex<double> a1 = 0.0, a2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        a1 += (double)em1[3 * i + k] * S[3 * i + k];
        a2 += (double)em2[3 * i + k] * S[3 * i + k];
      }
      f1[i] = make_float2((float)(term * a1.real()), (float)(term * a1.imag()));
      f2[i] = make_float2((float)(term * a2.real()), (float)(term * a2.imag()));
    }
  }
  return 0;
}

// Amplitude and Mueller matrices of incident-polarization pairs.  The
// amplitudes are SCAT's (cu:841-845), f_ij = k^3/|E0| e_i . S_j with
// e_1 = e_par_s, e_2 = e_perp the scattered basis and j the incident
// polarization; the conversion that DDSCAT's GETMUELLER makes on the host
// (DDSCAT_cuda.f90:3294-3300, routine not in the reference repository) is
// restated from Bohren & Huffman ch. 3:
//   [[S2, S3], [S4, S1]] = -i F M,  M_jk = conj(e0j) . (e_par_i, e_perp)_k,
//   e_par_i = khat x e_perp,
// and the Mueller matrix is their eq. 3.16.
static void amplitude_to_mueller(const std::complex<double> f[4], const double khat[3], const double ep[3],
                                 const std::complex<double> e01[3], const std::complex<double> e02[3], double* M) {
  using cd = std::complex<double>;
  const double pi_[3] = {khat[1] * ep[2] - khat[2] * ep[1], khat[2] * ep[0] - khat[0] * ep[2],
                         khat[0] * ep[1] - khat[1] * ep[0]};
  cd m11 = 0.0, m12 = 0.0, m21 = 0.0, m22 = 0.0;
  for (int k = 0; k < 3; ++k) {
    m11 += std::conj(e01[k]) * pi_[k];
    m12 += std::conj(e01[k]) * ep[k];
    m21 += std::conj(e02[k]) * pi_[k];
    m22 += std::conj(e02[k]) * ep[k];
  }
  const cd mi(0.0, -1.0);
  const cd f11 = f[0], f21 = f[1], f12 = f[2], f22 = f[3];
  const cd S2 = mi * (f11 * m11 + f12 * m21), S3 = mi * (f11 * m12 + f12 * m22);
  const cd S4 = mi * (f21 * m11 + f22 * m21), S1 = mi * (f21 * m12 + f22 * m22);
  const double a1 = std::norm(S1), a2 = std::norm(S2), a3 = std::norm(S3), a4 = std::norm(S4);
  const cd s23 = S2 * std::conj(S3), s14 = S1 * std::conj(S4), s24 = S2 * std::conj(S4), s13 = S1 * std::conj(S3);
  const cd s12 = S1 * std::conj(S2), s34 = S3 * std::conj(S4);
  M[0] = 0.5 * (a2 + a3 + a4 + a1);
  M[1] = 0.5 * (a2 - a3 + a4 - a1);
  M[2] = (s23 + s14).real();
  M[3] = (s23 - s14).imag();
  M[4] = 0.5 * (a2 - a4 + a3 - a1);
  M[5] = 0.5 * (a2 - a3 - a4 + a1);
  M[6] = (s23 - s14).real();
  M[7] = (s23 + s14).imag();
  M[8] = (s24 + s13).real();
  M[9] = (s24 - s13).real();
  M[10] = (s12 + s34).real();
  M[11] = (std::conj(s12) + std::conj(s34)).imag();
  M[12] = (std::conj(s24) + s13).imag();  // Im(S2* S4 + S3* S1)
  M[13] = (std::conj(s24) - s13).imag();  // Im(S2* S4 - S3* S1)
  M[14] = (s12 - s34).imag();
  M[15] = (s12 - s34).real();
}

// Radiative torque of the resident solution (DDSCAT's DOTORQ sums), as torque
// cross sections C = Gamma k / u_rad, u_rad = E0^2 / 8 pi, about the origin of
// the site coordinates r_j = ((i+1) + x0) dx:
//   ext: the incident wave's torque on the dipoles,
//        sum_j [ r_j x F_j + (1/2) Re(P_j x E_j^*) ],  F_j = -(1/2) Im(P_j^* . E_j) k;
//   rad: the recoil of the radiated field.  The radiative part of the
//        interaction is G_I(r) = (k^3 / 4 pi) int dOmega (I - n n) e^{i k n.r}
//        (radiation reaction (2/3) k^3 at r = 0), so its torque on the dipoles is
//        (k^3 / 4 pi) int dOmega [ (1/2) Re(-i S x S_perp^*) - (k/2) Re(conj(R) S_perp) x n ]
//        with S_b = sum_j P_jb e^{-i k n.r_j}, R_ab = sum_j r_ja P_jb e^{-i k n.r_j},
//        S_perp = S - n (n.S), integrated over SCAT's ETASCA quadrature.
// The reactive part of the interaction exerts no net torque, so ext + rad is
// the torque on the target; for an absorbing sphere in circularly polarised
// light its component along k is +-C_abs.  E is the resident right-hand side
// (the incident field at the sites, as EVALQ's resident mode uses it).
int torque(Context& c, const float* ak, const float2* e01, float e02, float etasca, const float* dx, const float* x0,
           double* cext, double* crad, int* navg) {
  clear_error();
  if (!c.ready || !c.solution_valid || !ak || !e01 || !dx || !x0 || !cext || !crad || !navg || !(e02 > 0.f) ||
      !(etasca > 0.f)) {
    set_error("torque", "invalid state/argument");
    return 350;
  }
  const double kv[3] = {ak[0], ak[1], ak[2]};
  const double k = std::sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  if (!(k > 0.0)) { set_error("torque", "zero wave vector"); return 350; }
  hipStream_t st = c.stream;
  // incident-beam part: per-workgroup partials in the reduction buffer
  const int nb = std::min(RED_MAX_BLOCKS, std::max(1, (c.nat0 + TQ_THREADS - 1) / TQ_THREADS));
  hipLaunchKernelGGL(k_torque_local, dim3(nb), dim3(TQ_THREADS), 0, st, c.sx, c.sb, c.occ, c.nx, c.ny, c.nat0,
                     (double)dx[0], (double)dx[1], (double)dx[2], (double)x0[0], (double)x0[1], (double)x0[2], c.red);
  if (hipGetLastError() != hipSuccess) { set_error("torque", "local-sum kernel launch failed"); return 351; }
  std::vector<double> part((size_t)nb * 6);
  if (hipMemcpyAsync(part.data(), c.red, part.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    set_error("torque", "local sums D2H failed");
    return 351;
  }
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int i = 0; i < 6; ++i) v[i] += part[(size_t)b * 6 + i];
  const double u = (double)e02 / (8.0 * M_PI);
  const double cf = k / u;
  // sum_j r_j x F_j = -(1/2) (sum_j Im(P^*.E) r_j) x k
  cext[0] = cf * (-0.5 * (v[1] * kv[2] - v[2] * kv[1]) + 0.5 * v[3]);
  cext[1] = cf * (-0.5 * (v[2] * kv[0] - v[0] * kv[2]) + 0.5 * v[4]);
  cext[2] = cf * (-0.5 * (v[0] * kv[1] - v[1] * kv[0]) + 0.5 * v[5]);
  // radiated part over the quadrature directions
  std::vector<ScatDir> dirs;
  if (scat_integration_dirs(ak, e01, etasca, c.nat0, dirs)) return 350;
  *navg = (int)dirs.size();
  std::vector<float> kd(3 * dirs.size());
  for (size_t i = 0; i < dirs.size(); ++i)
    for (int a = 0; a < 3; ++a) kd[3 * i + a] = dirs[i].k[a];
  std::vector<std::complex<double>> SR;
  if (farfield_sums(c, c.sx, kd, dx, x0, SR, nullptr, true)) return 351;
  using cd = std::complex<double>;
  double acc[3] = {0, 0, 0};
  for (size_t i = 0; i < dirs.size(); ++i) {
    const cd* q = SR.data() + 12 * i;
    double n[3];
    double kk = 0.0;
    for (int a = 0; a < 3; ++a) { n[a] = kd[3 * i + a]; kk += n[a] * n[a]; }
    kk = std::sqrt(kk);
    for (int a = 0; a < 3; ++a) n[a] /= kk;
    const cd S[3] = {q[0], q[1], q[2]};
    const cd ns = n[0] * S[0] + n[1] * S[1] + n[2] * S[2];
    const cd Sp[3] = {S[0] - n[0] * ns, S[1] - n[1] * ns, S[2] - n[2] * ns};
    double t[3];
    for (int a = 0; a < 3; ++a) {
      const int b = (a + 1) % 3, c2 = (a + 2) % 3;
      // (1/2) Re(-i (S x Sp^*)_a)
      const cd cr = S[b] * std::conj(Sp[c2]) - S[c2] * std::conj(Sp[b]);
      t[a] = 0.5 * std::imag(cr);
    }
    // m_a = Re(sum_b conj(R_ab) Sp_b); orbital -(k/2) m x n
    double m[3];
    for (int a = 0; a < 3; ++a) {
      cd z = 0.0;
      for (int b = 0; b < 3; ++b) z += std::conj(q[3 + 3 * a + b]) * Sp[b];
      m[a] = z.real();
    }
    t[0] -= 0.5 * k * (m[1] * n[2] - m[2] * n[1]);
    t[1] -= 0.5 * k * (m[2] * n[0] - m[0] * n[2]);
    t[2] -= 0.5 * k * (m[0] * n[1] - m[1] * n[0]);
    for (int a = 0; a < 3; ++a) acc[a] += dirs[i].w[0] * t[a];
  }
  const double pre = k * k * k / (4.0 * M_PI) * cf;
  for (int a = 0; a < 3; ++a) crad[a] = pre * acc[a];
  return 0;
}

int scat_matrix(Context& c, int npair, const float2* p, int ndir, const float* ak, const float2* e0,
                const float* aks, const float* em1, const float* em2, const float* dx, const float* x0, float e02,
                float2* famp, double* mueller) {
  clear_error();
  if (!c.ready || npair <= 0 || ndir <= 0 || !p || !ak || !e0 || !aks || !em1 || !em2 || !dx || !x0 || !famp ||
      !(e02 > 0.f)) {
    set_error("scat_matrix", "invalid state/argument");
    return 340;
  }
  const size_t n = (size_t)3 * c.nat0;
  for (int q = 0; q < npair; ++q) {
    const float* akq = ak + 3 * q;
    const double k2 = (double)akq[0] * akq[0] + (double)akq[1] * akq[1] + (double)akq[2] * akq[2], k = std::sqrt(k2);
    if (!(k > 0.0)) { set_error("scat_matrix", "zero wave vector"); return 340; }
    const double term = k * k2 / std::sqrt((double)e02);
    const double khat[3] = {akq[0] / k, akq[1] / k, akq[2] / k};
    std::complex<double> e01v[3], e02v[3];
    for (int i = 0; i < 3; ++i) {
      e01v[i] = {e0[6 * q + i].x, e0[6 * q + i].y};
      e02v[i] = {e0[6 * q + 3 + i].x, e0[6 * q + 3 + i].y};
    }
    const size_t o = (size_t)q * ndir * 3;
    std::vector<float> kd(aks + o, aks + o + 3 * (size_t)ndir);
    std::vector<std::complex<double>> S[2];
    for (int j = 0; j < 2; ++j)
      if (farfield_sums(c, p + (2 * (size_t)q + j) * n, kd, dx, x0, S[j], nullptr)) return 341;
    for (int d = 0; d < ndir; ++d) {
      const float* a = em1 + o + 3 * d;
      const float* b = em2 + o + 3 * d;
      std::complex<double> f[4];  // f11, f21, f12, f22
      for (int j = 0; j < 2; ++j) {
        std::complex<double> s1 = 0.0, s2 = 0.0;
        for (int kk = 0; kk < 3; ++kk) {
          s1 += (double)a[kk] * S[j][3 * d + kk];
          s2 += (double)b[kk] * S[j][3 * d + kk];
        }
        f[2 * j] = term * s1;
        f[2 * j + 1] = term * s2;
      }
      float2* fo = famp + ((size_t)q * ndir + d) * 4;
      for (int i = 0; i < 4; ++i) fo[i] = make_float2((float)f[i].real(), (float)f[i].imag());
      if (mueller) {
        const double ep[3] = {b[0], b[1], b[2]};
        amplitude_to_mueller(f, khat, ep, e01v, e02v, mueller + ((size_t)q * ndir + d) * 16);
      }
    }
  }
  return 0;
}

}  // namespace ddhip

using namespace ddhip;

extern "C" DDSCAT_HIP_API int ddscat_hip_scat_matrix(int npair, const void* p, int ndir, const float* ak,
                                                     const void* e0, const float* aks, const float* em1,
                                                     const float* em2, const float* dx, const float* x0, float e02,
                                                     void* famp, double* mueller) {
  return scat_matrix(ctx(), npair, (const float2*)p, ndir, ak, (const float2*)e0, aks, em1, em2, dx, x0, e02,
                     (float2*)famp, mueller);
}

extern "C" DDSCAT_HIP_API int ddscat_cuda_evalq_f32(int use_resident, const void* e_compact, const void* p_compact,
                                                    const void* adia_compact, const void* aoff_compact, float akx,
                                                    float aky, float akz, float e02, double* cabs, double* cext,
                                                    double* cpha) {
  return evalq(ctx(), use_resident, e_compact, p_compact, adia_compact, aoff_compact, akx, aky, akz, e02, cabs, cext,
               cpha);
}

extern "C" DDSCAT_HIP_API int ddscat_cuda_scat_f32(int jpbc, int ndir, const float* ak, const float* aks,
                                                   const float* dx, const float* em1, const float* em2, float e02,
                                                   float etasca, const void* e01, const float* x0, double* cbksca,
                                                   double* csca, double* cscag, double* cscag2, int* navg, void* f1,
                                                   void* f2) {
  return scat(ctx(), jpbc, ndir, ak, aks, dx, em1, em2, e02, etasca, e01, x0, cbksca, csca, cscag, cscag2, navg, f1,
              f2, nullptr);
}

extern "C" DDSCAT_HIP_API int ddscat_hip_torque(const float* ak, const void* e01, float e02, float etasca,
                                                const float* dx, const float* x0, double* cext, double* crad,
                                                int* navg) {
  return torque(ctx(), ak, (const float2*)e01, e02, etasca, dx, x0, cext, crad, navg);
}

// ---- allocation-free SCAT ---------------------------------------------------
// The far-field workspace of ddscat_hip_scat_f32_noalloc lives in this
// file-scope device array, one copy per GPU, held until the process ends.
// Host code never uses the array's name as an address: its device address
// is looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_addr.  Every call on a GPU shares that GPU's copy (calls that
// overlap on one GPU race); nothing in it needs to start at zero (every
// kernel writes what the next one reads).
namespace {
constexpr size_t SCAT_SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[SCAT_SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
void* g_scratch_addr[SCRATCH_MAX_DEVICES] = {};

int scratch_region(FarfieldRegion& r) {
  int dev = 0;
  DDH_TRY(hipGetDevice(&dev), "SCAT scratch", 1);
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) { set_error("SCAT scratch", "device index out of range"); return 1; }
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    DDH_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "SCAT scratch address", 1);
    if (!ptr) { set_error("SCAT scratch", "no device address for the scratch array"); return 1; }
    g_scratch_addr[dev] = ptr;
  }
  r.base = (unsigned char*)g_scratch_addr[dev];
  r.bytes = SCAT_SCRATCH_BYTES;
  return 0;
}
}  // namespace

extern "C" DDSCAT_HIP_API int ddscat_hip_scat_f32_noalloc(int jpbc, int ndir, const float* ak, const float* aks,
                                                          const float* dx, const float* em1, const float* em2,
                                                          float e02, float etasca, const void* e01, const float* x0,
                                                          double* cbksca, double* csca, double* cscag,
                                                          double* cscag2, int* navg, void* f1, void* f2) {
  Context& c = ctx();
  if (!c.ready) {
    clear_error();
    set_error("GPU SCAT", "invalid state/argument");
    return 320;
  }
  FarfieldRegion r;
  if (scratch_region(r)) return 335;
  return scat(c, jpbc, ndir, ak, aks, dx, em1, em2, e02, etasca, e01, x0, cbksca, csca, cscag, cscag2, navg, f1, f2,
              &r);
}
