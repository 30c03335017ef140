// Reference-compatible C ABI: prepare / apply / release / last_error and the
// ddscat_hip_* extensions.  Argument checks, return codes and lifetime follow
// ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:913-916, 1005-1029.
#include "solver_common.hpp"
#include "../../include/ddscat_matvec_hip.h"
#include <cmath>
#include <vector>

using namespace ddhip;

#ifndef DDSCAT_HIP_DEFAULT_BACKEND
#define DDSCAT_HIP_DEFAULT_BACKEND 0
#endif

namespace {
struct DefaultBackend {
  DefaultBackend() { ctx().backend = DDSCAT_HIP_DEFAULT_BACKEND; }
} g_default_backend;

bool grid_args_ok(int nx, int ny, int nz, int nat, int nat0, int ipbc) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nat <= 0 || nat0 <= 0 || nat0 > nat) return false;
  if (ipbc != 0 && ipbc != 1) return false;
  if ((long long)nx * ny * nz != (long long)nat) return false;
  if ((long long)3 * nat > 0x7fffffffLL) return false;
  return true;
}
}  // namespace

extern "C" DDSCAT_HIP_API int ddscat_cuda_release(void) {
  release_context(ctx());
  clear_error();
  return 0;
}

extern "C" DDSCAT_HIP_API const char* ddscat_cuda_last_error(void) { return error_text(); }

extern "C" DDSCAT_HIP_API int ddscat_cuda_prepare_green_f32(int nx, int ny, int nz, int ipbc, int nat, int nat0,
                                                            int idipint, float gamma, float pyd, float pzd, float akx,
                                                            float aky, float akz, float dx1, float dx2, float dx3,
                                                            const int16_t* iocc) {
  clear_error();
  if (!grid_args_ok(nx, ny, nz, nat, nat0, ipbc) || (idipint != 0 && idipint != 1) || !iocc) {
    set_error("prepare_green", "invalid argument");
    return 201;
  }
  if (ipbc && gamma <= 0.f) {
    set_error("prepare_green", "GAMMA must be >0 for PBC");
    return 202;
  }
  if (ipbc && !(pyd > 0.f) && !(pzd > 0.f)) {
    // the replica window divides by the period of every periodic axis
    set_error("prepare_green", "IPBC=1 needs PYD>0 or PZD>0");
    return 201;
  }
  GreenBuild gb;
  gb.ipbc = ipbc; gb.idipint = idipint;
  gb.gamma = gamma; gb.pyd = pyd; gb.pzd = pzd;
  gb.ak[0] = akx; gb.ak[1] = aky; gb.ak[2] = akz;
  gb.dx[0] = dx1; gb.dx[1] = dx2; gb.dx[2] = dx3;
  Context& c = ctx();
  release_context(c);
  if (int rc = ensure_stream(c)) return rc;
  c.nx = nx; c.ny = ny; c.nz = nz; c.ipbc = ipbc; c.nat = nat; c.nat0 = nat0;
  if (ipbc) {
    const double kp2 = pyd <= 0.f ? (double)akx * akx + (double)aky * aky
                     : pzd <= 0.f ? (double)akx * akx + (double)akz * akz : (double)akx * akx;
    if (!(kp2 > 0.0)) {
      set_error("CUDA Green", "Green kernel status=13 (11=CISI,12=IDIPINT,13=KPERP)");
      release_context(c);
      return 234;
    }
    // bound the replica sums so that one prepare cannot occupy the GPU for long
    const double work = green_pbc_work(c, gb), limit = 2.0e11;
    if (work > limit) {
      char msg[200];
      std::snprintf(msg, sizeof(msg), "PBC replica sums need ~%.3g interactions (limit %.3g); increase GAMMA", work,
                    limit);
      set_error("prepare_green", msg);
      release_context(c);
      return 235;
    }
  }
  c.green_count = 6 * green_points(nx, ny, nz, ipbc);
  if (hipMalloc((void**)&c.green, c.green_count * sizeof(float2)) != hipSuccess) {
    set_error("prepare_green", "malloc resident Green failed");
    release_context(c);
    return 204;
  }
  int grc = green_build(c, gb);
  if (grc) { release_context(c); return 220 + grc; }
  int rc = finish_prepare(c, iocc, 204, 208);
  if (rc) { release_context(c); return rc; }
  c.ready = true;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_cuda_prepare_f32(int nx, int ny, int nz, int ipbc, int nat, int nat0,
                                                      const void* green, const int16_t* iocc) {
  clear_error();
  if (!grid_args_ok(nx, ny, nz, nat, nat0, ipbc) || !green || !iocc) {
    set_error("prepare", "invalid argument");
    return 1;
  }
  Context& c = ctx();
  release_context(c);
  if (ensure_stream(c)) return 2;
  c.nx = nx; c.ny = ny; c.nz = nz; c.ipbc = ipbc; c.nat = nat; c.nat0 = nat0;
  // CXZC: [6][NZ+1][NY+1][NX+1] for IPBC=0, [6][2NZ][2NY][2NX] for IPBC=1 (cu:1020)
  c.green_count = 6 * green_points(nx, ny, nz, ipbc);
  if (hipMalloc((void**)&c.green, c.green_count * sizeof(float2)) != hipSuccess) {
    set_error("prepare", "malloc green failed");
    release_context(c);
    return 2;
  }
  // CXZC arrives component-major; the resident copy is point-interleaved.  It
  // is uploaded before finish_prepare so that the autotuner times the pass-3
  // candidates on the real tensor.
  float2* tmp = nullptr;
  if (hipMalloc((void**)&tmp, c.green_count * sizeof(float2)) != hipSuccess) {
    set_error("prepare", "malloc green staging failed");
    release_context(c);
    return 2;
  }
  const bool ok = hipMemcpy(tmp, green, c.green_count * sizeof(float2), hipMemcpyHostToDevice) == hipSuccess &&
                  green_interleave(c, tmp, c.green, c.stream) == 0 && hipStreamSynchronize(c.stream) == hipSuccess;
  hipFree(tmp);
  if (!ok) {
    set_error("prepare", "green H2D failed");
    release_context(c);
    return 4;
  }
  int rc = finish_prepare(c, iocc, 2, 6);
  if (rc) { release_context(c); return rc == 6 ? 6 : (rc >= 200 ? rc : 2); }
  c.ready = true;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_cuda_apply_f32(int cwhat, const void* x, void* y, const void* adia,
                                                    const void* aoff) {
  Context& c = ctx();
  if (!c.ready) { set_error("apply", "not prepared"); return 10; }
  if ((cwhat != 'N' && cwhat != 'C') || !x || !y || !adia || !aoff) { set_error("apply", "invalid argument"); return 11; }
  clear_error();
  const size_t b = (size_t)3 * c.nat * sizeof(float2);
  DDH_TRY(hipMemcpyAsync(c.xfull, x, b, hipMemcpyHostToDevice, c.stream), "x H2D", 12);
  DDH_TRY(hipMemcpyAsync(c.adia, adia, b, hipMemcpyHostToDevice, c.stream), "adia H2D", 13);
  DDH_TRY(hipMemcpyAsync(c.aoff, aoff, b, hipMemcpyHostToDevice, c.stream), "aoff H2D", 14);
  int rc = (c.backend == BACKEND_SLICES) ? slices_matvec(c, cwhat, c.xfull, c.yfull, c.adia, c.aoff, 0, c.stream)
                                         : matvec_launch(c, cwhat, c.xfull, c.yfull, c.adia, c.aoff, 0, c.stream);
  if (rc) return rc;
  DDH_TRY(hipMemcpyAsync(y, c.yfull, b, hipMemcpyDeviceToHost, c.stream), "y D2H", 15);
  DDH_TRY(hipStreamSynchronize(c.stream), "y D2H", 15);
  return 0;
}

// ------------------------------------------------------------- extensions --
extern "C" DDSCAT_HIP_API int ddscat_hip_set_backend(int backend) {
  if (backend != BACKEND_FFT3D && backend != BACKEND_SLICES) { set_error("set_backend", "unknown backend"); return 1; }
  ctx().backend = backend;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_set_precision(int precision) {
  if (precision != PREC_MIXED && precision != PREC_FP32) { set_error("set_precision", "unknown precision"); return 1; }
  ctx().precision = precision;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_set_reliable_mode(int mode) {
  if (mode != RELIABLE_REPLACE && mode != RELIABLE_RESTART) {
    set_error("set_reliable_mode", "expected DDSCAT_HIP_RELIABLE_REPLACE or DDSCAT_HIP_RELIABLE_RESTART");
    return 1;
  }
  ctx().reliable_mode = mode;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_set_fft_kernels(int generic) {
  if (generic != 0 && generic != 1) { set_error("set_fft_kernels", "expected 0 or 1"); return 1; }
  ctx().fft_generic = generic;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_set_autotune(int on) {
  if (on != 0 && on != 1) { set_error("set_autotune", "expected 0 or 1"); return 1; }
  ctx().autotune = on;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_set_pass_kernel(int pass, int kind) {
  Context& c = ctx();
  if (!c.ready || c.backend != BACKEND_FFT3D) { set_error("set_pass_kernel", "not prepared (FFT3D)"); return 10; }
  if (pass < 1 || pass > 5) { set_error("set_pass_kernel", "pass must be 1..5"); return 11; }
  const int lens[6] = {0, c.mv.gx, c.mv.gy, c.mv.gz, c.mv.gy, c.mv.gx};
  if (kind < -1 || kind >= reg_candidates(lens[pass], pass)) { set_error("set_pass_kernel", "no such kernel"); return 11; }
  c.mv.kind[pass] = kind;
  c.mv.tuned = 0;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_apply_device(int cwhat, const void* x, void* y, const void* adia,
                                                      const void* aoff, int layout, void* stream) {
  Context& c = ctx();
  if (!c.ready) { set_error("apply_device", "not prepared"); return 10; }
  if ((cwhat != 'N' && cwhat != 'C') || !x || !y || !adia || !aoff || (layout != 0 && layout != 1)) {
    set_error("apply_device", "invalid argument");
    return 11;
  }
  hipStream_t st = stream ? (hipStream_t)stream : c.stream;
  return (c.backend == BACKEND_SLICES)
             ? slices_matvec(c, cwhat, (const float2*)x, (float2*)y, (const float2*)adia, (const float2*)aoff, layout, st)
             : matvec_launch(c, cwhat, (const float2*)x, (float2*)y, (const float2*)adia, (const float2*)aoff, layout, st);
}

extern "C" DDSCAT_HIP_API int ddscat_hip_get_green(void* out) {
  Context& c = ctx();
  if (!c.ready || !out) { set_error("get_green", "not prepared"); return 10; }
  float2* tmp = nullptr;
  DDH_TRY(hipMalloc((void**)&tmp, c.green_count * sizeof(float2)), "get_green staging", 12);
  const bool ok = green_deinterleave(c, c.green, tmp, c.stream) == 0 && hipStreamSynchronize(c.stream) == hipSuccess &&
                  hipMemcpy(out, tmp, c.green_count * sizeof(float2), hipMemcpyDeviceToHost) == hipSuccess;
  hipFree(tmp);
  if (!ok) { set_error("get_green", "D2H failed"); return 12; }
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_info(long long* out, int n) {
  Context& c = ctx();
  if (!out) return 1;
  const long long v[] = {c.nx, c.ny, c.nz, c.nat, c.nat0, (long long)c.green_count, (long long)c.matvec_count,
                         c.backend, c.precision, c.mv.gx, c.mv.gy, c.mv.gz, c.mv.rx_fwd, c.mv.cy, c.mv.cz,
                         c.mv.rx_inv, (long long)(c.green_build_ms * 1000.0), c.mv.kind[1], c.mv.kind[2],
                         c.mv.kind[3], c.mv.kind[4], c.mv.kind[5], c.mv.tuned, c.restarts};
  const int m = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < n && i < m; ++i) out[i] = v[i];
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_solve_log(int rhs, int* restarts, double* resid, int n, double* checks,
                                                   int nchecks, int* nchecks_out) {
  const SolveLog& L = ctx().log;
  if (rhs < 0 || rhs >= (int)L.resid.size()) return -1;
  const std::vector<double>& h = L.resid[rhs];
  const std::vector<double>& k = L.checks[rhs];
  if (restarts) *restarts = L.restarts[rhs];
  if (resid)
    for (int i = 0; i < n && i < (int)h.size(); ++i) resid[i] = h[i];
  const int nk = (int)(k.size() / 4);
  if (checks)
    for (int i = 0; i < 4 * nchecks && i < 4 * nk; ++i) checks[i] = k[i];
  if (nchecks_out) *nchecks_out = nk;
  return (int)h.size();
}

extern "C" DDSCAT_HIP_API int ddscat_hip_force_restart(void) {
  ctx().forced_gap = 1.0;
  return 0;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_poison_workspace(void) {
  Context& c = ctx();
  if (!c.ready) { set_error("poison_workspace", "not prepared"); return 10; }
  if (int rc = ensure_solver_workspace(c, 14)) return rc;
  if (ensure_stream(c)) return 12;
  // all-ones bytes: every float of the workspace reads as NaN
  DDH_TRY(hipMemsetAsync(c.wrk, 0xFF, (size_t)c.wrk_vectors * 3 * c.nat0 * sizeof(float2), c.stream),
          "poison_workspace", 12);
  if (int rc = batch_poison(c)) return rc;
  DDH_TRY(hipStreamSynchronize(c.stream), "poison_workspace", 12);
  return 0;
}

extern "C" DDSCAT_HIP_API void* ddscat_hip_stream(void) {
  Context& c = ctx();
  if (ensure_stream(c)) return nullptr;
  return (void*)c.stream;
}

extern "C" DDSCAT_HIP_API int ddscat_hip_sync(void) {
  Context& c = ctx();
  if (!c.stream) return 0;
  DDH_TRY(hipStreamSynchronize(c.stream), "sync", 1);
  return 0;
}

namespace {
// out = e_inc - y (E_inc + E_sca with y = -E_sca), or -y without e_inc
__global__ void k_nearfield_finish(const float2* __restrict__ y, const float2* __restrict__ einc,
                                   float2* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = y[i];
  out[i] = einc ? make_float2(einc[i].x - v.x, einc[i].y - v.y) : make_float2(-v.x, -v.y);
}
}  // namespace

// Near field on the whole lattice (DDSCAT NRFLD): with the local operator
// switched off, the MATVEC returns y = -E_sca where E_sca,j = sum_k c_jk P_k is
// the field the dipoles radiate at site j (c = -A_jk, cu:625-626), and pass 5
// without an IOCC mask evaluates it on empty sites too.
extern "C" DDSCAT_HIP_API int ddscat_hip_nearfield(const void* p_compact, const void* einc_full, void* out_full) {
  Context& c = ctx();
  if (!c.ready) { set_error("nearfield", "not prepared"); return 10; }
  if (!p_compact || !out_full) { set_error("nearfield", "invalid argument"); return 11; }
  clear_error();
  hipStream_t st = c.stream;
  const size_t n = (size_t)3 * c.nat;
  DDH_TRY(hipMemsetAsync(c.adia, 0, n * sizeof(float2), st), "nearfield", 12);
  DDH_TRY(hipMemsetAsync(c.aoff, 0, n * sizeof(float2), st), "nearfield", 12);
  if (scatter_full(c, (const float2*)p_compact, c.xfull, st)) return 12;
  int16_t* keep = c.iocc;
  c.iocc = nullptr;  // full-layout finish without the occupancy mask
  const int rc = (c.backend == BACKEND_SLICES) ? slices_matvec(c, 'N', c.xfull, c.yfull, c.adia, c.aoff, 0, st)
                                               : matvec_launch(c, 'N', c.xfull, c.yfull, c.adia, c.aoff, 0, st);
  c.iocc = keep;
  if (rc) return rc;
  hipLaunchKernelGGL(k_nearfield_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c.yfull,
                     (const float2*)einc_full, (float2*)out_full, n);
  DDH_LAUNCH_OK("near-field finish", 13);
  DDH_TRY(hipStreamSynchronize(st), "nearfield", 13);
  return 0;
}

// Runs `reps` MATVECs on device pointers with HIP events between the five
// FFT3D passes; ms_out[0..4] = mean per-pass time, ms_out[5] = mean MATVEC
// time (first to last event).  For the SLICES backend only ms_out[5] is set.
extern "C" DDSCAT_HIP_API int ddscat_hip_time_matvec(int cwhat, const void* x, void* y, const void* adia,
                                                     const void* aoff, int layout, int reps, double* ms_out) {
  Context& c = ctx();
  if (!c.ready) { set_error("time_matvec", "not prepared"); return 10; }
  if (!x || !y || !adia || !aoff || !ms_out || reps <= 0) { set_error("time_matvec", "invalid argument"); return 11; }
  hipEvent_t ev[6];
  for (int i = 0; i < 6; ++i) DDH_TRY(hipEventCreate(&ev[i]), "event create", 12);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int rc = 0;
  for (int r = 0; r < reps && !rc; ++r) {
    if (c.backend == BACKEND_SLICES) {
      hipEventRecord(ev[0], c.stream);
      rc = slices_matvec(c, cwhat, (const float2*)x, (float2*)y, (const float2*)adia, (const float2*)aoff, layout,
                         c.stream);
      hipEventRecord(ev[5], c.stream);
    } else {
      rc = matvec_launch(c, cwhat, (const float2*)x, (float2*)y, (const float2*)adia, (const float2*)aoff, layout,
                         c.stream, ev);
    }
    if (rc) break;
    if (hipEventSynchronize(ev[5]) != hipSuccess) { set_error("time_matvec", "event sync"); rc = 12; break; }
    float t = 0.f;
    if (c.backend != BACKEND_SLICES) {
      for (int i = 0; i < 5; ++i) {
        hipEventElapsedTime(&t, ev[i], ev[i + 1]);
        acc[i] += t;
      }
    }
    hipEventElapsedTime(&t, ev[0], ev[5]);
    acc[5] += t;
  }
  for (int i = 0; i < 6; ++i) {
    ms_out[i] = acc[i] / reps;
    hipEventDestroy(ev[i]);
  }
  return rc;
}
