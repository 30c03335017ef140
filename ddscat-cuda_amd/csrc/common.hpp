// Shared definitions for the MI355X (gfx950) DDSCAT hot-path library.
// Complex values are interleaved float2 (complex64) or double2 (complex128).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <cstdio>
#include <string>

namespace ddhip {

// ---------------------------------------------------------------- complex --
__host__ __device__ __forceinline__ float2 cmk(float r, float i) { return make_float2(r, i); }
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by s*i (s = +1 or -1)
template <int S>
__host__ __device__ __forceinline__ float2 cmul_si(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

__host__ __device__ __forceinline__ double2 zmk(double r, double i) { return make_double2(r, i); }
__host__ __device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__host__ __device__ __forceinline__ double zabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ __forceinline__ bool zzero(double2 a) { return a.x == 0.0 && a.y == 0.0; }
// float vector times double scalar, FP64 product rounded to FP32
__host__ __device__ __forceinline__ float2 zmulf(double2 a, float2 b) {
  double x = a.x * (double)b.x - a.y * (double)b.y;
  double y = a.x * (double)b.y + a.y * (double)b.x;
  return make_float2((float)x, (float)y);
}

// ------------------------------------------------------------------ errors --
void set_error(const char* where, const char* msg);
void set_hip_error(const char* where, hipError_t e);
const char* error_text();
void clear_error();

#define DDH_TRY(expr, where, code)                         \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) { ::ddhip::set_hip_error(where, _e); return (code); } \
  } while (0)

#define DDH_LAUNCH_OK(where, code)                         \
  do {                                                     \
    hipError_t _e = hipGetLastError();                     \
    if (_e != hipSuccess) { ::ddhip::set_hip_error(where, _e); return (code); } \
  } while (0)

// -------------------------------------------------------------------- FFT --
constexpr int FFT_MAX_STAGES = 16;
struct FftPlan {
  int L;                      // transform length
  int nst;                    // number of radix stages
  int rad[FFT_MAX_STAGES];    // radices (2,3,4,5,7,8), applied in order
};

bool fft_make_plan(int L, FftPlan& p);
// e^{+2 pi i m / L}, m = 0..L-1, computed in FP64 and rounded to FP32
void fft_twiddles_host(int L, float2* out);

}  // namespace ddhip
