// Device reductions and vector kernels shared by the Krylov solvers.
//
// Dot products of complex64 vectors are accumulated in FP64 (PREC_MIXED, the
// reference's policy: ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:260-307)
// or in FP32 (PREC_FP32, the plain-FP32 comparison mode).  Several dots over
// the same vectors are computed in one pass; partial sums per workgroup go to
// a fixed-size buffer and a single-workgroup kernel adds them in a fixed
// order, so every result is deterministic and no partial visits the host.
#pragma once
#include "state.hpp"

namespace ddhip {

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 2048;
constexpr int RED_MAX_TERMS = 32;   // real-valued sums per launch
constexpr int DOT_MAX = 12;         // complex dots per launch
constexpr int SCAL_COUNT = 512;     // doubles in the device scalar block

// kind: 0 = sum conj(a) b, 1 = sum a b (unconjugated), 2 = sum |a|^2,
//       3 = sum |a - b|^2, 4 = sum |a - b - c|^2  (kinds 2-4 real)
enum { DOT_C = 0, DOT_U = 1, DOT_N2 = 2, DOT_D2 = 3, DOT_G2 = 4 };
struct DotTerm {
  const float2* a;
  const float2* b;
  const float2* c;
  int kind;
};
struct DotSet {
  DotTerm t[DOT_MAX];
  int n = 0;
  DotSet& add(const float2* a, const float2* b, int kind, const float2* c = nullptr) {
    t[n++] = DotTerm{a, b, c, kind};
    return *this;
  }
};

// Computes the dots of `s` over n complex elements; result i is written as
// (re, im) to c.scal[dst + 2 i], c.scal[dst + 2 i + 1].  Stream-ordered, no sync.
int dots(Context& c, const DotSet& s, size_t n, int dst, hipStream_t st);
// Adds the per-workgroup partials c.red[k][0..nblocks) (stride RED_MAX_BLOCKS)
// of nsums real sums into c.scal[dst + k], in the precision of c.precision.
int finalize_partials(Context& c, int nblocks, int nsums, int dst, hipStream_t st);

// Generic linear combination d = sum_i coef_i v_i over n elements.  A
// coefficient is either a device scalar (complex double at c.scal[idx]) or a
// literal; pointers may alias d.
constexpr int LC_MAX = 6;
struct LinTerm {
  const float2* v;
  int sidx;        // >= 0: complex scalar at scal[sidx], scal[sidx+1]; < 0: literal
  double re, im;   // literal (used when sidx < 0)
  int negate;      // multiply the device scalar by -1
};
struct LinComb {
  LinTerm t[LC_MAX];
  int n = 0;
  LinComb& lit(const float2* v, double re, double im = 0.0) {
    t[n++] = LinTerm{v, -1, re, im, 0};
    return *this;
  }
  LinComb& dev(const float2* v, int sidx, int neg = 0) {
    t[n++] = LinTerm{v, sidx, 0.0, 0.0, neg};
    return *this;
  }
};
int lincomb(Context& c, float2* d, const LinComb& lc, size_t n, hipStream_t st);

// copy scal[first .. first+count) to the pinned host mirror and wait
int fetch_scalars(Context& c, int first, int count, hipStream_t st);

// compact <-> full layout (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:123-134)
int gather_compact(Context& c, const float2* full, float2* compact, hipStream_t st);
int scatter_full(Context& c, const float2* compact, float2* full, hipStream_t st);

// SLICES backend (low-memory MATVEC)
int slices_setup(Context& c);
void slices_release(Context& c);
// act: the right-hand side's device-side active flag when the MATVEC is one
// of a batch (every kernel exits at once when it reads 0), else nullptr
int slices_matvec(Context& c, int cwhat, const float2* x, float2* y, const float2* adia,
                  const float2* aoff, int layout, hipStream_t st, const int* act = nullptr);
// batched MATVEC on either backend: FFT3D runs the B right-hand sides in each
// launch (matvec_launch), SLICES runs them one after another through the
// context's slice buffers, each honouring its active flag
int batch_matvec(Context& c, int cwhat, const float2* x, float2* y, const float2* adia, const float2* aoff,
                 const MvBatch& bt, hipStream_t st);

int finish_prepare(Context& c, const int16_t* iocc, int code_alloc, int code_occ);
// frees the batched-solver buffers (bsolve.hip)
void batch_release(Context& c);
// fills the batched-solver vectors, when allocated, with NaN bit patterns (test hook)
int batch_poison(Context& c);
int ensure_stream(Context& c);

// MATVEC on the compact solver vectors with the context's operators
int run_mv(Context& c, int cwhat, const float2* x, float2* y, int& nc);

}  // namespace ddhip
