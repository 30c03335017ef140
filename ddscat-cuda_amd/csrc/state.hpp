// Device-resident state of one prepared target (Green tensor, FFT work
// buffers, occupancy maps, solver workspace).  The reference C ABI is
// stateful and single-instance (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:20-56);
// the library keeps one default Context for it.
#pragma once
#include "common.hpp"
#include <vector>

namespace ddhip {

enum Backend { BACKEND_FFT3D = 0, BACKEND_SLICES = 1 };
enum Precision { PREC_MIXED = 0, PREC_FP32 = 1 };

struct MatvecPlan {
  int nx = 0, ny = 0, nz = 0, nat = 0, nat0 = 0;
  int gx = 0, gy = 0, gz = 0;
  size_t ngrid = 0;
  FftPlan px{}, py{}, pz{};
  float2 *twx = nullptr, *twy = nullptr, *twz = nullptr;  // device twiddle tables
  // tile parameters
  int rx_fwd = 1, rx_inv = 1, cy = 1, cz = 1;
  int ty = 256, tz = 256;  // workgroup size of the LDS-staged y passes and z pass
  size_t p1 = 0, p2 = 0;  // plane strides (elements) of a1 [3][nz][ny][gx] and a2 [3][nz][plane p2]
  // FFT kernel per pass 1..5: index into the register shapes of that pass's
  // axis length, or -1 for the LDS-staged kernel (chosen by the autotuner)
  int kind[6] = {-1, -1, -1, -1, -1, -1};
  int tuned = 0;
};

// Convergence record of the last solve (batched: one entry per RHS;
// single-RHS solvers: RHS 0), read by ddscat_hip_solve_log.
// What a failed reliable-residual check (gap >= 1e-3 or false convergence,
// cu:543-556) does to QMRCCG: REPLACE takes the true residual as the
// recursive one and keeps the Lanczos recurrence (default); RESTART restarts
// the recurrence from the true residual, as the reference does.
enum ReliableMode { RELIABLE_REPLACE = 0, RELIABLE_RESTART = 1 };

struct SolveLog {
  enum { RESTART = 1, REPLACE = 2 };  // action column of a check row
  std::vector<std::vector<double>> resid;   // recursive residual after each iteration
  std::vector<std::vector<double>> checks;  // (iteration, true residual, gap, action) per reliable check
  std::vector<int> restarts;
  void reset(int n) {
    resid.assign(n, {});
    checks.assign(n, {});
    restarts.assign(n, 0);
  }
  void iter(int r, double v) {
    if (r >= 0 && r < (int)resid.size()) resid[r].push_back(v);
  }
  void check(int r, double true_res, double gap, int action) {
    if (r < 0 || r >= (int)resid.size()) return;
    checks[r].insert(checks[r].end(), {(double)resid[r].size(), true_res, gap, (double)action});
    if (action == RESTART) ++restarts[r];
  }
};

// Prepared targets the library holds at once: slot 0 is the reference ABI's
// single target; slots 1..MAX_TARGETS serve multi-target batched solves
// (ddscat_hip_select_target, ddscat_hip_solve_targets).
constexpr int MAX_TARGETS = 64;

struct Context {
  int slot = 0;  // index of this context among the held targets
  bool ready = false;
  int nx = 0, ny = 0, nz = 0, ipbc = 0, nat = 0, nat0 = 0;
  int backend = BACKEND_FFT3D;
  int precision = PREC_MIXED;
  MatvecPlan mv;
  hipStream_t stream = nullptr;
  // Spectral Green tensor, point-interleaved complex64: the trimmed octant
  // [nz+1][ny+1][nx+1][6] for IPBC=0, the whole grid [2nz][2ny][2nx][6] for IPBC=1
  float2* green = nullptr;
  size_t green_count = 0;
  // FFT3D work: a1 [3][nz][ny][gx], a2 [3][nz][gy][gx]
  float2 *a1 = nullptr, *a2 = nullptr;
  // occupancy
  int16_t* iocc = nullptr;     // [nat]
  int* occ = nullptr;          // [nat0] occupied site j, increasing
  int* rowptr = nullptr;       // [ny*nz+1] compact range of each (iy,iz) row
  int* rowlo = nullptr;        // [ny*nz] first occupied ix of each row
  int* qmap = nullptr;         // [nat] full site -> compact index (-1 void); built when rows are not intervals
  int rows_contiguous = 0;     // every row's occupied sites form one interval
  int fft_generic = 0;         // 1: force the LDS-staged generic FFT kernels
  int autotune = 1;            // 1: time the FFT kernel candidates per pass at prepare
  // full-layout operator and vector staging
  float2 *adia = nullptr, *aoff = nullptr;      // [3*nat]
  float2 *adia_c = nullptr, *aoff_c = nullptr;  // [3*nat0]
  float2 *xfull = nullptr, *yfull = nullptr;    // [3*nat]
  // solver storage (compact, 3*nat0 each)
  float2 *sx = nullptr, *sy = nullptr, *sb = nullptr, *wrk = nullptr;
  int wrk_vectors = 0;
  double* red = nullptr;        // reduction partials
  double* scal = nullptr;       // device scalar block (doubles)
  double* hscal = nullptr;      // pinned host mirror
  bool solution_valid = false;
  unsigned long long matvec_count = 0;
  int restarts = 0;               // reliable-residual restarts of the last single-RHS solve
  SolveLog log;                   // convergence record of the last solve
  int reliable_mode = RELIABLE_REPLACE;  // QMRCCG's response to a failed reliable check
  double forced_gap = 0.0;        // added once to the next reliable check's gap (test hook)
  double green_build_ms = 0.0;
};

Context& ctx();                 // the selected target's context
Context& ctx_at(int slot);      // target slot 0..MAX_TARGETS
int selected_target();
void select_target(int slot);
void release_context(Context& c);
int  ensure_solver_workspace(Context& c, int nvec);
void release_solver_workspace(Context& c);

// Several right-hand sides through one MATVEC launch sequence: RHS r uses
// x + r*xs, y + r*xs, adia + r*os, aoff + r*os (os = 0: shared operators) and
// its own FFT work a1 + r*a1s, a2 + r*a2s; RHS with active[r] == 0 are
// skipped (every workgroup of theirs exits at once).  Kernel-side view:
struct BatchArgs {
  size_t xs = 0, os = 0, a1s = 0, a2s = 0;
  const int* active = nullptr;
  // SLICES z-forward input read straight from the MATVEC's x: element j of
  // component c is gsrc[c gcs + gmap[j]] (compact; 0 where gmap[j] < 0) or
  // gsrc[c gcs + j] (gmap == nullptr), conjugated when gconj
  const float2* gsrc = nullptr;
  const int* gmap = nullptr;
  size_t gcs = 0;
  int gconj = 0;
};
struct MvBatch {
  int nrhs = 1;
  BatchArgs k;
  float2* a1 = nullptr;  // nullptr: the context's single-RHS buffers
  float2* a2 = nullptr;
};

// MATVEC: y = A x.  layout 0: full [3][nat] vectors (iocc mask applied);
// layout 1: compact [3][nat0] vectors (adia_c/aoff_c).  conj_mode: 'N' or 'C'.
// ev (optional): 6 events recorded before passes 1..5 and after pass 5.
int matvec_launch(Context& c, int cwhat, const float2* x, float2* y, const float2* adia,
                  const float2* aoff, int layout, hipStream_t st, hipEvent_t* ev = nullptr,
                  const MvBatch* bt = nullptr);
int matvec_setup(Context& c);
void matvec_release(Context& c);
// FFT work elements per RHS: a1 [3][nz][ny][gx], a2 [3][nz][gy][gx]
size_t matvec_a1_count(const Context& c);
size_t matvec_a2_count(const Context& c);
// register-FFT shapes available for pass `which` (1..5) on a padded axis length (0: none)
int reg_candidates(int L, int which);
// register-FFT pass `which` (1..5) with shape `cand` of its axis length; 1 = launched
int reg_pass(Context& c, int which, int cand, int cj, int compact, const float2* x, float2* y, const float2* adia,
             const float2* aoff, hipStream_t st, const MvBatch& bt);
// SLICES y-FFT + Green + inverse y-FFT of nb slice planes from kz = k0 with
// the register kernels when 2NY has a register shape; 1 = launched.  act (all
// SLICES launchers): the right-hand side's active flag in a batch, or nullptr.
int slices_y_green_reg(Context& c, float2* sl, int k0, int nb, hipStream_t st, const int* act);
// SLICES x-FFT (dir +1: zwork rows -> slice) or inverse x-FFT (dir -1:
// slice -> zwork rows) of nb planes from kz = k0 with the register kernels
// when 2NX has a register shape; 1 = launched
int slices_x_reg(Context& c, int dir, float2* zw, float2* sl, int k0, int nb, hipStream_t st, const int* act);
// SLICES z-FFT (dir +1: [3][NZ][NY][NX] -> zwork) or inverse z-FFT (dir -1:
// zwork -> [3][NZ][NY][NX]) with the register column kernels when 2NZ has a
// register shape; 1 = launched.  slices_z_reg_length: whether L has one.
// gather (dir +1 only): `in` is the MATVEC's x, compact through the site map
// when `compact`, conjugated when `cj`, read inside the z-FFT's loads.
int slices_z_reg(Context& c, int dir, const float2* in, float2* out, hipStream_t st, const int* act,
                 bool gather = false, int compact = 0, int cj = 0);
int slices_z_reg_length(int L);
// times every FFT kernel candidate of each pass and stores the fastest in c.mv.kind
int matvec_autotune(Context& c);

// Green tensor build on device, result in c.green (octant for IPBC=0, whole
// grid for IPBC=1).  gamma, pyd, pzd and the wave-vector components matter
// only for IPBC=1 (replica sums, cu:634-643).
struct GreenBuild {
  int ipbc = 0, idipint = 0;
  double gamma = 0.0, pyd = 0.0, pzd = 0.0;
  double ak[3] = {0.0, 0.0, 0.0};
  double dx[3] = {1.0, 1.0, 1.0};
};
int green_build(Context& c, const GreenBuild& b);
// number of direct interactions the IPBC=1 replica sums would evaluate (upper
// bound over the lag grid); 0 for IPBC=0
double green_pbc_work(const Context& c, const GreenBuild& b);

// The resident spectral Green tensor c.green is stored point-interleaved:
// the six components (xx, xy, xz, yy, yz, zz) of octant point
// p = ir + (NX+1)(jr + (NY+1) kr) sit at green[6p .. 6p+5], so a MATVEC reads
// them as three 16-byte loads from one 48-byte run.  The reference layout
// ([6][NZ+1][NY+1][NX+1], CXZC) is converted at the ABI edges.
struct Green6 {
  float2 xx, xy, xz, yy, yz, zz;
};
__device__ __forceinline__ Green6 green_at(const float2* __restrict__ g, size_t p) {
  const float4* q = reinterpret_cast<const float4*>(g) + 3 * p;
  const float4 a = q[0], b = q[1], c = q[2];
  return Green6{make_float2(a.x, a.y), make_float2(a.z, a.w), make_float2(b.x, b.y),
                make_float2(b.z, b.w), make_float2(c.x, c.y), make_float2(c.z, c.w)};
}
// Where the Green entry of spectral point (kx, ky, kz) lives, and its signs.
// IPBC=0 stores the octant and rebuilds the rest by mirror symmetry
// (cu:109-115): index min(k, 2N - k) per axis, the off-diagonal components
// flip sign with the mirrored axes' parities.  IPBC=1 stores the whole grid.
// Both use the same instructions:
//   r = min(k, m - k),  sign = (k <= t) ? +1 : -1,  point = rx + ox (ry + oy rz)
// octant: m = 2N, t = N, ox = NX+1, oy = NY+1; full grid: m = t = 2^30
// (no mirror, all signs +1), ox = 2NX, oy = 2NY.
struct GreenGeo {
  int ox, oy;
  int mx, my, mz;
  int tx, ty, tz;
};
inline GreenGeo green_geo(const Context& c) {
  constexpr int NOMIRROR = 1 << 30;
  if (c.ipbc) return GreenGeo{2 * c.nx, 2 * c.ny, NOMIRROR, NOMIRROR, NOMIRROR, NOMIRROR, NOMIRROR, NOMIRROR};
  return GreenGeo{c.nx + 1, c.ny + 1, 2 * c.nx, 2 * c.ny, 2 * c.nz, c.nx, c.ny, c.nz};
}
// number of spectral Green points stored: octant or whole padded grid
inline size_t green_points(int nx, int ny, int nz, int ipbc) {
  return ipbc ? (size_t)(2 * nx) * (2 * ny) * (2 * nz) : (size_t)(nx + 1) * (ny + 1) * (nz + 1);
}

// component-major [6][npts] <-> point-interleaved [npts][6], device to device
int green_interleave(Context& c, const float2* comp_major, float2* interleaved, hipStream_t st);
int green_deinterleave(Context& c, const float2* interleaved, float2* comp_major, hipStream_t st);

}  // namespace ddhip
