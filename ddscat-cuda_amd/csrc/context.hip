// Context lifetime, error text and the shared prepare step.
// Lifetime semantics follow the reference plugin: every prepare* starts by
// releasing the previous state, the Green tensor and FFT work stay resident
// until release or the next prepare
// (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:66-79, 1005-1026).
#include "state.hpp"
#include "solver_common.hpp"
#include <vector>
#include <cstring>
#include <mutex>

namespace ddhip {

// error text: one per process, as the reference's (several targets may
// solve at once from worker threads, so writes are serialised)
static std::string g_last_error;
static std::mutex g_error_mutex;

void set_error(const char* where, const char* msg) {
  char b[1024];
  std::snprintf(b, sizeof(b), "%s: %s", where, msg);
  std::lock_guard<std::mutex> lock(g_error_mutex);
  g_last_error = b;
}
void set_hip_error(const char* where, hipError_t e) { set_error(where, hipGetErrorString(e)); }
const char* error_text() { return g_last_error.c_str(); }
void clear_error() {
  std::lock_guard<std::mutex> lock(g_error_mutex);
  g_last_error.clear();
}

static Context g_ctx[MAX_TARGETS + 1];
static int g_selected = 0;
Context& ctx_at(int slot) {
  Context& c = g_ctx[slot];
  c.slot = slot;
  return c;
}
Context& ctx() { return ctx_at(g_selected); }
int selected_target() { return g_selected; }
void select_target(int slot) { g_selected = slot; }

// hipFree only what was allocated: a context that never reached the device
// releases without touching the HIP runtime
template <typename T>
static void dfree(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

void release_solver_workspace(Context& c) {
  dfree(c.wrk);
  c.wrk_vectors = 0;
}

int ensure_solver_workspace(Context& c, int nvec) {
  if (c.wrk && c.wrk_vectors >= nvec) return 0;
  release_solver_workspace(c);
  const size_t bytes = (size_t)nvec * 3 * c.nat0 * sizeof(float2);
  DDH_TRY(hipMalloc((void**)&c.wrk, bytes), "solver Krylov workspace", 29);
  c.wrk_vectors = nvec;
  return 0;
}

void release_context(Context& c) {
  if (c.stream) hipStreamSynchronize(c.stream);
  matvec_release(c);
  slices_release(c);
  batch_release(c);
  dfree(c.green);
  dfree(c.iocc); dfree(c.occ); dfree(c.rowptr); dfree(c.rowlo); dfree(c.qmap);
  dfree(c.adia); dfree(c.aoff); dfree(c.adia_c); dfree(c.aoff_c);
  dfree(c.xfull); dfree(c.yfull);
  dfree(c.sx); dfree(c.sy); dfree(c.sb);
  release_solver_workspace(c);
  dfree(c.red); dfree(c.scal);
  if (c.hscal) hipHostFree(c.hscal);
  hipStream_t keep = c.stream;
  int backend = c.backend, prec = c.precision, gen = c.fft_generic, tune = c.autotune, rmode = c.reliable_mode;
  const int slot = c.slot;
  c = Context{};
  c.slot = slot;
  c.stream = keep;
  c.reliable_mode = rmode;
  c.backend = backend;
  c.precision = prec;
  c.fft_generic = gen;
  c.autotune = tune;
}

// Common tail of prepare_green / prepare: dimensions already set, Green
// resident.  Allocates MATVEC and solver buffers, uploads IOCC and builds
// the occupied-site list and per-row compact ranges.
int finish_prepare(Context& c, const int16_t* iocc, int code_alloc, int code_occ) {
  int rc = (c.backend == BACKEND_SLICES) ? slices_setup(c) : matvec_setup(c);
  if (rc) return rc;
  const size_t n = (size_t)3 * c.nat, cn = (size_t)3 * c.nat0;
  DDH_TRY(hipMalloc((void**)&c.iocc, (size_t)c.nat * sizeof(int16_t)), "malloc iocc", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.occ, (size_t)c.nat0 * sizeof(int)), "malloc occupied map", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.rowptr, ((size_t)c.ny * c.nz + 1) * sizeof(int)), "malloc row map", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.adia, n * sizeof(float2)), "malloc adia", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.aoff, n * sizeof(float2)), "malloc aoff", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.adia_c, cn * sizeof(float2)), "malloc compact adia", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.aoff_c, cn * sizeof(float2)), "malloc compact aoff", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.xfull, n * sizeof(float2)), "malloc full x", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.yfull, n * sizeof(float2)), "malloc full y", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.sx, cn * sizeof(float2)), "malloc compact x", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.sy, cn * sizeof(float2)), "malloc compact y", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.sb, cn * sizeof(float2)), "malloc compact b", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.red, (size_t)RED_MAX_BLOCKS * RED_MAX_TERMS * sizeof(double)), "malloc reduction partials", code_alloc);
  DDH_TRY(hipMalloc((void**)&c.scal, (size_t)SCAL_COUNT * sizeof(double)), "malloc scalar block", code_alloc);
  DDH_TRY(hipHostMalloc((void**)&c.hscal, (size_t)SCAL_COUNT * sizeof(double), hipHostMallocDefault), "pinned scalar block", code_alloc);
  DDH_TRY(hipMemcpy(c.iocc, iocc, (size_t)c.nat * sizeof(int16_t), hipMemcpyHostToDevice), "iocc H2D", code_alloc);
  std::vector<int> occ;
  occ.reserve((size_t)c.nat0);
  const int nrows = c.ny * c.nz;
  std::vector<int> rowptr((size_t)nrows + 1, 0), rowlo((size_t)nrows, 0);
  bool contiguous = true;
  for (int g = 0; g < nrows; ++g) {
    rowptr[g] = (int)occ.size();
    const int16_t* row = iocc + (size_t)g * c.nx;
    int first = -1, last = -1;
    for (int ix = 0; ix < c.nx; ++ix)
      if (row[ix] != 0) {
        occ.push_back(g * c.nx + ix);
        if (first < 0) first = ix;
        last = ix;
      }
    rowlo[g] = first < 0 ? 0 : first;
    if (first >= 0 && last - first + 1 != (int)occ.size() - rowptr[g]) contiguous = false;
  }
  rowptr[nrows] = (int)occ.size();
  c.rows_contiguous = contiguous ? 1 : 0;
  DDH_TRY(hipMalloc((void**)&c.rowlo, (size_t)nrows * sizeof(int)), "malloc row map", code_alloc);
  DDH_TRY(hipMemcpy(c.rowlo, rowlo.data(), (size_t)nrows * sizeof(int), hipMemcpyHostToDevice), "row map H2D", code_alloc);
  if (!contiguous || c.backend == BACKEND_SLICES) {
    std::vector<int> qmap((size_t)c.nat, -1);
    for (size_t q = 0; q < occ.size() && q < (size_t)c.nat0; ++q) qmap[(size_t)occ[q]] = (int)q;
    DDH_TRY(hipMalloc((void**)&c.qmap, (size_t)c.nat * sizeof(int)), "malloc site map", code_alloc);
    DDH_TRY(hipMemcpy(c.qmap, qmap.data(), (size_t)c.nat * sizeof(int), hipMemcpyHostToDevice), "site map H2D", code_alloc);
  }
  if ((int)occ.size() != c.nat0) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "IOCC contains %zu occupied sites but NAT0=%d", occ.size(), c.nat0);
    set_error("occupied-index map", msg);
    return code_occ;
  }
  DDH_TRY(hipMemcpy(c.occ, occ.data(), (size_t)c.nat0 * sizeof(int), hipMemcpyHostToDevice), "occupied map H2D", code_occ);
  DDH_TRY(hipMemcpy(c.rowptr, rowptr.data(), ((size_t)nrows + 1) * sizeof(int), hipMemcpyHostToDevice), "row map H2D", code_occ);
  DDH_TRY(hipMemset(c.adia, 0, n * sizeof(float2)), "clear adia", code_alloc);
  DDH_TRY(hipMemset(c.aoff, 0, n * sizeof(float2)), "clear aoff", code_alloc);
  if (c.backend == BACKEND_FFT3D && matvec_autotune(c)) return code_alloc;
  return 0;
}

int ensure_stream(Context& c) {
  if (!c.stream) DDH_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "create stream", 203);
  return 0;
}

}  // namespace ddhip
