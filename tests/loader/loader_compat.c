/*
 * Test harness: binds a DDSCAT matvec plugin the way the unchanged DDSCAT
 * loader does (behaviour of ddscat_cuda_unified/cuda/ddscat_cuda_loader.c:41-102,
 * restated): library path from DDSCAT_CUDA_DLL, else libddscat_matvec_cuda.so;
 * 13 required symbols (missing -> 1000+2, dlopen failure -> 1000+1);
 * ddscat_cuda_prepare_green_f32 optional (absent -> -777 so the caller falls
 * back to the CPU Green build); nonzero plugin codes pass through with
 * last_error().  The probe only makes calls that the plugin answers on the
 * host (argument checks), so it needs no GPU.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char *required[] = {
    "ddscat_cuda_prepare_f32", "ddscat_cuda_apply_f32", "ddscat_cuda_solve_gpbicg_f32",
    "ddscat_cuda_solve_qmrccg_f32", "ddscat_cuda_solve_pbcgst_f32", "ddscat_cuda_solve_pbcgs2_f32",
    "ddscat_cuda_solve_petrkp_f32", "ddscat_cuda_solve_bicgstab4_f32", "ddscat_cuda_solve_gpbicgstab2_f32",
    "ddscat_cuda_solve_gpbicgstab4_f32", "ddscat_cuda_evalq_f32", "ddscat_cuda_scat_f32",
    "ddscat_cuda_release", "ddscat_cuda_last_error"};

typedef int (*prepare_green_t)(int, int, int, int, int, int, int, float, float, float, float, float, float, float,
                               float, float, const int16_t *);
typedef int (*apply_t)(int, const void *, void *, const void *, const void *);
typedef int (*solver_t)(const void *, void *, const void *, const void *, int, int, float, int *, float *, int *);
typedef const char *(*err_t)(void);

static void *lib;
static void *sym[32];
static prepare_green_t prepare_green;

static int load(void) {
  if (lib) return 0;
  const char *path = getenv("DDSCAT_CUDA_DLL");
  if (!path || !*path) path = "libddscat_matvec_cuda.so";
  lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!lib) return 1;
  for (size_t i = 0; i < sizeof(required) / sizeof(required[0]); ++i) {
    sym[i] = dlsym(lib, required[i]);
    if (!sym[i]) {
      fprintf(stdout, "missing %s\n", required[i]);
      return 2;
    }
  }
  prepare_green = (prepare_green_t)dlsym(lib, "ddscat_cuda_prepare_green_f32");
  return 0;
}

static int loader_prepare_green(int nx) {
  int r = load();
  if (r) return 1000 + r;
  if (!prepare_green) return -777;
  int16_t iocc[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  return prepare_green(nx, 2, 2, 0, 8, 8, 0, 0.f, 0.f, 0.f, 0.1f, 0.f, 0.f, 1.f, 1.f, 1.f, iocc);
}

int main(void) {
  int r = loader_prepare_green(0);
  printf("prepare_green_invalid %d\n", r);
  if (r >= 1000 || r == -777) return 0;
  err_t last = (err_t)sym[13];
  printf("last_error %s\n", last());
  r = ((apply_t)sym[1])('N', NULL, NULL, NULL, NULL);
  printf("apply_unprepared %d\n", r);
  int it = 0, nc = 0;
  float te = 0.f;
  r = ((solver_t)sym[5])(NULL, NULL, NULL, NULL, 24, 10, 1e-5f, &it, &te, &nc);
  printf("pbcgs2_unprepared %d\n", r);
  return 0;
}
