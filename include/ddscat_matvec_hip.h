/*
 * C ABI of the MI355X (gfx950) DDSCAT hot-path library.
 *
 * The first block is a drop-in for the reference plugin's export set
 * (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.h:11-51): same names, argument
 * order, meaning, return codes and lifetime, so the unchanged dlopen loader
 * (ddscat_cuda_unified/cuda/ddscat_cuda_loader.c:41-102) and Fortran bridge
 * (ddscat_cuda_unified/fortran/cuda_matvec_bridge.f90) bind it as they bind
 * libddscat_matvec_cuda.so.  All pointers in that block are caller-owned host
 * memory; complex values are interleaved complex64; full vectors are
 * component-major [3][NAT] with site j = ix + NX*(iy + NY*iz).
 *
 * The second block (ddscat_hip_*) holds extensions the reference loader never
 * resolves: device-pointer MATVEC, backend / precision selection, inspection.
 */
#ifndef DDSCAT_MATVEC_HIP_H
#define DDSCAT_MATVEC_HIP_H
#include <stdint.h>

#define DDSCAT_HIP_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference export set ------------------------------------------------ */

/* Build the spectral Green tensor on the device (FP64) and allocate the
 * MATVEC / solver state.  Replaces ddscat_matvec_cuda.h:11-13
 * (implementation ddscat_matvec_cuda.cu:1005-1017).  IPBC=1 (periodic
 * target, period PYD*DX2 in y and/or PZD*DX3 in z) sums damped replicas per
 * lag (cu:634-643) and keeps the whole spectral grid.  Codes: 201 invalid
 * argument (also IPBC=1 with neither PYD>0 nor PZD>0), 202 GAMMA<=0 with PBC,
 * 203 device setup, 204 allocation, 206 unsupported grid, 208 IOCC count !=
 * NAT0, 220+g Green build failure (234: k has no component across the
 * periodic axes, 235: the replica sums would exceed 2e11 interactions). */
DDSCAT_HIP_API int ddscat_cuda_prepare_green_f32(int nx, int ny, int nz, int ipbc, int nat, int nat0,
                                                 int idipint, float gamma, float pyd, float pzd,
                                                 float akx, float aky, float akz, float dx1,
                                                 float dx2, float dx3, const int16_t *iocc);

/* Same lifetime with a host-supplied transformed Green tensor CXZC
 * [6][NZ+1][NY+1][NX+1] complex64 (IPBC=0) or [6][2NZ][2NY][2NX] (IPBC=1).
 * Replaces ddscat_matvec_cuda.h:14-15
 * (ddscat_matvec_cuda.cu:1018-1026).  Codes 1-6 as the reference. */
DDSCAT_HIP_API int ddscat_cuda_prepare_f32(int nx, int ny, int nz, int ipbc, int nat, int nat0,
                                           const void *green, const int16_t *iocc);

/* y = A x ('N', cwhat=78) or conj(A conj(x)) ('C', cwhat=67) on full host
 * vectors.  Replaces ddscat_matvec_cuda.h:16-17 (ddscat_matvec_cuda.cu:1027-1029).
 * Codes: 10 not prepared, 11 bad argument, 12-15 copies, 20-29 MATVEC. */
DDSCAT_HIP_API int ddscat_cuda_apply_f32(int cwhat, const void *x, void *y, const void *adia,
                                         const void *aoff);

/* Krylov solvers, ddscat_matvec_cuda.h:18-41.  b, x full [3*NAT] host vectors
 * (x: initial guess in, solution out), ndim must be 3*NAT.  itno = iterations,
 * tole = achieved relative residual, ncompte = MATVEC count.  Non-convergence
 * returns 0 with the achieved residual except PBCGST (87), as the reference. */
DDSCAT_HIP_API int ddscat_cuda_solve_gpbicg_f32(const void *b, void *x, const void *adia,
                                                const void *aoff, int ndim, int maxit, float tol,
                                                int *itno, float *tole, int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_qmrccg_f32(const void *b, void *x, const void *adia,
                                                const void *aoff, int ndim, int maxit, float tol,
                                                int *itno, float *tole, int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_pbcgst_f32(const void *b, void *x, const void *adia,
                                                const void *aoff, int ndim, int maxit, float tol,
                                                int *itno, float *tole, int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_pbcgs2_f32(const void *b, void *x, const void *adia,
                                                const void *aoff, int ndim, int maxit, float tol,
                                                int *itno, float *tole, int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_petrkp_f32(const void *b, void *x, const void *adia,
                                                const void *aoff, int ndim, int maxit, float tol,
                                                int *itno, float *tole, int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_bicgstab4_f32(const void *b, void *x, const void *adia,
                                                   const void *aoff, int ndim, int maxit, float tol,
                                                   int *itno, float *tole, int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_gpbicgstab2_f32(const void *b, void *x, const void *adia,
                                                     const void *aoff, int ndim, int maxit,
                                                     float tol, int *itno, float *tole,
                                                     int *ncompte);
DDSCAT_HIP_API int ddscat_cuda_solve_gpbicgstab4_f32(const void *b, void *x, const void *adia,
                                                     const void *aoff, int ndim, int maxit,
                                                     float tol, int *itno, float *tole,
                                                     int *ncompte);

/* Cabs, Cext, Cpha from the resident solution (use_resident=1) or compact
 * host arrays [3][NAT0].  Replaces ddscat_matvec_cuda.h:42-44
 * (ddscat_matvec_cuda.cu:919-955).  Codes 300-307. */
DDSCAT_HIP_API int ddscat_cuda_evalq_f32(int use_resident, const void *e_compact,
                                         const void *p_compact, const void *adia_compact,
                                         const void *aoff_compact, float akx, float aky, float akz,
                                         float e02, double *cabs, double *cext, double *cpha);

/* Far-field sums over the resident solution.  Replaces ddscat_matvec_cuda.h:45-49
 * (ddscat_matvec_cuda.cu:957-1002).  Codes 320-333. */
DDSCAT_HIP_API int ddscat_cuda_scat_f32(int jpbc, int ndir, const float *ak, const float *aks,
                                        const float *dx, const float *em1, const float *em2,
                                        float e02, float etasca, const void *e01, const float *x0,
                                        double *cbksca, double *csca, double *cscag,
                                        double *cscag2, int *navg, void *f1, void *f2);

/* Free all device state.  Replaces ddscat_matvec_cuda.h:50 (ddscat_matvec_cuda.cu:915). */
DDSCAT_HIP_API int ddscat_cuda_release(void);
/* Text of the last error.  Replaces ddscat_matvec_cuda.h:51 (ddscat_matvec_cuda.cu:916). */
DDSCAT_HIP_API const char *ddscat_cuda_last_error(void);

/* ---- extensions (no counterpart in the reference) ------------------------- */

enum { DDSCAT_HIP_BACKEND_FFT3D = 0, DDSCAT_HIP_BACKEND_SLICES = 1 };
enum { DDSCAT_HIP_PREC_MIXED = 0, DDSCAT_HIP_PREC_FP32 = 1 };

/* MATVEC backend used by the next prepare call (default: FFT3D, or SLICES
 * for libddscat_matvec_cuda_slice.so). */
DDSCAT_HIP_API int ddscat_hip_set_backend(int backend);
/* Solver reduction precision: MIXED = FP64 dots/scalars (reference policy),
 * FP32 = plain single precision throughout. */
DDSCAT_HIP_API int ddscat_hip_set_precision(int precision);

/* What QMRCCG (single-RHS and batched) does when a reliable-residual check
 * (every 20 iterations, cu:543-556) finds a gap >= 1e-3 between the
 * recursive and the true residual, or a false convergence:
 *   REPLACE (default): the true residual replaces the recursive one and the
 *     Lanczos recurrence goes on (QMR's residual vector only monitors
 *     convergence).  With FP32 vectors the two residuals drift apart by
 *     about 1e-6 |b|, so below a residual of about 1e-3 every check fails;
 *   RESTART: the reference's policy, a restart of the recurrence from the
 *     true residual at every failed check.
 * The other solvers keep the reference's restart.  Code 1: unknown mode. */
enum { DDSCAT_HIP_RELIABLE_REPLACE = 0, DDSCAT_HIP_RELIABLE_RESTART = 1 };
DDSCAT_HIP_API int ddscat_hip_set_reliable_mode(int mode);

/* FFT kernel family: 0 = register-resident two-level FFTs where the padded
 * axis length allows (64, 128, 256), LDS-staged Stockham otherwise
 * (default); 1 = LDS-staged Stockham for every axis. */
DDSCAT_HIP_API int ddscat_hip_set_fft_kernels(int generic);

/* FFT kernel autotuning at prepare (FFT3D backend): 1 (default) times every
 * candidate kernel of each MATVEC pass (register shapes of the pass's axis
 * length and the LDS-staged kernel) and keeps the fastest, cached per grid
 * for the process; 0 uses each length's default register shape. */
DDSCAT_HIP_API int ddscat_hip_set_autotune(int on);

/* Override the FFT kernel of MATVEC pass 1..5 of the prepared grid: kind -1
 * = LDS-staged, k >= 0 = register candidate k of the pass's axis length
 * (ddscat_hip_info reports the current choice).  For testing and manual
 * tuning; the next prepare re-tunes.  Codes 10 not prepared, 11 no such kernel. */
DDSCAT_HIP_API int ddscat_hip_set_pass_kernel(int pass, int kind);

/* y = A x on DEVICE pointers, stream-ordered on `stream` (NULL = library
 * stream), no synchronisation.  layout 0: full [3][NAT] vectors and
 * operators; layout 1: compact [3][NAT0] vectors and compact operators. */
DDSCAT_HIP_API int ddscat_hip_apply_device(int cwhat, const void *x, void *y, const void *adia,
                                           const void *aoff, int layout, void *stream);

/* Solver methods for ddscat_hip_solve_device. */
enum {
  DDSCAT_HIP_PBCGS2 = 0, DDSCAT_HIP_QMRCCG = 1, DDSCAT_HIP_PETRKP = 2, DDSCAT_HIP_GPBICG = 3,
  DDSCAT_HIP_PBCGST = 4, DDSCAT_HIP_BICGSTAB4 = 5, DDSCAT_HIP_GPBICGSTAB2 = 6, DDSCAT_HIP_GPBICGSTAB4 = 7
};

/* The ddscat_cuda_solve_*_f32 solvers on compact [3][NAT0] DEVICE vectors
 * (b, x in/out, adia, aoff; ndim = 3*NAT0), with the same iteration,
 * convergence and return-code behaviour as the host entry points; blocks
 * until the solve ends.  The solution stays resident for EVALQ / SCAT. */
DDSCAT_HIP_API int ddscat_hip_solve_device(int method, const void *b, void *x, const void *adia,
                                           const void *aoff, int ndim, int maxit, float tol, int *itno,
                                           float *tole, int *ncompte);

/* nrhs MATVECs in one launch sequence on compact DEVICE vectors x, y
 * [nrhs][3][NAT0]; adia, aoff per RHS (op_per_rhs = 1) or shared (0).
 * Stream-ordered on `stream` (NULL = library stream).  FFT3D runs the nrhs
 * right-hand sides in each kernel launch; SLICES runs them one after another
 * through its slice buffers. */
DDSCAT_HIP_API int ddscat_hip_apply_batch(int cwhat, int nrhs, const void *x, void *y, const void *adia,
                                          const void *aoff, int op_per_rhs, void *stream);

/* Batched solve of nrhs right-hand sides of the prepared target (either
 * backend; method DDSCAT_HIP_QMRCCG or DDSCAT_HIP_PBCGS2).  b, x (initial guess in, solution out):
 * compact DEVICE vectors [nrhs][3][NAT0]; adia, aoff: [nrhs][3][NAT0] when
 * op_per_rhs = 1 (per-RHS LDR operators), one shared [3][NAT0] when 0.
 * itno, tole, ncompte: host arrays [nrhs] (per-RHS iterations, achieved
 * residual, MATVEC count).  Every RHS follows the single-RHS algorithm and
 * reliable-residual policy; finished RHS stop while the rest continue.
 * use_graph = 1 replays one captured HIP graph per iteration from the fifth
 * iteration on (the first four run as direct launches); iterations run
 * in chunks of up to 20 between host synchronisations, with the per-RHS
 * iteration counts and stopping tests kept on the device.  Returns 0, or
 * the breakdown code of the first RHS that broke down (the others are
 * still solved); 1 invalid argument, 2 method not batched, 29 memory. */
DDSCAT_HIP_API int ddscat_hip_solve_batch(int method, int nrhs, const void *b, void *x, const void *adia,
                                          const void *aoff, int op_per_rhs, int maxit, float tol, int *itno,
                                          float *tole, int *ncompte, int use_graph);

/* Target slots.  The library holds up to 64 prepared targets besides the
 * reference ABI's one (slot 0, selected by default).  Every entry point,
 * ddscat_cuda_* and ddscat_hip_*, acts on the selected slot: select a slot,
 * then prepare it (ddscat_cuda_prepare_green_f32 / _prepare_f32), build its
 * inputs (ddscat_hip_setup_rhs) and read its results (ddscat_hip_evalq_batch)
 * as for the single target; ddscat_cuda_release frees the selected slot.
 * The selection is process-wide (not per thread).  Code 1: slot outside
 * 0..64. */
DDSCAT_HIP_API int ddscat_hip_select_target(int slot);

/* Batched solves of several prepared targets in one call (the uneven
 * multi-target batch): target t = slots[t] with nrhs[t] right-hand sides and
 * its own compact DEVICE arrays b[t], x[t] (initial guess in, solution out),
 * adia[t], aoff[t] laid out as for ddscat_hip_solve_batch (per-RHS operators
 * when op_per_rhs = 1).  Each target runs the batched QMRCCG or PBCGS2 of
 * ddscat_hip_solve_batch on its own stream, driven from its own host thread,
 * so the targets' kernels share the GPU; the call returns when all are done.
 * itno, tole, ncompte: host arrays of sum(nrhs) entries, target after
 * target.  Every slot must be prepared (either backend) and appear once.
 * Returns 0 or the first target's nonzero code (1 invalid argument, 2
 * method/backend, 29 memory, else the solver's). */
DDSCAT_HIP_API int ddscat_hip_solve_targets(int method, int ntargets, const int *slots, const int *nrhs,
                                            void *const *b, void *const *x, void *const *adia,
                                            void *const *aoff, int op_per_rhs, int maxit, float tol, int *itno,
                                            float *tole, int *ncompte, int use_graph);

/* EVALQ for nrhs solutions on DEVICE: e (incident field), p (polarization)
 * [nrhs][3][NAT0], adia / aoff per RHS (op_per_rhs = 1) or shared (0);
 * ak [nrhs][3] host floats; out [nrhs][3] host doubles = Cabs, Cext, Cpha
 * per RHS with the formulas of ddscat_cuda_evalq_f32.  Codes 300-307. */
DDSCAT_HIP_API int ddscat_hip_evalq_batch(int nrhs, const void *e, const void *p, const void *adia,
                                          const void *aoff, int op_per_rhs, const float *ak, float e02,
                                          double *out);

/* Adjoint sensitivity of Cabs (the EVALQ absorption cross section) to the
 * operators, for nrhs solutions p [nrhs][3][NAT0] on DEVICE with their adia /
 * aoff (per RHS or shared) and ak [nrhs][3]:
 *   dCabs = sum Re(g_adia * d adia) + sum Re(g_aoff * d aoff)
 * g_adia (and g_aoff unless NULL) are written as complex64 [nrhs][3][NAT0]
 * DEVICE arrays.  One batched adjoint solve A mu = W conj(p) (method
 * DDSCAT_HIP_QMRCCG or DDSCAT_HIP_PBCGS2, tolerance tol) with the forward
 * MATVEC (A is complex symmetric); itno / tole [nrhs] report it.  Codes: 1
 * invalid argument, 2 method, 3 launch, 29 memory, else the solver's. */
DDSCAT_HIP_API int ddscat_hip_cabs_adjoint(int method, int nrhs, const void *p, const void *adia,
                                           const void *aoff, int op_per_rhs, const float *ak, float e02,
                                           int maxit, float tol, void *g_adia, void *g_aoff, int *itno,
                                           float *tole);

/* Radiative torque of the resident solution (DDSCAT's DOTORQ sums, which the
 * reference's GPU SCAT leaves to the CPU SCAT:
 * ddscat_cuda_unified/GPU_FINAL_SECTIONS.md:75-80, dispatch
 * fortran/getfml_cuda.f90:1135-1150), as torque cross sections
 * C = Gamma k / u_rad (u_rad = e02 / 8 pi; Q_Gamma = C / (pi a_eff^2)) about
 * the origin of the site coordinates r_j = ((i+1) + x0) dx:
 *   cext[3]  the incident wave's torque on the dipoles, with the resident
 *            right-hand side as the incident field (FP64 site sums);
 *   crad[3]  the recoil of the radiated field: the far-field sums S(n) and the
 *            position-weighted sums R(n) = sum_j r_j P_j exp(-i k n.r_j)
 *            integrated over SCAT's ETASCA quadrature (*navg directions).
 * cext + crad is the total torque (DDSCAT's QTRQAB + QTRQSC split it
 * differently into absorption and scattering parts).  ak, e01 (complex64
 * [3]), e02, etasca, dx, x0 as for ddscat_cuda_scat_f32.  Codes: 350
 * invalid state (no resident solution) or argument, 351 device failure. */
DDSCAT_HIP_API int ddscat_hip_torque(const float *ak, const void *e01, float e02, float etasca,
                                     const float *dx, const float *x0, double *cext, double *crad,
                                     int *navg);

/* ddscat_cuda_scat_f32 without device allocation: same arguments, results
 * and codes.  No counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The far-field workspace
 * (directions, phase tables, row partials, sums) lives in a 64 MiB
 * file-scope device array per GPU, held until the process ends, instead of
 * per-call allocations; directions are processed in slabs that fit it.
 * Calls that overlap on one GPU share that array and race.  Code 334: the
 * target has too many lattice rows for one direction's workspace to fit
 * (NY*NZ above about 1.4 M); 335: the array's device address could not be
 * obtained. */
DDSCAT_HIP_API int ddscat_hip_scat_f32_noalloc(int jpbc, int ndir, const float *ak, const float *aks,
                                               const float *dx, const float *em1, const float *em2,
                                               float e02, float etasca, const void *e01, const float *x0,
                                               double *cbksca, double *csca, double *cscag,
                                               double *cscag2, int *navg, void *f1, void *f2);

/* LDR polarizability prescription used by ddscat_hip_setup_rhs (DDSCAT's
 * CALPHA, ddscat_cuda_unified/fortran/reapar_cuda.f90:497-513): LATTDR
 * (Draine & Goodman 1993, default) or GKDLDR (Gutkowicz-Krusin & Draine
 * 2004: a diagonal tensor, component j with S_j = khat_j^2).  FLTRCD's
 * filtered interaction is IDIPINT=1 of ddscat_cuda_prepare_green_f32.
 * Code 1: unknown prescription. */
enum { DDSCAT_HIP_LATTDR = 0, DDSCAT_HIP_GKDLDR = 1 };
DDSCAT_HIP_API int ddscat_hip_set_polarizability(int prescription);

/* Per-RHS inputs of a batched solve built on the DEVICE for the prepared
 * target (what DDSCAT's EVALE and ALPHADIAG/LATTDR produce on the host):
 * for r < nrhs with k d = ak[3r..3r+2] and polarization e0[6r..6r+5]
 * (3 interleaved complex64),
 *   b[r]    = e0 exp(i k.x_q), x_q = ((ix+1) + x0) dx per axis (FP64 phase),
 *   adia[r] = 1/alpha of the lattice dispersion relation for eps(q)
 *             (per component; see ddscat_hip_set_polarizability),
 *   aoff[r] = 0,
 * all compact [nrhs][3][NAT0] DEVICE arrays.  eps: complex64 DEVICE array,
 * [NAT0] (isotropic, eps_per_comp = 0) or [3][NAT0] (eps_per_comp = 1).
 * Blocks until done.  Codes: 1 invalid argument, 2 copy, 3 launch, 29 memory. */
DDSCAT_HIP_API int ddscat_hip_setup_rhs(int nrhs, const float *ak, const float *e0, const float *x0,
                                        const float *dx, const void *eps, int eps_per_comp, void *b,
                                        void *adia, void *aoff);

/* Near field on every lattice site of the prepared grid (DDSCAT NRFLD):
 * out_full [3][NAT] (DEVICE) = einc_full + E_sca, or E_sca alone when
 * einc_full is NULL, where E_sca,j = sum_k G_jk P_k is radiated by the compact
 * polarization p_compact [3][NAT0] (DEVICE); computed with the MATVEC's FFT
 * convolution and no occupancy mask (either backend); blocks until done.
 * Codes 10 not prepared, 11 argument, 12-13 launch or copy. */
DDSCAT_HIP_API int ddscat_hip_nearfield(const void *p_compact, const void *einc_full, void *out_full);

/* Amplitude and Mueller matrices for npair incident-polarization pairs (the
 * S-matrix output path: the amplitudes of ddscat_cuda_scat_f32,
 * ddscat_matvec_cuda.cu:841-845, for both polarizations at once, converted
 * as DDSCAT's GETMUELLER does on the host, DDSCAT_cuda.f90:3294-3300;
 * formulas of Bohren & Huffman ch. 3, eq. 3.16).
 *   p      device, compact [2*npair][3][NAT0] complex64: pair q solved for
 *          incident polarizations e01 (2q) and e02 (2q+1)
 *   ak     [npair][3] k d of the incident wave of each pair (target frame)
 *   e0     [npair][2][3] complex64: e01, e02 of each pair (orthonormal)
 *   aks    [npair][ndir][3] scattering wave vectors k d khat_s
 *   em1    [npair][ndir][3] e_par_s = khat_s x e_perp (in the scattering plane)
 *   em2    [npair][ndir][3] e_perp, the scattering-plane normal shared by the
 *          incident and scattered bases (Bohren & Huffman: sin(phi) x - cos(phi) y
 *          for incidence along z)
 *   dx, x0 lattice ratios and origin offset as for ddscat_cuda_scat_f32
 *   famp   host [npair][ndir][4] complex64: f11 f21 f12 f22
 *   mueller host [npair][ndir][4][4] doubles (may be NULL)
 * Codes: 340 invalid state or argument, 341 far-field failure. */
DDSCAT_HIP_API int ddscat_hip_scat_matrix(int npair, const void *p, int ndir, const float *ak, const void *e0,
                                          const float *aks, const float *em1, const float *em2, const float *dx,
                                          const float *x0, float e02, void *famp, double *mueller);

/* Test hook for the reliable-residual policy (cu:543-556): the next
 * reliable check of a single-RHS solve sees its gap raised by 1 (above the
 * 1e-3 restart threshold) and restarts from the true residual.  The restart
 * count of the last single-RHS solve is ddscat_hip_info field 23. */
DDSCAT_HIP_API int ddscat_hip_force_restart(void);

/* Test hook for uninitialised memory: fills the selected slot's single-RHS
 * solver workspace (sized for the largest solver) and, when a batched solve
 * has allocated them, its batched solver vectors with NaN bit patterns, as
 * memory that held other data can; every solver must write a work vector
 * before it reads it.  Codes 10 not prepared, 12 launch, 29 memory. */
DDSCAT_HIP_API int ddscat_hip_poison_workspace(void);

/* Convergence record of the last solve: RHS `rhs` of the last
 * ddscat_hip_solve_batch / ddscat_hip_cabs_adjoint, or rhs 0 of the last
 * single-RHS solve.  *restarts = reliable-residual restarts (cu:543-556
 * policy); resid[0..n-1] = the recursive residual after each iteration (as
 * many as fit; PBCGS2 and QMRCCG); checks[4*i..4*i+3] = (iteration, true
 * residual, gap, action: 0 none, 1 restart, 2 residual replaced) of
 * reliable check i (up to nchecks of them);
 * *nchecks_out = the number of checks.  Any output pointer may be NULL.
 * Returns the number of iterations recorded, -1 if there is no such RHS. */
DDSCAT_HIP_API int ddscat_hip_solve_log(int rhs, int *restarts, double *resid, int n, double *checks,
                                        int nchecks, int *nchecks_out);

/* Copy the resident spectral Green tensor (ddscat_hip_info field 5 complex64
 * values) to host memory in the CXZC layout ddscat_cuda_prepare_f32 takes. */
DDSCAT_HIP_API int ddscat_hip_get_green(void *out);

/* Fill up to n int64 fields: nx ny nz nat nat0 green_count matvec_count
 * backend precision gx gy gz rx_fwd cy cz rx_inv green_build_us, then the FFT
 * kernel of passes 1..5 (-1 LDS-staged, k >= 0 register shape k of the axis
 * length), whether the autotuner chose them, and the reliable-residual
 * restarts of the last single-RHS solve. */
DDSCAT_HIP_API int ddscat_hip_info(long long *out, int n);

/* Run `reps` MATVECs on device pointers with HIP events between the five
 * FFT3D passes: ms_out[0..4] mean per-pass milliseconds, ms_out[5] mean
 * MATVEC milliseconds (SLICES backend: only ms_out[5]). */
DDSCAT_HIP_API int ddscat_hip_time_matvec(int cwhat, const void *x, void *y, const void *adia,
                                          const void *aoff, int layout, int reps, double *ms_out);

/* The library's HIP stream (hipStream_t) and a blocking wait on it. */
DDSCAT_HIP_API void *ddscat_hip_stream(void);
DDSCAT_HIP_API int ddscat_hip_sync(void);

#ifdef __cplusplus
}
#endif
#endif
