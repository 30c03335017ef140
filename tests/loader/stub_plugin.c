/* Minimal plugins for the loader tests: STUB_FULL exports the 13 required
 * symbols without ddscat_cuda_prepare_green_f32; otherwise only one symbol. */
int ddscat_cuda_prepare_f32(void) { return 0; }
#ifdef STUB_FULL
int ddscat_cuda_apply_f32(void) { return 0; }
int ddscat_cuda_solve_gpbicg_f32(void) { return 0; }
int ddscat_cuda_solve_qmrccg_f32(void) { return 0; }
int ddscat_cuda_solve_pbcgst_f32(void) { return 0; }
int ddscat_cuda_solve_pbcgs2_f32(void) { return 0; }
int ddscat_cuda_solve_petrkp_f32(void) { return 0; }
int ddscat_cuda_solve_bicgstab4_f32(void) { return 0; }
int ddscat_cuda_solve_gpbicgstab2_f32(void) { return 0; }
int ddscat_cuda_solve_gpbicgstab4_f32(void) { return 0; }
int ddscat_cuda_evalq_f32(void) { return 0; }
int ddscat_cuda_scat_f32(void) { return 0; }
int ddscat_cuda_release(void) { return 0; }
const char *ddscat_cuda_last_error(void) { return ""; }
#endif
