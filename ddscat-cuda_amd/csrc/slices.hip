// SLICES (low-memory) MATVEC backend -- placeholder until the z-then-xy
// streamed implementation lands.
#include "solver_common.hpp"

namespace ddhip {

int slices_setup(Context& c) {
  (void)c;
  set_error("SLICES backend", "not available in this build");
  return 206;
}
void slices_release(Context& c) { (void)c; }
int slices_matvec(Context& c, int cwhat, const float2* x, float2* y, const float2* adia, const float2* aoff,
                  int layout, hipStream_t st) {
  (void)c; (void)cwhat; (void)x; (void)y; (void)adia; (void)aoff; (void)layout; (void)st;
  set_error("SLICES backend", "not available in this build");
  return 20;
}

}  // namespace ddhip
