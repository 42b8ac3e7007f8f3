// The FP32 instantiation of the C-grid plane's device code (plane_core.h): a * b + c within one expression of the
// stencil or the RK4 update contracts into one FMA.
#define LMW_PLANE_CONTRACT _Pragma("clang fp contract(on)")
#include "plane_core.h"

namespace lmw {
std::unique_ptr<PlaneCoreBase> make_plane_core_f32(const lmw_plane_params &p, int members) {
  return std::make_unique<PlaneCore<float>>(p, members);
}
}  // namespace lmw
