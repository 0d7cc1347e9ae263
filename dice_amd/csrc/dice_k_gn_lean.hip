// gn kernels, lean batched mode: template instantiations of Solver::run_lean in
// dice_device.h for each pixel storage / interpolation combination.
#include "dice_device.h"

namespace dice_hip {

hipError_t launch_gn_lean(int storage, int interp, const Launch &L, const KParams &kp, const DImage &ref,
                          const DImage &def, const DSubsets &subs, double *fields, const PassArgs &pa,
                          hipStream_t stream) {
#define V(T, I) dev::launch_lean_variant<T, I>(L, kp, ref, def, subs, fields, pa, stream)
  if (storage == 0) return interp == 0 ? V(float, 0) : V(float, 2);
  return interp == 0 ? V(uint16_t, 0) : V(uint16_t, 2);
#undef V
}

}  // namespace dice_hip
