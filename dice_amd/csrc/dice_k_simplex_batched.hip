// simplex kernels, batched mode: template instantiations of the solver in
// dice_device.h for each pixel storage / interpolation / gradient-source combination.
#include "dice_device.h"

namespace dice_hip {

hipError_t launch_simplex_batched(int storage, int interp, int defgrad, const Launch &L, const KParams &kp,
                              const DImage &ref, const DImage &def, const DSubsets &subs, double *fields,
                              const PassArgs &pa, hipStream_t stream) {
#define V(T, I, D) dev::launch_variant<T, I, D, true, 1>(L, kp, ref, def, subs, fields, pa, stream)
  (void)defgrad;
  if (storage == 0) return interp == 0 ? V(float, 0, false) : V(float, 2, false);
  return interp == 0 ? V(uint16_t, 0, false) : V(uint16_t, 2, false);
#undef V
}

}  // namespace dice_hip
