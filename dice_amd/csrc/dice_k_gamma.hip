// Differentiation path kernels: template instantiations of the forward gamma
// and the adjoint of dice_device.h for each pixel storage / interpolation.
#include "dice_device.h"

namespace dice_hip {

hipError_t launch_gamma(const Launch &L, int storage, const KParams &kp, const DImage &ref, const DImage &def,
                        const DSubsets &subs, const GammaArgs &ga, int *work, hipStream_t stream) {
  if (subs.n <= 0) return hipSuccess;
  int *ov_list = work + kWorkHeader;
  hipError_t e = launch_zero_ints(work, kWorkHeader, stream);
  if (e != hipSuccess) return e;
#define V(T, I, M) dev::launch_gamma_variant<T, I, M>(L, kp, ref, def, subs, ga, pa, stream)
  PassArgs pa{0, nullptr, nullptr, nullptr, nullptr, ov_list, work + kWorkOv0Count, work};
  if (L.mode == 1) {
    if (storage == 0) e = kp.interp == 0 ? V(float, 0, 1) : V(float, 2, 1);
    else e = kp.interp == 0 ? V(uint16_t, 0, 1) : V(uint16_t, 2, 1);
    if (e != hipSuccess) return e;
    pa = PassArgs{0, ov_list, work + kWorkOv0Count, nullptr, nullptr, nullptr, nullptr, work};
  }
  if (storage == 0) return kp.interp == 0 ? V(float, 0, 0) : V(float, 2, 0);
  return kp.interp == 0 ? V(uint16_t, 0, 0) : V(uint16_t, 2, 0);
#undef V
}

}  // namespace dice_hip
