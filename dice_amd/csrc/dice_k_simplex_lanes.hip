// simplex_lanes_kernel: BATCHED Nelder-Mead iterations with one subset per lane
// (template instantiations of dice_device.h per pixel storage / interpolation).
#include "dice_device.h"

namespace dice_hip {

namespace {
// Persistent grid: every CU at the occupancy the kernel's registers allow.
template <typename T, int INTERP, bool SQUARE>
int lane_groups() {
  static int groups = 0;
  if (groups == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::simplex_lanes_kernel<T, INTERP, SQUARE>, 256, 0) !=
            hipSuccess ||
        per_cu <= 0)
      per_cu = 2;
    groups = cus * per_cu;
  }
  return groups;
}
}  // namespace

hipError_t launch_simplex_lanes(int storage, int interp, const KParams &kp, const DImage &ref, const DImage &def,
                                const DSubsets &subs, const double *fields, const PassArgs &pa, float *nm_out,
                                int *counter, hipStream_t stream) {
  const int need = (subs.n + 255) / 256;  // at most one lane per subset
#define V(T, I, Q)                                                                                         \
  dev::launch_simplex_lanes_variant<T, I, Q>(kp, ref, def, subs, fields, pa, nm_out, counter,              \
                                             need < lane_groups<T, I, Q>() ? need : lane_groups<T, I, Q>(), stream)
  if (subs.S > 0) {
    if (storage == 0) return interp == 0 ? V(float, 0, true) : V(float, 2, true);
    return interp == 0 ? V(uint16_t, 0, true) : V(uint16_t, 2, true);
  }
  if (storage == 0) return interp == 0 ? V(float, 0, false) : V(float, 2, false);
  return interp == 0 ? V(uint16_t, 0, false) : V(uint16_t, 2, false);
#undef V
}

}  // namespace dice_hip
