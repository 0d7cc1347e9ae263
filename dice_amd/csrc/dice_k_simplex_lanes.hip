// simplex_groups_kernel: BATCHED Nelder-Mead iterations with 8 or 16 lanes per
// subset (template instantiations of dice_device.h per pixel storage /
// interpolation / subset shape, and per bound on the pixels of one lane).
#include "dice_device.h"

namespace dice_hip {

namespace {
// Persistent grid: every CU at the occupancy the kernel's registers and LDS allow.
template <typename T, int INTERP, bool SQUARE, int W, int KPER>
int group_blocks() {
  static int blocks = 0;
  if (blocks == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::simplex_groups_kernel<T, INTERP, SQUARE, W, KPER>,
                                                     256, (size_t)(256 / W) * dev::group_slot_floats<W>() * sizeof(float)) !=
            hipSuccess ||
        per_cu <= 0)
      per_cu = 2;
    blocks = cus * per_cu;
  }
  return blocks;
}

template <typename T, int INTERP, bool SQUARE, int W, int KPER>
hipError_t launch(const KParams &kp, const DImage &ref, const DImage &def, const DSubsets &subs,
                  const double *fields, const PassArgs &pa, float *nm_out, int *counter, hipStream_t stream) {
  const int need = (subs.n + 256 / W - 1) / (256 / W);  // at most one row of W lanes per subset
  const int have = group_blocks<T, INTERP, SQUARE, W, KPER>();
  return dev::launch_simplex_lanes_variant<T, INTERP, SQUARE, W, KPER>(kp, ref, def, subs, fields, pa, nm_out,
                                                                       counter, need < have ? need : have, stream);
}

// subsets of at most kNarrowSimplexMaxPixels pixels (9 x 9 and smaller): rows
// of 8 lanes, 11 pixels per lane; larger ones: rows of 16 lanes, up to 16 pixels
// per lane.  The width must be the one the objective order of the other modes
// uses (KParams::small_w).
template <typename T, int INTERP, bool SQUARE>
hipError_t launch_by_size(int max_pixels, const KParams &kp, const DImage &ref, const DImage &def,
                          const DSubsets &subs, const double *fields, const PassArgs &pa, float *nm_out, int *counter,
                          hipStream_t stream) {
  if (max_pixels <= kNarrowSimplexMaxPixels) {
    if (kp.small_w != 8) return hipErrorInvalidValue;
    return launch<T, INTERP, SQUARE, 8, kNarrowSimplexMaxPixels / 8>(kp, ref, def, subs, fields, pa, nm_out,
                                                                     counter, stream);
  }
  if (kp.small_w != 16) return hipErrorInvalidValue;
  return launch<T, INTERP, SQUARE, 16, kLaneSimplexMaxPixels / 16>(kp, ref, def, subs, fields, pa, nm_out, counter,
                                                                   stream);
}
}  // namespace

hipError_t launch_simplex_lanes(int storage, int interp, int max_pixels, const KParams &kp, const DImage &ref,
                                const DImage &def, const DSubsets &subs, const double *fields, const PassArgs &pa,
                                float *nm_out, int *counter, hipStream_t stream) {
  if (max_pixels <= 0 || max_pixels > kLaneSimplexMaxPixels) return hipErrorInvalidValue;
#define V(T, I, Q) launch_by_size<T, I, Q>(max_pixels, kp, ref, def, subs, fields, pa, nm_out, counter, stream)
  if (subs.S > 0) {
    if (storage == 0) return interp == 0 ? V(float, 0, true) : V(float, 2, true);
    return interp == 0 ? V(uint16_t, 0, true) : V(uint16_t, 2, true);
  }
  if (storage == 0) return interp == 0 ? V(float, 0, false) : V(float, 2, false);
  return interp == 0 ? V(uint16_t, 0, false) : V(uint16_t, 2, false);
#undef V
}

}  // namespace dice_hip
