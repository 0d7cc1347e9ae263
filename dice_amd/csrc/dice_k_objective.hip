// Objective_ZNSSD calls for one subset (DICe_Objective.h:82-242): gamma, sigma,
// computeUpdateFast, computeUpdateRobust, on one 256-thread workgroup.
#include "dice_device.h"

namespace dice_hip {
namespace dev {

// io[0..5]: shape parameters in (u, v, theta, ex, ey, gxy) and, for the
// updates, out; io[6], io[7]: gamma | sigma, noise | status, iterations.
template <typename T, int INTERP, bool DEFGRAD>
__global__ __launch_bounds__(kPlainThreads) void objective_kernel(KParams kp, DImage ref, DImage def, DSubsets ss,
                                                                  double *fields, int gid, int call, double *io) {
  __shared__ double red[4 * kMaxRed];
  BlockGroup g{(lds_double *)red};
  Solver<T, BlockGroup, INTERP, DEFGRAD> s(kp, ref, def, g, (lds_float *)red, 0);
  s.set_subset(ss, gid);
  Affine A;
  float p[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) p[k] = A.p[k] = (float)io[k] * s.en[k];
  double out0 = 0.0, out1 = 0.0;
  if (call == 0) {  // Objective::gamma (DICe_Objective.cpp:58-76)
    A.prepare();
    int active;
    float contrast;
    out0 = s.gamma_eval(A, active, contrast);
  } else if (call == 1) {  // Objective::sigma (:150-190)
    float noise = 0.0f;
    out0 = s.sigma(p, A, noise);
    out1 = noise;
  } else if (call == 2) {  // computeUpdateFast (:430-657): F-bar once, then the Gauss-Newton loop
    double acc[2] = {0.0, 0.0};
    s.for_pixels([&](int x, int y, bool a) {
      if (!a) return;
      acc[0] += (double)s.ref_value(x - ref.ox, y - ref.oy);
      acc[1] += 1.0;
    });
    s.grp.sum(acc);
    s.meanF = acc[1] > 0.0 ? (float)(acc[0] / acc[1]) : 0.0f;
    int iters = -1;
    float p_state[6];
    out0 = s.gauss_newton(A, iters, p_state, fields, ss.n, gid);
    out1 = iters;
  } else {  // computeUpdateRobust (:369-388)
    int iters = -1;
    out0 = s.simplex(A, iters);
    out1 = iters;
  }
  if (threadIdx.x == 0) {
    if (call >= 2)
#pragma unroll
      for (int k = 0; k < 6; ++k) io[k] = A.p[k];
    io[6] = out0;
    io[7] = out1;
  }
}

template <typename T, int INTERP, bool DEFGRAD>
hipError_t launch_objective_t(const KParams &kp, const DImage &ref, const DImage &def, const DSubsets &ss,
                              double *fields, int gid, int call, double *io, hipStream_t stream) {
  hipLaunchKernelGGL((objective_kernel<T, INTERP, DEFGRAD>), dim3(1), dim3(kPlainThreads), 0, stream, kp, ref, def, ss,
                     fields, gid, call, io);
  return hipGetLastError();
}

}  // namespace dev

hipError_t launch_objective(int storage, const KParams &kp, const DImage &ref, const DImage &def, const DSubsets &ss,
                            double *fields, int gid, int call, double *io, hipStream_t stream) {
#define V(T, I, D) dev::launch_objective_t<T, I, D>(kp, ref, def, ss, fields, gid, call, io, stream)
  const bool d = kp.def_grads != 0;
  if (storage == 0) {
    if (kp.interp == 0) return d ? V(float, 0, true) : V(float, 0, false);
    return d ? V(float, 2, true) : V(float, 2, false);
  }
  if (kp.interp == 0) return d ? V(uint16_t, 0, true) : V(uint16_t, 0, false);
  return d ? V(uint16_t, 2, true) : V(uint16_t, 2, false);
#undef V
}

}  // namespace dice_hip
