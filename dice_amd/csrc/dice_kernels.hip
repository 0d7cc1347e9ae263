// Image-level kernels of the subset-correlation library and the dispatch of
// the correlation passes (the solver kernels live in dice_k_*.hip).
#include "dice_device.h"

namespace dice_hip {

typedef hipError_t (*PassLauncher)(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                                   const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_gn_plain(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                           const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_gn_batched(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                             const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_simplex_plain(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                                const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_simplex_batched(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                                  const DSubsets &, double *, const PassArgs &, hipStream_t);

namespace {
using dev::ldpix;

// FD gradients (Image_::compute_gradients_finite_difference, DICe_Image.cpp:1064-1094):
// 4th-order central differences, first-order one-sided within 2 px of the edge.
template <typename T>
__global__ void gradients_kernel(const void *pix, int w, int h, float *gx, float *gy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const int o = y * w + x;
  const float c1 = 1.0f / 12.0f, c2 = -8.0f / 12.0f;
  float vx, vy;
  if (x < 2) vx = ldpix<T>(pix, o + 1) - ldpix<T>(pix, o);
  else if (x >= w - 2) vx = ldpix<T>(pix, o) - ldpix<T>(pix, o - 1);
  else vx = c1 * ldpix<T>(pix, o - 2) + c2 * ldpix<T>(pix, o - 1) - c2 * ldpix<T>(pix, o + 1) - c1 * ldpix<T>(pix, o + 2);
  if (y < 2) vy = ldpix<T>(pix, o + w) - ldpix<T>(pix, o);
  else if (y >= h - 2) vy = ldpix<T>(pix, o) - ldpix<T>(pix, o - w);
  else vy = c1 * ldpix<T>(pix, o - 2 * w) + c2 * ldpix<T>(pix, o - w) - c2 * ldpix<T>(pix, o + w) - c1 * ldpix<T>(pix, o + 2 * w);
  gx[o] = vx;
  gy[o] = vy;
}

// 5x5 binomial smoothing of the gradients (DICe_Image.cpp:1032-1062), interior only.
__global__ void smooth_kernel(const float *inx, const float *iny, int w, int h, float *gx, float *gy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const int o = y * w + x;
  if (x < 2 || x >= w - 2 || y < 2 || y >= h - 2) { gx[o] = inx[o]; gy[o] = iny[o]; return; }
  const float k1[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
  float vx = 0.0f, vy = 0.0f;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      const float c = k1[i] * k1[j];
      vx += c * inx[o + (i - 2) * w + (j - 2)];
      vy += c * iny[o + (i - 2) * w + (j - 2)];
    }
  gx[o] = vx;
  gy[o] = vy;
}

__global__ void zero_ints_kernel(int *p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

__global__ void copy_doubles_kernel(double *dst, const double *src, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

// One optimizer pass over `items` (nullptr: all subsets).  BATCHED mode runs the
// persistent kernel, then the PLAIN kernel over the subsets it handed over.
// A Nelder-Mead pass with L.lane_simplex first runs the iterations one subset
// per lane (results in nm); the persistent kernel then does the checks and the
// epilogue of each subset around them.
static hipError_t launch_pass(bool simplex, int storage, const Launch &L, const KParams &kp, const DImage &ref,
                              const DImage &def, const DSubsets &subs, double *fields, PassArgs pa, int *ov_list,
                              int *ov_count, float *nm, hipStream_t stream) {
  const int defgrad = simplex ? 0 : kp.def_grads;
  PassLauncher plain = simplex ? launch_simplex_plain : launch_gn_plain;
  pa.nm_result = nullptr;
  if (L.mode == 0) {
    pa.ov_list = pa.ov_count = nullptr;
    return plain(storage, kp.interp, defgrad, L, kp, ref, def, subs, fields, pa, stream);
  }
  PassLauncher batched = simplex ? launch_simplex_batched : launch_gn_batched;
  hipError_t e;
  if (simplex && L.lane_simplex) {
    int *lane_counter = pa.counter + kWorkLaneCount;
    e = launch_zero_ints(lane_counter, 1, stream);
    if (e != hipSuccess) return e;
    e = launch_simplex_lanes(storage, kp.interp, kp, ref, def, subs, fields, pa, nm, lane_counter, stream);
    if (e != hipSuccess) return e;
    pa.nm_result = nm;
  }
  e = launch_zero_ints(pa.counter, kWorkRanges, stream);
  if (e != hipSuccess) return e;
  pa.ov_list = ov_list;
  pa.ov_count = ov_count;
  e = batched(storage, kp.interp, defgrad, L, kp, ref, def, subs, fields, pa, stream);
  if (e != hipSuccess) return e;
  PassArgs po = pa;
  po.items = ov_list;
  po.item_count = ov_count;
  po.ov_list = po.ov_count = nullptr;
  return plain(storage, kp.interp, defgrad, L, kp, ref, def, subs, fields, po, stream);
}

hipError_t launch_correlate(const Launch &L, int storage, const KParams &kp, const DImage &ref, const DImage &def,
                            const DSubsets &subs, double *fields, int *work, hipStream_t stream) {
  if (subs.n <= 0) return hipSuccess;
  // work: the header (range counters, fallback and overflow counts), then the
  // fallback list and the overflow lists of pass 0 and pass 1 (n each)
  const size_t n = (size_t)subs.n;
  int *fb_list = work + kWorkHeader, *ov0 = fb_list + n, *ov1 = fb_list + 2 * n;
  float *nm = reinterpret_cast<float *>(fb_list + 3 * n);
  hipError_t e = launch_zero_ints(work, kWorkHeader, stream);
  if (e != hipSuccess) return e;
  const bool simplex_first = kp.opt == 0 || kp.opt == 3;
  PassArgs pa{0, nullptr, nullptr, fb_list, work + kWorkFbCount, nullptr, nullptr, work, nullptr};
  e = launch_pass(simplex_first, storage, L, kp, ref, def, subs, fields, pa, ov0, work + kWorkOv0Count, nm, stream);
  if (e != hipSuccess || !(kp.opt == 3 || kp.opt == 4)) return e;
  PassArgs pb{1, fb_list, work + kWorkFbCount, nullptr, nullptr, nullptr, nullptr, work, nullptr};
  return launch_pass(!simplex_first, storage, L, kp, ref, def, subs, fields, pb, ov1, work + kWorkOv1Count, nm,
                     stream);
}

hipError_t launch_gradients(int storage, const void *pix, int w, int h, float *gx, float *gy, int smooth,
                            float *tmpx, float *tmpy, hipStream_t stream) {
  dim3 block(64, 4), grid((w + 63) / 64, (h + 3) / 4);
  float *ox = smooth ? tmpx : gx, *oy = smooth ? tmpy : gy;
  if (storage == 0)
    hipLaunchKernelGGL(gradients_kernel<float>, grid, block, 0, stream, pix, w, h, ox, oy);
  else
    hipLaunchKernelGGL(gradients_kernel<uint16_t>, grid, block, 0, stream, pix, w, h, ox, oy);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !smooth) return e;
  hipLaunchKernelGGL(smooth_kernel, grid, block, 0, stream, tmpx, tmpy, w, h, gx, gy);
  return hipGetLastError();
}

hipError_t launch_zero_ints(int *p, int count, hipStream_t stream) {
  hipLaunchKernelGGL(zero_ints_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, p, count);
  return hipGetLastError();
}

hipError_t launch_copy_doubles(double *dst, const double *src, size_t count, hipStream_t stream) {
  size_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_doubles_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, dst, src, count);
  return hipGetLastError();
}

}  // namespace dice_hip
