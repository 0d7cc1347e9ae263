// Renders sums of anisotropic Gaussian blobs into a float image (host, OpenMP).
// Used to build the synthetic speckle workloads; not part of the correlation
// path.  Blob k adds amp[k] * exp(-0.5 * d^T A_k d), d = (x - cx[k], y - cy[k]),
// A_k = [[ixx, ixy], [ixy, iyy]] (inverse covariance), over the pixels within
// radius[k] of its centre.  Blobs must be sorted by cy - radius.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" int dice_render_blobs(int w, int h, int nb, const double *cx, const double *cy, const double *ixx,
                                 const double *ixy, const double *iyy, const double *amp, const double *radius,
                                 float *out, int threads) {
  if (w < 1 || h < 1 || nb < 0 || !out) return -1;
  std::vector<double> acc((size_t)w * h, 0.0);
  // y-ranges of the blobs (sorted by their top edge)
  std::vector<double> top(nb);
  double rmax = 0.0;
  for (int k = 0; k < nb; ++k) {
    top[k] = cy[k] - radius[k];
    rmax = std::max(rmax, radius[k]);
  }
  const int band = 16;
  const int nbands = (h + band - 1) / band;
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#endif
  for (int b = 0; b < nbands; ++b) {
    const int y0 = b * band, y1 = std::min(h, y0 + band);
    // blobs whose vertical extent reaches [y0, y1)
    const double lo = y0 - 2.0 * rmax - 1.0;
    const int k0 = (int)(std::lower_bound(top.begin(), top.end(), lo) - top.begin());
    for (int k = k0; k < nb && top[k] < y1; ++k) {
      const int by0 = std::max(y0, (int)std::floor(cy[k] - radius[k]));
      const int by1 = std::min(y1 - 1, (int)std::ceil(cy[k] + radius[k]));
      if (by0 > by1) continue;
      const int bx0 = std::max(0, (int)std::floor(cx[k] - radius[k]));
      const int bx1 = std::min(w - 1, (int)std::ceil(cx[k] + radius[k]));
      for (int y = by0; y <= by1; ++y) {
        const double dy = y - cy[k];
        double *row = acc.data() + (size_t)y * w;
        for (int x = bx0; x <= bx1; ++x) {
          const double dx = x - cx[k];
          const double q = ixx[k] * dx * dx + 2.0 * ixy[k] * dx * dy + iyy[k] * dy * dy;
          if (q < 25.0) row[x] += amp[k] * std::exp(-0.5 * q);
        }
      }
    }
  }
  for (size_t i = 0; i < acc.size(); ++i) out[i] = (float)acc[i];
  (void)threads;
  return 0;
}
