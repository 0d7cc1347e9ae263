// CPU oracle for the DICe local subset-correlation path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
// CPU-baseline leg of bench.py.  The product path never links or calls it.
//
// A from-scratch restatement of the reference algorithm (DICe, NTESS, BSD-3),
// written from its mathematics and control flow.  Each function cites the
// reference file:line it follows.  Templated on S (scalar_t) and T
// (storage_t): S=float reproduces the reference default build, S=double its
// DICE_USE_DOUBLE build.  Expressions keep the reference's operand types
// (double literals promote, stores round to S) so that the float build is
// restated faithfully.  No static state: subsets can run on OpenMP threads.
#include "dice_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <set>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846;     // DICe.h:45
const double kTwoPi = 6.28318530717958647692;  // DICe.h:46

// Status_Flag values, DICe.h:609-677.
enum Status {
  CORRELATION_SUCCESSFUL = 0,
  INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL = 1,
  INITIALIZE_USING_NEIGHBOR_VALUE_SUCCESSFUL = 3,
  INITIALIZE_SUCCESSFUL = 4,
  INITIALIZE_FAILED = 5,
  CORRELATION_FAILED = 8,
  MAX_ITERATIONS_REACHED = 11,
  INITIALIZE_FAILED_BY_EXCEPTION = 12,
  JUMP_TOLERANCE_EXCEEDED = 16,
  HESSIAN_SINGULAR = 21,
  FRAME_FAILED_DUE_TO_HIGH_GAMMA = 23,
  FRAME_FAILED_DUE_TO_NEGATIVE_SIGMA = 24,
};

enum { BILINEAR = 0, BICUBIC = 1, KEYS_FOURTH = 2 };
enum { SIMPLEX = 0, GRADIENT_BASED = 1, GRADIENT_THEN_SEARCH = 2,
       SIMPLEX_THEN_GRADIENT_BASED = 3, GRADIENT_BASED_THEN_SIMPLEX = 4 };

// Keys 4th-order kernel pieces with the reference's truncated decimals
// (DICe_Image.cpp:114-122).  Evaluated in double, returned as S.
template <typename S> inline S keys_f0(const S &s) {
  return 1.33333333333333 * s * s * s - 2.33333333333333 * s * s + 1.0;
}
template <typename S> inline S keys_f1(const S &s) {
  return -0.58333333333333 * s * s * s + 3.0 * s * s - 4.91666666666666 * s + 2.5;
}
template <typename S> inline S keys_f2(const S &s) {
  return 0.08333333333333 * s * s * s - 0.66666666666666 * s * s + 1.75 * s - 1.5;
}

// ---------------------------------------------------------------- Image_ ---
template <typename S, typename T>
struct Image {
  int w = 0, h = 0, ox = 0, oy = 0;
  const T *I = nullptr;
  std::vector<S> gx, gy;
  bool has_gradients = false;

  // Out-of-bounds returns pixel 0 (DICe_Image.h:186-192).
  T at(int x, int y) const {
    if (x < 0 || y < 0 || x >= w || y >= h) return I[0];
    return I[y * w + x];
  }
  S grad_x(int x, int y) const { return gx[y * w + x]; }  // DICe_Image.h:337-340
  S grad_y(int x, int y) const { return gy[y * w + x]; }

  // compute_gradients_finite_difference, DICe_Image.cpp:1064-1094 (c1, c2 at :351-352).
  void compute_gradients_fd() {
    gx.assign((size_t)w * h, S(0));
    gy.assign((size_t)w * h, S(0));
    const S c1 = 1.0 / 12.0;
    const S c2 = -8.0 / 12.0;
    for (int y = 0; y < h; ++y) {
      for (int x = 0; x < w; ++x) {
        if (x < 2)
          gx[y * w + x] = I[y * w + x + 1] - I[y * w + x];
        else if (x >= w - 2)
          gx[y * w + x] = I[y * w + x] - I[y * w + x - 1];
        else
          gx[y * w + x] = c1 * I[y * w + x - 2] + c2 * I[y * w + x - 1] - c2 * I[y * w + x + 1] -
                          c1 * I[y * w + x + 2];
        if (y < 2)
          gy[y * w + x] = I[(y + 1) * w + x] - I[y * w + x];
        else if (y >= h - 2)
          gy[y * w + x] = I[y * w + x] - I[(y - 1) * w + x];
        else
          gy[y * w + x] = c1 * I[(y - 2) * w + x] + c2 * I[(y - 1) * w + x] -
                          c2 * I[(y + 1) * w + x] - c1 * I[(y + 2) * w + x];
      }
    }
    has_gradients = true;
  }

  // smooth_gradients_convolution_5_point, DICe_Image.cpp:1032-1062.
  void smooth_gradients() {
    static const S k[5][5] = {{0.00390625, 0.015625, 0.0234375, 0.015625, 0.00390625},
                              {0.015625, 0.0625, 0.09375, 0.0625, 0.015625},
                              {0.0234375, 0.09375, 0.140625, 0.09375, 0.0234375},
                              {0.015625, 0.0625, 0.09375, 0.0625, 0.015625},
                              {0.00390625, 0.015625, 0.0234375, 0.015625, 0.00390625}};
    std::vector<S> tx = gx, ty = gy;
    for (int y = 2; y < h - 2; ++y)
      for (int x = 2; x < w - 2; ++x) {
        S vx = 0.0, vy = 0.0;
        for (int i = 0; i < 5; ++i)
          for (int j = 0; j < 5; ++j) {
            vx += k[i][j] * tx[(y + i - 2) * w + x + j - 2];
            vy += k[i][j] * ty[(y + i - 2) * w + x + j - 2];
          }
        gx[y * w + x] = vx;
        gy[y * w + x] = vy;
      }
  }

  template <typename A>
  S bilinear_of(const A *arr, const S &lx, const S &ly) const {
    const int x1 = (int)lx, x2 = x1 + 1, y1 = (int)ly, y2 = y1 + 1;
    return arr[y1 * w + x1] * (x2 - lx) * (y2 - ly) + arr[y1 * w + x2] * (lx - x1) * (y2 - ly) +
           arr[y2 * w + x2] * (lx - x1) * (ly - y1) + arr[y2 * w + x1] * (x2 - lx) * (ly - y1);
  }

  // interpolate_bilinear, DICe_Image.cpp:432-445 (zero outside [0, w-1.5) x [0, h-1.5)).
  S bilinear(const S &lx, const S &ly) const {
    if (lx < 0.0 || lx >= w - 1.5 || ly < 0.0 || ly >= h - 1.5) return 0.0;
    return bilinear_of(I, lx, ly);
  }
  // interpolate_bilinear_all, DICe_Image.cpp:392-425.
  void bilinear_all(S &iv, S &gxv, S &gyv, bool cg, const S &lx, const S &ly) const {
    if (lx < 0.0 || lx >= w - 1.5 || ly < 0.0 || ly >= h - 1.5) {
      iv = 0.0;
      if (cg) { gxv = 0.0; gyv = 0.0; }
      return;
    }
    iv = bilinear_of(I, lx, ly);
    if (cg) {
      gxv = bilinear_of(gx.data(), lx, ly);
      gyv = bilinear_of(gy.data(), lx, ly);
    }
  }
  S grad_bilinear(const std::vector<S> &g, const S &lx, const S &ly) const {
    if (lx < 0.0 || lx >= w - 1.5 || ly < 0.0 || ly >= h - 1.5) return 0.0;
    return bilinear_of(g.data(), lx, ly);
  }

  void keys_coeffs(const S &lx, const S &ly, int &ix, int &iy, S cx[6], S cy[6]) const {
    ix = (int)lx;
    iy = (int)ly;
    const S dx = lx - ix, dy = ly - iy;
    cx[0] = keys_f2<S>(dx + 2.0); cx[1] = keys_f1<S>(dx + 1.0); cx[2] = keys_f0<S>(dx);
    cx[3] = keys_f0<S>(1.0 - dx); cx[4] = keys_f1<S>(2.0 - dx); cx[5] = keys_f2<S>(3.0 - dx);
    cy[0] = keys_f2<S>(dy + 2.0); cy[1] = keys_f1<S>(dy + 1.0); cy[2] = keys_f0<S>(dy);
    cy[3] = keys_f0<S>(1.0 - dy); cy[4] = keys_f1<S>(2.0 - dy); cy[5] = keys_f2<S>(3.0 - dy);
  }

  // interpolate_keys_fourth_all, DICe_Image.cpp:832-883.  Near the border the
  // reference writes the bilinear value and then ADDS the 6x6 Keys sum (no
  // return); that fall-through is kept.  Taps outside the image (undefined
  // reads in the reference) contribute 0 here.  Subset::initialize never
  // reaches the border branch because of its 4-px guard.
  void keys_all(S &iv, S &gxv, S &gyv, bool cg, const S &lx, const S &ly) const {
    iv = 0.0;
    if (cg) { gxv = 0.0; gyv = 0.0; }
    if (lx <= 2.5 || lx >= w - 3.5 || ly <= 2.5 || ly >= h - 3.5) {
      iv = bilinear(lx, ly);
      if (cg) { gxv = grad_bilinear(gx, lx, ly); gyv = grad_bilinear(gy, lx, ly); }
    }
    int ix, iy;
    S cx[6], cy[6];
    keys_coeffs(lx, ly, ix, iy, cx, cy);
    for (int m = 0; m < 6; ++m) {
      for (int n = 0; n < 6; ++n) {
        const int xx = ix - 2 + n, yy = iy - 2 + m;
        if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
        const S cc = cy[m] * cx[n];
        iv += cc * I[yy * w + xx];
        if (cg) {
          gxv += cc * gx[yy * w + xx];
          gyv += cc * gy[yy * w + xx];
        }
      }
    }
  }

  // interpolate_keys_fourth, DICe_Image.cpp:890-924 (bilinear near the border).
  S keys(const S &lx, const S &ly) const {
    if (lx <= 2.5 || lx >= w - 3.5 || ly <= 2.5 || ly >= h - 3.5) return bilinear(lx, ly);
    int ix, iy;
    S cx[6], cy[6];
    keys_coeffs(lx, ly, ix, iy, cx, cy);
    S value = 0.0;
    for (int m = 0; m < 6; ++m)
      for (int n = 0; n < 6; ++n) value += cy[m] * cx[n] * I[(iy - 2 + m) * w + ix - 2 + n];
    return value;
  }

  void interp_all(int method, S &iv, S &gxv, S &gyv, bool cg, const S &lx, const S &ly) const {
    if (method == BILINEAR) bilinear_all(iv, gxv, gyv, cg, lx, ly);
    else keys_all(iv, gxv, gyv, cg, lx, ly);  // BICUBIC is not on this path
  }
};

// ------------------------------------------------- Affine_Shape_Function ---
template <typename S>
struct Affine {
  S p[6] = {0, 0, 0, 0, 0, 0};
  S deltas[6] = {0, 0, 0, 0, 0, 0};
  int n = 2;
  int iu = 0, iv = 1, it = -1, iex = -1, iey = -1, ig = -1;

  // Affine_Shape_Function::init, DICe_LocalShapeFunction.cpp:170-218:
  // parameter slots follow insertion order u, v, [theta], [ex, ey], [gxy].
  void init(bool rot, bool normal, bool shear, S delta_disp, S delta_theta) {
    n = 2;
    if (rot) it = n++;
    if (normal) { iex = n++; iey = n++; }
    if (shear) ig = n++;
    for (int i = 0; i < n; ++i) { p[i] = 0.0; deltas[i] = delta_theta; }
    deltas[iu] = delta_disp;
    deltas[iv] = delta_disp;
  }
  S theta() const { return it >= 0 ? p[it] : S(0.0); }
  S ex() const { return iex >= 0 ? p[iex] : S(0.0); }
  S ey() const { return iey >= 0 ? p[iey] : S(0.0); }
  S gxy() const { return ig >= 0 ? p[ig] : S(0.0); }

  // map, DICe_LocalShapeFunction.cpp:220-248.
  void map(const S &x, const S &y, const S &cx, const S &cy, S &out_x, S &out_y) const {
    const S dispx = p[iu], dispy = p[iv];
    const S th = theta(), dudx = ex(), dvdy = ey(), g = gxy();
    const S cost = std::cos(th), sint = std::sin(th);
    const S dx = x - cx, dy = y - cy;
    const S Dx = (1.0 + dudx) * dx + g * dy;
    const S Dy = (1.0 + dvdy) * dy + g * dx;
    out_x = cost * Dx - sint * Dy + dispx + cx;
    out_y = sint * Dx + cost * Dy + dispy + cy;
  }

  // residuals, DICe_LocalShapeFunction.cpp:350-396.
  void residuals(const S &x, const S &y, const S &cx, const S &cy, const S &gx, const S &gy, S *r,
                 bool use_ref_grads) const {
    const S th = theta(), dudx = ex(), dvdy = ey(), g = gxy();
    const S c = std::cos(th), s = std::sin(th);
    const S dx = x - cx, dy = y - cy;
    const S Dx = (1.0 + dudx) * (dx) + g * (dy);
    const S Dy = (1.0 + dvdy) * (dy) + g * (dx);
    const S Gx = use_ref_grads ? c * gx - s * gy : gx;
    const S Gy = use_ref_grads ? s * gx + c * gy : gy;
    r[iu] = Gx;
    r[iv] = Gy;
    if (it >= 0) r[it] = Gx * (-s * Dx - c * Dy) + Gy * (c * Dx - s * Dy);
    if (iex >= 0) r[iex] = Gx * dx * c + Gy * dx * s;
    if (iey >= 0) r[iey] = -Gx * dy * s + Gy * dy * c;
    if (ig >= 0) r[ig] = Gx * (c * dy - s * dx) + Gy * (s * dy + c * dx);
  }

  // test_for_convergence, DICe_LocalShapeFunction.cpp:433-459: only u, v, theta.
  bool converged(const S *old, const S &tol) const {
    bool conv = true;
    if (std::abs(p[iu] - old[iu]) >= tol) conv = false;
    if (std::abs(p[iv] - old[iv]) >= tol) conv = false;
    if (it >= 0 && std::abs(p[it] - old[it]) >= tol) conv = false;
    return conv;
  }
};

// ---------------------------------------------------------------- Subset ---
template <typename S>
struct Subset {
  int cx = 0, cy = 0, n = 0;
  std::vector<int> x, y;
  std::vector<S> ref, def, gx, gy;
  std::vector<char> active, deact;
  bool has_gradients = false;
  // obstructed_coords_ and pixels_blocked_by_other_subsets_ (DICe_Subset.h), merged: image pixels
  // (y, x) that deactivate a warped pixel rounding onto them
  std::set<std::pair<int, int>> blocked;

  void alloc() {
    ref.assign(n, S(0)); def.assign(n, S(0)); gx.assign(n, S(0)); gy.assign(n, S(0));
    active.assign(n, 1); deact.assign(n, 0);
  }
  // Square constructor, DICe_Subset.cpp:78-113 (odd size, negative coords clamp to 0).
  void make_square(int cx_, int cy_, int size) {
    cx = cx_; cy = cy_;
    const int hw = size / 2, hh = size / 2;
    n = (2 * hw + 1) * (2 * hh + 1);
    x.resize(n); y.resize(n);
    int idx = 0;
    for (int yy = cy - hh; yy <= cy + hh; ++yy)
      for (int xx = cx - hw; xx <= cx + hw; ++xx) {
        x[idx] = xx < 0 ? 0 : xx;
        y[idx] = yy < 0 ? 0 : yy;
        ++idx;
      }
    alloc();
  }
  // Conformal constructor from an explicit pixel list, DICe_Subset.cpp:115-168.
  void make_list(int cx_, int cy_, int count, const int *px, const int *py, const unsigned char *act) {
    cx = cx_; cy = cy_; n = count;
    x.resize(n); y.resize(n);
    for (int i = 0; i < n; ++i) { x[i] = px[i] < 0 ? 0 : px[i]; y[i] = py[i] < 0 ? 0 : py[i]; }
    alloc();
    if (act) for (int i = 0; i < n; ++i) active[i] = act[i] ? 1 : 0;
  }
  bool on(int i) const { return active[i] && !deact[i]; }

  // initialize(REF_INTENSITIES, no map), DICe_Subset.cpp:387-391,428-436.
  template <typename T>
  void init_ref(const Image<S, T> &im) {
    for (int i = 0; i < n; ++i) ref[i] = im.at(x[i] - im.ox, y[i] - im.oy);
    if (im.has_gradients) {
      for (int i = 0; i < n; ++i) {
        gx[i] = im.grad_x(x[i] - im.ox, y[i] - im.oy);
        gy[i] = im.grad_y(x[i] - im.ox, y[i] - im.oy);
      }
      has_gradients = true;
    }
  }
  // initialize(DEF_INTENSITIES, no map): plain copy, DICe_Subset.cpp:387-391.
  template <typename T>
  void init_def_nomap(const Image<S, T> &im) {
    for (int i = 0; i < n; ++i) def[i] = im.at(x[i] - im.ox, y[i] - im.oy);
  }
  // initialize(DEF_INTENSITIES, map, interp), DICe_Subset.cpp:392-427.
  template <typename T>
  void init_def(const Image<S, T> &im, const Affine<S> &sf, int interp) {
    const int ox = im.ox, oy = im.oy, w = im.w, h = im.h;
    const S sox = (S)ox, soy = (S)oy;
    const bool cg = im.has_gradients;
    S mx = 0.0, my = 0.0;
    for (int i = 0; i < n; ++i) {
      sf.map((S)x[i], (S)y[i], (S)cx, (S)cy, mx, my);
      const int px = ((int)(mx + 0.5) == (int)(mx)) ? (int)(mx) : (int)(mx) + 1;
      const int py = ((int)(my + 0.5) == (int)(my)) ? (int)(my) : (int)(my) + 1;
      if (px < ox + 4 || px >= ox + w - 4 || py < oy + 4 || py >= oy + h - 4) {
        deact[i] = 1;
        continue;
      }
      if (!blocked.empty()) {
        // is_obstructed_pixel (:447-460): c = (int)x, +1 when the fraction is >= 0.5; the blocked-pixel
        // test (:412-419) uses (px, py), the same pixel for the positions the guard lets through
        int c_x = (int)mx, c_y = (int)my;
        if (mx - (int)mx >= 0.5) c_x++;
        if (my - (int)my >= 0.5) c_y++;
        if (blocked.count(std::make_pair(c_y, c_x)) || blocked.count(std::make_pair(py, px))) {
          deact[i] = 1;
          continue;
        }
      }
      deact[i] = 0;
      im.interp_all(interp, def[i], gx[i], gy[i], cg, mx - sox, my - soy);
    }
  }
  int num_active() const {  // DICe_Subset.cpp:597-605
    int c = 0;
    for (int i = 0; i < n; ++i) if (on(i)) ++c;
    return c;
  }
  // mean(target), DICe_Subset.cpp:285-301.
  S mean(bool use_def) const {
    S m = 0.0;
    const int na = num_active();
    const std::vector<S> &v = use_def ? def : ref;
    for (int i = 0; i < n; ++i) if (on(i)) m += v[i];
    return na != 0 ? m / na : 0.0;
  }
  // mean(target, sum), DICe_Subset.cpp:303-321.
  S mean(bool use_def, S &sum) const {
    const S m = mean(use_def);
    const std::vector<S> &v = use_def ? def : ref;
    sum = 0.0;
    for (int i = 0; i < n; ++i) if (on(i)) sum += (v[i] - m) * (v[i] - m);
    sum = std::sqrt(sum);
    return m;
  }
  // gamma (ZNSSD), DICe_Subset.cpp:323-340.
  S gamma() const {
    S sr = 0.0, sd = 0.0;
    const S mr = mean(false, sr);
    const S md = mean(true, sd);
    if (sr == 0.0 || sd == 0.0) return -1.0;
    S g = 0.0, v = 0.0;
    for (int i = 0; i < n; ++i)
      if (on(i)) {
        v = (def[i] - md) / sd - (ref[i] - mr) / sr;
        g += v * v;
      }
    return g;
  }
  // contrast_std_dev, DICe_Subset.cpp:607-620.
  S contrast_std_dev() const {
    const S m = mean(true);
    S sd = 0.0;
    int na = 0;
    for (int i = 0; i < n; ++i)
      if (on(i)) { ++na; sd += (def[i] - m) * (def[i] - m); }
    sd = na == 0 ? 0.0 : std::sqrt(sd / na);
    return sd;
  }
  // noise_std_dev, DICe_Subset.cpp:622-703 (Immerkaer 3x3 mask over the bbox
  // shifted by (u, v); note the reference resets min_x/min_y, not max, when
  // the box leaves the image, and reads pixel 0 out of bounds).
  template <typename T>
  S noise_std_dev(const Image<S, T> &im, const Affine<S> &sf) const {
    static const S mask[3][3] = {{1, -2, 1}, {-2, 4, -2}, {1, -2, 1}};
    int min_x = cx, max_x = cx, min_y = cy, max_y = cy, ndeact = 0;
    for (int i = 0; i < n; ++i) {
      if (deact[i] || !active[i]) { ++ndeact; continue; }
      if (x[i] < min_x) min_x = x[i];
      if (x[i] > max_x) max_x = x[i];
      if (y[i] < min_y) min_y = y[i];
      if (y[i] > max_y) max_y = y[i];
    }
    if (ndeact == n) return -1.0;
    const S u = sf.p[sf.iu], v = sf.p[sf.iv];
    min_x += u; max_x += u;
    min_y += v; max_y += v;
    const int ih = im.h, iw = im.w, ox = im.ox, oy = im.oy;
    if (min_x < ox) min_x = ox;
    if (min_y < oy) min_y = oy;
    if (max_x >= iw + ox) min_x = 0;
    if (max_y >= ih + oy) min_y = 0;
    int hh = max_y - min_y + 1, ww = max_x - min_x + 1;
    if (ww < 3) ww = 3;
    if (hh < 3) hh = 3;
    S variance = 0.0, conv = 0.0;
    for (int yy = min_y; yy < max_y; ++yy)
      for (int xx = min_x; xx < max_x; ++xx) {
        if (xx - ox < 1 || xx - ox >= iw - 1 || yy - oy < 1 || yy - oy >= ih - 1) {
          variance += std::sqrt(im.at(xx - ox, yy - oy) * im.at(xx - ox, yy - oy));
        } else {
          conv = 0.0;
          for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) conv += im.at(xx - ox + (i - 1), yy - oy + (j - 1)) * mask[i][j];
          variance += std::abs(conv);
        }
      }
    variance *= std::sqrt(0.5 * kPi) / (6.0 * (ww - 2) * (hh - 2));
    return variance;
  }
};

// ------------------------------------------------------- GETRF / GETRI ----
// LU with partial pivoting (first maximal |pivot|), unblocked like LAPACK
// dgetf2, then the explicit inverse.  INFO>0 (exactly zero pivot) leaves the
// matrix holding its LU factors, as GETRI does (DICe_Objective.cpp:571,591;
// INFO is never checked there).
int lu_inverse(int n, double A[6][6]) {
  int piv[6];
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = std::abs(A[j][j]);
    for (int i = j + 1; i < n; ++i)
      if (std::abs(A[i][j]) > best) { best = std::abs(A[i][j]); p = i; }
    piv[j] = p;
    if (A[p][j] != 0.0) {
      if (p != j) for (int k = 0; k < n; ++k) std::swap(A[p][k], A[j][k]);
      const double r = 1.0 / A[j][j];
      for (int i = j + 1; i < n; ++i) A[i][j] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int i = j + 1; i < n; ++i)
      for (int k = j + 1; k < n; ++k) A[i][k] -= A[i][j] * A[j][k];
  }
  if (info != 0) return info;
  // Solve A X = I with the factors: P A = L U.
  double X[6][6];
  for (int c = 0; c < n; ++c) {
    double b[6];
    for (int i = 0; i < n; ++i) b[i] = 0.0;
    b[c] = 1.0;
    for (int j = 0; j < n; ++j) if (piv[j] != j) std::swap(b[j], b[piv[j]]);
    for (int i = 0; i < n; ++i) {  // L y = P b (unit lower)
      double s = b[i];
      for (int k = 0; k < i; ++k) s -= A[i][k] * b[k];
      b[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {  // U x = y
      double s = b[i];
      for (int k = i + 1; k < n; ++k) s -= A[i][k] * b[k];
      b[i] = s / A[i][i];
    }
    for (int i = 0; i < n; ++i) X[i][c] = b[i];
  }
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) A[i][k] = X[i][k];
  return 0;
}

// ------------------------------------------------------------- Objective ---
// Stop rules of nelder_mead (bits of *stop).
enum { NM_EARLY = 1, NM_RTOL = 2, NM_RTOL_DIFF = 4, NM_REPEATS = 8, NM_MAX_ITERATIONS = 16 };

// Simplex::minimize, DICe_Simplex.cpp:60-277, over nd variables with a generic
// objective f(variables).  As in the reference the variables array is the one
// the objective reads: every evaluation first writes its point into it, so on
// exit it holds the LAST evaluated point, not necessarily the best vertex.
// stop (may be null) receives the NM_* bits of the rules that ended the loop.
template <typename S, typename F>
int nelder_mead(int nd, S *var, const S *deltas, F &&f, int max_iterations, double tolerance, S threshold,
                int &num_iterations, int *stop) {
  const S tiny = 1.0E-10;
  if (stop) *stop = 0;
  S init[6], pts[7][6], gv[7], sums[6], ptry[6];
  auto eval = [&](const S *pt) {
    for (int j = 0; j < nd; ++j) var[j] = pt[j];
    return f(var);
  };
  for (int i = 0; i < nd; ++i) init[i] = var[i];
  const int mpts = nd + 1;
  for (int i = 0; i < mpts; ++i) {
    for (int j = 0; j < nd; ++j) pts[i][j] = var[j];
    if (i > 0) pts[i][i - 1] += deltas[i - 1];
    gv[i] = eval(pts[i]);
    if (i == 0 && gv[i] < threshold && gv[i] >= 0.0) {
      num_iterations = 0;
      if (stop) *stop = NM_EARLY;
      return CORRELATION_SUCCESSFUL;
    }
    for (int j = 0; j < nd; ++j) var[j] = init[j];
  }
  for (int j = 0; j < nd; ++j) {
    S s = 0.0;
    for (int i = 0; i < mpts; ++i) s += pts[i][j];
    sums[j] = s;
  }
  S old_rtol = 0.0, gamma_new = 0.0, gamma_old = -1.0;
  int repeats = 0;
  const int repeats_allowed = 15;
  const S gamma_tol = 1.0E-8;
  // the trial point sums * (1 - fac) / nd - pts[ihi] * ((1 - fac) / nd - fac), evaluated;
  // it replaces the worst vertex when better
  auto trial = [&](int ihi, S fac) {
    const S fac1 = (1.0 - fac) / nd, fac2 = fac1 - fac;
    for (int j = 0; j < nd; j++) ptry[j] = sums[j] * fac1 - pts[ihi][j] * fac2;
    const S ytry = eval(ptry);
    if (ytry < gv[ihi]) {
      gv[ihi] = ytry;
      for (int j = 0; j < nd; j++) { sums[j] += ptry[j] - pts[ihi][j]; pts[ihi][j] = ptry[j]; }
    }
    return ytry;
  };
  int iteration = 0;
  for (iteration = 0; iteration < max_iterations; iteration++) {
    if (iteration >= max_iterations - 1) {
      num_iterations = iteration;
      if (stop) *stop = NM_MAX_ITERATIONS;
      return MAX_ITERATIONS_REACHED;
    }
    int inhi;
    int ilo = 0;
    int ihi = gv[0] > gv[1] ? (inhi = 1, 0) : (inhi = 0, 1);
    for (int i = 0; i < mpts; i++) {
      if (gv[i] <= gv[ilo]) ilo = i;
      if (gv[i] > gv[ihi]) { inhi = ihi; ihi = i; }
      else if (gv[i] > gv[inhi] && i != ihi) inhi = i;
    }
    const S rtol = 2.0 * std::abs(gv[ihi] - gv[ilo]) / (std::abs(gv[ihi]) + std::abs(gv[ilo]) + tiny);
    const S rtol_diff = std::abs(rtol - old_rtol);
    old_rtol = rtol;
    if (std::abs(gamma_new - gamma_old) < gamma_tol) repeats++;
    gamma_old = gamma_new;
    if (rtol < tolerance || rtol_diff < tolerance || repeats > repeats_allowed) {
      if (stop) *stop = (rtol < tolerance ? NM_RTOL : 0) | (rtol_diff < tolerance ? NM_RTOL_DIFF : 0) |
                        (repeats > repeats_allowed ? NM_REPEATS : 0);
      std::swap(gv[0], gv[ilo]);
      for (int i = 0; i < nd; i++) std::swap(pts[0][i], pts[ilo][i]);
      break;
    }
    S ytry = trial(ihi, -1.0);       // reflect
    if (ytry <= gv[ilo]) {
      trial(ihi, 2.0);               // expand
    } else if (ytry >= gv[inhi]) {
      const S ysave = gv[ihi];
      ytry = trial(ihi, 0.5);        // contract
      if (ytry >= ysave) {           // shrink toward the best vertex
        for (int i = 0; i < mpts; i++) {
          if (i != ilo) {
            for (int j = 0; j < nd; j++) pts[i][j] = sums[j] = 0.5 * (pts[i][j] + pts[ilo][j]);
            gv[i] = eval(sums);
          }
        }
        for (int j = 0; j < nd; j++) {
          S s = 0.0;
          for (int i = 0; i < mpts; i++) s += pts[i][j];
          sums[j] = s;
        }
      }
    }
    gamma_new = f(var);  // the extra evaluation at the last point (DICe_Simplex.cpp:264)
  }
  num_iterations = iteration;
  return CORRELATION_SUCCESSFUL;
}

template <typename S, typename T>
struct Objective {
  const oracle_params *P;
  const Image<S, T> *refi, *defi;
  Subset<S> *sub;
  double *F;  // field store
  int nsub, gid;

  double &field(int slot) { return F[(size_t)slot * nsub + gid]; }

  // Objective::gamma, DICe_Objective.cpp:58-76.
  S gamma(const Affine<S> &sf) {
    sub->init_def(*defi, sf, P->interpolation_method);
    S g = sub->gamma();
    if (P->normalize_gamma_with_active_pixels) {
      int na = 0;
      for (int i = 0; i < sub->n; ++i) if (sub->active[i]) ++na;
      if (na > 0) g /= na == 0.0 ? 1.0 : na;
    }
    return g;
  }

  // Objective::sigma, DICe_Objective.cpp:150-184.
  S sigma(const Affine<S> &sf, S &noise) {
    if (!sub->has_gradients) return 0.0;
    noise = sub->noise_std_dev(*defi, sf);
    if (noise < 0.0) return -1.0;
    S sgx = 0.0, sgy = 0.0;
    for (int i = 0; i < sub->n; ++i) {
      if (!sub->active[i] || sub->deact[i]) continue;
      sgx += sub->gx[i] * sub->gx[i];
      sgy += sub->gy[i] * sub->gy[i];
    }
    const S sg = sgx > sgy ? sgy : sgx;
    const S sig = sg > 0.0 ? std::sqrt(2.0 * noise * noise / sg) : -1.0;
    return sig;
  }

  // computeUncertaintyFields (no image deformer), DICe_Objective.cpp:193-299.
  void uncertainty(const Affine<S> &sf) {
    S n_uhat2 = 0.0, n_err_g2 = 0.0, n_err_jg2 = 0.0, n_uhat_g2 = 0.0, r_tot2 = 0.0;
    S n_ut2 = 0.0, n_ut_g2 = 0.0;
    S sssig = 1.0;
    const S u = sf.p[sf.iu], v = sf.p[sf.iv];
    const int ox = refi->ox, oy = refi->oy;
    for (int i = 0; i < sub->n; ++i) {
      const S x = sub->x[i], y = sub->y[i];
      const S bx = 0.0, by = 0.0;
      const S gx = sub->gx[i], gy = sub->gy[i];
      sssig += gx * gx + gy * gy;
      const S mag_ut2 = bx * bx + by * by;
      n_ut2 += mag_ut2;
      const S mag_uhat2 = u * u + v * v;
      n_uhat2 += mag_uhat2;
      const S inv_g2 = gx * gx + gy * gy < 1.0E-4 ? 0.0 : 1.0 / (gx * gx + gy * gy);
      n_ut_g2 += (bx * gx + by * gy) * (bx * gx + by * gy) * inv_g2;
      n_uhat_g2 += (u * gx + v * gy) * (u * gx + v * gy) * inv_g2;
      n_err_g2 += ((u - bx) * gx + (v - by) * gy) * ((u - bx) * gx + (v - by) * gy) * inv_g2;
      n_err_jg2 += ((v - by) * gx - (u - bx) * gy) * ((v - by) * gx - (u - bx) * gy) * inv_g2;
      const S refv = refi->at((int)(x - ox), (int)(y - oy));
      const S sub_r = defi->keys(x - ox + u, y - oy + v) - refv;
      const S sub_r_exact = defi->keys(x - ox + bx, y - oy + by) - refv;
      r_tot2 += (sub_r - sub_r_exact) * (sub_r - sub_r_exact) * inv_g2;
    }
    sssig /= sub->n == 0.0 ? 1.0 : sub->n;
    const S cos_hat = n_uhat2 == 0.0 ? 1.0 : std::sqrt(n_uhat_g2 / n_uhat2);
    field(OF_FIELD_1) = cos_hat;
    const S cos_t = n_ut2 == 0.0 ? 1.0 : std::sqrt(n_ut_g2 / n_ut2);
    field(OF_FIELD_2) = cos_t;
    field(OF_FIELD_3) = std::sqrt(n_err_g2 + n_err_jg2);
    field(OF_FIELD_4) = cos_hat == 0.0 ? 0.0 : 1.0 / cos_hat * std::sqrt(r_tot2);
    field(OF_FIELD_5) = std::sqrt(n_err_g2);
    field(OF_FIELD_6) = std::sqrt(r_tot2);
    field(OF_FIELD_7) = sssig;
  }

  // Objective_ZNSSD::computeUpdateFast, DICe_Objective.cpp:430-657.
  int update_fast(Affine<S> &sf, int &num_iterations) {
    const int N = sf.n;
    const S tol = P->fast_solver_tolerance;
    const int max_its = P->max_solver_iterations_fast;
    double H[6][6], q[6];
    S r[6], def_old[6], def_update[6], def_update_old[6];
    for (int i = 0; i < 6; ++i) { q[i] = 0.0; def_update_old[i] = 0.0; for (int j = 0; j < 6; ++j) H[i][j] = 0.0; }
    const S cx = sub->cx, cy = sub->cy;
    const S meanF = sub->mean(false);
    int it = 0;
    for (; it <= max_its; ++it) {
      num_iterations = it;
      sub->init_def(*defi, sf, P->interpolation_method);
      const S meanG = sub->mean(true);
      const bool use_ref_grads = defi->has_gradients ? false : true;
      S GmF = 0.0;
      for (int k = 0; k < sub->n; ++k) {
        if (sub->deact[k] || !sub->active[k]) continue;
        GmF = (sub->def[k] - meanG) - (sub->ref[k] - meanF);
        for (int i = 0; i < N; ++i) r[i] = 0.0;
        sf.residuals((S)sub->x[k], (S)sub->y[k], cx, cy, sub->gx[k], sub->gy[k], r, use_ref_grads);
        for (int i = 0; i < N; ++i) {
          q[i] += GmF * r[i];
          for (int j = 0; j < N; ++j) H[i][j] += r[i] * r[j];
        }
      }
      if (P->levenberg_marquardt_regularization_factor > 0.0) {
        const S alpha = P->levenberg_marquardt_regularization_factor;
        H[0][0] += alpha;
        H[1][1] += alpha;
      }
      const S det_h = H[0][0] * H[1][1] - H[1][0] * H[0][1];
      const S norm_H = std::sqrt(H[0][0] * H[0][0] + H[0][1] * H[0][1] + H[1][0] * H[1][0] + H[1][1] * H[1][1]);
      S cond = -1.0;
      if (det_h != 0.0) {
        const S norm_Hi = det_h == 0.0 ? 0.0
            : std::sqrt((1.0 / (det_h * det_h)) * (H[0][0] * H[0][0] + H[0][1] * H[0][1] + H[1][0] * H[1][0] + H[1][1] * H[1][1]));
        cond = norm_H * norm_Hi;
      }
      field(OF_CONDITION) = cond;
      if (cond > 1.0E12) return HESSIAN_SINGULAR;
      lu_inverse(N, H);
      for (int i = 0; i < N; ++i) def_old[i] = sf.p[i];
      for (int i = 0; i < N; ++i) def_update[i] = 0.0;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) def_update[i] += H[i][j] * (-1.0) * q[j];
      if (P->momentum_factor != 0.0) {
        const S mom = P->momentum_factor;
        for (int i = 0; i < N; ++i) { def_update[i] += mom * def_update_old[i]; def_update_old[i] = def_update[i]; }
      }
      for (int i = 0; i < N; ++i) sf.p[i] += def_update[i];
      if (sf.converged(def_old, tol)) {
        if (P->compute_uncertainty_fields) uncertainty(sf);
        break;
      }
      for (int i = 0; i < 6; ++i) { q[i] = 0.0; for (int j = 0; j < 6; ++j) H[i][j] = 0.0; }
    }
    if (it > max_its) return MAX_ITERATIONS_REACHED;
    return CORRELATION_SUCCESSFUL;
  }

  // Simplex::minimize on Subset_Simplex (objective = Objective::gamma),
  // DICe_Simplex.cpp:60-277,279-302, via computeUpdateRobust (DICe_Objective.cpp:369-388).
  // The variables are the shape-function parameters themselves, so on exit
  // they hold the LAST evaluated point, not necessarily the best vertex.
  int update_robust(Affine<S> &sf, int &num_iterations) {
    auto objective = [&](const S *) { return gamma(sf); };
    return nelder_mead<S>(sf.n, sf.p, sf.deltas, objective, P->max_solver_iterations_robust,
                          P->robust_solver_tolerance, (S)P->skip_solve_gamma_threshold, num_iterations, nullptr);
  }
};

template <typename S>
void init_from_fields(Affine<S> &sf, double *F, int n, int gid) {
  // Affine initialize_parameters_from_fields (DISPLACEMENT_BASED), DICe_LocalShapeFunction.cpp:251-312.
  auto f = [&](int slot) { return F[(size_t)slot * n + gid]; };
  sf.p[sf.iu] = f(OF_DISP_X);
  sf.p[sf.iv] = f(OF_DISP_Y);
  if (sf.it >= 0) sf.p[sf.it] = f(OF_ROT_Z);
  if (sf.iex >= 0) { sf.p[sf.iex] = f(OF_NS_XX); sf.p[sf.iey] = f(OF_NS_YY); }
  if (sf.ig >= 0) sf.p[sf.ig] = f(OF_SS_XY);
}

template <typename S>
void save_fields(const Affine<S> &sf, double *F, int n, int gid) {
  // Local_Shape_Function::save_fields, DICe_LocalShapeFunction.cpp:65-74.
  F[(size_t)OF_DISP_X * n + gid] = sf.p[sf.iu];
  F[(size_t)OF_DISP_Y * n + gid] = sf.p[sf.iv];
  if (sf.it >= 0) F[(size_t)OF_ROT_Z * n + gid] = sf.p[sf.it];
  if (sf.iex >= 0) { F[(size_t)OF_NS_XX * n + gid] = sf.p[sf.iex]; F[(size_t)OF_NS_YY * n + gid] = sf.p[sf.iey]; }
  if (sf.ig >= 0) F[(size_t)OF_SS_XY * n + gid] = sf.p[sf.ig];
}

void record_failed(double *F, int n, int gid, int status, int iters) {
  // Schema::record_failed_step, DICe_Schema.cpp:1958-1984.
  auto f = [&](int slot) -> double & { return F[(size_t)slot * n + gid]; };
  f(OF_SIGMA) = -1.0; f(OF_MATCH) = -1.0; f(OF_GAMMA) = -1.0; f(OF_BETA) = -1.0;
  f(OF_NOISE) = -1.0; f(OF_CONTRAST) = -1.0; f(OF_ACTIVE_PIXELS) = -1.0;
  f(OF_STATUS) = status; f(OF_ITERATIONS) = iters;
}

// Schema::generic_correlation_routine (GENERIC_ROUTINE), DICe_Schema.cpp:2025-2368.  sid: the subset
// whose fields give the initial guess (Field_Value_Initializer::initial_guess,
// DICe_Initializer.cpp:398-434): gid itself, or its neighbour with USE_NEIGHBOR_VALUES.
// Per-subset options beyond the subset set: Schema::set_force_simplex and the blocking ids
// (generic_correlation_routine, DICe_Schema.cpp:2160-2185) as force[i] != 0, and the blocked image
// pixels of subset i, mask_x/mask_y[mask_off[i] .. mask_off[i + 1]) (Subset::obstructed_coords_ and
// pixels_blocked_by_other_subsets_).  Either may be null.
struct Extras {
  const unsigned char *force = nullptr;
  const int *mask_off = nullptr, *mask_x = nullptr, *mask_y = nullptr;
};

template <typename S, typename T>
void correlate_one(const oracle_params *P, const Image<S, T> &refi, const Image<S, T> &defi,
                   Subset<S> &sub, double *F, int n, int gid, int subset_size, bool first_frame, int sid,
                   bool force_simplex = false) {
  auto f = [&](int slot) -> double & { return F[(size_t)slot * n + gid]; };
  // Field-of-view check on the first frame for square subsets (:2033-2049).
  if (subset_size > 0 && first_frame) {
    const S px = f(OF_COORD_X) + f(OF_DISP_X);
    const S py = f(OF_COORD_Y) + f(OF_DISP_Y);
    if (px < defi.ox + subset_size / 2 || px > defi.w + defi.ox - subset_size / 2 ||
        py < defi.oy + subset_size / 2 || py > defi.h + defi.oy - subset_size / 2) {
      record_failed(F, n, gid, INITIALIZE_FAILED_BY_EXCEPTION, -1);
      return;
    }
  }
  Objective<S, T> obj{P, &refi, &defi, &sub, F, n, gid};
  Affine<S> sf;
  sf.init(P->enable_rotation != 0, P->enable_normal_strain != 0, P->enable_shear_strain != 0,
          (S)P->robust_delta_disp, (S)P->robust_delta_theta);
  int num_iterations = -1;
  // Field_Value_Initializer::initial_guess, DICe_Initializer.cpp:398-434.
  const int guess_status = sid == gid ? INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL
                                      : INITIALIZE_USING_NEIGHBOR_VALUE_SUCCESSFUL;
  if (F[(size_t)OF_SIGMA * n + sid] == -1.0) { record_failed(F, n, gid, INITIALIZE_FAILED, num_iterations); return; }
  init_from_fields(sf, F, n, sid);
  int init_status = guess_status;
  if (P->initial_gamma_threshold != -1.0) {
    const S g0 = obj.gamma(sf);
    if (g0 > P->initial_gamma_threshold || g0 < 0.0) { record_failed(F, n, gid, INITIALIZE_FAILED, num_iterations); return; }
  }
  const S prev_u = f(OF_DISP_X), prev_v = f(OF_DISP_Y), prev_t = f(OF_ROT_Z);
  const int om = P->optimization_method;
  int corr = CORRELATION_FAILED;
  if (om == SIMPLEX || om == SIMPLEX_THEN_GRADIENT_BASED || force_simplex) corr = obj.update_robust(sf, num_iterations);
  else corr = obj.update_fast(sf, num_iterations);
  auto jump_fail = [&]() {
    const S du = sf.p[sf.iu] - prev_u, dv = sf.p[sf.iv] - prev_v, dt = sf.theta() - prev_t;
    return std::abs(du) > P->disp_jump_tol || std::abs(dv) > P->disp_jump_tol || std::abs(dt) > P->theta_jump_tol;
  };
  if (corr != CORRELATION_SUCCESSFUL || jump_fail()) {
    bool second_failed = false;
    if (om == SIMPLEX || om == GRADIENT_BASED || force_simplex) {
      second_failed = true;
    } else if (om == GRADIENT_BASED_THEN_SIMPLEX || om == GRADIENT_THEN_SEARCH) {
      init_from_fields(sf, F, n, sid);
      init_status = guess_status;
      if (om == GRADIENT_BASED_THEN_SIMPLEX) corr = obj.update_robust(sf, num_iterations);
      else corr = CORRELATION_FAILED;  // search initializer: out of scope
      if (corr != CORRELATION_SUCCESSFUL) second_failed = true;
    } else if (om == SIMPLEX_THEN_GRADIENT_BASED) {
      init_from_fields(sf, F, n, sid);
      init_status = guess_status;
      corr = obj.update_fast(sf, num_iterations);
      if (corr != CORRELATION_SUCCESSFUL) second_failed = true;
    }
    if (second_failed && corr != CORRELATION_SUCCESSFUL) { record_failed(F, n, gid, corr, num_iterations); return; }
  }
  S noise = 0.0;
  const S sig = obj.sigma(sf, noise);
  if (sig < 0.0) { record_failed(F, n, gid, FRAME_FAILED_DUE_TO_NEGATIVE_SIGMA, num_iterations); return; }
  const S gam = obj.gamma(sf);
  const S beta = 0.0;  // output_beta is off
  if ((P->final_gamma_threshold != -1.0 && gam > P->final_gamma_threshold) || gam < 0.0) {
    record_failed(F, n, gid, FRAME_FAILED_DUE_TO_HIGH_GAMMA, num_iterations);
    return;
  }
  if (jump_fail()) { record_failed(F, n, gid, JUMP_TOLERANCE_EXCEEDED, num_iterations); return; }
  const S contrast = sub.contrast_std_dev();
  const int active = sub.num_active();
  // Schema::record_step, DICe_Schema.cpp:1986-2009 (status = init_status on success).
  save_fields(sf, F, n, gid);
  f(OF_SIGMA) = sig; f(OF_MATCH) = 0.0; f(OF_GAMMA) = gam; f(OF_BETA) = beta;
  f(OF_NOISE) = noise; f(OF_CONTRAST) = contrast; f(OF_ACTIVE_PIXELS) = active;
  f(OF_STATUS) = init_status; f(OF_ITERATIONS) = num_iterations;
}

// Image_::gauss_filter, DICe_Image.cpp:1192-1240.  Its masks (DICe_Image.cpp:54-126) are
// the outer products c_i c_j of the 1-D factors below; each product is formed here in S
// from the S-rounded factors, which reproduces the reference's filtered-image goldens
// (tests/component/images/outFilter*.rawi float and outFilter*_d.rawi double) bit for bit
// (the source tables hold the same products written to nine digits, an ulp away for a few
// entries).  Pixels at least M/2+1 from every edge are replaced by
// sum_i sum_j c[i*M+j] I(x+i-M/2, y+j-M/2) (x offset outer) in S, cast to the storage
// type; uint16 results are clamped to 65535 first (the 13-point mask sums to 1.0002).
template <typename S>
bool gauss_table(int M, std::vector<S> &c) {
  static const double f5[3] = {0.0014, 0.1574, 0.62825};
  static const double f7[4] = {0.006, 0.0606, 0.2418, 0.3831};
  static const double f9[5] = {0.0007, 0.0108, 0.0748, 0.2384, 0.3505};
  static const double f11[6] = {0.0001, 0.0017, 0.0168, 0.087, 0.2328, 0.3231};
  static const double f13[7] = {0.0001, 0.0012, 0.0085, 0.038, 0.1109, 0.2108, 0.2611};
  const double *f = M == 5 ? f5 : M == 7 ? f7 : M == 9 ? f9 : M == 11 ? f11 : M == 13 ? f13 : nullptr;
  if (!f) return false;
  c.assign((size_t)M * M, S(0));
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j) c[(size_t)i * M + j] = (S)f[i <= M / 2 ? i : M - 1 - i] * (S)f[j <= M / 2 ? j : M - 1 - j];
  return true;
}

template <typename S, typename T>
int gauss_filter_t(int M, int w, int h, T *img) {
  std::vector<S> c;
  if (!gauss_table<S>(M, c) || w < M || h < M) return -1;
  const int hm = M / 2 + 1;
  const std::vector<T> tmp(img, img + (size_t)w * h);
  for (int y = hm; y < h - hm; ++y)
    for (int x = hm; x < w - hm; ++x) {
      S value = 0.0;
      for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j)
          value += c[(size_t)i * M + j] * (S)tmp[(size_t)(y + j - hm + 1) * w + x + (i - hm + 1)];
      if (std::is_same<T, uint16_t>::value && value > (S)65535) value = (S)65535;
      img[(size_t)y * w + x] = static_cast<T>(value);
    }
  return 0;
}

template <typename S, typename T>
int correlate_t(const oracle_params *P, int w, int h, const T *ref, const T *def, int n, int subset_size,
                const int *co, const int *cxs, const int *cys, const unsigned char *cact, double *F,
                int first_frame, int threads, const Extras &ex) {
  if (!P || !ref || !def || !F || n < 1 || w < 1 || h < 1) return -1;
  if (!co && subset_size < 1) return -1;
  Image<S, T> refi, defi;
  refi.w = defi.w = w;
  refi.h = defi.h = h;
  refi.I = ref;
  defi.I = def;
  // both images are filtered as they are built, before their gradients (DICe_Image.cpp:303-318,
  // DICe_Schema.cpp:313-314,429-431)
  std::vector<T> ref_f, def_f;
  if (P->gauss_filter_images) {
    ref_f.assign(ref, ref + (size_t)w * h);
    def_f.assign(def, def + (size_t)w * h);
    if (gauss_filter_t<S, T>(P->gauss_filter_mask_size, w, h, ref_f.data()) ||
        gauss_filter_t<S, T>(P->gauss_filter_mask_size, w, h, def_f.data()))
      return -1;
    refi.I = ref_f.data();
    defi.I = def_f.data();
  }
  // Ref gradients are always on for gradient methods (DICe_Schema.cpp:759-761) and by
  // default otherwise (compute_ref_gradients = true, :628); def gradients only on request.
  refi.compute_gradients_fd();
  if (P->gradient_method == 1) refi.smooth_gradients();
  if (P->compute_def_gradients) {
    defi.compute_gradients_fd();
    if (P->gradient_method == 1) defi.smooth_gradients();
  }
  auto one = [&](int gid, int sid) {
    Subset<S> sub;
    const int cx = (int)F[(size_t)OF_COORD_X * n + gid];
    const int cy = (int)F[(size_t)OF_COORD_Y * n + gid];
    if (co) {
      const int b = co[gid], e = co[gid + 1];
      sub.make_list(cx, cy, e - b, cxs + b, cys + b, cact ? cact + b : nullptr);
    } else {
      sub.make_square(cx, cy, subset_size);
    }
    sub.init_ref(refi);
    if (ex.mask_off)
      for (int k = ex.mask_off[gid]; k < ex.mask_off[gid + 1]; ++k) sub.blocked.insert({ex.mask_y[k], ex.mask_x[k]});
    correlate_one<S, T>(P, refi, defi, sub, F, n, gid, co ? 0 : subset_size, first_frame != 0, sid,
                        ex.force && ex.force[gid]);
  };
  const bool neighbors = P->initialization_method == 1 || (P->initialization_method == 2 && first_frame);
  if (neighbors) {
    // USE_NEIGHBOR_VALUES: the subsets one after another in id order (execute_correlation,
    // DICe_Schema.cpp:1720-1735), each starting from NEIGHBOR_ID's fields as they stand
    // (this frame's result for a lower id), or its own when the id is -1.
    for (int gid = 0; gid < n; ++gid) {
      const double nb = F[(size_t)OF_NEIGHBOR_ID * n + gid];
      if (nb != (double)(int)nb || nb < -1.0 || nb >= n) return -1;
      one(gid, nb == -1.0 ? gid : (int)nb);
    }
    return 0;
  }
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
  for (int gid = 0; gid < n; ++gid) one(gid, gid);
  (void)threads;
  return 0;
}

// One Objective_ZNSSD call on one square subset (DICe_Objective.h:82-242): 0 gamma, 1 sigma,
// 2 computeUpdateFast, 3 computeUpdateRobust.  params: 6 (u, v, theta, ex, ey, gxy; disabled
// ones ignored), updated by the updates; result: 10 (value | sigma | status, noise | iterations,
// then CONDITION_NUMBER and FIELD_1..7 after computeUpdateFast).
template <typename S, typename T>
int objective_t(const oracle_params *P, int w, int h, const T *ref, const T *def, int subset_size, int cx, int cy,
                int call, double *params, double *result) {
  if (!P || !ref || !def || !params || !result || subset_size < 1 || call < 0 || call > 3) return -1;
  Image<S, T> refi, defi;
  refi.w = defi.w = w;
  refi.h = defi.h = h;
  refi.I = ref;
  defi.I = def;
  std::vector<T> ref_f, def_f;
  if (P->gauss_filter_images) {
    ref_f.assign(ref, ref + (size_t)w * h);
    def_f.assign(def, def + (size_t)w * h);
    if (gauss_filter_t<S, T>(P->gauss_filter_mask_size, w, h, ref_f.data()) ||
        gauss_filter_t<S, T>(P->gauss_filter_mask_size, w, h, def_f.data()))
      return -1;
    refi.I = ref_f.data();
    defi.I = def_f.data();
  }
  refi.compute_gradients_fd();
  if (P->gradient_method == 1) refi.smooth_gradients();
  if (P->compute_def_gradients) {
    defi.compute_gradients_fd();
    if (P->gradient_method == 1) defi.smooth_gradients();
  }
  Subset<S> sub;
  sub.make_square(cx, cy, subset_size);
  sub.init_ref(refi);
  double F[OF_COUNT] = {0};
  F[OF_COORD_X] = cx;
  F[OF_COORD_Y] = cy;
  Objective<S, T> obj{P, &refi, &defi, &sub, F, 1, 0};
  Affine<S> sf;
  sf.init(P->enable_rotation != 0, P->enable_normal_strain != 0, P->enable_shear_strain != 0,
          (S)P->robust_delta_disp, (S)P->robust_delta_theta);
  const int idx[6] = {sf.iu, sf.iv, sf.it, sf.iex, sf.iey, sf.ig};
  for (int k = 0; k < 6; ++k)
    if (idx[k] >= 0) sf.p[idx[k]] = (S)params[k];
  for (int k = 0; k < 10; ++k) result[k] = 0.0;
  if (call == 0) {
    result[0] = obj.gamma(sf);
  } else if (call == 1) {
    S noise = 0.0;
    result[0] = obj.sigma(sf, noise);
    result[1] = noise;
  } else {
    int iters = -1;
    result[0] = call == 2 ? obj.update_fast(sf, iters) : obj.update_robust(sf, iters);
    result[1] = iters;
    for (int k = 0; k < 6; ++k) params[k] = idx[k] >= 0 ? (double)sf.p[idx[k]] : 0.0;
    if (call == 2) {
      result[2] = F[OF_CONDITION];
      for (int k = 0; k < 7; ++k) result[3 + k] = F[OF_FIELD_1 + k];
    }
  }
  return 0;
}

// storage: 0 = float32 pixels, 1 = uint16 pixels, 2 = float64 pixels.  With
// S = double, float32 pixels are widened to double storage (the
// DICE_USE_DOUBLE build stores intensities as scalar_t).
template <typename S>
int correlate_any(const oracle_params *P, int w, int h, int storage, const void *ref, const void *def, int n,
                  int subset_size, const int *co, const int *cx, const int *cy, const unsigned char *ca,
                  double *F, int first_frame, int threads, const Extras &ex = Extras{}) {
  if (!ref || !def || w < 1 || h < 1) return -1;
  if (storage == 1)
    return correlate_t<S, uint16_t>(P, w, h, (const uint16_t *)ref, (const uint16_t *)def, n, subset_size, co,
                                    cx, cy, ca, F, first_frame, threads, ex);
  if (storage == 0 && sizeof(S) == sizeof(float))
    return correlate_t<S, S>(P, w, h, (const S *)ref, (const S *)def, n, subset_size, co, cx, cy, ca, F,
                             first_frame, threads, ex);
  if (storage == 2 && sizeof(S) == sizeof(double))
    return correlate_t<S, S>(P, w, h, (const S *)ref, (const S *)def, n, subset_size, co, cx, cy, ca, F,
                             first_frame, threads, ex);
  if (storage == 0) {
    const size_t np = (size_t)w * h;
    std::vector<S> r(np), d(np);
    for (size_t i = 0; i < np; ++i) { r[i] = ((const float *)ref)[i]; d[i] = ((const float *)def)[i]; }
    return correlate_t<S, S>(P, w, h, r.data(), d.data(), n, subset_size, co, cx, cy, ca, F, first_frame, threads,
                             ex);
  }
  return -1;
}

// angle_2d, DICe_Shape.cpp:88-103, in scalar_t S: std::atan2 of S, the wrap against
// the double constants rounded back to S.
template <typename S>
S angle_2d(S x1, S y1, S x2, S y2) {
  S d = std::atan2(y2, x2) - std::atan2(y1, x1);
  while (d > kPi) d -= kTwoPi;
  while (d < -kPi) d += kTwoPi;
  return d;
}

// Polygon ctor + get_owned_pixels (no map), DICe_Shape.cpp:50-80,136-219, in S: the
// winding angle of each pixel of the vertex bbox, |sum| >= pi inside.
template <typename S>
int polygon_pixels_t(int nv, const int *vx, const int *vy, int cap, int *out_x, int *out_y) {
  if (nv < 3) return -1;
  std::vector<int> X(vx, vx + nv), Y(vy, vy + nv);
  X.push_back(vx[0]); Y.push_back(vy[0]);
  int mnx = X[0], mny = Y[0], mxx = X[0], mxy = Y[0];
  for (int i = 0; i < nv; ++i) {
    mnx = std::min(mnx, X[i]); mny = std::min(mny, Y[i]);
    mxx = std::max(mxx, X[i]); mxy = std::max(mxy, Y[i]);
  }
  int count = 0;
  for (int y = mny; y <= mxy; ++y)
    for (int x = mnx; x <= mxx; ++x) {
      S angle = 0.0;
      for (int i = 0; i < nv; ++i)
        angle += angle_2d<S>((S)(X[i] - x), (S)(Y[i] - y), (S)(X[i + 1] - x), (S)(Y[i + 1] - y));
      if (std::abs(angle) >= kPi) {
        if (count < cap) { out_x[count] = x; out_y[count] = y; }
        ++count;
      }
    }
  return count;
}

}  // namespace

extern "C" {

void oracle_default_params(oracle_params *p) {
  // dice_default_params, DICe_ParameterUtilities.cpp:360-420.
  std::memset(p, 0, sizeof(*p));
  p->interpolation_method = KEYS_FOURTH;
  p->optimization_method = GRADIENT_BASED_THEN_SIMPLEX;
  p->correlation_routine = 0;
  p->initialization_method = 0;
  p->enable_translation = 1;
  p->max_solver_iterations_fast = 25;
  p->max_solver_iterations_robust = 1000;
  p->fast_solver_tolerance = 1.0E-4;
  p->robust_solver_tolerance = 1.0E-6;
  p->robust_delta_disp = 1.0;
  p->robust_delta_theta = 0.1;
  p->skip_solve_gamma_threshold = 1.0E-10;
  p->initial_gamma_threshold = -1.0;
  p->final_gamma_threshold = -1.0;
  p->disp_jump_tol = 10000.0;
  p->theta_jump_tol = 100.0;
  p->levenberg_marquardt_regularization_factor = -1.0;
  p->momentum_factor = 0.0;
  p->compute_uncertainty_fields = 1;
  p->gauss_filter_mask_size = 7;  // DICe_Schema.cpp:465
}

void oracle_tracking_default_params(oracle_params *p) {
  // tracking_default_params, DICe_ParameterUtilities.cpp:297-358.
  oracle_default_params(p);
  p->correlation_routine = 1;
  p->optimization_method = SIMPLEX;
  p->enable_rotation = 1;
  p->disp_jump_tol = 8.0;
  p->theta_jump_tol = 0.1;
}

void oracle_fd_gradients_f(int w, int h, const float *img, float *gx, float *gy) {
  Image<float, float> im; im.w = w; im.h = h; im.I = img;
  im.compute_gradients_fd();
  std::copy(im.gx.begin(), im.gx.end(), gx);
  std::copy(im.gy.begin(), im.gy.end(), gy);
}
void oracle_fd_gradients_d(int w, int h, const double *img, double *gx, double *gy) {
  Image<double, double> im; im.w = w; im.h = h; im.I = img;
  im.compute_gradients_fd();
  std::copy(im.gx.begin(), im.gx.end(), gx);
  std::copy(im.gy.begin(), im.gy.end(), gy);
}
void oracle_smooth_gradients_f(int w, int h, float *gx, float *gy) {
  Image<float, float> im; im.w = w; im.h = h;
  im.gx.assign(gx, gx + (size_t)w * h); im.gy.assign(gy, gy + (size_t)w * h);
  im.smooth_gradients();
  std::copy(im.gx.begin(), im.gx.end(), gx);
  std::copy(im.gy.begin(), im.gy.end(), gy);
}

}  // extern "C"
template <typename S>
static void interp_impl(int method, int w, int h, const S *img, int n, const S *x, const S *y, S *out) {
  Image<S, S> im; im.w = w; im.h = h; im.I = img;
  for (int i = 0; i < n; ++i) {
    S a = 0, b = 0, c = 0;
    if (method == 3) out[i] = im.keys(x[i], y[i]);  // interpolate_keys_fourth (single)
    else { im.interp_all(method, a, b, c, false, x[i], y[i]); out[i] = a; }
  }
}
extern "C" {
void oracle_interp_f(int method, int w, int h, const float *img, int n, const float *x, const float *y, float *out) {
  interp_impl<float>(method, w, h, img, n, x, y, out);
}
void oracle_interp_d(int method, int w, int h, const double *img, int n, const double *x, const double *y, double *out) {
  interp_impl<double>(method, w, h, img, n, x, y, out);
}

}  // extern "C"
template <typename S>
static void apply_transform_impl(int w, int h, const S *in, S *out, int cx, int cy, S u, S v, S theta) {
  // apply_transform, DICe_ImageUtils.cpp:50-77, with the default (all-enabled) affine map.
  Image<S, S> im; im.w = w; im.h = h; im.I = in;
  Affine<S> sf;
  sf.init(true, true, true, 1.0, 0.1);
  sf.p[sf.iu] = -u; sf.p[sf.iv] = -v; sf.p[sf.it] = theta;
  const S CX = cx + u, CY = cy + v;
  S mx = 0, my = 0;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      sf.map((S)x, (S)y, CX, CY, mx, my);
      out[y * w + x] = im.keys(mx, my);
    }
}
extern "C" {
void oracle_apply_transform_d(int w, int h, const double *in, double *out, int cx, int cy, double u, double v, double theta) {
  apply_transform_impl<double>(w, h, in, out, cx, cy, u, v, theta);
}
void oracle_apply_transform_f(int w, int h, const float *in, float *out, int cx, int cy, float u, float v, float theta) {
  apply_transform_impl<float>(w, h, in, out, cx, cy, u, v, theta);
}

double oracle_subset_gamma_f(int w, int h, const float *ref, const float *def, int cx, int cy, int size) {
  Image<float, float> ri, di;
  ri.w = di.w = w; ri.h = di.h = h; ri.I = ref; di.I = def;
  Subset<float> s;
  s.make_square(cx, cy, size);
  s.init_ref(ri);
  s.init_def_nomap(di);
  return s.gamma();
}

int oracle_polygon_pixels(int nv, const int *vx, const int *vy, int cap, int *out_x, int *out_y) {
  return polygon_pixels_t<double>(nv, vx, vy, cap, out_x, out_y);
}
int oracle_polygon_pixels_f(int nv, const int *vx, const int *vy, int cap, int *out_x, int *out_y) {
  return polygon_pixels_t<float>(nv, vx, vy, cap, out_x, out_y);
}

int oracle_deformed_polygon_pixels(int nv, const int *vx, const int *vy, const double *params, int cx, int cy,
                                   double skin_factor, int cap, int *out_x, int *out_y) {
  // Polygon::get_owned_pixels(shape_function, cx, cy, skin_factor), DICe_Shape.cpp:136-219, in float
  if (nv < 3 || !params) return -1;
  Affine<float> sf;
  sf.init(true, true, true, 1.0f, 0.1f);
  for (int k = 0; k < 6; ++k) sf.p[k] = (float)params[k];
  std::vector<int> X(nv + 1), Y(nv + 1);
  for (int i = 0; i <= nv; ++i) {
    const int x = vx[i % nv], y = vy[i % nv];
    float mx = 0.0f, my = 0.0f;
    sf.map((float)x, (float)y, (float)cx, (float)cy, mx, my);
    int nx = (int)mx, ny = (int)my;
    if (mx - (int)mx >= 0.5) nx++;
    if (my - (int)my >= 0.5) ny++;
    X[i] = nx;
    Y[i] = ny;
  }
  int ccx = 0, ccy = 0;
  for (int i = 0; i < nv; ++i) { ccx += X[i]; ccy += Y[i]; }
  ccx = (int)(ccx / (double)nv);
  ccy = (int)(ccy / (double)nv);
  int mnx = 0, mxx = 0, mny = 0, mxy = 0;
  for (int i = 0; i <= nv; ++i) {
    X[i] = (int)((float)skin_factor * (X[i] - ccx) + ccx);
    Y[i] = (int)((float)skin_factor * (Y[i] - ccy) + ccy);
    if (i == 0) { mnx = mxx = X[i]; mny = mxy = Y[i]; }
    mnx = std::min(mnx, X[i]); mxx = std::max(mxx, X[i]);
    mny = std::min(mny, Y[i]); mxy = std::max(mxy, Y[i]);
  }
  int count = 0;
  for (int y = mny; y <= mxy; ++y)
    for (int x = mnx; x <= mxx; ++x) {
      float angle = 0.0f;
      for (int i = 0; i < nv; ++i)
        angle += angle_2d<float>((float)(X[i] - x), (float)(Y[i] - y), (float)(X[i + 1] - x), (float)(Y[i + 1] - y));
      if (std::abs(angle) >= kPi) {
        if (count < cap) { out_x[count] = x; out_y[count] = y; }
        ++count;
      }
    }
  return count;
}

int oracle_correlate_f(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int n, int subset_size, const int *co, const int *cx, const int *cy,
                       const unsigned char *ca, double *fields, int first_frame, int threads) {
  return correlate_any<float>(p, w, h, storage, ref, def, n, subset_size, co, cx, cy, ca, fields, first_frame, threads);
}
int oracle_correlate_d(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int n, int subset_size, const int *co, const int *cx, const int *cy,
                       const unsigned char *ca, double *fields, int first_frame, int threads) {
  return correlate_any<double>(p, w, h, storage, ref, def, n, subset_size, co, cx, cy, ca, fields, first_frame, threads);
}
int oracle_correlate_ex_f(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                          int n, int subset_size, const int *co, const int *cx, const int *cy,
                          const unsigned char *ca, double *fields, int first_frame, int threads,
                          const unsigned char *force, const int *mask_off, const int *mask_x, const int *mask_y) {
  Extras ex;
  ex.force = force;
  ex.mask_off = mask_off;
  ex.mask_x = mask_x;
  ex.mask_y = mask_y;
  return correlate_any<float>(p, w, h, storage, ref, def, n, subset_size, co, cx, cy, ca, fields, first_frame, threads,
                              ex);
}
int oracle_correlate_ex_d(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                          int n, int subset_size, const int *co, const int *cx, const int *cy,
                          const unsigned char *ca, double *fields, int first_frame, int threads,
                          const unsigned char *force, const int *mask_off, const int *mask_x, const int *mask_y) {
  Extras ex;
  ex.force = force;
  ex.mask_off = mask_off;
  ex.mask_x = mask_x;
  ex.mask_y = mask_y;
  return correlate_any<double>(p, w, h, storage, ref, def, n, subset_size, co, cx, cy, ca, fields, first_frame,
                               threads, ex);
}

int oracle_gauss_filter_f(int w, int h, int storage, void *img, int mask_size) {
  if (!img) return -1;
  if (storage == 0) return gauss_filter_t<float, float>(mask_size, w, h, (float *)img);
  if (storage == 1) return gauss_filter_t<float, uint16_t>(mask_size, w, h, (uint16_t *)img);
  return -1;
}
int oracle_gauss_filter_d(int w, int h, double *img, int mask_size) {
  if (!img) return -1;
  return gauss_filter_t<double, double>(mask_size, w, h, img);
}

int oracle_vsg_strain(int n, double *F, int window, double *vsg) {
  // VSG_Strain_Post_Processor: pre_execution_tasks + initialize_neighborhood (radius search,
  // DICe_PostProcessor.cpp:140-236, 908-929) and execute (:931-1136).  Neighbours by brute
  // force, sorted by distance as nanoflann's sorted radius search returns them (ties by id);
  // squared distances and the radius test in float, as the float kd-tree has them.
  // window < 0: the -window nearest subsets (knnSearch, :186-206), ties by id; nanoflann's own
  // order among equidistant points depends on its tree, so only untied cases are pinned
  if (n < 1 || !F || !vsg || window == 0) return -1;
  auto f = [&](int slot, int i) -> double & { return F[(size_t)slot * n + i]; };
  const float rad = (float)window / 2.0f;
  const float r2 = rad * rad + (float)1.0E-5;
  const size_t knn = window < 0 ? (size_t)(-window) : 0;
  for (int i = 0; i < n; ++i) {
    const float xi = (float)f(OF_COORD_X, i), yi = (float)f(OF_COORD_Y, i);
    std::vector<std::pair<float, int>> nb;
    for (int j = 0; j < n; ++j) {
      const float dx = (float)f(OF_COORD_X, j) - xi, dy = (float)f(OF_COORD_Y, j) - yi;
      const float d2 = dx * dx + dy * dy;
      if (knn || d2 < r2) nb.push_back({d2, j});
    }
    std::sort(nb.begin(), nb.end());
    if (knn && nb.size() > knn) nb.resize(knn);
    double out[7] = {0, 0, 0, 0, 0, 0, 0};
    int valid = 0;
    for (const auto &e : nb) valid += f(OF_SIGMA, e.second) >= 0.0;
    bool ok = valid >= 3 && !nb.empty() && f(OF_SIGMA, nb[0].second) >= 0.0;
    if (ok) {
      double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, bu[3] = {0, 0, 0}, bv[3] = {0, 0, 0};
      for (const auto &e : nb) {
        const int j = e.second;
        if (!(f(OF_SIGMA, j) >= 0.0)) continue;
        // neighbour offsets and displacements as scalar_t (float), DICe_PostProcessor.cpp:219-220,1011-1018
        const double row[3] = {1.0, (double)((float)f(OF_COORD_X, j) - xi), (double)((float)f(OF_COORD_Y, j) - yi)};
        const double u = (float)f(OF_DISP_X, j), v = (float)f(OF_DISP_Y, j);
        for (int k = 0; k < 3; ++k) {
          for (int m = 0; m < 3; ++m) A[k][m] += row[k] * row[m];
          bu[k] += row[k] * u;
          bv[k] += row[k] * v;
        }
      }
      // GETRF + GECON + GETRI (:1047-1085); the condition number exactly in the 1-norm
      double anorm = 0.0;
      for (int j = 0; j < 3; ++j) anorm = std::max(anorm, std::abs(A[0][j]) + std::abs(A[1][j]) + std::abs(A[2][j]));
      double inv[6][6];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inv[r][c] = A[r][c];
      const int info = lu_inverse(3, inv);
      double inorm = 0.0;
      for (int j = 0; j < 3; ++j)
        inorm = std::max(inorm, std::abs(inv[0][j]) + std::abs(inv[1][j]) + std::abs(inv[2][j]));
      ok = info == 0 && 1.0 / (anorm * inorm) >= 1.0E-12;
      if (ok) {
        double cu[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            cu[r] += inv[r][c] * bu[c];
            cv[r] += inv[r][c] * bv[c];
          }
        const double dudx = cu[1], dudy = cu[2], dvdx = cv[1], dvdy = cv[2];
        // the Green-Lagrange strains (:1118-1124), stored in scalar_t fields
        out[0] = (float)(0.5 * (2.0 * dudx + dudx * dudx + dvdx * dvdx));
        out[1] = (float)(0.5 * (2.0 * dvdy + dudy * dudy + dvdy * dvdy));
        out[2] = (float)(0.5 * (dudy + dvdx + dudx * dudy + dvdx * dvdy));
        out[3] = (float)dudx;
        out[4] = (float)dudy;
        out[5] = (float)dvdx;
        out[6] = (float)dvdy;
      }
    }
    for (int k = 0; k < 7; ++k) vsg[(size_t)k * n + i] = out[k];
    if (!ok) f(OF_MATCH, i) = -1.0;
  }
  return 0;
}

int oracle_objective_f(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int subset_size, int cx, int cy, int call, double *params, double *result) {
  if (storage == 0)
    return objective_t<float, float>(p, w, h, (const float *)ref, (const float *)def, subset_size, cx, cy, call,
                                     params, result);
  if (storage == 1)
    return objective_t<float, uint16_t>(p, w, h, (const uint16_t *)ref, (const uint16_t *)def, subset_size, cx, cy,
                                        call, params, result);
  return -1;
}
int oracle_objective_d(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int subset_size, int cx, int cy, int call, double *params, double *result) {
  if (storage == 2)
    return objective_t<double, double>(p, w, h, (const double *)ref, (const double *)def, subset_size, cx, cy, call,
                                       params, result);
  if (storage == 1)
    return objective_t<double, uint16_t>(p, w, h, (const uint16_t *)ref, (const uint16_t *)def, subset_size, cx, cy,
                                         call, params, result);
  return -1;
}

int oracle_uncertainty(int n, const double *F, double *out) {
  // Uncertainty_Post_Processor::execute, DICe_PostProcessor.cpp:1507-1537 (2-D: max_m = 0).
  if (n < 1 || !F || !out) return -1;
  for (int i = 0; i < n; ++i) {
    const float angle = (float)F[(size_t)OF_FIELD_1 * n + i];
    const float sig = (float)F[(size_t)OF_SIGMA * n + i];
    out[(size_t)n + i] = angle;
    out[i] = sig < 0.0f ? 0.0f : (angle == 0.0f ? 0.0f : (float)(1.0 / (double)angle * (double)sig));
  }
  return 0;
}

int oracle_nelder_mead(int kind, int nd, double *p, const double *deltas, const double *coef, int max_iterations,
                       double tolerance, double threshold, int cap, double *trace, int *n_evals, int *num_iterations,
                       int *stop) {
  if (nd < 1 || nd > 6 || !p || !deltas || !n_evals || !num_iterations || (kind == 0 && !coef) || kind < 0 ||
      kind > 1)
    return -1;
  int evals = 0;
  auto f = [&](const double *v) {
    double r = 0.0;
    for (int i = 0; i < nd; ++i) {
      const double d = kind == 0 ? v[i] - coef[nd + i] : v[i] * v[i] - 1.0;
      r += (kind == 0 ? coef[i] : 1.0) * d * d;
    }
    if (trace && evals < cap) {
      for (int i = 0; i < nd; ++i) trace[(size_t)evals * (nd + 1) + i] = v[i];
      trace[(size_t)evals * (nd + 1) + nd] = r;
    }
    ++evals;
    return r;
  };
  const int status = nelder_mead<double>(nd, p, deltas, f, max_iterations, tolerance, threshold, *num_iterations, stop);
  *n_evals = evals;
  return status;
}

int oracle_lu_inverse(int n, double *a) {
  if (n < 1 || n > 6) return -1;
  double A[6][6];
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) A[i][j] = a[i * n + j];
  const int info = lu_inverse(n, A);
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) a[i * n + j] = A[i][j];
  return info;
}

}  // extern "C"
