// Device-side building blocks of the subset-correlation kernels (included by
// the kernel translation units only).
//
// One cooperating group of lanes owns one subset from initial guess to the
// recorded field values, running the reference's per-subset driver
// (Schema::generic_correlation_routine, DICe_Schema.cpp:2025-2368) on the
// device:
//   * Gauss-Newton (Objective_ZNSSD::computeUpdateFast, DICe_Objective.cpp:430-657)
//   * Nelder-Mead (Simplex::minimize, DICe_Simplex.cpp:60-277) as primary or fallback
//   * sigma / noise (DICe_Objective.cpp:150-184, DICe_Subset.cpp:622-703),
//     final ZNSSD gamma (DICe_Subset.cpp:323-340), contrast, active pixels,
//     uncertainty FIELD_1..7 (DICe_Objective.cpp:193-299).
//
// Two group shapes share the code:
//   PLAIN   : a 256-thread workgroup per subset, reductions through LDS and
//             barriers, deformed intensities gathered straight from HBM/L2.
//   BATCHED : one 64-lane wavefront per subset; waves of a persistent grid pull
//             subset indices from a global work counter (work stealing), stage
//             the warped subset's deformed window in an LDS slot once per
//             iteration with coalesced loads, gather the 6x6 Keys taps from LDS
//             and reduce with cross-lane shuffles (no barriers).
//
// Precision: pixel work in FP32; per-pixel products accumulated in FP64
// (Hessian, gradient, ZNSSD sums); the 6x6 solve in FP64.
#pragma once
#include "dice_kernels.h"

#include <math.h>

namespace dice_hip {
namespace dev {

// Pointers into LDS carry their address space so that no flat-address checks
// are generated for them.
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) double lds_double;

// ------------------------------------------------------------ utilities ---
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

// One wavefront.
struct WaveGroup {
  int lane;
  static constexpr int kSize = kWave;
  __device__ __forceinline__ int rank() const { return lane; }
  __device__ __forceinline__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  template <int K>
  __device__ __forceinline__ void sum(double (&v)[K]) const {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  }
  __device__ __forceinline__ int imin(int v) const { return wave_min(v); }
  __device__ __forceinline__ int imax(int v) const { return wave_max(v); }
};

// A 256-thread workgroup.  red: LDS scratch of 4 * kMaxRed doubles.
constexpr int kMaxRed = 40;
struct BlockGroup {
  lds_double *red;
  static constexpr int kSize = kPlainThreads;
  __device__ __forceinline__ int rank() const { return threadIdx.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
  template <int K>
  __device__ __forceinline__ void sum(double (&v)[K]) const {
    static_assert(K <= kMaxRed, "reduction too wide");
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double s = wave_sum(v[k]);
      if (l == 0) red[w * kMaxRed + k] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k)
      v[k] = (red[k] + red[kMaxRed + k]) + (red[2 * kMaxRed + k] + red[3 * kMaxRed + k]);
    __syncthreads();
  }
  __device__ __forceinline__ int ired(int v, bool take_min) const {
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    int s = take_min ? wave_min(v) : wave_max(v);
    __attribute__((address_space(3))) int *ri = reinterpret_cast<__attribute__((address_space(3))) int *>(red);
    if (l == 0) ri[w] = s;
    __syncthreads();
    int r = ri[0];
    for (int k = 1; k < kPlainThreads / kWave; ++k) r = take_min ? min(r, ri[k]) : max(r, ri[k]);
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ int imin(int v) const { return ired(v, true); }
  __device__ __forceinline__ int imax(int v) const { return ired(v, false); }
};

template <typename T> __device__ __forceinline__ float ldpix(const void *p, int i);
template <> __device__ __forceinline__ float ldpix<float>(const void *p, int i) {
  return static_cast<const float *>(p)[i];
}
template <> __device__ __forceinline__ float ldpix<uint16_t>(const void *p, int i) {
  return (float)static_cast<const uint16_t *>(p)[i];
}

// Keys 4th-order cubic-convolution weights (Keys 1981; DICe_Image.cpp:114-122,
// 860-871) of the six taps at offsets -2..+3 from floor(x), written as cubics
// in the fractional offset t in [0, 1).  Same function as the reference's
// f2(t+2), f1(t+1), f0(t), f0(1-t), f1(2-t), f2(3-t).
__device__ __forceinline__ void keys_weights(float t, float w[6]) {
  const float t2 = t * t, t3 = t2 * t;
  w[0] = (t3 - 2.0f * t2 + t) * (1.0f / 12.0f);
  w[1] = (-7.0f / 12.0f) * t3 + 1.25f * t2 - (2.0f / 3.0f) * t;
  w[2] = (4.0f / 3.0f) * t3 - (7.0f / 3.0f) * t2 + 1.0f;
  w[3] = (-4.0f / 3.0f) * t3 + (5.0f / 3.0f) * t2 + (2.0f / 3.0f) * t;
  w[4] = (7.0f / 12.0f) * t3 - 0.5f * t2 - (1.0f / 12.0f) * t;
  w[5] = (t2 - t3) * (1.0f / 12.0f);
}

// Affine map and its Jacobian rows (Affine_Shape_Function::map / residuals,
// DICe_LocalShapeFunction.cpp:220-248, 350-396), parameters in the fixed slot
// order u, v, theta, ex, ey, gxy.
struct Affine {
  float p[6];
  float c, s;            // cos, sin theta
  float a11, a12, a21, a22;
  __device__ __forceinline__ void prepare() {
    c = cosf(p[2]);
    s = sinf(p[2]);
    const float sx = 1.0f + p[3], sy = 1.0f + p[4], g = p[5];
    // x' - cx = c*Dx - s*Dy + u, Dx = sx*dx + g*dy, Dy = sy*dy + g*dx
    a11 = c * sx - s * g;
    a12 = c * g - s * sy;
    a21 = s * sx + c * g;
    a22 = s * g + c * sy;
  }
  // displacement of pixel offset (dx, dy) from the centroid: x' - cx, y' - cy
  __device__ __forceinline__ void disp(float dx, float dy, float &ox, float &oy) const {
    ox = p[0] + (a11 * dx + a12 * dy);
    oy = p[1] + (a21 * dx + a22 * dy);
  }
  // residual row; (gx, gy) is the gradient to use (rotated ref or interpolated def)
  __device__ __forceinline__ void residual(float dx, float dy, float Gx, float Gy, float r[6]) const {
    const float Dx = (1.0f + p[3]) * dx + p[5] * dy;
    const float Dy = (1.0f + p[4]) * dy + p[5] * dx;
    r[0] = Gx;
    r[1] = Gy;
    r[2] = Gx * (-s * Dx - c * Dy) + Gy * (c * Dx - s * Dy);
    r[3] = (Gx * c + Gy * s) * dx;
    r[4] = (-Gx * s + Gy * c) * dy;
    r[5] = Gx * (c * dy - s * dx) + Gy * (s * dy + c * dx);
  }
};

// Gauss-Newton update du = -H^-1 q of the reference (GETRF + GETRI, then
// def_update = H^-1 * (-q), DICe_Objective.cpp:571-603) in FP64.  The inverse
// is never formed: LU with partial pivoting (first maximal pivot, as dgetf2)
// followed by two triangular solves gives the same vector.  An exactly zero
// pivot (INFO > 0) makes GETRI return early and leave the LU factors in H, so
// the reference's update is then -(factors) * q; that case is reproduced.
__device__ __forceinline__ void lu_update6(double A[6][6], const double q[6], double du[6]) {
  int piv[6];
  bool singular = false;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    int p = j;
    double best = fabs(A[j][j]);
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      const double a = fabs(A[i][j]);
      if (a > best) { best = a; p = i; }
    }
    piv[j] = p;
    double pv = A[j][j];
#pragma unroll
    for (int i = j + 1; i < 6; ++i) pv = (i == p) ? A[i][j] : pv;
    if (pv != 0.0) {
      if (p != j) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const double tj = A[j][k];
          double tp = tj;
#pragma unroll
          for (int i = j + 1; i < 6; ++i) tp = (i == p) ? A[i][k] : tp;
          A[j][k] = tp;
#pragma unroll
          for (int i = j + 1; i < 6; ++i) A[i][k] = (i == p) ? tj : A[i][k];
        }
      }
      const double r = 1.0 / A[j][j];
#pragma unroll
      for (int i = j + 1; i < 6; ++i) A[i][j] *= r;
    } else {
      singular = true;
    }
#pragma unroll
    for (int i = j + 1; i < 6; ++i)
#pragma unroll
      for (int k = j + 1; k < 6; ++k) A[i][k] = fma(-A[i][j], A[j][k], A[i][k]);
  }
  if (singular) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) s = fma(A[i][k], -q[k], s);
      du[i] = s;
    }
    return;
  }
  double b[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) b[i] = -q[i];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const double bj = b[j];
    double bp = bj;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) bp = (i == piv[j]) ? b[i] : bp;
    b[j] = bp;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) b[i] = (i == piv[j]) ? bj : b[i];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s = fma(-A[i][k], b[k], s);
    b[i] = s;
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s = fma(-A[i][k], b[k], s);
    b[i] = s / A[i][i];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) du[i] = b[i];
}

// ------------------------------------------------------ subset solver ---
template <typename T, typename G, int INTERP, bool DEFGRAD>
struct Solver {
  const KParams kp;
  const DImage ref;
  const DImage def;
  G grp;
  // LDS window (BATCHED only)
  lds_float *win;
  int win_cap;
  bool use_win;
  int wx0, wy0, wpitch;
  // subset
  int cx, cy, n, S, hw, off;
  int bx0, by0, bx1, by1;
  int n_def_active;       // pixels flagged active in the subset definition
  const int *px, *py;
  const uint8_t *pact;
  float invS;
  float meanF;
  float en[6];            // 1 for enabled parameters

  __device__ __forceinline__ Solver(const KParams &k, const DImage &r, const DImage &d, G g, lds_float *w, int cap)
      : kp(k), ref(r), def(d), grp(g), win(w), win_cap(cap), use_win(false) {
    en[0] = en[1] = 1.0f;
    en[2] = k.en_rot ? 1.0f : 0.0f;
    en[3] = en[4] = k.en_normal ? 1.0f : 0.0f;
    en[5] = k.en_shear ? 1.0f : 0.0f;
  }

  __device__ __forceinline__ void set_subset(const DSubsets &ss, int i) {
    cx = ss.cx[i];
    cy = ss.cy[i];
    S = ss.S;
    if (S > 0) {
      hw = S / 2;
      S = 2 * hw + 1;     // square subsets are odd (DICe_Subset.cpp:90-92)
      n = S * S;
      invS = 1.0f / (float)S;
      bx0 = max(cx - hw, 0); bx1 = max(cx + hw, 0);
      by0 = max(cy - hw, 0); by1 = max(cy + hw, 0);
      n_def_active = n;
      px = py = nullptr;
      pact = nullptr;
      off = 0;
    } else {
      off = ss.off[i];
      n = ss.off[i + 1] - off;
      px = ss.px + off;
      py = ss.py + off;
      pact = ss.act ? ss.act + off : nullptr;
      const int4 b = ss.bbox[i];
      bx0 = b.x; by0 = b.y; bx1 = b.z; by1 = b.w;
      n_def_active = ss.n_active_def[i];
    }
  }

  // pixel coordinates (negative coordinates clamp to 0, DICe_Subset.cpp:65-68,100-101)
  __device__ __forceinline__ void pixel(int i, int &x, int &y, bool &a) const {
    if (S > 0) {
      const int row = (int)(((float)i + 0.5f) * invS);
      const int col = i - row * S;
      x = max(cx - hw + col, 0);
      y = max(cy - hw + row, 0);
      a = true;
    } else {
      x = max(px[i], 0);
      y = max(py[i], 0);
      a = pact ? (pact[i] != 0) : true;
    }
  }

  // Image_::operator() (DICe_Image.h:186-192): out of bounds reads pixel 0
  __device__ __forceinline__ float img_at(const DImage &im, int x, int y) const {
    if (x < 0 || y < 0 || x >= im.w || y >= im.h) return ldpix<T>(im.pix, 0);
    return ldpix<T>(im.pix, y * im.w + x);
  }
  __device__ __forceinline__ float grad_at(const float *g, const DImage &im, int x, int y) const {
    if (x < 0 || y < 0 || x >= im.w || y >= im.h) return 0.0f;
    return g[y * im.w + x];
  }

  // Is the warped pixel kept?  Rounding + 4-px guard of Subset::initialize
  // (DICe_Subset.cpp:404-411); (ddx, ddy) = x' - cx, y' - cy.
  __device__ __forceinline__ bool kept(float ddx, float ddy) const {
    const int rx = cx + (int)floorf(ddx + 0.5f);
    const int ry = cy + (int)floorf(ddy + 0.5f);
    return !(rx < def.ox + 4 || rx >= def.ox + def.w - 4 || ry < def.oy + 4 || ry >= def.oy + def.h - 4);
  }

  // Deformed intensity (and gradients) at x' = cx + ddx, y' = cy + ddy.
  // Only called for kept pixels, so all taps lie inside the image.
  __device__ __forceinline__ float sample(float ddx, float ddy, bool want_grad, float &gxo, float &gyo) const {
    const float flx = floorf(ddx), fly = floorf(ddy);
    const float tx = ddx - flx, ty = ddy - fly;
    const int ix = cx - def.ox + (int)flx;
    const int iy = cy - def.oy + (int)fly;
    if (INTERP == 0) {
      // interpolate_bilinear_all, DICe_Image.cpp:392-425
      const float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty);
      const float w11 = tx * ty, w01 = (1.0f - tx) * ty;
      float v;
      if (use_win) {
        const lds_float *b = win + (iy - wy0) * wpitch + (ix - wx0);
        v = b[0] * w00 + b[1] * w10 + b[wpitch + 1] * w11 + b[wpitch] * w01;
      } else {
        const int o = iy * def.w + ix;
        v = ldpix<T>(def.pix, o) * w00 + ldpix<T>(def.pix, o + 1) * w10 +
            ldpix<T>(def.pix, o + def.w + 1) * w11 + ldpix<T>(def.pix, o + def.w) * w01;
      }
      if (want_grad) {
        const int o = iy * def.w + ix;
        gxo = def.gx[o] * w00 + def.gx[o + 1] * w10 + def.gx[o + def.w + 1] * w11 + def.gx[o + def.w] * w01;
        gyo = def.gy[o] * w00 + def.gy[o + 1] * w10 + def.gy[o + def.w + 1] * w11 + def.gy[o + def.w] * w01;
      }
      return v;
    }
    // interpolate_keys_fourth_all, DICe_Image.cpp:832-883 (interior branch)
    float wx[6], wy[6];
    keys_weights(tx, wx);
    keys_weights(ty, wy);
    float v = 0.0f;
    if (use_win) {
      const lds_float *b = win + (iy - 2 - wy0) * wpitch + (ix - 2 - wx0);
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const lds_float *row = b + m * wpitch;
        float rs = row[0] * wx[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) rs = fmaf(row[k], wx[k], rs);
        v = fmaf(rs, wy[m], v);
      }
    } else {
      const int o = (iy - 2) * def.w + ix - 2;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const int ro = o + m * def.w;
        float rs = ldpix<T>(def.pix, ro) * wx[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) rs = fmaf(ldpix<T>(def.pix, ro + k), wx[k], rs);
        v = fmaf(rs, wy[m], v);
      }
    }
    if (want_grad) {
      const int o = (iy - 2) * def.w + ix - 2;
      float ax = 0.0f, ay = 0.0f;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const int ro = o + m * def.w;
        float rx = 0.0f, ry = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          rx = fmaf(def.gx[ro + k], wx[k], rx);
          ry = fmaf(def.gy[ro + k], wx[k], ry);
        }
        ax = fmaf(rx, wy[m], ax);
        ay = fmaf(ry, wy[m], ay);
      }
      gxo = ax;
      gyo = ay;
    }
    return v;
  }

  // Stage the deformed window covering the warped bbox (+ interpolation
  // support) into this group's LDS slot.  Falls back to global gathers when
  // the window does not fit.
  __device__ __forceinline__ void stage_window(const Affine &A) {
    use_win = false;
    if (win_cap <= 0) return;
    float lo_x = 1e30f, hi_x = -1e30f, lo_y = 1e30f, hi_y = -1e30f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float dx = (float)((k & 1) ? bx1 : bx0) - (float)cx;
      const float dy = (float)((k & 2) ? by1 : by0) - (float)cy;
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      lo_x = fminf(lo_x, ox); hi_x = fmaxf(hi_x, ox);
      lo_y = fminf(lo_y, oy); hi_y = fmaxf(hi_y, oy);
    }
    const int lead = INTERP == 0 ? 0 : 2, trail = INTERP == 0 ? 1 : 3;
    int x0 = cx - def.ox + (int)floorf(lo_x) - lead - 1;
    int x1 = cx - def.ox + (int)floorf(hi_x) + trail + 1;
    int y0 = cy - def.oy + (int)floorf(lo_y) - lead - 1;
    int y1 = cy - def.oy + (int)floorf(hi_y) + trail + 1;
    x0 = max(x0, 0); y0 = max(y0, 0);
    x1 = min(x1, def.w - 1); y1 = min(y1, def.h - 1);
    const int ww = x1 - x0 + 1, hh = y1 - y0 + 1;
    if (ww <= 0 || hh <= 0 || ww * hh > win_cap) return;
    wx0 = x0; wy0 = y0; wpitch = ww;
    grp.sync();
    const int tot = ww * hh;
    for (int k = grp.rank(); k < tot; k += G::kSize) {
      const int r = k / ww;
      const int c = k - r * ww;
      win[k] = ldpix<T>(def.pix, (y0 + r) * def.w + x0 + c);
    }
    grp.sync();
    use_win = true;
  }

  // Objective::gamma (DICe_Objective.cpp:58-76) at parameters A: warp, then
  // the ZNSSD of Subset::gamma from FP64 sums; also active count and contrast.
  __device__ __forceinline__ float gamma_eval(const Affine &A, int &active, float &contrast) {
    stage_window(A);
    double acc[6] = {0, 0, 0, 0, 0, 0};  // n, sd, sdd, sr, srr, sdr
    #pragma unroll 1
    for (int i = grp.rank(); i < n; i += G::kSize) {
      int x, y;
      bool a;
      pixel(i, x, y, a);
      if (!a) continue;
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      if (!kept(ox, oy)) continue;
      float gdum, gdum2;
      const double d = sample(ox, oy, false, gdum, gdum2);
      const double r = img_at(ref, x - ref.ox, y - ref.oy);
      acc[0] += 1.0;
      acc[1] += d;
      acc[2] += d * d;
      acc[3] += r;
      acc[4] += r * r;
      acc[5] += d * r;
    }
    grp.sum(acc);
    const double cnt = acc[0];
    active = (int)cnt;
    if (cnt <= 0.0) { contrast = 0.0f; return -1.0f; }
    const double md = acc[1] / cnt, mr = acc[3] / cnt;
    const double sdd = fmax(acc[2] - cnt * md * md, 0.0);
    const double srr = fmax(acc[4] - cnt * mr * mr, 0.0);
    const double sdr = acc[5] - cnt * md * mr;
    contrast = (float)sqrt(sdd / cnt);
    // Subset::gamma returns -1 when either norm is zero (flat subset)
    if (sdd <= 1e-13 * acc[2] || srr <= 1e-13 * acc[4]) return -1.0f;
    double gam = 2.0 - 2.0 * sdr / (sqrt(sdd) * sqrt(srr));
    if (gam < 0.0) gam = 0.0;
    float g = (float)gam;
    if (kp.normalize_gamma && n_def_active > 0) g = (float)((double)g / (double)n_def_active);
    return g;
  }

  // Objective_ZNSSD::computeUpdateFast.  Returns the solver status; p_state
  // receives the parameters of the last warp (for sigma).
  __device__ __forceinline__ int gauss_newton(Affine &A, int &iters, float p_state[6], double *fields, int nsub, int gid) {
    const int max_its = kp.max_fast;
    float upd_old[6] = {0, 0, 0, 0, 0, 0};
    float shift = meanF;  // running estimate of meanG (one-pass ZSSD gradient)
    int it = 0;
    for (; it <= max_its; ++it) {
      iters = it;
      A.prepare();
#pragma unroll
      for (int k = 0; k < 6; ++k) p_state[k] = A.p[k];
      stage_window(A);
      // acc: q[6], H[21] (upper triangle), sr[6], sG, cnt
      double acc[35];
#pragma unroll
      for (int k = 0; k < 35; ++k) acc[k] = 0.0;
      #pragma unroll 1
      for (int i = grp.rank(); i < n; i += G::kSize) {
        int x, y;
        bool a;
        pixel(i, x, y, a);
        if (!a) continue;
        const float dx = (float)(x - cx), dy = (float)(y - cy);
        float ox, oy;
        A.disp(dx, dy, ox, oy);
        if (!kept(ox, oy)) continue;
        float Gx, Gy;
        const float Gv = sample(ox, oy, DEFGRAD, Gx, Gy);
        const int rx = x - ref.ox, ry = y - ref.oy;
        const float Fv = img_at(ref, rx, ry);
        if (!DEFGRAD) {
          const float gx = grad_at(ref.gx, ref, rx, ry), gy = grad_at(ref.gy, ref, rx, ry);
          Gx = A.c * gx - A.s * gy;
          Gy = A.s * gx + A.c * gy;
        }
        float r[6];
        A.residual(dx, dy, Gx, Gy, r);
        double rd[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) rd[k] = (double)(r[k] * en[k]);
        const double gmf = (double)((Gv - shift) - (Fv - meanF));
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] = fma(gmf, rd[k], acc[k]);
        int h = 6;
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
          for (int l = k; l < 6; ++l) { acc[h] = fma(rd[k], rd[l], acc[h]); ++h; }
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[27 + k] += rd[k];
        acc[33] += (double)Gv;
        acc[34] += 1.0;
      }
      grp.sum(acc);
      const double cnt = acc[34];
      const float meanG = cnt > 0.0 ? (float)(acc[33] / cnt) : 0.0f;
      const double corr = (double)meanG - (double)shift;
      double q[6], H[6][6];
#pragma unroll
      for (int k = 0; k < 6; ++k) q[k] = acc[k] - corr * acc[27 + k];
      {
        int h = 6;
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
          for (int l = k; l < 6; ++l) { H[k][l] = acc[h]; H[l][k] = acc[h]; ++h; }
      }
      // disabled parameters: identity rows, zero gradient (same solution as the
      // reduced N x N system of the reference)
#pragma unroll
      for (int k = 2; k < 6; ++k)
        if (en[k] == 0.0f) {
#pragma unroll
          for (int l = 0; l < 6; ++l) { H[k][l] = 0.0; H[l][k] = 0.0; }
          H[k][k] = 1.0;
          q[k] = 0.0;
        }
      if (kp.lm_alpha > 0.0f) { H[0][0] += kp.lm_alpha; H[1][1] += kp.lm_alpha; }
      // 2x2 condition estimate in scalar_t (float) as the reference stores it
      const float det_h = (float)(H[0][0] * H[1][1] - H[1][0] * H[0][1]);
      const float norm_H = (float)sqrt(H[0][0] * H[0][0] + H[0][1] * H[0][1] + H[1][0] * H[1][0] + H[1][1] * H[1][1]);
      float cond = -1.0f;
      if (det_h != 0.0f) {
        const float dd = det_h * det_h;
        const float norm_Hi = (float)sqrt((1.0 / (double)dd) *
                                          (H[0][0] * H[0][0] + H[0][1] * H[0][1] + H[1][0] * H[1][0] + H[1][1] * H[1][1]));
        cond = norm_H * norm_Hi;
      }
      if (grp.rank() == 0) fields[(size_t)F_CONDITION * nsub + gid] = cond;
      if (cond > 1.0e12f) return HESSIAN_SINGULAR;
      float old[6], du[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) old[k] = A.p[k];
      {
        double dd[6];
        lu_update6(H, q, dd);
#pragma unroll
        for (int k = 0; k < 6; ++k) du[k] = (float)dd[k];
      }
      if (kp.momentum != 0.0f) {
#pragma unroll
        for (int k = 0; k < 6; ++k) { du[k] += kp.momentum * upd_old[k]; upd_old[k] = du[k]; }
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) A.p[k] += du[k] * en[k];
      // test_for_convergence (DICe_LocalShapeFunction.cpp:433-459): u, v, theta
      bool conv = fabsf(A.p[0] - old[0]) < kp.fast_tol && fabsf(A.p[1] - old[1]) < kp.fast_tol;
      if (kp.en_rot) conv = conv && fabsf(A.p[2] - old[2]) < kp.fast_tol;
      if (conv) {
        if (kp.uncertainty) uncertainty(A, fields, nsub, gid);
        break;
      }
      shift = meanG;
    }
    if (it > max_its) return MAX_ITERATIONS_REACHED;
    return CORRELATION_SUCCESSFUL;
  }

  // computeUncertaintyFields without an image deformer (DICe_Objective.cpp:193-299):
  // exact displacement 0, so the "exact" residual reads the deformed pixel itself.
  __device__ __forceinline__ void uncertainty(const Affine &A, double *fields, int nsub, int gid) {
    const float u = A.p[0], v = A.p[1];
    double acc[6] = {0, 0, 0, 0, 0, 0};  // sssig, n_uhat_g2, n_err_g2, n_err_jg2, r_tot2, n_uhat2
    #pragma unroll 1
    for (int i = grp.rank(); i < n; i += G::kSize) {
      int x, y;
      bool a;
      pixel(i, x, y, a);
      const int rx = x - ref.ox, ry = y - ref.oy;
      // the subset's gradient arrays: reference gradients (or, in def-gradient mode,
      // the reference value for pixels this sketch does not re-interpolate)
      const float gx = grad_at(ref.gx, ref, rx, ry), gy = grad_at(ref.gy, ref, rx, ry);
      const float g2 = gx * gx + gy * gy;
      acc[0] += (double)g2;
      const float inv = g2 < 1.0e-4f ? 0.0f : 1.0f / g2;
      const float ug = u * gx + v * gy;
      const float jg = v * gx - u * gy;
      acc[1] += (double)(ug * ug * inv);
      acc[2] += (double)(ug * ug * inv);
      acc[3] += (double)(jg * jg * inv);
      const float refv = img_at(ref, rx, ry);
      const float sub_r = keys_single((float)rx + u, (float)ry + v) - refv;
      const float sub_e = keys_single((float)rx, (float)ry) - refv;
      acc[4] += (double)((sub_r - sub_e) * (sub_r - sub_e) * inv);
      acc[5] += (double)(u * u + v * v);
    }
    grp.sum(acc);
    if (grp.rank() != 0) return;
    const double sssig = (1.0 + acc[0]) / (double)(n == 0 ? 1 : n);
    const double cos_hat = acc[5] == 0.0 ? 1.0 : sqrt(acc[1] / acc[5]);
    fields[(size_t)F_FIELD_1 * nsub + gid] = (float)cos_hat;
    fields[(size_t)F_FIELD_2 * nsub + gid] = 1.0;
    fields[(size_t)F_FIELD_3 * nsub + gid] = (float)sqrt(acc[2] + acc[3]);
    fields[(size_t)F_FIELD_4 * nsub + gid] = cos_hat == 0.0 ? 0.0 : (float)(1.0 / cos_hat * sqrt(acc[4]));
    fields[(size_t)F_FIELD_5 * nsub + gid] = (float)sqrt(acc[2]);
    fields[(size_t)F_FIELD_6 * nsub + gid] = (float)sqrt(acc[4]);
    fields[(size_t)F_FIELD_7 * nsub + gid] = (float)sssig;
  }

  // interpolate_keys_fourth (DICe_Image.cpp:890-924) at local coordinates:
  // bilinear near the border (and 0 outside [0, w-1.5)).
  __device__ __forceinline__ float keys_single(float lx, float ly) const {
    const int w = def.w, h = def.h;
    if (lx <= 2.5f || lx >= w - 3.5f || ly <= 2.5f || ly >= h - 3.5f) {
      if (lx < 0.0f || lx >= w - 1.5f || ly < 0.0f || ly >= h - 1.5f) return 0.0f;
      const int x1 = (int)lx, y1 = (int)ly;
      const float tx = lx - x1, ty = ly - y1;
      const int o = y1 * w + x1;
      return ldpix<T>(def.pix, o) * (1.0f - tx) * (1.0f - ty) + ldpix<T>(def.pix, o + 1) * tx * (1.0f - ty) +
             ldpix<T>(def.pix, o + w + 1) * tx * ty + ldpix<T>(def.pix, o + w) * (1.0f - tx) * ty;
    }
    const int ix = (int)lx, iy = (int)ly;
    float wx[6], wy[6];
    keys_weights(lx - ix, wx);
    keys_weights(ly - iy, wy);
    float v = 0.0f;
    const int o = (iy - 2) * w + ix - 2;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      float rs = 0.0f;
#pragma unroll
      for (int k = 0; k < 6; ++k) rs = fmaf(ldpix<T>(def.pix, o + m * w + k), wx[k], rs);
      v = fmaf(rs, wy[m], v);
    }
    return v;
  }

  // Simplex::minimize via Objective::gamma (DICe_Simplex.cpp:60-277).  The
  // variables are the shape-function parameters: on exit A.p holds the LAST
  // evaluated point.  Vertex storage stays in registers: every vertex index
  // (ilo, ihi, inhi) is wave-uniform and is resolved with unrolled selects.
  __device__ __forceinline__ float eval_here(Affine &A) {
    int active;
    float contrast;
    A.prepare();
    return gamma_eval(A, active, contrast);
  }
  __device__ __forceinline__ int simplex(Affine &A, int &iters) {
    // slot of the k-th enabled parameter in reference order u, v, theta, ex, ey, gxy
    int slot[6] = {0, 1, 2, 3, 4, 5};
    int nd = 2;
    if (kp.en_rot) slot[nd++] = 2;
    if (kp.en_normal) { slot[nd++] = 3; slot[nd++] = 4; }
    if (kp.en_shear) slot[nd++] = 5;
    const int mpts = nd + 1;
    float pts[7][6], gv[7], sums[6], init[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) init[j] = A.p[j];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      gv[i] = 0.0f;
#pragma unroll
      for (int j = 0; j < 6; ++j) pts[i][j] = init[j];
    }
#pragma unroll
    for (int i = 1; i < 7; ++i) {
      // vertex i moves the (i-1)-th enabled parameter by its delta (DICe_LocalShapeFunction.cpp:189-194)
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (i - 1 < nd && j == slot[i - 1]) pts[i][j] += (j < 2) ? kp.delta_disp : kp.delta_theta;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (i < mpts) {
#pragma unroll
        for (int j = 0; j < 6; ++j) A.p[j] = pts[i][j];
        gv[i] = eval_here(A);
        if (i == 0 && gv[0] < kp.skip_thr && gv[0] >= 0.0f) { iters = 0; return CORRELATION_SUCCESSFUL; }
#pragma unroll
        for (int j = 0; j < 6; ++j) A.p[j] = init[j];
      }
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      float acc = 0.0f;
#pragma unroll
      for (int i = 0; i < 7; ++i) if (i < mpts) acc += pts[i][j];
      sums[j] = acc;
    }
    const float tiny = 1.0e-10f;
    float old_rtol = 0.0f, gamma_new = 0.0f, gamma_old = -1.0f;
    int repeats = 0;
    int iteration = 0;
    for (iteration = 0; iteration < kp.max_robust; iteration++) {
      if (iteration >= kp.max_robust - 1) { iters = iteration; return MAX_ITERATIONS_REACHED; }
      int ilo = 0, ihi, inhi;
      if (gv[0] > gv[1]) { inhi = 1; ihi = 0; } else { inhi = 0; ihi = 1; }
      float g_lo = gv[0], g_hi = gv[ihi == 0 ? 0 : 1], g_nhi = gv[inhi == 0 ? 0 : 1];
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        if (i < mpts) {
          if (gv[i] <= g_lo) { ilo = i; g_lo = gv[i]; }
          if (gv[i] > g_hi) { inhi = ihi; g_nhi = g_hi; ihi = i; g_hi = gv[i]; }
          else if (gv[i] > g_nhi && i != ihi) { inhi = i; g_nhi = gv[i]; }
        }
      }
      const float rtol = (float)(2.0 * fabsf(g_hi - g_lo) / (double)(fabsf(g_hi) + fabsf(g_lo) + tiny));
      const float rtol_diff = fabsf(rtol - old_rtol);
      old_rtol = rtol;
      if (fabsf(gamma_new - gamma_old) < 1.0e-8f) repeats++;
      gamma_old = gamma_new;
      if (rtol < kp.robust_tol || rtol_diff < kp.robust_tol || repeats > 15) break;
      float phi[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        float v = 0.0f;
#pragma unroll
        for (int k = 0; k < 7; ++k) if (k == ihi) v = pts[k][j];
        phi[j] = v;
      }
      // reflect / expand / contract (fac -1, 2, 0.5) with replacement of the high vertex
      float ytry = 0.0f;
      const float facs[3] = {-1.0f, 2.0f, 0.5f};
      int stage = 0;  // 0 reflect, 1 expand, 2 contract
      float ysave = 0.0f;
      bool shrink = false;
#pragma unroll
      for (int st = 0; st < 3; ++st) {
        if (st == stage) {
          const float fac = facs[st];
          const float fac1 = (float)((1.0 - fac) / nd);
          const float fac2 = fac1 - fac;
          if (st == 2) ysave = g_hi;
#pragma unroll
          for (int j = 0; j < 6; ++j) A.p[j] = sums[j] * fac1 - phi[j] * fac2;
#pragma unroll
          for (int j = 2; j < 6; ++j) if (en[j] == 0.0f) A.p[j] = init[j];
          ytry = eval_here(A);
          if (ytry < g_hi) {
#pragma unroll
            for (int k = 0; k < 7; ++k)
              if (k == ihi) {
                gv[k] = ytry;
#pragma unroll
                for (int j = 0; j < 6; ++j) { sums[j] += A.p[j] - pts[k][j]; pts[k][j] = A.p[j]; }
              }
#pragma unroll
            for (int j = 0; j < 6; ++j) phi[j] = A.p[j];
            g_hi = ytry;
          }
          if (st == 0) {
            if (ytry <= g_lo) stage = 1;
            else if (ytry >= g_nhi) stage = 2;
            else stage = 3;
          } else if (st == 1) {
            stage = 3;
          } else {
            shrink = ytry >= ysave;
            stage = 3;
          }
        }
      }
      if (shrink) {
        float plo[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          float v = 0.0f;
#pragma unroll
          for (int k = 0; k < 7; ++k) if (k == ilo) v = pts[k][j];
          plo[j] = v;
        }
#pragma unroll
        for (int i = 0; i < 7; ++i) {
          if (i < mpts && i != ilo) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
              pts[i][j] = sums[j] = (float)(0.5 * ((double)pts[i][j] + (double)plo[j]));
              A.p[j] = sums[j];
            }
            gv[i] = eval_here(A);
          }
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          float acc = 0.0f;
#pragma unroll
          for (int i = 0; i < 7; ++i) if (i < mpts) acc += pts[i][j];
          sums[j] = acc;
        }
      }
      // the extra objective evaluation at the current variables (DICe_Simplex.cpp:264)
      gamma_new = eval_here(A);
    }
    iters = iteration;
    return CORRELATION_SUCCESSFUL;
  }

  // Immerkaer noise estimate (Subset::noise_std_dev, DICe_Subset.cpp:622-703)
  // over the bbox of the kept pixels at p_state, shifted by the final (u, v).
  __device__ __forceinline__ float noise_std_dev(int min_x, int max_x, int min_y, int max_y, float u, float v) {
    min_x = (int)((float)min_x + u); max_x = (int)((float)max_x + u);
    min_y = (int)((float)min_y + v); max_y = (int)((float)max_y + v);
    const int ih = def.h, iw = def.w, ox = def.ox, oy = def.oy;
    if (min_x < ox) min_x = ox;
    if (min_y < oy) min_y = oy;
    if (max_x >= iw + ox) min_x = 0;
    if (max_y >= ih + oy) min_y = 0;
    int hh = max_y - min_y + 1, ww = max_x - min_x + 1;
    if (ww < 3) ww = 3;
    if (hh < 3) hh = 3;
    const int bw = max_x - min_x, bh = max_y - min_y;
    double acc[1] = {0.0};
    if (bw > 0 && bh > 0) {
      const int tot = bw * bh;
      for (int k = grp.rank(); k < tot; k += G::kSize) {
        const int r = k / bw;
        const int xx = min_x + (k - r * bw), yy = min_y + r;
        const int lx = xx - ox, ly = yy - oy;
        if (lx < 1 || lx >= iw - 1 || ly < 1 || ly >= ih - 1) {
          const float p = img_at(def, lx, ly);
          acc[0] += (double)sqrtf(p * p);
        } else {
          const int o = ly * iw + lx;
          const void *P = def.pix;
          const float conv = ldpix<T>(P, o - iw - 1) - 2.0f * ldpix<T>(P, o - iw) + ldpix<T>(P, o - iw + 1) -
                             2.0f * ldpix<T>(P, o - 1) + 4.0f * ldpix<T>(P, o) - 2.0f * ldpix<T>(P, o + 1) +
                             ldpix<T>(P, o + iw - 1) - 2.0f * ldpix<T>(P, o + iw) + ldpix<T>(P, o + iw + 1);
          acc[0] += (double)fabsf(conv);
        }
      }
    }
    grp.sum(acc);
    const float variance = (float)acc[0];
    return (float)((double)variance * (sqrt(0.5 * 3.14159265358979323846) / (6.0 * (ww - 2) * (hh - 2))));
  }

  // Objective::sigma (DICe_Objective.cpp:150-184) with the subset state of the
  // last warp (parameters p_state) and the final parameters A.
  __device__ __forceinline__ float sigma(const float p_state[6], const Affine &A, float &noise) {
    Affine Ps;
#pragma unroll
    for (int k = 0; k < 6; ++k) Ps.p[k] = p_state[k];
    Ps.prepare();
    int mnx = cx, mxx = cx, mny = cy, mxy = cy;
    double acc[3] = {0, 0, 0};  // sum gx^2, sum gy^2, kept count
    #pragma unroll 1
    for (int i = grp.rank(); i < n; i += G::kSize) {
      int x, y;
      bool a;
      pixel(i, x, y, a);
      if (!a) continue;
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      Ps.disp(dx, dy, ox, oy);
      if (!kept(ox, oy)) continue;
      mnx = min(mnx, x); mxx = max(mxx, x);
      mny = min(mny, y); mxy = max(mxy, y);
      float gx, gy;
      if (DEFGRAD) {
        const bool keep_win = use_win;
        use_win = false;
        sample(ox, oy, true, gx, gy);
        use_win = keep_win;
      } else {
        gx = grad_at(ref.gx, ref, x - ref.ox, y - ref.oy);
        gy = grad_at(ref.gy, ref, x - ref.ox, y - ref.oy);
      }
      acc[0] += (double)(gx * gx);
      acc[1] += (double)(gy * gy);
      acc[2] += 1.0;
    }
    grp.sum(acc);
    mnx = grp.imin(mnx); mxx = grp.imax(mxx);
    mny = grp.imin(mny); mxy = grp.imax(mxy);
    if (acc[2] <= 0.0) { noise = -1.0f; return -1.0f; }
    noise = noise_std_dev(mnx, mxx, mny, mxy, A.p[0], A.p[1]);
    if (noise < 0.0f) return -1.0f;
    const float sgx = (float)acc[0], sgy = (float)acc[1];
    const float sg = sgx > sgy ? sgy : sgx;
    return sg > 0.0f ? (float)sqrt(2.0 * (double)noise * (double)noise / (double)sg) : -1.0f;
  }

  __device__ __forceinline__ void record_failed(double *fields, int nsub, int gid, int status, int iters) {
    if (grp.rank() != 0) return;
    auto f = [&](int s) -> double & { return fields[(size_t)s * nsub + gid]; };
    f(F_SIGMA) = -1.0; f(F_MATCH) = -1.0; f(F_GAMMA) = -1.0; f(F_BETA) = -1.0;
    f(F_NOISE) = -1.0; f(F_CONTRAST) = -1.0; f(F_ACTIVE_PIXELS) = -1.0;
    f(F_STATUS) = status; f(F_ITERATIONS) = iters;
  }

  __device__ __forceinline__ void init_params(Affine &A, const double *fields, int nsub, int gid) {
    auto f = [&](int s) { return fields[(size_t)s * nsub + gid]; };
    A.p[0] = (float)f(F_DISP_X);
    A.p[1] = (float)f(F_DISP_Y);
    A.p[2] = kp.en_rot ? (float)f(F_ROT_Z) : 0.0f;
    A.p[3] = kp.en_normal ? (float)f(F_NS_XX) : 0.0f;
    A.p[4] = kp.en_normal ? (float)f(F_NS_YY) : 0.0f;
    A.p[5] = kp.en_shear ? (float)f(F_SS_XY) : 0.0f;
  }

  // generic_correlation_routine (DICe_Schema.cpp:2025-2368) for one subset,
  // split in two passes so that each kernel holds one optimizer:
  //   pass 0: FOV / initial-guess checks, the primary optimizer, and either
  //           the epilogue (sigma, gamma, record) or a fallback request;
  //   pass 1: the fallback optimizer of GRADIENT_BASED_THEN_SIMPLEX /
  //           SIMPLEX_THEN_GRADIENT_BASED for the subsets that asked for it.
  // SIMPLEX_OPT selects which optimizer this instantiation runs.
  template <bool SIMPLEX_OPT>
  __device__ __forceinline__ void run(const DSubsets &ss, int gid, double *fields, int pass, int *fb_list,
                                      int *fb_count) {
    const int nsub = ss.n;
    set_subset(ss, gid);
    auto fv = [&](int s) { return fields[(size_t)s * nsub + gid]; };
    if (pass == 0) {
      // field-of-view test on the first frame when a subset size is set (:2033-2049)
      if (ss.fov > 0 && kp.first_frame) {
        const float cpx = (float)(fv(F_COORD_X) + fv(F_DISP_X));
        const float cpy = (float)(fv(F_COORD_Y) + fv(F_DISP_Y));
        const int h2 = ss.fov / 2;
        if (cpx < def.ox + h2 || cpx > def.w + def.ox - h2 || cpy < def.oy + h2 || cpy > def.h + def.oy - h2) {
          record_failed(fields, nsub, gid, INITIALIZE_FAILED_BY_EXCEPTION, -1);
          return;
        }
      }
      // Field_Value_Initializer (DICe_Initializer.cpp:398-434)
      if (fv(F_SIGMA) == -1.0) { record_failed(fields, nsub, gid, INITIALIZE_FAILED, -1); return; }
    }
    Affine A;
    init_params(A, fields, nsub, gid);
    const int init_status = INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL;
    int iters = -1;
    if (!SIMPLEX_OPT) {
      // mean of the reference subset over defined-active pixels (DICe_Objective.cpp:464)
      double acc[2] = {0.0, 0.0};
      #pragma unroll 1
      for (int i = grp.rank(); i < n; i += G::kSize) {
        int x, y;
        bool a;
        pixel(i, x, y, a);
        if (!a) continue;
        acc[0] += (double)img_at(ref, x - ref.ox, y - ref.oy);
        acc[1] += 1.0;
      }
      grp.sum(acc);
      meanF = acc[1] > 0.0 ? (float)(acc[0] / acc[1]) : 0.0f;
    }
    if (pass == 0 && kp.init_gamma_thr != -1.0) {
      int act;
      float con;
      A.prepare();
      const float g0 = gamma_eval(A, act, con);
      if (g0 > kp.init_gamma_thr || g0 < 0.0f) { record_failed(fields, nsub, gid, INITIALIZE_FAILED, iters); return; }
    }
    const float prev_u = (float)fv(F_DISP_X), prev_v = (float)fv(F_DISP_Y), prev_t = (float)fv(F_ROT_Z);
    float p_state[6];
    int corr;
    if (SIMPLEX_OPT) {
      corr = simplex(A, iters);
#pragma unroll
      for (int k = 0; k < 6; ++k) p_state[k] = A.p[k];
    } else {
      corr = gauss_newton(A, iters, p_state, fields, nsub, gid);
    }
    const float du = A.p[0] - prev_u, dv = A.p[1] - prev_v, dt = (kp.en_rot ? A.p[2] : 0.0f) - prev_t;
    const bool jump = fabsf(du) > kp.disp_jump_tol || fabsf(dv) > kp.disp_jump_tol || fabsf(dt) > kp.theta_jump_tol;
    if (pass == 0) {
      if (corr != CORRELATION_SUCCESSFUL || jump) {
        if (kp.opt == 3 || kp.opt == 4) {
          // the fallback starts again from the initial guess (:2238, :2266)
          if (grp.rank() == 0) {
            const int slot = atomicAdd(fb_count, 1);
            fb_list[slot] = gid;
          }
          return;
        }
        if (corr != CORRELATION_SUCCESSFUL) { record_failed(fields, nsub, gid, corr, iters); return; }
        // a successful solve that jumped goes on and fails the later jump test
      }
    } else if (corr != CORRELATION_SUCCESSFUL) {
      record_failed(fields, nsub, gid, corr, iters);
      return;
    }
    float noise = 0.0f;
    const float sig = sigma(p_state, A, noise);
    if (sig < 0.0f) { record_failed(fields, nsub, gid, FRAME_FAILED_DUE_TO_NEGATIVE_SIGMA, iters); return; }
    A.prepare();
    int active;
    float contrast;
    const float gam = gamma_eval(A, active, contrast);
    if ((kp.final_gamma_thr != -1.0 && gam > kp.final_gamma_thr) || gam < 0.0f) {
      record_failed(fields, nsub, gid, FRAME_FAILED_DUE_TO_HIGH_GAMMA, iters);
      return;
    }
    if (jump) { record_failed(fields, nsub, gid, JUMP_TOLERANCE_EXCEEDED, iters); return; }
    if (grp.rank() == 0) {
      auto f = [&](int s) -> double & { return fields[(size_t)s * nsub + gid]; };
      f(F_DISP_X) = A.p[0];
      f(F_DISP_Y) = A.p[1];
      if (kp.en_rot) f(F_ROT_Z) = A.p[2];
      if (kp.en_normal) { f(F_NS_XX) = A.p[3]; f(F_NS_YY) = A.p[4]; }
      if (kp.en_shear) f(F_SS_XY) = A.p[5];
      f(F_SIGMA) = sig; f(F_MATCH) = 0.0; f(F_GAMMA) = gam; f(F_BETA) = 0.0;
      f(F_NOISE) = noise; f(F_CONTRAST) = contrast; f(F_ACTIVE_PIXELS) = active;
      f(F_STATUS) = init_status; f(F_ITERATIONS) = iters;
    }
  }
};

// ---------------------------------------------------------------- kernels ---

// Work items: pass 0 walks all subsets; pass 1 walks the fallback list.
__device__ __forceinline__ int work_item(int k, int pass, const int *fb_list) { return pass == 0 ? k : fb_list[k]; }
__device__ __forceinline__ int work_count(int n, int pass, const int *fb_count) { return pass == 0 ? n : *fb_count; }

template <typename T, int INTERP, bool DEFGRAD, bool SIMPLEX_OPT>
__global__ __launch_bounds__(kPlainThreads) void correlate_plain_kernel(KParams kp, DImage ref, DImage def,
                                                                        DSubsets ss, double *fields, int pass,
                                                                        int *fb_list, int *fb_count) {
  __shared__ double red[4 * kMaxRed];
  const int k = blockIdx.x;
  lds_double *red_l = (lds_double *)red;
  if (k >= work_count(ss.n, pass, fb_count)) return;
  const int gid = work_item(k, pass, fb_list);
  BlockGroup g{red_l};
  Solver<T, BlockGroup, INTERP, DEFGRAD> s(kp, ref, def, g, (lds_float *)red, 0);
  s.template run<SIMPLEX_OPT>(ss, gid, fields, pass, fb_list, fb_count);
}

template <typename T, int INTERP, bool DEFGRAD, bool SIMPLEX_OPT>
__global__ __launch_bounds__(256) void correlate_batched_kernel(KParams kp, DImage ref, DImage def, DSubsets ss,
                                                                double *fields, int pass, int *fb_list,
                                                                int *fb_count, int *counter, int lds_per_wave) {
  extern __shared__ float lds[];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  WaveGroup g{lane};
  Solver<T, WaveGroup, INTERP, DEFGRAD> s(kp, ref, def, g, (lds_float *)(lds + (size_t)wave * lds_per_wave),
                                          lds_per_wave);
  const int total = work_count(ss.n, pass, fb_count);
  while (true) {
    int k = 0;
    if (lane == 0) k = atomicAdd(counter, 1);
    k = __shfl(k, 0, kWave);
    if (k >= total) break;
    s.template run<SIMPLEX_OPT>(ss, work_item(k, pass, fb_list), fields, pass, fb_list, fb_count);
  }
}

// Launch one pass of one kernel variant.
template <typename T, int INTERP, bool DEFGRAD, bool SX, int MODE>
hipError_t launch_variant(const Launch &L, const KParams &kp, const DImage &ref, const DImage &def,
                          const DSubsets &subs, double *fields, int pass, int *fb_list, int *fb_count,
                          int *counter, hipStream_t stream) {
  if (MODE == 0) {
    hipLaunchKernelGGL((correlate_plain_kernel<T, INTERP, DEFGRAD, SX>), dim3(subs.n), dim3(kPlainThreads), 0, stream,
                       kp, ref, def, subs, fields, pass, fb_list, fb_count);
    return hipGetLastError();
  }
  const size_t shm = (size_t)L.waves_per_group * L.lds_floats_per_wave * sizeof(float);
  hipLaunchKernelGGL((correlate_batched_kernel<T, INTERP, DEFGRAD, SX>), dim3(L.groups),
                     dim3(L.waves_per_group * kWave), shm, stream, kp, ref, def, subs, fields, pass, fb_list,
                     fb_count, counter, L.lds_floats_per_wave);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace dice_hip
