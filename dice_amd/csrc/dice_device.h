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
//   PLAIN   : a 256-thread workgroup per subset.  Deformed intensities are
//             gathered straight from HBM/L2, every per-pixel product is added
//             in FP64 (as the reference does) and reductions go through LDS
//             with barriers.
//   BATCHED : one 64-lane wavefront per subset.  Waves of a persistent grid
//             pull subset indices from a global work counter (work stealing).
//             Before every warp evaluation the wave stages the deformed window
//             that covers the warped subset (+ interpolation support) into its
//             LDS slot with coalesced loads and gathers all interpolation taps
//             from LDS.  Per-pixel products go to FP32 per-lane partial sums
//             (at most ceil(N/64) pixels each) that are added across lanes in
//             FP64 through an LDS transpose; the 6x6 system is solved in FP64
//             with one row per lane.  No barriers: a wave synchronises itself.
//             A subset whose window does not fit the slot is handed over to the
//             PLAIN kernel (overflow list) and restarted there.
#pragma once
#include "dice_kernels.h"

#include <math.h>

#include <type_traits>

namespace dice_hip {
namespace dev {

// Pointers carry their address space so that loads compile to ds_read (LDS)
// and global_load (HBM) with immediate offsets instead of flat accesses.
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) double lds_double;
typedef __attribute__((address_space(1))) const float gfloat;
typedef __attribute__((address_space(1))) const uint16_t gu16;
typedef __attribute__((address_space(1))) const int gint;
typedef __attribute__((address_space(1))) const uint8_t gu8;
typedef __attribute__((address_space(1))) const uint32_t gu32;

// ------------------------------------------------------------ utilities ---
// Cross-lane movement without LDS addresses: DPP within rows of 16 lanes and
// v_readlane to combine the four rows.  Every lane of the wave must be active.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // lane i <-> 7 - i within 8
constexpr int kDppMirror = 0x140;     // lane i <-> 15 - i within 16
template <int CTRL, typename V>
__device__ __forceinline__ V dpp(V v) {
  return __builtin_amdgcn_update_dpp(V(0), v, CTRL, 0xF, 0xF, false);
}
// v_readlane / v_readfirstlane move 32 bits: 64-bit values go as two halves
// (the builtins take int, so a double argument would be converted).
__device__ __forceinline__ int lane_value(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
// A value that is the same in every lane, made wave-uniform (scalar register).
__device__ __forceinline__ double uniform(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
// Sum over a row of 16 lanes (every lane of the row active).  Each step adds a
// value and its partner's, so all lanes of the row hold bit-identical sums: the
// pairwise tree (((v0 + v1) + (v2 + v3)) + ((v4 + v5) + (v6 + v7))) + (same
// for v8..v15).
__device__ __forceinline__ double row_sum16(double v) {
  v += dpp<kDppXor1>(v);
  v += dpp<kDppXor2>(v);
  v += dpp<kDppHalfMirror>(v);
  v += dpp<kDppMirror>(v);
  return v;
}
// The same tree over each half-row of 8 lanes (the first three steps).
__device__ __forceinline__ double row_sum8(double v) {
  v += dpp<kDppXor1>(v);
  v += dpp<kDppXor2>(v);
  v += dpp<kDppHalfMirror>(v);
  return v;
}
template <int W>
__device__ __forceinline__ double row_sum(double v) {
  static_assert(W == 8 || W == 16, "rows of 8 or 16 lanes");
  if constexpr (W == 8) return row_sum8(v);
  else return row_sum16(v);
}
// Sum over the wave: the four row sums added by v_readlane; the result is uniform.
__device__ __forceinline__ double wave_sum(double v) {
  v = row_sum16(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, dpp<kDppXor1>(v));
  v = min(v, dpp<kDppXor2>(v));
  v = min(v, dpp<kDppHalfMirror>(v));
  v = min(v, dpp<kDppMirror>(v));
  return min(min(lane_value(v, 0), lane_value(v, 16)), min(lane_value(v, 32), lane_value(v, 48)));
}
__device__ __forceinline__ int wave_max(int v) {
  v = max(v, dpp<kDppXor1>(v));
  v = max(v, dpp<kDppXor2>(v));
  v = max(v, dpp<kDppHalfMirror>(v));
  v = max(v, dpp<kDppMirror>(v));
  return max(max(lane_value(v, 0), lane_value(v, 16)), max(lane_value(v, 32), lane_value(v, 48)));
}

// The lane id, re-read at every use: values derived from it cannot be hoisted
// out of the pixel loops and stay live across the whole kernel (which pushes
// the register budget into spills).
__device__ __forceinline__ int opaque_lane(int lane) {
  asm volatile("" : "+v"(lane));
  return lane;
}

// One wavefront.
struct WaveGroup {
  int lane;
  static constexpr int kSize = kWave;
  __device__ __forceinline__ int rank() const { return opaque_lane(lane); }
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
  // rank 0's values to every rank
  template <int K>
  __device__ __forceinline__ void bcast0(double (&v)[K]) const {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = lane_value(v[k], 0);
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
  template <int K>
  __device__ __forceinline__ void bcast0(double (&v)[K]) const {
    static_assert(K <= kMaxRed, "broadcast too wide");
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < K; ++k) red[k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = red[k];
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

// float element i (i < 2^30) of a global array through a 32-bit byte offset
typedef __attribute__((address_space(1))) const char gchar;
__device__ __forceinline__ float ldf_u(const void *p, unsigned i) {
  return *(const gfloat *)((const gchar *)p + (i << 2));
}

// Pixel storage of the images (storage_t, DICe.h:105-112): float or uint16.
template <typename T> struct Pix;
// at_u: element i (i < 2^30) through a 32-bit byte offset, which the load
// takes as an unsigned VGPR offset from the scalar base (no 64-bit address math).
template <> struct Pix<float> {
  typedef gfloat G;
  __device__ __forceinline__ static float at(const void *p, int i) { return ((const gfloat *)p)[i]; }
  __device__ __forceinline__ static float at_u(const void *p, unsigned i) { return ldf_u(p, i); }
  __device__ __forceinline__ static const gfloat *ptr(const void *p) { return (const gfloat *)p; }
};
template <> struct Pix<uint16_t> {
  typedef gu16 G;
  __device__ __forceinline__ static float at(const void *p, int i) { return (float)((const gu16 *)p)[i]; }
  __device__ __forceinline__ static float at_u(const void *p, unsigned i) {
    return (float)*(const gu16 *)((const gchar *)p + (i << 1));
  }
  __device__ __forceinline__ static const gu16 *ptr(const void *p) { return (const gu16 *)p; }
};
// Kept for the image-level kernels.
template <typename T> __device__ __forceinline__ float ldpix(const void *p, int i) { return Pix<T>::at(p, i); }

// Keys 4th-order cubic-convolution weights (Keys 1981; DICe_Image.cpp:114-122,
// 860-871) of the six taps at offsets -2..+3 from floor(x), written as cubics
// in the fractional offset t in [0, 1).  Same function as the reference's
// f2(t+2), f1(t+1), f0(t), f0(1-t), f1(2-t), f2(3-t).
__device__ __forceinline__ void keys_weights(float t, float w[6]) {
  // Horner forms: 17 operations for the six weights
  const float t2 = t * t;
  w[0] = fmaf(fmaf(t, 1.0f / 12.0f, -1.0f / 6.0f), t, 1.0f / 12.0f) * t;
  w[1] = fmaf(fmaf(t, -7.0f / 12.0f, 1.25f), t, -2.0f / 3.0f) * t;
  w[2] = fmaf(fmaf(t, 4.0f / 3.0f, -7.0f / 3.0f), t2, 1.0f);
  w[3] = fmaf(fmaf(t, -4.0f / 3.0f, 5.0f / 3.0f), t, 2.0f / 3.0f) * t;
  w[4] = fmaf(fmaf(t, 7.0f / 12.0f, -0.5f), t, -1.0f / 12.0f) * t;
  w[5] = fmaf(t, -1.0f / 12.0f, 1.0f / 12.0f) * t2;
}

// A wave-uniform double as a scalar value (both halves through readfirstlane),
// so that it takes SGPRs rather than two VGPRs.
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ float wave_uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// d/dt of keys_weights (the slopes of the six cubics).
__device__ __forceinline__ void keys_weight_slopes(float t, float w[6]) {
  const float t2 = t * t;
  w[0] = (3.0f * t2 - 4.0f * t + 1.0f) * (1.0f / 12.0f);
  w[1] = -1.75f * t2 + 2.5f * t - (2.0f / 3.0f);
  w[2] = 4.0f * t2 - (14.0f / 3.0f) * t;
  w[3] = -4.0f * t2 + (10.0f / 3.0f) * t + (2.0f / 3.0f);
  w[4] = 1.75f * t2 - t - (1.0f / 12.0f);
  w[5] = (2.0f * t - 3.0f * t2) * (1.0f / 12.0f);
}

// 6x6 Keys sum from rows of a 2-D array (LDS window or image) at `b`, pitch p.
// Rows are reached by adding the pitch (no per-row multiply).
template <typename P>
__device__ __forceinline__ float keys_sum(const P *b, int p, const float wx[6], const float wy[6]) {
  float v = 0.0f;
  const P *row = b;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    float rs = (float)row[0] * wx[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) rs = fmaf((float)row[k], wx[k], rs);
    v = m == 0 ? rs * wy[0] : fmaf(rs, wy[m], v);
    row += p;
  }
  return v;
}
template <typename P>
__device__ __forceinline__ float bilinear_sum(const P *b, int p, float tx, float ty) {
  // explicit products and fmaf: no multiply-add is left for the compiler to
  // contract, so every kernel that inlines this rounds it the same way
  const float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty), w11 = tx * ty, w01 = (1.0f - tx) * ty;
  float v = (float)b[0] * w00;
  v = fmaf((float)b[1], w10, v);
  v = fmaf((float)b[p + 1], w11, v);
  return fmaf((float)b[p], w01, v);
}

// Affine map and its Jacobian rows (Affine_Shape_Function::map / residuals,
// DICe_LocalShapeFunction.cpp:220-248, 350-396), parameters in the fixed slot
// order u, v, theta, ex, ey, gxy.
struct Affine {
  float p[6];
  float c, s;  // cos, sin theta
  float a11, a12, a21, a22;
  __device__ __forceinline__ void prepare() {
    c = cosf(p[2]);
    s = sinf(p[2]);
    const float sx = 1.0f + p[3], sy = 1.0f + p[4], g = p[5];
    // x' - cx = c*Dx - s*Dy + u, Dx = sx*dx + g*dy, Dy = sy*dy + g*dx.  Here and in
    // disp the multiply-adds are explicit fmaf: nothing is left for the compiler
    // to contract, so every kernel computes the same warp (the Nelder-Mead paths
    // of the execution modes must agree bit for bit)
    a11 = fmaf(c, sx, -(s * g));
    a12 = fmaf(c, g, -(s * sy));
    a21 = fmaf(s, sx, c * g);
    a22 = fmaf(s, g, c * sy);
  }
  // displacement of pixel offset (dx, dy) from the centroid: x' - cx, y' - cy
  __device__ __forceinline__ void disp(float dx, float dy, float &ox, float &oy) const {
    // (the dx products first: a pixel walk that keeps dx and steps dy computes
    // them once per column)
    ox = p[0] + fmaf(a12, dy, a11 * dx);
    oy = p[1] + fmaf(a22, dy, a21 * dx);
  }
  // residual row; (Gx, Gy) is the gradient to use (rotated ref or interpolated def)
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

// Same update with the work spread over a wavefront: lane i (< 6) owns row i
// of [H | -q].  Per column: first-maximal pivot among the remaining rows, row
// exchange, scaled multipliers and the rank-1 update of the rows below -- the
// arithmetic of the unblocked LU above element for element -- with the
// right-hand side carried along (forward substitution), then back
// substitution.  Pivot column, pivot rows and the solution are read with
// v_readlane, so they are wave-uniform.  Exactly-zero pivot: du = -(factors) q.
// hrow / qv hold row `lane` of H and q on lanes 0..5 (anything on other lanes).
__device__ __forceinline__ void wave_lu_update6(double (&hrow)[6], double qv, double (&du)[6]) {
  const int lane = opaque_lane((int)(threadIdx.x % kWave));
  const bool mine = lane < 6;
  double b = -qv;
  bool singular = false;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    // pivot: first row p >= j with the largest |A[p][j]|
    double pv = lane_value(hrow[j], j);
    double best = fabs(pv);
    int p = j;
#pragma unroll
    for (int r = j + 1; r < 6; ++r) {
      const double a = lane_value(hrow[j], r);
      if (fabs(a) > best) { best = fabs(a); p = r; pv = a; }
    }
    if (pv != 0.0) {
      if (p != j) {
        // exchange rows j and p (right-hand side included)
        double rj[6], rp[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) { rj[k] = lane_value(hrow[k], j); rp[k] = lane_value(hrow[k], p); }
        const double bj = lane_value(b, j), bp = lane_value(b, p);
        if (lane == j) {
#pragma unroll
          for (int k = 0; k < 6; ++k) hrow[k] = rp[k];
          b = bp;
        } else if (lane == p) {
#pragma unroll
          for (int k = 0; k < 6; ++k) hrow[k] = rj[k];
          b = bj;
        }
      }
      const double r = 1.0 / pv;
      if (mine && lane > j) hrow[j] *= r;
    } else {
      singular = true;
    }
    // rank-1 update of the rows below j with (the new) row j
    if (j < 5) {
      double rowj[6];
#pragma unroll
      for (int k = j + 1; k < 6; ++k) rowj[k] = lane_value(hrow[k], j);
      const double bj = lane_value(b, j);
      if (mine && lane > j) {
        const double l = hrow[j];
#pragma unroll
        for (int k = j + 1; k < 6; ++k) hrow[k] = fma(-l, rowj[k], hrow[k]);
        if (pv != 0.0) b = fma(-l, bj, b);
      }
    }
  }
  if (singular) {
    // GETRI returned early: du = (stored factors) * (-q), q in its original order
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) s = fma(hrow[k], -lane_value(qv, k), s);
#pragma unroll
    for (int i = 0; i < 6; ++i) du[i] = lane_value(s, i);
    return;
  }
  // back substitution U x = y
#pragma unroll
  for (int k = 5; k >= 0; --k) {
    const double xk = lane_value(b, k) / lane_value(hrow[k], k);
    du[k] = xk;
    if (lane < k) b = fma(-hrow[k], xk, b);
  }
}

// The same update du = -H^-1 q by H = L D L^T without pivoting, for the
// symmetric positive definite Gauss-Newton matrix: 6 divisions and about a
// hundred FP64 multiply-adds, every lane computing the same values (no
// cross-lane traffic).  It agrees with the pivoted LU above to rounding in FP64
// (the update is then rounded to FP32).  Returns false, leaving du unset, when
// a pivot is not safely positive (a singular or indefinite system, or NaN):
// the caller then takes the pivoted LU, which reproduces the reference there.
// H = L D L^T of a symmetric 6 x 6 (lower triangle read): L strictly lower,
// inv = 1 / diag(D).  False when a pivot is not safely positive (or NaN).
__device__ __forceinline__ bool ldlt_factor6(const double (&H)[6][6], double (&L)[6][6], double (&inv)[6]) {
  double d[6];
  double hmax = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) hmax = fmax(hmax, H[i][i]);
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double v[6];
    double dj = H[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) {
      v[k] = L[j][k] * d[k];
      dj = fma(-L[j][k], v[k], dj);
    }
    d[j] = dj;
    ok = ok && dj > 1.0e-13 * hmax;  // false for NaN too
    inv[j] = 1.0 / dj;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double s = H[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s = fma(-L[i][k], v[k], s);
      L[i][j] = s * inv[j];
    }
  }
  return ok;
}
__device__ __forceinline__ bool ldlt_update6(const double (&H)[6][6], const double (&q)[6], double (&du)[6]) {
  double L[6][6], inv[6];
  if (!ldlt_factor6(H, L, inv)) return false;
  // L y = -q, then z = D^-1 y, then L^T du = z
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = -q[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s = fma(-L[i][k], y[k], s);
    y[i] = s;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) y[i] *= inv[i];
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s = fma(-L[k][i], du[k], s);
    du[i] = s;
  }
  return true;
}

// 2x2 condition estimate of the (u, v) block in scalar_t, as the reference
// stores it (DICe_Objective.cpp:558-564); -1 when the determinant is 0.
__device__ __forceinline__ float cond2x2(double h00, double h01, double h10, double h11) {
  const float det_h = (float)(h00 * h11 - h10 * h01);
  const float norm_H = (float)sqrt(h00 * h00 + h01 * h01 + h10 * h10 + h11 * h11);
  float c = -1.0f;
  if (det_h != 0.0f) {
    const float dd = det_h * det_h;
    const float norm_Hi = (float)sqrt((1.0 / (double)dd) * (h00 * h00 + h01 * h01 + h10 * h10 + h11 * h11));
    c = norm_H * norm_Hi;
  }
  return c;
}

// Status of a wave that handed its subset over (never stored in a field).
constexpr int kOverflow = -1000;

// ------------------------------------------------------ subset solver ---
template <typename T, typename G, int INTERP, bool DEFGRAD>
struct Solver {
  static constexpr bool kIsWave = G::kSize == kWave;
  const KParams kp;
  const DImage ref;
  const DImage def;
  G grp;
  // LDS slot of this wave (BATCHED only): deformed window / reduction scratch
  lds_float *win;
  int win_cap;
  // BATCHED fast Gauss-Newton pass (reference gradients): per-subset constants
  // sum(phi phi^T) at pers[0, 36) and sum(phi) at pers[36, 42) over the active
  // pixels, phi = (gx, gy, dx gx, dy gx, dx gy, dy gy); last_fast: the last
  // Gauss-Newton warp kept every active pixel
  static constexpr bool kFast = kIsWave && !DEFGRAD;
  lds_double *pers;
  bool last_fast;
  int wx0, wy0, wpitch, wrows;
  bool win_ok;  // the LDS slot holds the deformed image over the window above
  bool overflow;
  // subset
  int cx, cy, n, S, hw;
  int bx0, by0, bx1, by1;
  bool ref_in;       // every subset pixel lies inside the reference image
  int n_def_active;  // pixels flagged active in the subset definition
  const gint *px;
  const gint *py;
  const gu8 *pact;
  const gu32 *pk;       // conformal pixels packed (x | y << 15 | inactive << 30), or null
  int abx0, aby0, abx1, aby1;  // conformal: bbox of the active pixels
  // blocked image pixels of the subset: box (mx0, my0, mw, mh), bits at mb (mw = 0: none)
  int mx0, my0, mw, mh;
  const gu32 *mb;
  // a workgroup's LDS tables for the objective of subsets of at most
  // kLaneSimplexMaxPixels pixels (3 kLaneSimplexMaxPixels floats), or null
  lds_float *small_tab = nullptr;
  bool forced;  // Nelder-Mead only
  float meanF;
  float en[6];  // 1 for enabled parameters

  __device__ __forceinline__ Solver(const KParams &k, const DImage &r, const DImage &d, G g, lds_float *w, int cap)
      : kp(k), ref(r), def(d), grp(g), win(w), win_cap(kFast ? cap - kPersistFloats : cap), wx0(0), wy0(0),
        wpitch(0), wrows(0), win_ok(false), overflow(false) {
    pers = kFast ? (lds_double *)(w + win_cap) : nullptr;
    last_fast = false;
    en[0] = en[1] = 1.0f;
    en[2] = k.en_rot ? 1.0f : 0.0f;
    en[3] = en[4] = k.en_normal ? 1.0f : 0.0f;
    en[5] = k.en_shear ? 1.0f : 0.0f;
  }

  __device__ __forceinline__ void set_subset(const DSubsets &ss, int i) {
    cx = ss.cx[i];
    cy = ss.cy[i];
    S = ss.S;
    overflow = false;
    win_ok = false;
    if (S > 0) {
      hw = S / 2;
      S = 2 * hw + 1;  // square subsets are odd (DICe_Subset.cpp:90-92)
      n = S * S;
      bx0 = max(cx - hw, 0);
      bx1 = max(cx + hw, 0);
      by0 = max(cy - hw, 0);
      by1 = max(cy + hw, 0);
      n_def_active = n;
      px = py = nullptr;
      pact = nullptr;
      pk = nullptr;
    } else {
      const int off = ss.off[i];
      n = ss.off[i + 1] - off;
      px = (const gint *)ss.px + off;
      py = (const gint *)ss.py + off;
      pact = ss.act ? (const gu8 *)ss.act + off : nullptr;
      pk = ss.pxy ? (const gu32 *)ss.pxy + off : nullptr;
      const int4 b = ss.bbox[i];
      bx0 = b.x;
      by0 = b.y;
      bx1 = b.z;
      by1 = b.w;
      const int4 ab = ss.abox[i];
      abx0 = ab.x;
      aby0 = ab.y;
      abx1 = ab.z;
      aby1 = ab.w;
      n_def_active = ss.n_active_def[i];
    }
    ref_in = bx0 - ref.ox >= 0 && by0 - ref.oy >= 0 && bx1 - ref.ox < ref.w && by1 - ref.oy < ref.h;
    mw = 0;
    mx0 = my0 = mh = 0;
    mb = nullptr;
    if (ss.mbox) {
      const int4 b = ss.mbox[i];
      mx0 = b.x;
      my0 = b.y;
      mw = b.z;
      mh = b.w;
      mb = (const gu32 *)ss.mbits + ss.moff[i];
    }
    forced = ss.force ? ss.force[i] != 0 : false;
  }

  // Visit the group's share of the subset pixels: f(x, y, active).  Negative
  // coordinates clamp to 0 (DICe_Subset.cpp:65-68, 100-101).  Square subsets
  // walk (row, col) incrementally in a loop of their own.
  template <class F>
  __device__ __forceinline__ void for_square(F &&f) const {
    const int r0 = grp.rank();
    // (k + 0.5) / S in float has the integer part of k / S for k < 2^8 and S < 2^12
    const float inv = 1.0f / (float)S;
    const int drow = S < 4096 ? (int)(((float)G::kSize + 0.5f) * inv) : G::kSize / S;
    const int dcol = G::kSize - drow * S;
    int row = S < 4096 ? (int)(((float)r0 + 0.5f) * inv) : r0 / S;
    int col = r0 - row * S;
    const int x0 = cx - hw, y0 = cy - hw;
#pragma unroll 1
    for (int i = r0; i < n; i += G::kSize) {
      f(max(x0 + col, 0), max(y0 + row, 0), true);
      col += dcol;
      row += drow;
      if (col >= S) { col -= S; ++row; }
    }
  }
  // The lean kernel's walk of a square subset (wave groups) whose first row and
  // column lie in the image: strips of 32 columns, each half-wave taking 32
  // consecutive pixels of one row, so that the taps it reads from the LDS
  // window are consecutive words (no bank conflicts while the warp is close to
  // a translation); the last S mod 32 columns R at a time, 64 / R rows per
  // step.  Each lane keeps one column per strip and steps down its rows:
  // f(x, y, dx, dy, o) with (dx, dy) = (x - cx, y - cy) and o the reference
  // pixel offset (y - ref.oy) * ref.w + x - ref.ox.
  template <class F>
  __device__ __forceinline__ void for_strips(F &&f) const {
    const int lane = grp.rank();
    const int x0 = cx - hw, y0 = cy - hw;
    const int nfull = S >> 5;
    auto column = [&](int x, int r0, int step) {
      int o = (y0 + r0 - ref.oy) * ref.w + (x - ref.ox);
      const int ostep = step * ref.w;
      const float dx = (float)(x - cx), fstep = (float)step;
      float dy = (float)(y0 + r0 - cy);
#pragma unroll 1
      for (int r = r0; r < S; r += step) {
        f(x, y0 + r, dx, dy, o);
        o += ostep;
        dy += fstep;
      }
    };
#pragma unroll 1
    for (int st = 0; st < nfull; ++st) column(x0 + (st << 5) + (lane & 31), lane >> 5, 2);
    const int R = S & 31;
    if (R) {
      const int rows = kWave / R;
      // lane / R in float: exact for lane < 64 and R < 32
      const int lr = (int)(((float)lane + 0.5f) / (float)R), lc = lane - lr * R;
      if (lr < rows) column(x0 + (nfull << 5) + lc, lr, rows);
    }
  }
  // reference intensity and gradients at pixel offset o of the reference image
  __device__ __forceinline__ void ref_at(int o, float &F, float &gx, float &gy) const {
    F = Pix<T>::at_u(ref.pix, (unsigned)o);
    gx = ldf_u(ref.gx, (unsigned)o);
    gy = ldf_u(ref.gy, (unsigned)o);
  }
  template <class F>
  __device__ __forceinline__ void for_pixels(F &&f) const {
    const int r0 = grp.rank();
    if (S > 0) {
      for_square(f);
    } else if (pk) {
      // one packed word per pixel; the next one is loaded before this pixel's work
      uint32_t next = r0 < n ? pk[r0] : 0u;
#pragma unroll 1
      for (int i = r0; i < n; i += G::kSize) {
        const uint32_t w = next;
        if (i + G::kSize < n) next = pk[i + G::kSize];
        f((int)(w & 0x7fffu), (int)((w >> 15) & 0x7fffu), (w >> 30) == 0u);
      }
    } else {
#pragma unroll 1
      for (int i = r0; i < n; i += G::kSize) f(max(px[i], 0), max(py[i], 0), pact ? (pact[i] != 0) : true);
    }
  }

  // for_pixels with the body instantiated twice, once for subsets inside the
  // reference image (tag true: reference reads need no bounds test) and once
  // for the others, selected by a uniform branch: f(tag, x, y, active).
  template <class F>
  __device__ __forceinline__ void for_pixels_ref(F &&f) const {
    if (ref_in)
      for_pixels([&](int x, int y, bool a) { f(std::true_type{}, x, y, a); });
    else
      for_pixels([&](int x, int y, bool a) { f(std::false_type{}, x, y, a); });
  }
  // LEAN: a square subset inside the reference image (the lean kernel's case)
  // walked by one loop; otherwise for_pixels_ref
  template <bool LEAN, class F>
  __device__ __forceinline__ void walk(F &&f) const {
    if constexpr (LEAN)
      for_square([&](int x, int y, bool a) { f(std::true_type{}, x, y, a); });
    else
      for_pixels_ref(f);
  }

  // Reference intensity and gradients at local (rx, ry); outside the image the
  // intensity is pixel 0 (Image_::operator(), DICe_Image.h:186-192) and the
  // gradients 0.  Subsets inside the image skip the test (uniform branch).
  template <bool INSIDE = false>
  __device__ __forceinline__ void ref_pixel(int rx, int ry, float &F, float &gx, float &gy) const {
    if (INSIDE || ref_in || (rx >= 0 && ry >= 0 && rx < ref.w && ry < ref.h)) {
      const unsigned o = (unsigned)(ry * ref.w + rx);
      F = Pix<T>::at_u(ref.pix, o);
      gx = ldf_u(ref.gx, o);
      gy = ldf_u(ref.gy, o);
    } else {
      F = Pix<T>::at(ref.pix, 0);
      gx = gy = 0.0f;
    }
  }
  template <bool INSIDE = false>
  __device__ __forceinline__ float ref_value(int rx, int ry) const {
    if (INSIDE || ref_in || (rx >= 0 && ry >= 0 && rx < ref.w && ry < ref.h))
      return Pix<T>::at_u(ref.pix, (unsigned)(ry * ref.w + rx));
    return Pix<T>::at(ref.pix, 0);
  }

  // Image_::operator() (DICe_Image.h:186-192): out of bounds reads pixel 0
  __device__ __forceinline__ float img_at(const DImage &im, int x, int y) const {
    if (x < 0 || y < 0 || x >= im.w || y >= im.h) return Pix<T>::at(im.pix, 0);
    return Pix<T>::at(im.pix, y * im.w + x);
  }

  // Is the warped pixel kept?  Rounding + 4-px guard of Subset::initialize
  // (DICe_Subset.cpp:404-411); (ddx, ddy) = x' - cx, y' - cy.
  // A pixel that rounds onto a blocked image pixel of the subset is dropped as
  // well (is_obstructed_pixel and the blocked-pixel test, :412-419, 447-460).
  __device__ __forceinline__ bool kept(float ddx, float ddy) const {
    const int rx = cx + (int)floorf(ddx + 0.5f);
    const int ry = cy + (int)floorf(ddy + 0.5f);
    if (rx < def.ox + 4 || rx >= def.ox + def.w - 4 || ry < def.oy + 4 || ry >= def.oy + def.h - 4) return false;
    if (mw > 0) {
      const unsigned ux = (unsigned)(rx - mx0), uy = (unsigned)(ry - my0);
      if (ux < (unsigned)mw && uy < (unsigned)mh) {
        const unsigned bit = uy * (unsigned)mw + ux;
        if ((mb[bit >> 5] >> (bit & 31u)) & 1u) return false;
      }
    }
    return true;
  }

  // Deformed intensity at x' = cx + ddx, y' = cy + ddy (interior interpolant of
  // interpolate_keys_fourth_all / interpolate_bilinear_all, DICe_Image.cpp:392-425,
  // 832-883).  Only called for kept pixels, so all taps lie inside the image
  // and, in BATCHED mode, inside the staged window.
  __device__ __forceinline__ float sample(float ddx, float ddy) const {
    const float flx = floorf(ddx), fly = floorf(ddy);
    const float tx = ddx - flx, ty = ddy - fly;
    const int ix = cx - def.ox + (int)flx;
    const int iy = cy - def.oy + (int)fly;
    if constexpr (kIsWave) {
      // window origin and pitch are scalar; kept pixels index the window with
      // non-negative offsets below 2^24
      if (INTERP == 0) return bilinear_sum(win + (__umul24(iy - wy0, wpitch) + (ix - wx0)), wpitch, tx, ty);
      float wx[6], wy[6];
      keys_weights(tx, wx);
      keys_weights(ty, wy);
      return keys_sum(win + (__umul24(iy - 2 - wy0, wpitch) + (ix - 2 - wx0)), wpitch, wx, wy);
    } else {
      const auto *P = Pix<T>::ptr(def.pix);
      if (INTERP == 0) return bilinear_sum(P + iy * def.w + ix, def.w, tx, ty);
      float wx[6], wy[6];
      keys_weights(tx, wx);
      keys_weights(ty, wy);
      return keys_sum(P + (iy - 2) * def.w + ix - 2, def.w, wx, wy);
    }
  }
  // Deformed-image gradients at the same position (DEFGRAD: the deformed image
  // carries gradients and interpolate_*_all returns them too).
  __device__ __forceinline__ void sample_grad(float ddx, float ddy, float &gxo, float &gyo) const {
    const float flx = floorf(ddx), fly = floorf(ddy);
    const float tx = ddx - flx, ty = ddy - fly;
    const int ix = cx - def.ox + (int)flx;
    const int iy = cy - def.oy + (int)fly;
    const gfloat *GX = (const gfloat *)def.gx, *GY = (const gfloat *)def.gy;
    if (INTERP == 0) {
      gxo = bilinear_sum(GX + iy * def.w + ix, def.w, tx, ty);
      gyo = bilinear_sum(GY + iy * def.w + ix, def.w, tx, ty);
      return;
    }
    float wx[6], wy[6];
    keys_weights(tx, wx);
    keys_weights(ty, wy);
    gxo = keys_sum(GX + (iy - 2) * def.w + ix - 2, def.w, wx, wy);
    gyo = keys_sum(GY + (iy - 2) * def.w + ix - 2, def.w, wx, wy);
  }

  // Stage the deformed-image window [x0, x1] x [y0, y1] (local coordinates,
  // clamped to the image) into the wave's LDS slot.  Returns false (and marks
  // the subset as overflowing) when it does not fit.
  // ring false: no ring (a window read for one evaluation only).
  __device__ __forceinline__ bool stage(int x0, int x1, int y0, int y1, bool hand_over = true, bool ring = true) {
    x0 = max(x0, 0);
    y0 = max(y0, 0);
    x1 = min(x1, def.w - 1);
    y1 = min(y1, def.h - 1);
    const int ww = x1 - x0 + 1, hh = y1 - y0 + 1;
    if (ww <= 0 || hh <= 0) {
      wx0 = x0; wy0 = y0; wpitch = 1; wrows = 0;
      win_ok = false;
      return true;  // nothing is kept
    }
    // the deformed image does not change within a subset: a window already
    // staged that contains this one serves it (the taps are read at their
    // image coordinates through wx0, wy0, wpitch)
    if (win_ok && x0 >= wx0 && x1 < wx0 + wpitch && y0 >= wy0 && y1 < wy0 + wrows) return true;
    if (ww * hh > win_cap) {
      if (hand_over) overflow = true;
      return false;
    }
    // stage a ring of up to 2 px more where the slot has room (within the
    // image), so that the windows of the next steps of this subset, which move
    // by a fraction of a pixel, are usually inside this one
    {
      int e = ring ? 2 : 0;
      while (e > 0 && (ww + 2 * e) * (hh + 2 * e) > win_cap) --e;
      x0 = max(x0 - e, 0);
      y0 = max(y0 - e, 0);
      x1 = min(x1 + e, def.w - 1);
      y1 = min(y1 + e, def.h - 1);
      // an odd pitch where there is room: the 32 words of a column that a
      // half-wave reads at 32 consecutive rows then fall in 32 different banks
      if (((x1 - x0 + 1) & 1) == 0 && (x1 - x0 + 2) * (y1 - y0 + 1) <= win_cap) {
        if (x1 + 1 < def.w) ++x1;
        else if (x0 > 0) --x0;
      }
    }
    const int ww_s = x1 - x0 + 1, hh_s = y1 - y0 + 1;
    wx0 = __builtin_amdgcn_readfirstlane(x0);  // uniform already: keep it scalar
    wy0 = __builtin_amdgcn_readfirstlane(y0);
    wpitch = __builtin_amdgcn_readfirstlane(ww_s);
    wrows = __builtin_amdgcn_readfirstlane(hh_s);
    grp.sync();
    const auto *P = Pix<T>::ptr(def.pix) + y0 * def.w + x0;
    const float inv = 1.0f / (float)ww_s;
    const int tot = ww_s * hh_s;
    if constexpr (kIsWave && std::is_same<T, float>::value) {
      // direct global -> LDS loads (global_load_lds_dword): the 64 lanes of one
      // load fill 64 consecutive floats of the slot, so window element k =
      // 64 j + lane comes from batch j.  No VGPRs hold the data and every batch
      // is in flight before the single wait.
      const int lane = grp.rank();
#pragma unroll 4
      for (int k0 = 0; k0 < tot; k0 += kWave) {
        const int k = k0 + lane;
        if (k < tot) {
          const int r = (int)(((float)k + 0.5f) * inv);
          const int c = k - r * ww_s;
          __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)(P + (r * def.w + c)),
                                           (__attribute__((address_space(3))) void *)(win + k0), 4, 0, 0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
    } else {
#pragma unroll 4
      for (int k = grp.rank(); k < tot; k += G::kSize) {
        const int r = (int)(((float)k + 0.5f) * inv);
        const int c = k - r * ww_s;
        win[k] = (float)P[r * def.w + c];
      }
    }
    grp.sync();
    win_ok = true;
    return true;
  }

  // Extremes of the displacement (x' - cx, y' - cy) over the subset bbox at
  // parameters A: the map is affine, so they are taken at the bbox corners.
  __device__ __forceinline__ void warp_bounds(const Affine &A, float &lo_x, float &hi_x, float &lo_y,
                                              float &hi_y) const {
    lo_x = 1e30f, hi_x = -1e30f, lo_y = 1e30f, hi_y = -1e30f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float dx = (float)((k & 1) ? bx1 : bx0) - (float)cx;
      const float dy = (float)((k & 2) ? by1 : by0) - (float)cy;
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      lo_x = fminf(lo_x, ox);
      hi_x = fmaxf(hi_x, ox);
      lo_y = fminf(lo_y, oy);
      hi_y = fmaxf(hi_y, oy);
    }
  }

  // Window covering the warped subset bbox (+ interpolation support and a
  // 1-px rounding margin) at parameters A.
  __device__ __forceinline__ bool stage_affine(const Affine &A, bool ring = true) {
    if constexpr (!kIsWave) {
      (void)A;
      (void)ring;
      return true;
    } else {
      // the map is affine, so every pixel's position lies between the corner
      // extremes; a 1e-3 px margin covers the rounding of the per-pixel map
      float lo_x, hi_x, lo_y, hi_y;
      warp_bounds(A, lo_x, hi_x, lo_y, hi_y);
      const int lead = INTERP == 0 ? 0 : 2, trail = INTERP == 0 ? 1 : 3;
      const float m = 1e-3f;
      return stage(cx - def.ox + (int)floorf(lo_x - m) - lead, cx - def.ox + (int)floorf(hi_x + m) + trail,
                   cy - def.oy + (int)floorf(lo_y - m) - lead, cy - def.oy + (int)floorf(hi_y + m) + trail, true,
                   ring);
    }
  }

  // Does every pixel pass the 4-px guard of kept() at parameters A?  The
  // rounded positions are monotone in the displacement, so the bbox extremes
  // decide; a 1e-3 px margin covers the rounding of the per-pixel map.
  __device__ __forceinline__ bool all_kept(const Affine &A) const {
    if (mw > 0) return false;  // blocked pixels: the per-pixel test decides
    float lo_x, hi_x, lo_y, hi_y;
    warp_bounds(A, lo_x, hi_x, lo_y, hi_y);
    const int rx0 = cx + (int)floorf(lo_x + 0.5f - 1e-3f), rx1 = cx + (int)floorf(hi_x + 0.5f + 1e-3f);
    const int ry0 = cy + (int)floorf(lo_y + 0.5f - 1e-3f), ry1 = cy + (int)floorf(hi_y + 0.5f + 1e-3f);
    return rx0 >= def.ox + 4 && rx1 < def.ox + def.w - 4 && ry0 >= def.oy + 4 && ry1 < def.oy + def.h - 4;
  }

  // Sum the 35 per-lane FP32 Gauss-Newton partials over the wave into FP64
  // totals in LDS.  Lane values go to a 35 x 64 table with row stride 65 (the
  // padding keeps both the column writes and the row reads free of bank
  // conflicts), lanes < 35 each add one row in FP64 and store the total in the
  // layout the solve reads: q at [0, 6), H as a full 6 x 6 matrix at [8, 44),
  // sum r at [44, 50), sum G at 50, count at 51 (kGnTot* offsets).  Uses the
  // first kMinSlotFloats floats of the wave's LDS slot (no window is live).
  __device__ __forceinline__ const lds_double *lds_transpose_sum(const float (&v)[kGnSums]) {
    const int lane = grp.rank();
    lds_float *red = win;
    win_ok = false;
    grp.sync();
#pragma unroll
    for (int k = 0; k < kGnSums; ++k) red[k * kGnRedStride + lane] = v[k];
    grp.sync();
    lds_double *tot = (lds_double *)(red + kGnTotOffset);
    if (lane < kGnSums) {
      const lds_float *row = red + lane * kGnRedStride;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int j = 0; j < kWave; j += 4) {
        s0 += (double)row[j];
        s1 += (double)row[j + 1];
        s2 += (double)row[j + 2];
        s3 += (double)row[j + 3];
      }
      const double s = (s0 + s1) + (s2 + s3);
      // destination(s) of partial `lane`
      int d0, d1 = -1;
      if (lane < 6) {
        d0 = kGnTotQ + lane;
      } else if (lane < 27) {
        // packed upper triangle: row k holds 6 - k entries
        int h = lane - 6, k = 0, len = 6;
#pragma unroll
        for (int t = 0; t < 5; ++t)
          if (h >= len) { h -= len; --len; ++k; }
        const int l = k + h;
        d0 = kGnTotH + 6 * k + l;
        if (l != k) d1 = kGnTotH + 6 * l + k;
      } else {
        d0 = kGnTotR + (lane - 27);  // sum r [0, 6), then sum G, count
      }
      tot[d0] = s;
      if (d1 >= 0) tot[d1] = s;
    }
    grp.sync();
    return tot;
  }

  // Reference intensity and the gradient to use at a kept pixel: rotated
  // reference gradient (use_ref_grads, DICe_LocalShapeFunction.cpp:375-376) or
  // the interpolated deformed-image gradient.
  __device__ __forceinline__ void gn_ref(const Affine &A, int rx, int ry, float ox, float oy, float &Fv, float &Gx,
                                         float &Gy) const {
    if constexpr (DEFGRAD) {
      Fv = ref_value(rx, ry);
      sample_grad(ox, oy, Gx, Gy);
    } else {
      float gx, gy;
      ref_pixel(rx, ry, Fv, gx, gy);
      Gx = A.c * gx - A.s * gy;
      Gy = A.s * gx + A.c * gy;
    }
  }

  // One Gauss-Newton pass at parameters A: totals over the kept pixels,
  // acc = [q with the shift (6), H upper triangle (21), sum r (6), sum G, count].
  // PLAIN: every product added in FP64, workgroup reduction.
  __device__ __forceinline__ void accumulate_gn_block(const Affine &A, float shift, double (&acc)[35]) {
#pragma unroll
    for (int k = 0; k < 35; ++k) acc[k] = 0.0;
    for_pixels([&](int x, int y, bool a) {
      if (!a) return;
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      if (!kept(ox, oy)) return;
      const float Gv = sample(ox, oy);
      float Fv, Gx, Gy;
      gn_ref(A, x - ref.ox, y - ref.oy, ox, oy, Fv, Gx, Gy);
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
    });
    grp.sum(acc);
  }

  // BATCHED: FP32 products and per-lane partial sums, FP64 totals in LDS.
  __device__ __forceinline__ const lds_double *accumulate_gn_wave(const Affine &A, float shift) {
    float pa[kGnSums - 1];
#pragma unroll
    for (int k = 0; k < kGnSums - 1; ++k) pa[k] = 0.0f;
    int cntl = 0;
    for_pixels([&](int x, int y, bool a) {
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      if (!a || !kept(ox, oy)) return;
      const float Gv = sample(ox, oy);
      float Fv, Gx, Gy;
      gn_ref(A, x - ref.ox, y - ref.oy, ox, oy, Fv, Gx, Gy);
      float r[6];
      A.residual(dx, dy, Gx, Gy, r);
#pragma unroll
      for (int k = 2; k < 6; ++k) r[k] *= en[k];
      const float gmf = (Gv - shift) - (Fv - meanF);
#pragma unroll
      for (int k = 0; k < 6; ++k) pa[k] = fmaf(gmf, r[k], pa[k]);
      int h = 6;
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int l = k; l < 6; ++l) { pa[h] = fmaf(r[k], r[l], pa[h]); ++h; }
#pragma unroll
      for (int k = 0; k < 6; ++k) pa[27 + k] += r[k];
      pa[33] += Gv;
      ++cntl;
    });
    float part[kGnSums];
#pragma unroll
    for (int k = 0; k < kGnSums - 1; ++k) part[k] = pa[k];
    part[kGnSums - 1] = (float)cntl;
    return lds_transpose_sum(part);
  }

  // Sum K per-lane FP32 values over the wave in FP64, eight table rows at a
  // time: lane (lane / 8, lane % 8) adds eight entries of its row (start
  // rotated by the segment against bank conflicts) and DPP within the eight
  // lanes finishes the row, identically in all of them; dest(row, total) is
  // called by one lane per row.  The table uses the first K * kGnRedStride
  // floats of the slot.
  template <int K, class D>
  __device__ __forceinline__ void lds_chunk_sum(const float (&v)[K], D &&dest) {
    const int lane = grp.rank();
    lds_float *red = win;
    win_ok = false;
    grp.sync();
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * kGnRedStride + lane] = v[k];
    grp.sync();
    const int seg = lane & 7;
#pragma unroll
    for (int c = 0; c < (K + 7) / 8; ++c) {
      const int r = c * 8 + (lane >> 3);
      const lds_float *src = red + min(r, K - 1) * kGnRedStride + seg * 8;
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j) t += (double)src[(j + seg) & 7];
      t += dpp<kDppXor1>(t);
      t += dpp<kDppXor2>(t);
      t += dpp<kDppHalfMirror>(t);
      if (seg == 0 && r < K) dest(r, t);
    }
    grp.sync();
  }

  // Per-subset constants of the fast pass over the active pixels: sum(phi
  // phi^T) (a full symmetric 6 x 6) and sum(phi) into pers[], and the mean of
  // the reference intensities (computeUpdateFast's F-bar, DICe_Objective.cpp:464)
  // into meanF, from the same pass.
  template <bool LEAN = false>
  __device__ __forceinline__ void precompute_phi() {
    float pa[29];
#pragma unroll
    for (int k = 0; k < 29; ++k) pa[k] = 0.0f;
    walk<LEAN>([&](auto inside, int x, int y, bool a) {
      if (!a) return;
      float F, gx, gy;
      ref_pixel<decltype(inside)::value>(x - ref.ox, y - ref.oy, F, gx, gy);
      pa[27] += F;
      pa[28] += 1.0f;
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      const float ph[6] = {gx, gy, dx * gx, dy * gx, dx * gy, dy * gy};
      int h = 0;
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int l = k; l < 6; ++l) { pa[h] = fmaf(ph[k], ph[l], pa[h]); ++h; }
#pragma unroll
      for (int k = 0; k < 6; ++k) pa[21 + k] += ph[k];
    });
    lds_double *P = pers;
    lds_chunk_sum<29>(pa, [&](int r, double t) {
      if (r < 21) {
        int h = r, k = 0, len = 6;  // packed upper triangle: row k holds 6 - k entries
#pragma unroll
        for (int q = 0; q < 5; ++q)
          if (h >= len) { h -= len; --len; ++k; }
        const int l = k + h;
        P[6 * k + l] = t;
        P[6 * l + k] = t;
      } else {
        P[36 + (r - 21)] = t;  // sum(phi), then sum(F) and the pixel count at 42, 43
      }
    });
    const double cnt = uniform(P[43]);
    meanF = cnt > 0.0 ? (float)(uniform(P[42]) / cnt) : 0.0f;
  }

  // Fast pass: sum(gmf phi) (6) and sum(G) over the active pixels, all kept,
  // into pers[kPersistPass, +7).  The lane partials are summed in FP64 across
  // the wave without LDS, so the staged window stays valid for the next pass.
  template <bool LEAN = false>
  __device__ __forceinline__ void accumulate_gn_fast(const Affine &A, float shift, int out = kPersistPass) {
    float pa[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) pa[k] = 0.0f;
    auto body = [&](float dx, float dy, float Fv, float gx, float gy) {
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      const float Gv = sample(ox, oy);
      const float gmf = (Gv - shift) - (Fv - meanF);
      const float g1 = gmf * gx, g2 = gmf * gy;
      pa[0] += g1;
      pa[1] += g2;
      pa[2] = fmaf(g1, dx, pa[2]);
      pa[3] = fmaf(g1, dy, pa[3]);
      pa[4] = fmaf(g2, dx, pa[4]);
      pa[5] = fmaf(g2, dy, pa[5]);
      pa[6] += Gv;
    };
    if constexpr (LEAN) {
      for_strips([&](int, int, float dx, float dy, int o) {
        float Fv, gx, gy;
        ref_at(o, Fv, gx, gy);
        body(dx, dy, Fv, gx, gy);
      });
    } else {
      for_pixels_ref([&](auto inside, int x, int y, bool a) {
        if (!a) return;
        // the reference loads first: they are in flight during the interpolation
        float Fv, gx, gy;
        ref_pixel<decltype(inside)::value>(x - ref.ox, y - ref.oy, Fv, gx, gy);
        body((float)(x - cx), (float)(y - cy), Fv, gx, gy);
      });
    }
    double fs[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) fs[k] = (double)pa[k];
    grp.sum(fs);
    if (grp.rank() == 0) {
#pragma unroll
      for (int k = 0; k < 7; ++k) pers[out + k] = fs[k];
    }
    grp.sync();
  }

  // ------------------------------- objective sums of small subsets ---
  // Subsets of at most kLaneSimplexMaxPixels pixels sum the objective in one
  // fixed order in every execution mode, so that the Nelder-Mead paths (which
  // fork on rounding) agree between them: lane l < W of a row adds its pixels
  // l, l + W, l + 2 W, ... in that order (FP64), then the W partials are added
  // by row_sum<W>'s pairwise tree.  W = kp.small_w: 8 when no subset of the set
  // has more than kNarrowSimplexMaxPixels pixels, else 16.
  static constexpr int kSmallPerLane = kLaneSimplexMaxPixels / 16;
  // f(k, x, y, active) for the pixels i = l + W k of the subset (coordinates
  // as for_pixels gives them)
  // UNROLL: the loop unrolled (f may then index registers by k), else rolled.
  template <int W = 16, int KPER = kSmallPerLane, bool UNROLL = true, class F>
  __device__ __forceinline__ void small_walk(int l, F &&f) const {
    int row = 0, col = 0, drow = 0, dcol = 0;
    if (S > 0) {
      row = l / S;
      col = l - row * S;
      drow = W / S;
      dcol = W - drow * S;
    }
    auto pixel = [&](int k) {
      const int i = l + W * k;
      if (S > 0) {
        f(k, max(cx - hw + col, 0), max(cy - hw + row, 0), true);
        col += dcol;
        row += drow;
        if (col >= S) { col -= S; ++row; }
      } else if (pk) {
        const uint32_t w = pk[i];
        f(k, (int)(w & 0x7fffu), (int)((w >> 15) & 0x7fffu), (w >> 30) == 0u);
      } else {
        f(k, max(px[i], 0), max(py[i], 0), pact ? (pact[i] != 0) : true);
      }
    };
    if constexpr (UNROLL) {
#pragma unroll
      for (int k = 0; k < KPER; ++k) {
        if (l + W * k >= n) break;
        pixel(k);
      }
    } else {
#pragma unroll 1
      for (int k = 0; l + W * k < n; ++k) pixel(k);
    }
  }
  // Lane l's partial gamma sums (n, sd, sdd, sr, srr, sdr) of the kept pixels at
  // A; smp(ox, oy): the interpolant, refv(x, y, k): the reference value.  KPER:
  // a bound on the pixels per lane (the subset has at most W KPER pixels).
  template <int W = 16, int KPER = kSmallPerLane, bool UNROLL = true, class Smp, class Ref>
  __device__ __forceinline__ void small_partial(const Affine &A, int l, double (&acc)[6], Smp &&smp,
                                                Ref &&refv) const {
    small_walk<W, KPER, UNROLL>(l, [&](int k, int x, int y, bool a) {
      if (!a) return;
      const double r = refv(x, y, k);
      float ox, oy;
      A.disp((float)(x - cx), (float)(y - cy), ox, oy);
      if (!kept(ox, oy)) return;
      const double d = smp(ox, oy);
      acc[0] += 1.0;
      acc[1] += d;
      acc[2] = fma(d, d, acc[2]);
      acc[3] += r;
      acc[4] = fma(r, r, acc[4]);
      acc[5] = fma(d, r, acc[5]);
    });
  }

  // Objective::gamma (DICe_Objective.cpp:58-76) at parameters A: warp, then
  // the ZNSSD of Subset::gamma from FP64 sums; also active count and contrast.
  __device__ __forceinline__ float gamma_eval(const Affine &A, int &active, float &contrast) {
    active = 0;
    contrast = 0.0f;
    if (!stage_affine(A)) return -1.0f;
    double acc[6] = {0, 0, 0, 0, 0, 0};  // n, sd, sdd, sr, srr, sdr
    if constexpr (G::kSize >= 16) {
      if (n <= kLaneSimplexMaxPixels) {
        const int w16 = kp.small_w == 8 && n <= kNarrowSimplexMaxPixels ? 8 : 16;
        // every rank evaluates pixels rank, rank + G::kSize, ... into three LDS
        // tables (kept, d, r): after the window in a wave's slot, or small_tab
        // in a workgroup; then ranks 0..W-1 add them in small_partial's order --
        // the same values in the same order, with the interpolation (and, in a
        // workgroup, its global loads) spread over the whole group
        lds_float *kt = small_tab;
        if constexpr (kIsWave) {
          const int cells = wpitch * wrows;
          kt = cells + 3 * n <= win_cap ? win + cells : nullptr;
        }
        if (kt) {
          lds_float *const dt = kt + n, *const rt = dt + n;
          const int lane = grp.rank();
          const float invs = S > 0 ? 1.0f / (float)S : 0.0f;
          for (int i = lane; i < n; i += G::kSize) {
            int x, y;
            bool a = true;
            if (S > 0) {
              const int row = (int)(((float)i + 0.5f) * invs);  // exact: i < 2^8, S <= 16
              x = max(cx - hw + (i - row * S), 0);
              y = max(cy - hw + row, 0);
            } else if (pk) {
              const uint32_t w = pk[i];
              x = (int)(w & 0x7fffu);
              y = (int)((w >> 15) & 0x7fffu);
              a = (w >> 30) == 0u;
            } else {
              x = max(px[i], 0);
              y = max(py[i], 0);
              a = pact ? (pact[i] != 0) : true;
            }
            float kf = 0.0f, d = 0.0f, r = 0.0f;
            if (a) {
              r = ref_value(x - ref.ox, y - ref.oy);
              float ox, oy;
              A.disp((float)(x - cx), (float)(y - cy), ox, oy);
              if (kept(ox, oy)) {
                kf = 1.0f;
                d = sample(ox, oy);
              }
            }
            kt[i] = kf;
            dt[i] = d;
            rt[i] = r;
          }
          grp.sync();
          if (lane < w16) {
#pragma unroll 1
            for (int i = lane; i < n; i += w16) {
              if (kt[i] == 0.0f) continue;
              const double d = dt[i], r = rt[i];
              acc[0] += 1.0;
              acc[1] += d;
              acc[2] = fma(d, d, acc[2]);
              acc[3] += r;
              acc[4] = fma(r, r, acc[4]);
              acc[5] = fma(d, r, acc[5]);
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] = w16 == 8 ? row_sum8(acc[k]) : row_sum16(acc[k]);
          }
          grp.bcast0(acc);  // a workgroup: ends with a barrier
          if constexpr (kIsWave) grp.sync();
          return gamma_of(acc, active, contrast);
        }
        // the order of simplex_groups_kernel (small_partial), in ranks 0..W-1
        auto smp = [&](float ox, float oy) { return sample(ox, oy); };
        auto refv = [&](int x, int y, int) { return ref_value(x - ref.ox, y - ref.oy); };
        const int r0 = grp.rank();
        if (w16 == 8) {
          if (r0 < 8) {
            small_partial<8, kNarrowSimplexMaxPixels / 8, false>(A, r0, acc, smp, refv);
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] = row_sum8(acc[k]);
          }
        } else if (r0 < 16) {
          small_partial<16, kSmallPerLane, false>(A, r0, acc, smp, refv);
#pragma unroll
          for (int k = 0; k < 6; ++k) acc[k] = row_sum16(acc[k]);
        }
        grp.bcast0(acc);
        return gamma_of(acc, active, contrast);
      }
    }
    bool every = false;  // BATCHED: every active pixel is kept (no per-pixel guard test)
    if constexpr (kIsWave) every = all_kept(A);
    if (every) {
      gamma_sums_every<false>(A, acc);
    } else {
      for_pixels([&](int x, int y, bool a) {
        if (!a) return;
        const double r = ref_value(x - ref.ox, y - ref.oy);
        const float dx = (float)(x - cx), dy = (float)(y - cy);
        float ox, oy;
        A.disp(dx, dy, ox, oy);
        if (!kept(ox, oy)) return;
        const double d = sample(ox, oy);
        acc[0] += 1.0;
        acc[1] += d;
        acc[2] += d * d;
        acc[3] += r;
        acc[4] += r * r;
        acc[5] += d * r;
      });
    }
    grp.sum(acc);
    if (every) acc[0] = (double)n_def_active;
    return gamma_of(acc, active, contrast);
  }
  // The gamma sums (sd, sdd, sr, srr, sdr) when every active pixel is kept
  // (acc[0], the count, is left to the caller: the defined-active count).
  template <bool LEAN>
  __device__ __forceinline__ void gamma_sums_every(const Affine &A, double (&acc)[6]) {
    auto body = [&](float dx, float dy, float rv) {
      const double r = rv;
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      const double d = sample(ox, oy);
      acc[1] += d;
      acc[2] += d * d;
      acc[3] += r;
      acc[4] += r * r;
      acc[5] += d * r;
    };
    if constexpr (LEAN) {
      for_strips([&](int, int, float dx, float dy, int o) { body(dx, dy, Pix<T>::at_u(ref.pix, (unsigned)o)); });
    } else {
      for_pixels_ref([&](auto inside, int x, int y, bool a) {
        if (!a) return;
        // the reference load first: it is in flight during the interpolation
        body((float)(x - cx), (float)(y - cy), ref_value<decltype(inside)::value>(x - ref.ox, y - ref.oy));
      });
    }
  }
  // ZNSSD gamma from the FP64 sums (n, sd, sdd, sr, srr, sdr) of the kept pixels.
  __device__ __forceinline__ float gamma_of(const double (&acc)[6], int &active, float &contrast) const {
    const double cnt = acc[0];
    active = (int)cnt;
    if (cnt <= 0.0) return -1.0f;
    // explicit fma (no contraction left to the compiler, as in Affine)
    const double md = acc[1] / cnt, mr = acc[3] / cnt;
    const double sdd = fmax(fma(-(cnt * md), md, acc[2]), 0.0);
    const double srr = fmax(fma(-(cnt * mr), mr, acc[4]), 0.0);
    const double sdr = fma(-(cnt * md), mr, acc[5]);
    contrast = (float)sqrt(sdd / cnt);
    // Subset::gamma returns -1 when either norm is zero (flat subset)
    if (sdd <= 1e-13 * acc[2] || srr <= 1e-13 * acc[4]) return -1.0f;
    double gam = fma(-2.0, sdr / (sqrt(sdd) * sqrt(srr)), 2.0);
    if (gam < 0.0) gam = 0.0;
    float g = (float)gam;
    if (kp.normalize_gamma && n_def_active > 0) g = (float)((double)g / (double)n_def_active);
    return g;
  }

  // ---------------------------------------------- differentiation path ---
  // Objective of the adjoint: the ZNSSD gamma of the warped subset at p (the
  // quantity Subset::gamma returns after Subset::initialize, DICe_Subset.cpp:
  // 323-340, 373-440), gamma = 2 - 2 rho with rho = sdr / (|d - dbar| |r - rbar|)
  // over the kept pixels; no clamp and no normalisation by the pixel count.
  // ZStats: the FP64 sums that define it.  Returns false (gamma -1, gradient 0)
  // for a flat or empty subset.
  struct ZStats {
    double n, md, mr, nd, nr, rho;
  };
  __device__ __forceinline__ bool znssd_stats(const Affine &A, ZStats &z) {
    double acc[6] = {0, 0, 0, 0, 0, 0};  // n, sd, sdd, sr, srr, sdr
    for_pixels([&](int x, int y, bool a) {
      if (!a) return;
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      if (!kept(ox, oy)) return;
      const double d = sample(ox, oy);
      const double r = ref_value(x - ref.ox, y - ref.oy);
      acc[0] += 1.0;
      acc[1] += d;
      acc[2] += d * d;
      acc[3] += r;
      acc[4] += r * r;
      acc[5] += d * r;
    });
    grp.sum(acc);
    return zstats_from(acc, z);
  }
  // The ZNSSD terms from the sums (n, sd, sdd, sr, srr, sdr) over the kept pixels.
  __device__ __forceinline__ static bool zstats_from(const double (&acc)[6], ZStats &z) {
    z.n = acc[0];
    if (z.n <= 0.0) return false;
    z.md = acc[1] / z.n;
    z.mr = acc[3] / z.n;
    const double sdd = fmax(acc[2] - z.n * z.md * z.md, 0.0);
    const double srr = fmax(acc[4] - z.n * z.mr * z.mr, 0.0);
    if (sdd <= 1e-13 * acc[2] || srr <= 1e-13 * acc[4]) return false;
    z.nd = sqrt(sdd);
    z.nr = sqrt(srr);
    z.rho = (acc[5] - z.n * z.md * z.mr) / (z.nd * z.nr);
    return true;
  }

  // Deformed intensity and its x / y derivative at x' = cx + ddx, y' = cy + ddy
  // (the derivative of the interpolant itself, so that it matches finite
  // differences of the forward): Keys weights and their t-derivatives are
  // recomputed, nothing is stored.  tap(base offset, pitch, weights) gives the
  // scatter of the adjoint.
  __device__ __forceinline__ float sample_d(float ddx, float ddy, float &Ix, float &Iy) const {
    const float flx = floorf(ddx), fly = floorf(ddy);
    const float tx = ddx - flx, ty = ddy - fly;
    const int ix = cx - def.ox + (int)flx;
    const int iy = cy - def.oy + (int)fly;
    if (INTERP == 0) {
      float b00, b10, b01, b11;
      if constexpr (kIsWave) {
        const lds_float *b = win + (__umul24(iy - wy0, wpitch) + (ix - wx0));
        b00 = b[0]; b10 = b[1]; b01 = b[wpitch]; b11 = b[wpitch + 1];
      } else {
        const auto *b = Pix<T>::ptr(def.pix) + iy * def.w + ix;
        b00 = (float)b[0]; b10 = (float)b[1]; b01 = (float)b[def.w]; b11 = (float)b[def.w + 1];
      }
      Ix = (1.0f - ty) * (b10 - b00) + ty * (b11 - b01);
      Iy = (1.0f - tx) * (b01 - b00) + tx * (b11 - b10);
      return b00 * (1.0f - tx) * (1.0f - ty) + b10 * tx * (1.0f - ty) + b11 * tx * ty + b01 * (1.0f - tx) * ty;
    }
    float wx[6], wy[6], dwx[6], dwy[6];
    keys_weights(tx, wx);
    keys_weights(ty, wy);
    keys_weight_slopes(tx, dwx);
    keys_weight_slopes(ty, dwy);
    float v = 0.0f, vx = 0.0f, vy = 0.0f;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      float rs = 0.0f, rd = 0.0f;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        float t;
        if constexpr (kIsWave) t = win[__umul24(iy - 2 + m - wy0, wpitch) + (ix - 2 + k - wx0)];
        else t = (float)Pix<T>::ptr(def.pix)[(iy - 2 + m) * def.w + ix - 2 + k];
        rs = fmaf(t, wx[k], rs);
        rd = fmaf(t, dwx[k], rd);
      }
      v = fmaf(rs, wy[m], v);
      vx = fmaf(rd, wy[m], vx);
      vy = fmaf(rs, dwy[m], vy);
    }
    Ix = vx;
    Iy = vy;
    return v;
  }

  // Adds g * (interpolation weight) to every tap of the sample at (ddx, ddy):
  // into acc (LDS, the window's layout, pitch wpitch) when acc is set, else
  // into the global image gradient.
  // rmw: plain LDS read-modify-writes instead of LDS atomics, one tap at a
  // time in program order -- only for lanes whose samples lie in distinct
  // cells (no two lanes of the wave write one cell for the same tap); the
  // wave's LDS operations execute in order, and the compiler barrier after
  // each tap keeps that order.
  template <bool RMW = false>
  __device__ __forceinline__ void scatter_taps(float ddx, float ddy, float g, lds_float *acc, float *dimg) const {
    const float flx = floorf(ddx), fly = floorf(ddy);
    const float tx = ddx - flx, ty = ddy - fly;
    const int ix = cx - def.ox + (int)flx;
    const int iy = cy - def.oy + (int)fly;
    float wx[6], wy[6];
    int lead = 2, taps = 6;
    if (INTERP == 0) {
      wx[0] = 1.0f - tx; wx[1] = tx; wy[0] = 1.0f - ty; wy[1] = ty;
      lead = 0;
      taps = 2;
    } else {
      keys_weights(tx, wx);
      keys_weights(ty, wy);
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      if (m >= taps) break;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (k >= taps) break;
        const float c = g * wy[m] * wx[k];
        const int yy = iy - lead + m, xx = ix - lead + k;
        if constexpr (RMW) {
          lds_float *cell = acc + (__umul24(yy - wy0, wpitch) + (xx - wx0));
          *cell += c;
          asm volatile("" ::: "memory");
        } else if (acc) {
          __hip_atomic_fetch_add(acc + (__umul24(yy - wy0, wpitch) + (xx - wx0)), c, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WAVEFRONT);
        } else {
          unsafeAtomicAdd(dimg + (size_t)yy * def.w + xx, c);
        }
      }
    }
  }

  // Forward of the differentiation path: gamma of the subset at p.
  __device__ __forceinline__ bool gamma_at(const float (&p)[6], double &gamma_out) {
    Affine A;
#pragma unroll
    for (int k = 0; k < 6; ++k) A.p[k] = p[k];
    A.prepare();
    if (!stage_affine(A)) return false;
    ZStats z;
    gamma_out = znssd_stats(A, z) ? 2.0 - 2.0 * z.rho : -1.0;
    return true;
  }

  // Adjoint: dgamma/dd_j = (2 / nd) (rho dhat_j - rhat_j) for the kept pixels,
  // chained through the interpolant (d/dx', d/dy') and the affine map (the
  // residual rows of Affine::residual) into dL/dp, and through the tap weights
  // into dL/dI_def.  Returns false if the BATCHED window did not fit.
  __device__ __forceinline__ bool gamma_adjoint(const float (&p)[6], double dl_dgamma, double (&dp)[6],
                                                float *dimg) {
    Affine A;
#pragma unroll
    for (int k = 0; k < 6; ++k) A.p[k] = p[k];
    A.prepare();
    if (!stage_affine(A)) return false;
#pragma unroll
    for (int k = 0; k < 6; ++k) dp[k] = 0.0;
    ZStats z;
    if (!znssd_stats(A, z) || dl_dgamma == 0.0) return true;
    // a second copy of the window's layout in the slot accumulates dL/dI
    // with LDS atomics when it fits; otherwise taps go to global atomics
    lds_float *acc = nullptr;
    if constexpr (kIsWave) {
      const int cells = wpitch * wrows;
      if (cells > 0 && 2 * cells <= win_cap && dimg) {
        acc = win + cells;
        for (int k = grp.rank(); k < cells; k += G::kSize) acc[k] = 0.0f;
        grp.sync();
      }
    }
    const double s = dl_dgamma * 2.0 / z.nd;
    double g6[6] = {0, 0, 0, 0, 0, 0};
    // one pixel: dL/dp terms, and the taps of dL/dI (scatter(ox, oy, g))
    auto pixel = [&](int x, int y, auto &&scatter) {
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      A.disp(dx, dy, ox, oy);
      if (!kept(ox, oy)) return;
      float Ix, Iy;
      const double d = sample_d(ox, oy, Ix, Iy);
      const double r = ref_value(x - ref.ox, y - ref.oy);
      const double gj = s * (z.rho * (d - z.md) / z.nd - (r - z.mr) / z.nr);
      float row[6];
      A.residual(dx, dy, Ix, Iy, row);
#pragma unroll
      for (int k = 0; k < 6; ++k) g6[k] = fma(gj, (double)row[k], g6[k]);
      scatter(ox, oy, (float)gj);
    };
    // Square subsets with unclamped pixels whose warp moves two pixels of one
    // (x mod 2, y mod 2) class at least 1.1 px apart in x or y (|M v| >= |v|
    // (1 - ||M - I||) with |v| >= 2 and ||M - I||_inf <= 0.45): such samples
    // never share a cell, so the class's taps go to LDS with plain
    // read-modify-writes, class by class.
    bool classes = false;
    if constexpr (kIsWave) {
      const float mi = fmaxf(fabsf(A.a11 - 1.0f) + fabsf(A.a12), fabsf(A.a21) + fabsf(A.a22 - 1.0f));
      classes = acc != nullptr && S > 0 && cx - hw >= 0 && cy - hw >= 0 && mi <= 0.45f;
    }
    if (classes) {
      const int lane = grp.rank();
      const int x0 = cx - hw, y0 = cy - hw;
#pragma unroll 1
      for (int ph = 0; ph < 4; ++ph) {
        const int pa = ph >> 1, pb = ph & 1;
        const int nc = (S - pb + 1) >> 1, cnt = ((S - pa + 1) >> 1) * nc;
        const float inv = 1.0f / (float)nc;
#pragma unroll 1
        for (int k = lane; k < cnt; k += G::kSize) {
          const int rr = (int)(((float)k + 0.5f) * inv);  // exact: k < 2^20
          const int cc = k - rr * nc;
          pixel(x0 + pb + 2 * cc, y0 + pa + 2 * rr,
                [&](float ox, float oy, float g) { scatter_taps<true>(ox, oy, g, acc, dimg); });
        }
      }
    } else {
      for_pixels([&](int x, int y, bool a) {
        if (!a) return;
        pixel(x, y, [&](float ox, float oy, float g) {
          if (dimg) scatter_taps(ox, oy, g, acc, dimg);
        });
      });
    }
    grp.sum(g6);
#pragma unroll
    for (int k = 0; k < 6; ++k) dp[k] = g6[k];
    if constexpr (kIsWave) {
      if (acc) {
        grp.sync();
        const int cells = wpitch * wrows;
        for (int k = grp.rank(); k < cells; k += G::kSize) {
          const float v = acc[k];
          if (v != 0.0f) {
            const int r = k / wpitch, c = k - r * wpitch;
            unsafeAtomicAdd(dimg + (size_t)(wy0 + r) * def.w + (wx0 + c), v);
          }
        }
        grp.sync();
      }
    }
    return true;
  }

  // Square subsets of at most kGammaPixels pixels with no clamped pixel,
  // inside the reference image and without blocked pixels (BATCHED waves).
  // The slot holds a table of the subset's reference values (row-major pixel
  // order; direct global -> LDS loads issued before the window's, so one wait
  // covers both) and the window after it (no ring: one evaluation), so that
  // no pixel pass waits on a global load.  The gamma sums are those of
  // gamma_at in the same order.  The adjoint evaluates the interpolant with
  // its slopes in the same single pass and sums (d - c) row, (r - c) row and
  // row per lane in FP64 (c: a reference value, a shift), from which dL/dp
  // follows once the means are known; the interpolated values go to a second
  // table.  dL/dI_def then walks the (x mod 2, y mod 2) classes, so that the
  // taps can be plain read-modify-writes, into an LDS accumulator over the
  // window's space (the window's layout: the taps' offsets are immediates).
  static constexpr int kGammaPixels = 512;
  __device__ __forceinline__ bool gamma_square_ok() const {
    return kIsWave && S > 0 && n <= kGammaPixels && mw == 0 && ref_in && cx - hw >= 0 && cy - hw >= 0 &&
           2 * kGammaPixels + 2 * kWave <= win_cap;
  }

  template <bool ADJ>
  __device__ __forceinline__ bool gamma_square(const float (&p)[6], double dl_dgamma, double &gamma_out,
                                               double (&dp)[6], float *dimg) {
    const int lane = grp.rank();
    const int x0 = cx - hw, y0 = cy - hw;
    // reference table at the slot's start, (adjoint) the table of interpolated
    // values, then the window
    lds_float *const rtab = win;
    const int rpad = (n + kWave - 1) & ~(kWave - 1);
    lds_float *const dtab = rtab + rpad;
    const int tabs = ADJ ? 2 * rpad : rpad;
    {
      const float inv = 1.0f / (float)S;  // (k + 0.5) / S: the integer part of k / S for k < 2^9, S < 2^6
      const unsigned o0 = (unsigned)((y0 - ref.oy) * ref.w + (x0 - ref.ox));
#pragma unroll 1
      for (int k0 = 0; k0 < n; k0 += kWave) {
        const int k = k0 + lane;
        if (k < n) {
          const int r = (int)(((float)k + 0.5f) * inv), c = k - r * S;
          const unsigned o = o0 + (unsigned)(r * ref.w + c);
          if constexpr (std::is_same<T, float>::value)
            __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)(Pix<T>::ptr(ref.pix) + o),
                                             (__attribute__((address_space(3))) void *)(rtab + k0), 4, 0, 0);
          else
            rtab[k] = Pix<T>::at_u(ref.pix, o);
        }
      }
    }
    win += tabs;
    win_cap -= tabs;
    Affine A;
#pragma unroll
    for (int k = 0; k < 6; ++k) A.p[k] = p[k];
    A.prepare();
    const bool staged = stage_affine(A, false);  // its wait covers the table's loads
    if (staged) {
      // (k + 0.5) / S in float has the integer part of k / S for k < 2^8 and S < 2^12
      const float inv = 1.0f / (float)S;
      const int drow = (int)(((float)kWave + 0.5f) * inv), dcol = kWave - drow * S;
      const int row0 = (int)(((float)lane + 0.5f) * inv), col0 = lane - row0 * S;
      // every pixel passes kept()'s guard (no blocked pixels here): skip the test
      const bool allk = all_kept(A);
      double acc[6] = {0, 0, 0, 0, 0, 0};  // n, sd, sdd, sr, srr, sdr
      // adjoint: sum((d - c) row), sum((r - c) row), sum(row) over the kept
      // pixels, c = the first reference value (a shift that keeps the terms
      // small); dL/dp = a (sum((d - dbar) row)) - b (sum((r - rbar) row))
      double sd6[6] = {0, 0, 0, 0, 0, 0}, sr6[6] = {0, 0, 0, 0, 0, 0}, s16[6] = {0, 0, 0, 0, 0, 0};
      const double c = ADJ ? (double)wave_uniform_f(rtab[0]) : 0.0;
      {
        int row = row0, col = col0;
#pragma unroll 1
        for (int i = lane; i < n; i += kWave) {
          const float dx = (float)(col - hw), dy = (float)(row - hw);
          float ox, oy;
          A.disp(dx, dy, ox, oy);
          if (allk || kept(ox, oy)) {
            const float rf = rtab[i];
            const double r = rf;
            double d;
            if constexpr (ADJ) {
              float Ix, Iy;
              const float df = sample_d(ox, oy, Ix, Iy);
              dtab[i] = df;
              d = df;
              float rrow[6];
              A.residual(dx, dy, Ix, Iy, rrow);
              const double ed = d - c, er = r - c;
#pragma unroll
              for (int k = 0; k < 6; ++k) {
                const double rk = rrow[k];
                sd6[k] = fma(ed, rk, sd6[k]);
                sr6[k] = fma(er, rk, sr6[k]);
                s16[k] += rk;
              }
            } else {
              d = sample(ox, oy);
            }
            acc[0] += 1.0;
            acc[1] += d;
            acc[2] += d * d;
            acc[3] += r;
            acc[4] += r * r;
            acc[5] += d * r;
          }
          col += dcol;
          row += drow;
          if (col >= S) { col -= S; ++row; }
        }
      }
      grp.sum(acc);
      ZStats z;
      const bool ok = zstats_from(acc, z);
      gamma_out = ok ? 2.0 - 2.0 * z.rho : -1.0;
      if constexpr (ADJ) {
#pragma unroll
        for (int k = 0; k < 6; ++k) dp[k] = 0.0;
        if (ok && dl_dgamma != 0.0) {
          // dgamma/dd_j = (2 / nd) (rho (d_j - dbar) / nd - (r_j - rbar) / nr)
          const double s = dl_dgamma * 2.0 / z.nd;
          const double a = wave_uniform(s * z.rho / z.nd), b = wave_uniform(s / z.nr);
          const double md = wave_uniform(z.md), mr = wave_uniform(z.mr);
          double g6[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) g6[k] = a * (sd6[k] - (md - c) * s16[k]) - b * (sr6[k] - (mr - c) * s16[k]);
          grp.sum(g6);
#pragma unroll
          for (int k = 0; k < 6; ++k) dp[k] = g6[k];
          if (dimg) {
            // dL/dI_def: the accumulator over the window's space (dead now)
            const int cells = wpitch * wrows, apitch = wpitch;
            const int acells = cells;
            const bool in_lds = wrows > 0 && acells <= win_cap;
            lds_float *const accl = win;
            if (in_lds) {
              grp.sync();
              for (int k = lane; k < acells; k += kWave) accl[k] = 0.0f;
              grp.sync();
            }
            // the taps of one kept pixel (subset offsets x, y; table index i):
            // MODE 1 / 2 into the accumulator (1: plain read-modify-writes,
            // 2: LDS atomics), MODE 3 into the image (global atomics)
            auto pixel = [&](int x, int y, int i, auto mode) {
              constexpr int MODE = decltype(mode)::value;
              float ox, oy;
              A.disp((float)(x - hw), (float)(y - hw), ox, oy);
              if (!allk && !kept(ox, oy)) return;
              const float g = (float)(a * ((double)dtab[i] - md) - b * ((double)rtab[i] - mr));
              if constexpr (MODE == 3) {
                scatter_taps(ox, oy, g, nullptr, dimg);
              } else {
                const float flx = floorf(ox), fly = floorf(oy);
                const float tx = ox - flx, ty = oy - fly;
                constexpr int lead = INTERP == 0 ? 0 : 2, taps = INTERP == 0 ? 2 : 6;
                const int c0 = cx - def.ox + (int)flx - lead - wx0, r0 = cy - def.oy + (int)fly - lead - wy0;
                float wx[6], wy[6];
                if (INTERP == 0) {
                  wx[0] = 1.0f - tx; wx[1] = tx; wy[0] = 1.0f - ty; wy[1] = ty;
                } else {
                  keys_weights(tx, wx);
                  keys_weights(ty, wy);
                }
                lds_float *rowp = accl + (__umul24(r0, apitch) + c0);
#pragma unroll
                for (int m = 0; m < taps; ++m) {
                  const float gm = g * wy[m];
#pragma unroll
                  for (int k = 0; k < taps; ++k) {
                    const float cv = gm * wx[k];
                    if constexpr (MODE == 1) {
                      rowp[k] += cv;
                      asm volatile("" ::: "memory");
                    } else {
                      __hip_atomic_fetch_add(rowp + k, cv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    }
                  }
                  rowp += apitch;
                }
              }
            };
            if (!in_lds) {
              int row = row0, col = col0;
#pragma unroll 1
              for (int i = lane; i < n; i += kWave) {
                pixel(col, row, i, std::integral_constant<int, 3>{});
                col += dcol;
                row += drow;
                if (col >= S) { col -= S; ++row; }
              }
            } else {
              // (x mod 2, y mod 2) classes as in gamma_adjoint: with the warp
              // within ||M - I|| <= 0.45, two samples of a class never share a cell
              // MODE 1 takes two samples of a class per lane at a time when every
              // lane's first one is kept: the cells of all the class's samples
              // differ at each tap, so both reads of a tap go out before the one
              // wait.  A lane without a second sample points it at its first with
              // a zero contribution and writes it first, so the first's write is
              // the one that stays; no lane branches.
              auto pair = [&](int k0, int nc, int cnt, int pa, int pb, float invc) {
                lds_float *rp[2];
                float gv[2], wx[2][6], wy[2][6];
                bool on[2];
                int xs[2], ys[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                  const int kk = k0 + lane + j * kWave;
                  on[j] = kk < cnt;
                  const int kc = on[j] ? kk : 0;
                  const int rr = (int)(((float)kc + 0.5f) * invc);  // exact: kc < 2^20
                  ys[j] = pa + 2 * rr;
                  xs[j] = pb + 2 * (kc - rr * nc);
                  float ox, oy;
                  A.disp((float)(xs[j] - hw), (float)(ys[j] - hw), ox, oy);
                  if (!allk && on[j]) on[j] = kept(ox, oy);
                  const int i = ys[j] * S + xs[j];
                  gv[j] = on[j] ? (float)(a * ((double)dtab[i] - md) - b * ((double)rtab[i] - mr)) : 0.0f;
                  const float flx = floorf(ox), fly = floorf(oy);
                  const float tx = ox - flx, ty = oy - fly;
                  constexpr int lead = INTERP == 0 ? 0 : 2;
                  const int c0 = cx - def.ox + (int)flx - lead - wx0, r0 = cy - def.oy + (int)fly - lead - wy0;
                  if (INTERP == 0) {
                    wx[j][0] = 1.0f - tx; wx[j][1] = tx; wy[j][0] = 1.0f - ty; wy[j][1] = ty;
                  } else {
                    keys_weights(tx, wx[j]);
                    keys_weights(ty, wy[j]);
                  }
                  rp[j] = accl + (__umul24(r0, apitch) + c0);
                }
                if (__ballot(!on[0]) != 0ull) {
                  // a lane's first sample is missing or not kept: one at a time
                  if (on[0]) pixel(xs[0], ys[0], ys[0] * S + xs[0], std::integral_constant<int, 1>{});
                  if (on[1]) pixel(xs[1], ys[1], ys[1] * S + xs[1], std::integral_constant<int, 1>{});
                  return;
                }
                if (!on[1]) rp[1] = rp[0];
                constexpr int taps = INTERP == 0 ? 2 : 6;
#pragma unroll
                for (int m = 0; m < taps; ++m) {
                  const float gm0 = gv[0] * wy[0][m], gm1 = gv[1] * wy[1][m];
#pragma unroll
                  for (int t = 0; t < taps; ++t) {
                    const float v0 = rp[0][t], v1 = rp[1][t];
                    rp[1][t] = v1 + gm1 * wx[1][t];
                    rp[0][t] = v0 + gm0 * wx[0][t];
                    asm volatile("" ::: "memory");
                  }
                  rp[0] += apitch;
                  rp[1] += apitch;
                }
              };
              auto classes = [&](auto mode) {
                constexpr int MODE = decltype(mode)::value;
#pragma unroll 1
                for (int ph = 0; ph < 4; ++ph) {
                  const int pa = ph >> 1, pb = ph & 1;
                  const int nc = (S - pb + 1) >> 1, cnt = ((S - pa + 1) >> 1) * nc;
                  const float invc = 1.0f / (float)nc;
                  if constexpr (MODE == 1) {
#pragma unroll 1
                    for (int k0 = 0; k0 < cnt; k0 += 2 * kWave) pair(k0, nc, cnt, pa, pb, invc);
                  } else {
#pragma unroll 1
                    for (int k = lane; k < cnt; k += kWave) {
                      const int rr = (int)(((float)k + 0.5f) * invc);  // exact: k < 2^20
                      const int y = pa + 2 * rr, x = pb + 2 * (k - rr * nc);
                      pixel(x, y, y * S + x, mode);
                    }
                  }
                }
              };
              const float mi = fmaxf(fabsf(A.a11 - 1.0f) + fabsf(A.a12), fabsf(A.a21) + fabsf(A.a22 - 1.0f));
              if (mi <= 0.45f) classes(std::integral_constant<int, 1>{});
              else classes(std::integral_constant<int, 2>{});
              grp.sync();
              const float invw = 1.0f / (float)wpitch;
              for (int k = lane; k < cells; k += kWave) {
                const int r = (int)(((float)k + 0.5f) * invw);
                const int cc = k - r * wpitch;
                const float v = accl[k];
                if (v != 0.0f) unsafeAtomicAdd(dimg + (size_t)(wy0 + r) * def.w + (wx0 + cc), v);
              }
            }
          }
        }
      }
      grp.sync();
    }
    win -= tabs;
    win_cap += tabs;
    win_ok = false;
    return staged;
  }

  // Objective_ZNSSD::computeUpdateFast.  Returns the solver status (or
  // kOverflow); p_state receives the parameters of the last warp (for sigma).
  // Fast-pass system (every lane the same values): with reference gradients the
  // residual row is r = M phi, so H = M S M^T with the per-subset
  // S = sum(phi phi^T) (pers[0, 36)), and q = M (sum(gmf phi) - corr sum(phi))
  // from the pass sums at pers[kPersistPass, +6) and sum(phi) at pers[36, 42).
  // Rows of M over phi = (gx, gy, dx gx, dy gx, dx gy, dy gy): u (c, -s, 0, 0, 0, 0),
  // v (s, c, 0, 0, 0, 0), theta (0, 0, -g, -ey1, ex1, g), ex (0, 0, 1, 0, 0, 0),
  // ey (0, 0, 0, 0, 0, 1), gxy (0, 0, 0, 1, 1, 0).
  __device__ __forceinline__ void fast_system(const Affine &A, double corr, double (&H)[6][6], double (&q)[6]) const {
    const double c = A.c, sn = A.s, ex1 = 1.0 + (double)A.p[3], ey1 = 1.0 + (double)A.p[4], g = A.p[5];
    // (M t)_r for a vector t over phi
    auto apply = [&](int r, const double (&t)[6]) -> double {
      switch (r) {
        case 0: return c * t[0] - sn * t[1];
        case 1: return sn * t[0] + c * t[1];
        case 2: return ex1 * t[4] - ey1 * t[3] + g * (t[5] - t[2]);
        case 3: return t[2];
        case 4: return t[5];
        default: return t[3] + t[4];
      }
    };
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double t[6];  // S m_j (S is symmetric: row k of S against row j of M)
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double srow[6];
#pragma unroll
        for (int l = 0; l < 6; ++l) srow[l] = pers[6 * k + l];
        t[k] = apply(j, srow);
      }
#pragma unroll
      for (int i = 0; i <= j; ++i) {
        H[i][j] = apply(i, t);
        H[j][i] = H[i][j];
      }
    }
    double v[6];
#pragma unroll
    for (int l = 0; l < 6; ++l) v[l] = pers[kPersistPass + l] - corr * pers[36 + l];
#pragma unroll
    for (int i = 0; i < 6; ++i) q[i] = apply(i, v);
  }
  // Identity rows and columns and a zero gradient for disabled parameters (the
  // same solution as the reference's reduced N x N system), and the LM term.
  __device__ __forceinline__ void finish_system(double (&H)[6][6], double (&q)[6]) const {
#pragma unroll
    for (int k = 2; k < 6; ++k)
      if (en[k] == 0.0f) {
#pragma unroll
        for (int l = 0; l < 6; ++l) { H[k][l] = 0.0; H[l][k] = 0.0; }
        H[k][k] = 1.0;
        q[k] = 0.0;
      }
    const double lm = kp.lm_alpha > 0.0f ? (double)kp.lm_alpha : 0.0;
    H[0][0] += lm;
    H[1][1] += lm;
  }

  // Per-subset constants of the lean kernel in pers (kLean* layout): the
  // L D L^T factors of S = sum(phi phi^T), sum(phi), S00, S01, S11 and the mean
  // of the reference intensities.  They depend on the reference image and the
  // subset only, so they come from the per-subset cache entry when its tag is
  // the current generation; otherwise they are computed and the entry written.
  // Returns false when S is not safely positive definite.
  __device__ __forceinline__ bool lean_setup(int gid) {
    const int lane = grp.rank();
    double *entry = kp.lean_cache ? kp.lean_cache + (size_t)gid * kLeanCacheDoubles : nullptr;
    const double gen = (double)kp.lean_gen;
    if (entry) {
      const double e = lane < kLeanCacheDoubles ? entry[lane] : 0.0;
      const double tag = lane_value(e, kLeanTag);
      if (tag == gen || tag == -gen) {
        if (lane < kLeanCacheDoubles) pers[lane] = e;
        grp.sync();
        meanF = (float)uniform(pers[kLeanMeanF]);
        return tag > 0.0;
      }
    }
    precompute_phi<true>();  // S, sum(phi) and meanF
    double Sm[6][6], sphi[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
      for (int j = i; j < 6; ++j) {
        Sm[i][j] = pers[6 * i + j];
        Sm[j][i] = Sm[i][j];
      }
      sphi[i] = pers[36 + i];
    }
    grp.sync();
    double L[6][6], dinv[6];
    const bool ok = __builtin_amdgcn_readfirstlane((int)ldlt_factor6(Sm, L, dinv)) != 0;
    if (lane == 0) {
      int h = kLeanL;
#pragma unroll
      for (int i = 1; i < 6; ++i)
#pragma unroll
        for (int k = 0; k < i; ++k) pers[h++] = L[i][k];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        pers[kLeanDinv + i] = dinv[i];
        pers[kLeanSphi + i] = sphi[i];
      }
      pers[kLeanS00] = Sm[0][0];
      pers[kLeanS01] = Sm[0][1];
      pers[kLeanS11] = Sm[1][1];
      pers[kLeanMeanF] = (double)meanF;
      pers[kLeanTag] = ok ? gen : -gen;
    }
    grp.sync();
    if (entry && lane < kLeanCacheDoubles) entry[lane] = pers[lane];
    return ok;
  }

  // computeUncertaintyFields in the lean kernel: the square subset inside both
  // images with every sample x + u, y + v away from the border, so the Keys
  // weights are those of the fractional parts of (u, v) for every pixel and the
  // unshifted deformed pixel is in the window (otherwise false: hand over).
  __device__ __forceinline__ bool lean_uncertainty(const Affine &A, double *fields, int nsub, int gid) {
    const float u = A.p[0], v = A.p[1];
    const int fu = (int)floorf(u), fv = (int)floorf(v);
    const int lx0 = bx0 - ref.ox, lx1 = bx1 - ref.ox, ly0 = by0 - ref.oy, ly1 = by1 - ref.oy;
    const float m = 1e-3f;
    const bool uniform_w = (float)lx0 + u > 2.5f + m && (float)lx1 + u < (float)def.w - 3.5f - m &&
                           (float)ly0 + v > 2.5f + m && (float)ly1 + v < (float)def.h - 3.5f - m;
    if (!uniform_w || lx1 > def.w - 2 || ly1 > def.h - 2) return false;
    // the taps of x + u are x + fu - 2 .. x + fu + 3 (integer offsets: no rounding margin)
    if (!stage(lx0 + fu - 2, lx1 + fu + 3, ly0 + fv - 2, ly1 + fv + 3)) return false;
    if (!(wx0 <= lx0 && wx0 + wpitch > lx1 && wy0 <= ly0 && wy0 + wrows > ly1)) return false;
    float wxu[6], wyu[6];
    keys_weights(u - (float)fu, wxu);
    keys_weights(v - (float)fv, wyu);
    float pt[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // the walk of for_strips; the Keys sum at (x + u, y + v) is separable with
    // these fixed weights: each lane keeps the row sums of its column for the
    // six window rows under its current pixel and, stepping down two rows, adds
    // the sums of two new rows (the same products and order as keys_sum)
    const int lane = grp.rank();
    const int x0 = cx - hw, y0 = cy - hw;
    auto rowsum = [&](const auto *p) {
      float rs = p[0] * wxu[0];
#pragma unroll
      for (int k = 1; k < 6; ++k) rs = fmaf(p[k], wxu[k], rs);
      return rs;
    };
    auto pixel = [&](int o, float interp, const lds_float *e) {
      float refv, gx, gy;
      ref_at(o, refv, gx, gy);
      unc_terms(u, v, gx, gy, interp - refv, *e - refv, pt);
    };
    auto column = [&](int x, int r0, int step) {
      const int rx = x - ref.ox, ry0 = y0 + r0 - ref.oy;
      int o = ry0 * ref.w + rx;
      const int ostep = step * ref.w;
      // taps of (rx + u, ry + v): window rows ry + fv - 2 .. + 3, columns rx + fu - 2 .. + 3
      const lds_float *t = win + (__umul24(ry0 + fv - 2 - wy0, wpitch) + (rx + fu - 2 - wx0));
      const lds_float *e = win + (__umul24(ry0 - wy0, wpitch) + (rx - wx0));
      const int tstep = step * wpitch;
      {
        float h0 = rowsum(t), h1 = rowsum(t + wpitch), h2 = rowsum(t + 2 * wpitch), h3 = rowsum(t + 3 * wpitch),
              h4 = rowsum(t + 4 * wpitch), h5 = rowsum(t + 5 * wpitch);
        // this column in the last row of the window: the rows read ahead at the last step
        // are clamped to it (their sums are not used), so every step issues its
        // loads without a branch and they overlap the arithmetic
        const lds_float *tlast = win + (__umul24(wrows - 1, wpitch) + (rx + fu - 2 - wx0));
        float F0 = 0.0f, gx0 = 0.0f, gy0 = 0.0f;
        if (r0 < S) ref_at(o, F0, gx0, gy0);
#pragma unroll 1
        for (int r = r0; r < S; r += 2) {
          const lds_float *t4 = t + 6 * wpitch, *t5 = t + 7 * wpitch;
          t4 = t4 > tlast ? tlast : t4;
          t5 = t5 > tlast ? tlast : t5;
          float a[6], b[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            a[k] = t4[k];
            b[k] = t5[k];
          }
          float F1 = 0.0f, gx1 = 0.0f, gy1 = 0.0f;
          if (r + 2 < S) ref_at(o + ostep, F1, gx1, gy1);
          float vv = h0 * wyu[0];
          vv = fmaf(h1, wyu[1], vv);
          vv = fmaf(h2, wyu[2], vv);
          vv = fmaf(h3, wyu[3], vv);
          vv = fmaf(h4, wyu[4], vv);
          vv = fmaf(h5, wyu[5], vv);
          unc_terms(u, v, gx0, gy0, vv - F0, *e - F0, pt);
          F0 = F1;
          gx0 = gx1;
          gy0 = gy1;
          o += ostep;
          e += tstep;
          t += tstep;
          h0 = h2;
          h1 = h3;
          h2 = h4;
          h3 = h5;
          h4 = rowsum(a);
          h5 = rowsum(b);
        }
      }
    };
    const int nfull = S >> 5;
#pragma unroll 1
    for (int st = 0; st < nfull; ++st) column(x0 + (st << 5) + (lane & 31), lane >> 5, 2);
    const int R = S & 31;
    if (R) {
      // the last columns R at a time, 64 / R rows per step (row-major, so that
      // the reference loads of a half-wave stay in few cache lines)
      const int rows = kWave / R;
      const int lr = (int)(((float)lane + 0.5f) / (float)R), lc = lane - lr * R;
      if (lr < rows) {
        const int rx = x0 + (nfull << 5) + lc - ref.ox, ry0 = y0 + lr - ref.oy;
        int o = ry0 * ref.w + rx;
        const lds_float *t = win + (__umul24(ry0 + fv - 2 - wy0, wpitch) + (rx + fu - 2 - wx0));
        const lds_float *e = win + (__umul24(ry0 - wy0, wpitch) + (rx - wx0));
        const int ostep = rows * ref.w, tstep = rows * wpitch;
#pragma unroll 1
        for (int r = lr; r < S; r += rows) {
          pixel(o, keys_sum(t, wpitch, wxu, wyu), e);
          o += ostep;
          e += tstep;
          t += tstep;
        }
      }
    }
    unc_record(pt, u, v, fields, nsub, gid);
    return true;
  }

  // Gauss-Newton of the lean kernel (all six parameters, no LM term or
  // momentum: the host sends other runs to the general kernel).
  // Every step takes the fast pass, whose system is H = M S M^T, q = M v with
  // v = sum(gmf phi) - (meanG - shift) sum(phi) (see fast_system).  M is
  // invertible, so H du = -q is M^T du = -S^-1 v: w = S^-1 v from the cached
  // factors of S, then du from the closed form of M^T (a rotation block for
  // (u, v); for theta, ex, ey, gxy the rows phi3 - phi4 give theta, the others
  // one parameter each).  kOverflow hands the subset to the general kernel: a
  // window that does not fit, or a warped pixel that is not kept.
  __device__ __forceinline__ int gauss_newton_lean(Affine &A, int &iters, float p_state[6], double *fields,
                                                   int nsub, int gid) {
    const int max_its = kp.max_fast;
    float shift = meanF;  // running estimate of meanG (one-pass ZSSD gradient)
    int it = 0;
#pragma unroll 1
    for (; it <= max_its; ++it) {
      iters = it;
      A.prepare();
#pragma unroll
      for (int k = 0; k < 6; ++k) p_state[k] = A.p[k];
      if (!stage_affine(A) || !all_kept(A)) return kOverflow;
      accumulate_gn_fast<true>(A, shift, kLeanPass);
      const double cnt = (double)n_def_active;
      const float meanG = cnt > 0.0 ? (float)(uniform(pers[kLeanPass + 6]) / cnt) : 0.0f;
      const double corr = (double)meanG - (double)shift;
      const double c = A.c, sn = A.s, ex1 = 1.0 + (double)A.p[3], ey1 = 1.0 + (double)A.p[4], g = A.p[5];
      // the (u, v) block of H for the condition estimate, as fast_system forms it
      {
        const double s00 = pers[kLeanS00], s01 = pers[kLeanS01], s11 = pers[kLeanS11];
        const double t0a = c * s00 - sn * s01, t0b = c * s01 - sn * s11;
        const double t1a = sn * s00 + c * s01, t1b = sn * s01 + c * s11;
        const double h00 = c * t0a - sn * t0b, h01 = c * t1a - sn * t1b, h11 = sn * t1a + c * t1b;
        const float cond = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cond2x2(h00, h01, h01, h11))));
        if (grp.rank() == 0) fields[(size_t)F_CONDITION * nsub + gid] = cond;
        if (cond > 1.0e12f) return HESSIAN_SINGULAR;
      }
      // w = S^-1 v: L y = v, z = D^-1 y, L^T w = z
      double w[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double t = pers[kLeanPass + i] - corr * pers[kLeanSphi + i];
#pragma unroll
        for (int k = 0; k < i; ++k) t = fma(-pers[kLeanL + i * (i - 1) / 2 + k], w[k], t);
        w[i] = t;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) w[i] *= pers[kLeanDinv + i];
#pragma unroll
      for (int i = 4; i >= 0; --i) {
        double t = w[i];
#pragma unroll
        for (int k = i + 1; k < 6; ++k) t = fma(-pers[kLeanL + k * (k - 1) / 2 + i], w[k], t);
        w[i] = t;
      }
      // M^T du = -w
      double dd[6];
      const double det = c * c + sn * sn;
      dd[0] = -(c * w[0] - sn * w[1]) / det;
      dd[1] = -(sn * w[0] + c * w[1]) / det;
      dd[2] = (w[3] - w[4]) / (ex1 + ey1);
      dd[3] = g * dd[2] - w[2];
      dd[4] = -w[5] - g * dd[2];
      dd[5] = ey1 * dd[2] - w[3];
      float old[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const float du = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)dd[k])));
        old[k] = A.p[k];
        A.p[k] += du;
      }
      // test_for_convergence (DICe_LocalShapeFunction.cpp:433-459): u, v, theta
      bool conv = fabsf(A.p[0] - old[0]) < kp.fast_tol && fabsf(A.p[1] - old[1]) < kp.fast_tol;
      if (kp.en_rot) conv = conv && fabsf(A.p[2] - old[2]) < kp.fast_tol;
      if (conv) {
        if (kp.uncertainty && !lean_uncertainty(A, fields, nsub, gid)) return kOverflow;
        break;
      }
      shift = meanG;
    }
    if (it > max_its) return MAX_ITERATIONS_REACHED;
    return CORRELATION_SUCCESSFUL;
  }

  __device__ __forceinline__ int gauss_newton(Affine &A, int &iters, float p_state[6], double *fields, int nsub,
                                              int gid) {
    const int max_its = kp.max_fast;
    float upd_old[6] = {0, 0, 0, 0, 0, 0};
    float shift = meanF;  // running estimate of meanG (one-pass ZSSD gradient)
    int it = 0;
#pragma unroll 1
    for (; it <= max_its; ++it) {
      iters = it;
      A.prepare();
#pragma unroll
      for (int k = 0; k < 6; ++k) p_state[k] = A.p[k];
      if (!stage_affine(A)) return kOverflow;
      float du[6];
      float meanG;
      if constexpr (!kIsWave) {
        double acc[35];
        accumulate_gn_block(A, shift, acc);
        const double cnt = acc[34];
        meanG = cnt > 0.0 ? (float)(acc[33] / cnt) : 0.0f;
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
        const float cond = cond2x2(H[0][0], H[0][1], H[1][0], H[1][1]);
        if (grp.rank() == 0) fields[(size_t)F_CONDITION * nsub + gid] = cond;
        if (cond > 1.0e12f) return HESSIAN_SINGULAR;
        double dd[6];
        if (!ldlt_update6(H, q, dd)) lu_update6(H, q, dd);
#pragma unroll
        for (int k = 0; k < 6; ++k) du[k] = (float)dd[k];
      } else {
        // every lane builds the whole system (the same values in all lanes) and
        // solves it by L D L^T; the pivoted LU is the fallback
        bool fast = false;
        if constexpr (kFast) fast = all_kept(A);
        last_fast = fast;
        const lds_double *tot = nullptr;
        double corr;
        if (fast) {
          // every active pixel is kept: with reference gradients the residual
          // row is r = M phi, so H = M S M^T with the per-subset
          // S = sum(phi phi^T), and q = M (sum(gmf phi) - (meanG - shift) sum(phi))
          accumulate_gn_fast(A, shift);
          const double cnt = (double)n_def_active;
          meanG = cnt > 0.0 ? (float)(uniform(pers[kPersistPass + 6]) / cnt) : 0.0f;
        } else {
          // totals land in LDS
          tot = accumulate_gn_wave(A, shift);
          const double cnt = uniform(tot[kGnTotR + 7]);
          meanG = cnt > 0.0 ? (float)(uniform(tot[kGnTotR + 6]) / cnt) : 0.0f;
        }
        corr = (double)meanG - (double)shift;
        // the system from the LDS totals, with identity rows and columns and a
        // zero gradient for disabled parameters and the LM term on (u, v)
        auto system = [&](double (&H)[6][6], double (&q)[6]) {
          if (fast) {
            fast_system(A, corr, H, q);
          } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
#pragma unroll
              for (int j = i; j < 6; ++j) {
                H[i][j] = tot[kGnTotH + 6 * i + j];
                H[j][i] = H[i][j];
              }
              q[i] = tot[kGnTotQ + i] - corr * tot[kGnTotR + i];
            }
          }
          finish_system(H, q);
        };
        double H[6][6], q[6];
        system(H, q);
        const float cond =
            __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cond2x2(H[0][0], H[0][1], H[0][1], H[1][1]))));
        const int lane = grp.rank();
        if (lane == 0) fields[(size_t)F_CONDITION * nsub + gid] = cond;
        if (cond > 1.0e12f) return HESSIAN_SINGULAR;
        double dd[6];
        const bool spd = ldlt_update6(H, q, dd);
        if (!__builtin_amdgcn_readfirstlane((int)spd)) {
          // the system again from LDS (after the fence: H is not kept live
          // across the solve); lane i < 6 owns row i (selects, not a per-lane
          // index, which would move the array to scratch)
          grp.sync();
          double H2[6][6], q2[6];
          system(H2, q2);
          const int i = lane < 6 ? lane : 0;
          double hrow[6], qv = q2[0];
#pragma unroll
          for (int k = 0; k < 6; ++k) hrow[k] = H2[0][k];
#pragma unroll
          for (int r = 1; r < 6; ++r)
            if (i == r) {
#pragma unroll
              for (int k = 0; k < 6; ++k) hrow[k] = H2[r][k];
              qv = q2[r];
            }
          wave_lu_update6(hrow, qv, dd);
        }
        // the update is wave-uniform: keep it (and the parameters) scalar
#pragma unroll
        for (int k = 0; k < 6; ++k) du[k] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)dd[k])));
      }
      float old[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) old[k] = A.p[k];
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
        if (kp.uncertainty && !uncertainty(A, fields, nsub, gid)) return kOverflow;
        break;
      }
      shift = meanG;
    }
    if (it > max_its) return MAX_ITERATIONS_REACHED;
    return CORRELATION_SUCCESSFUL;
  }

  // interpolate_keys_fourth (DICe_Image.cpp:890-924) at local coordinates:
  // bilinear near the border (0 outside [0, w-1.5)), taps from base/pitch.
  template <typename P>
  __device__ __forceinline__ float keys_single(const P *base, int pitch, int ox0, int oy0, float lx, float ly) const {
    const int w = def.w, h = def.h;
    if (lx <= 2.5f || lx >= w - 3.5f || ly <= 2.5f || ly >= h - 3.5f) {
      if (lx < 0.0f || lx >= w - 1.5f || ly < 0.0f || ly >= h - 1.5f) return 0.0f;
      const int x1 = (int)lx, y1 = (int)ly;
      return bilinear_sum(base + (y1 - oy0) * pitch + (x1 - ox0), pitch, lx - x1, ly - y1);
    }
    const int ix = (int)lx, iy = (int)ly;
    float wx[6], wy[6];
    keys_weights(lx - ix, wx);
    keys_weights(ly - iy, wy);
    return keys_sum(base + (iy - 2 - oy0) * pitch + (ix - 2 - ox0), pitch, wx, wy);
  }

  // computeUncertaintyFields without an image deformer (DICe_Objective.cpp:193-299):
  // exact displacement 0, so the "exact" residual interpolates the deformed
  // image at the pixel itself.  At an integer position the Keys weights are
  // exactly (0, 0, 1, 0, 0, 0) and the bilinear ones (1, 0, 0, 0), so that
  // value is the pixel, or 0 where interpolate_keys_fourth returns 0 (x or y
  // outside [0, size - 1.5)); it is read directly.  Returns false if the
  // BATCHED window for the shifted positions did not fit.
  __device__ __forceinline__ bool uncertainty(const Affine &A, double *fields, int nsub, int gid) {
    const float u = A.p[0], v = A.p[1];
    if constexpr (kIsWave) {
      // window covering the subset bbox shifted by (u, v), Keys support and a
      // 1-px margin for the rounding of x + u
      const int fu = (int)floorf(u), fv = (int)floorf(v);
      if (!stage(bx0 - ref.ox + fu - 3, bx1 - ref.ox + fu + 4, by0 - ref.oy + fv - 3, by1 - ref.oy + fv + 4))
        return false;
    }
    float pt[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // lane partials of unc_terms
    // BATCHED, subset away from the image border: every sample x + u, y + v
    // takes the Keys branch of interpolate_keys_fourth and the fractional
    // offsets are the fractional parts of (u, v), so the weights are computed
    // once (the per-pixel float sum x + u only moves them at rounding level)
    bool uniform_w = false;
    float wxu[6], wyu[6];
    int fu = 0, fv = 0;
    if constexpr (kIsWave && INTERP != 0) {
      const float m = 1e-3f;
      uniform_w = (float)(bx0 - ref.ox) + u > 2.5f + m && (float)(bx1 - ref.ox) + u < (float)def.w - 3.5f - m &&
                  (float)(by0 - ref.oy) + v > 2.5f + m && (float)(by1 - ref.oy) + v < (float)def.h - 3.5f - m;
      fu = (int)floorf(u);
      fv = (int)floorf(v);
      keys_weights(u - (float)fu, wxu);
      keys_weights(v - (float)fv, wyu);
    }
    bool e_in_win = false;
    if constexpr (kIsWave)
      e_in_win = wx0 <= bx0 - ref.ox && wx0 + wpitch > bx1 - ref.ox && wy0 <= by0 - ref.oy &&
                 wy0 + wrows > by1 - ref.oy;
    for_pixels([&](int x, int y, bool a) {
      (void)a;
      const int rx = x - ref.ox, ry = y - ref.oy;
      // the subset's gradient arrays: the reference gradients
      float refv, gx, gy;
      ref_pixel(rx, ry, refv, gx, gy);
      float sub_r;
      if constexpr (kIsWave) {
        if (uniform_w)
          sub_r = keys_sum(win + (__umul24(ry + fv - 2 - wy0, wpitch) + (rx + fu - 2 - wx0)), wpitch, wxu, wyu) - refv;
        else
          sub_r = keys_single(win, wpitch, wx0, wy0, (float)rx + u, (float)ry + v) - refv;
      } else {
        sub_r = keys_single(Pix<T>::ptr(def.pix), def.w, 0, 0, (float)rx + u, (float)ry + v) - refv;
      }
      const bool inb = rx >= 0 && ry >= 0 && rx <= def.w - 2 && ry <= def.h - 2;
      float e = 0.0f;
      if (inb) {
        if constexpr (kIsWave) {
          // the unshifted pixel from the window when the window holds the subset bbox
          e = e_in_win ? (float)win[__umul24(ry - wy0, wpitch) + (rx - wx0)] : Pix<T>::at(def.pix, ry * def.w + rx);
        } else {
          e = Pix<T>::at(def.pix, ry * def.w + rx);
        }
      }
      unc_terms(u, v, gx, gy, sub_r, e - refv, pt);
    });
    unc_record(pt, u, v, fields, nsub, gid);
    return true;
  }

  // One pixel of computeUncertaintyFields (DICe_Objective.cpp:235-276; no image
  // deformer, so the exact motion is 0) into FP32 lane partials: pt[0] sum |g|^2,
  // pt[1] sum (u gx + v gy)^2 / |g|^2, pt[2] sum (v gx - u gy)^2 / |g|^2,
  // pt[3] sum (sub_r - sub_e)^2 / |g|^2 with 1 / |g|^2 = 0 below 1e-4.
  __device__ __forceinline__ static void unc_terms(float u, float v, float gx, float gy, float sub_r, float sub_e,
                                                   float (&pt)[4]) {
    const float g2 = gx * gx + gy * gy;
    const float inv = g2 < 1.0e-4f ? 0.0f : __builtin_amdgcn_rcpf(g2);
    const float ug = u * gx + v * gy;
    const float jg = v * gx - u * gy;
    const float d = sub_r - sub_e;
    pt[0] += g2;
    pt[1] = fmaf(ug * ug, inv, pt[1]);
    pt[2] = fmaf(jg * jg, inv, pt[2]);
    pt[3] = fmaf(d * d, inv, pt[3]);
  }
  // FIELD_1..7 from the lane partials of unc_terms (FP64 totals).
  __device__ __forceinline__ void unc_record(const float (&pt)[4], float u, float v, double *fields, int nsub,
                                             int gid) {
    double acc[4] = {pt[0], pt[1], pt[2], pt[3]};
    grp.sum(acc);
    if (grp.rank() != 0) return;
    // the error and estimate terms coincide (exact motion 0): n_err_g2 = n_uhat_g2;
    // n_uhat2 adds u^2 + v^2 once per pixel
    const double n_uhat_g2 = acc[1], n_err_jg2 = acc[2], r_tot2 = acc[3];
    const double n_uhat2 = (double)(u * u + v * v) * (double)n;
    const double sssig = (1.0 + acc[0]) / (double)(n == 0 ? 1 : n);
    const double cos_hat = n_uhat2 == 0.0 ? 1.0 : sqrt(n_uhat_g2 / n_uhat2);
    fields[(size_t)F_FIELD_1 * nsub + gid] = (float)cos_hat;
    fields[(size_t)F_FIELD_2 * nsub + gid] = 1.0;
    fields[(size_t)F_FIELD_3 * nsub + gid] = (float)sqrt(n_uhat_g2 + n_err_jg2);
    fields[(size_t)F_FIELD_4 * nsub + gid] = cos_hat == 0.0 ? 0.0 : (float)(1.0 / cos_hat * sqrt(r_tot2));
    fields[(size_t)F_FIELD_5 * nsub + gid] = (float)sqrt(n_uhat_g2);
    fields[(size_t)F_FIELD_6 * nsub + gid] = (float)sqrt(r_tot2);
    fields[(size_t)F_FIELD_7 * nsub + gid] = (float)sssig;
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
        if (overflow) return kOverflow;
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
    float last_g = 0.0f;  // gamma of the most recent evaluation (at the current A.p)
    int repeats = 0;
    int iteration = 0;
#pragma unroll 1
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
      int stage_no = 0;  // 0 reflect, 1 expand, 2 contract
      float ysave = 0.0f;
      bool shrink = false;
#pragma unroll
      for (int st = 0; st < 3; ++st) {
        if (st == stage_no) {
          const float fac = facs[st];
          const float fac1 = (float)((1.0 - fac) / nd);
          const float fac2 = fac1 - fac;
          if (st == 2) ysave = g_hi;
#pragma unroll
          for (int j = 0; j < 6; ++j) A.p[j] = fmaf(sums[j], fac1, -(phi[j] * fac2));
#pragma unroll
          for (int j = 2; j < 6; ++j) if (en[j] == 0.0f) A.p[j] = init[j];
          ytry = eval_here(A);
          last_g = ytry;
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
            if (ytry <= g_lo) stage_no = 1;
            else if (ytry >= g_nhi) stage_no = 2;
            else stage_no = 3;
          } else if (st == 1) {
            stage_no = 3;
          } else {
            shrink = ytry >= ysave;
            stage_no = 3;
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
            last_g = gv[i];
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
      // the extra objective evaluation at the current variables (DICe_Simplex.cpp:264):
      // the variables are the last point evaluated (the trial point, or the last
      // shrunk vertex) and the evaluation is deterministic, so its value is that
      // evaluation's gamma
      gamma_new = last_g;
      if (overflow) return kOverflow;
    }
    iters = iteration;
    return CORRELATION_SUCCESSFUL;
  }

  // Immerkaer noise estimate (Subset::noise_std_dev, DICe_Subset.cpp:622-703)
  // over the bbox of the kept pixels at p_state, shifted by the final (u, v).
  template <bool LEAN = false>
  __device__ __forceinline__ float noise_std_dev(int min_x, int max_x, int min_y, int max_y, float u, float v) {
    min_x = (int)((float)min_x + u);
    max_x = (int)((float)max_x + u);
    min_y = (int)((float)min_y + v);
    max_y = (int)((float)max_y + v);
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
    if constexpr (LEAN) {
      // lean kernel, the region and its 1-px ring inside the image and staged:
      // the mask is separable, conv = c(y - 1) - 2 c(y) + c(y + 1) with the row
      // filter c = p(x - 1) - 2 p(x) + p(x + 1); each lane walks one column of a
      // 32-column strip two rows at a time (or a remainder column) keeping the
      // row-filtered values of the rows around its pixel; FP32 lane partials
      const int lx0 = min_x - ox, ly0 = min_y - oy;
      if (bw > 0 && bh > 0 && lx0 >= 1 && ly0 >= 1 && lx0 + bw <= iw - 1 && ly0 + bh <= ih - 1 &&
          stage(lx0 - 1, lx0 + bw, ly0 - 1, ly0 + bh, false)) {
        const int lane = grp.rank();
        float part = 0.0f;
        auto rowf = [&](const lds_float *p) { return (p[-1] + p[1]) - 2.0f * p[0]; };
        auto column = [&](int c, int r0, int step) {
          const lds_float *t = win + (__umul24(ly0 + r0 - wy0, wpitch) + (lx0 + c - wx0));
          const int tstep = step * wpitch;
          {
            float cm = rowf(t - wpitch), c0 = rowf(t), cp = rowf(t + wpitch);
#pragma unroll 1
            for (int r = r0; r < bh; r += 2) {
              part += fabsf((cm + cp) - 2.0f * c0);
              t += tstep;
              cm = cp;
              if (r + 2 < bh) {
                c0 = rowf(t);
                cp = rowf(t + wpitch);
              }
            }
          }
        };
        const int nfull = bw >> 5;
#pragma unroll 1
        for (int st = 0; st < nfull; ++st) column((st << 5) + (lane & 31), lane >> 5, 2);
        const int R = bw & 31;
        if (R) {
          // the last columns column-major over the lanes (consecutive rows: a
          // window pitch apart, distinct banks for an odd pitch)
          int c = (int)(((float)lane + 0.5f) / (float)bh);
          int r = lane - c * bh;
          const int dc = kWave / bh, dr = kWave - dc * bh;
          const lds_float *t = win + (__umul24(ly0 + r - wy0, wpitch) + (lx0 + (nfull << 5) + c - wx0));
          const int tstep = dc + dr * wpitch, twrap = 1 - bh * wpitch;
#pragma unroll 1
          for (int k = lane; k < R * bh; k += kWave) {
            part += fabsf((rowf(t - wpitch) + rowf(t + wpitch)) - 2.0f * rowf(t));
            c += dc;
            r += dr;
            t += tstep;
            if (r >= bh) {
              r -= bh;
              ++c;
              t += twrap;
            }
          }
        }
        acc[0] = part;
        grp.sum(acc);
        const float variance = (float)acc[0];
        return (float)((double)variance * (sqrt(0.5 * 3.14159265358979323846) / (6.0 * (ww - 2) * (hh - 2))));
      }
    }
    if (bw > 0 && bh > 0) {
      const int tot = bw * bh;
      // BATCHED: the bbox and its 1-px ring staged in LDS when it fits (the
      // mask taps of the interior pixels then come from the window)
      bool in_lds = false;
      if constexpr (kIsWave) in_lds = stage(min_x - ox - 1, max_x - ox, min_y - oy - 1, max_y - oy, false);
      const bool fdiv = tot < (1 << 20);  // (k + 0.5) / bw in float is exact below this size
      const float inv = 1.0f / (float)bw;
      const auto *P = Pix<T>::ptr(def.pix);
#pragma unroll 1
      for (int k = grp.rank(); k < tot; k += G::kSize) {
        const int r = fdiv ? (int)(((float)k + 0.5f) * inv) : k / bw;
        const int xx = min_x + (k - r * bw), yy = min_y + r;
        const int lx = xx - ox, ly = yy - oy;
        if (lx < 1 || lx >= iw - 1 || ly < 1 || ly >= ih - 1) {
          const float p = img_at(def, lx, ly);
          acc[0] += (double)sqrtf(p * p);
        } else {
          float t[9];
          if (in_lds) {
            const lds_float *c0 = win + (__umul24(ly - 1 - wy0, wpitch) + (lx - 1 - wx0));
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
              for (int q = 0; q < 3; ++q) t[3 * m + q] = c0[m * wpitch + q];
          } else {
            const auto *c0 = P + (ly - 1) * iw + lx - 1;
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
              for (int q = 0; q < 3; ++q) t[3 * m + q] = (float)c0[m * iw + q];
          }
          const float conv = t[0] - 2.0f * t[1] + t[2] - 2.0f * t[3] + 4.0f * t[4] - 2.0f * t[5] + t[6] -
                             2.0f * t[7] + t[8];
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
    if constexpr (kFast) {
      if (last_fast && (S > 0 || n_def_active > 0)) {
        // the last warp kept every active pixel: the gradient sums are
        // sum(phi phi^T)[0][0] and [1][1], the kept bbox is the bbox of the
        // active pixels (with the centroid, where the general path starts)
        if (S > 0) noise = noise_std_dev(bx0, bx1, by0, by1, A.p[0], A.p[1]);
        else noise = noise_std_dev(min(cx, abx0), max(cx, abx1), min(cy, aby0), max(cy, aby1), A.p[0], A.p[1]);
        if (noise < 0.0f) return -1.0f;
        const float sgx = (float)uniform(pers[0]), sgy = (float)uniform(pers[7]);
        const float sg = sgx > sgy ? sgy : sgx;
        return sg > 0.0f ? (float)sqrt(2.0 * (double)noise * (double)noise / (double)sg) : -1.0f;
      }
    }
    Affine Ps;
#pragma unroll
    for (int k = 0; k < 6; ++k) Ps.p[k] = p_state[k];
    Ps.prepare();
    int mnx = cx, mxx = cx, mny = cy, mxy = cy;
    double acc[3] = {0, 0, 0};  // sum gx^2, sum gy^2, kept count
    for_pixels([&](int x, int y, bool a) {
      if (!a) return;
      const float dx = (float)(x - cx), dy = (float)(y - cy);
      float ox, oy;
      Ps.disp(dx, dy, ox, oy);
      if (!kept(ox, oy)) return;
      mnx = min(mnx, x);
      mxx = max(mxx, x);
      mny = min(mny, y);
      mxy = max(mxy, y);
      float gx, gy;
      if constexpr (DEFGRAD) {
        sample_grad(ox, oy, gx, gy);
      } else {
        float fdum;
        ref_pixel(x - ref.ox, y - ref.oy, fdum, gx, gy);
      }
      acc[0] += (double)(gx * gx);
      acc[1] += (double)(gy * gy);
      acc[2] += 1.0;
    });
    grp.sum(acc);
    mnx = grp.imin(mnx);
    mxx = grp.imax(mxx);
    mny = grp.imin(mny);
    mxy = grp.imax(mxy);
    if (acc[2] <= 0.0) {
      noise = -1.0f;
      return -1.0f;
    }
    noise = noise_std_dev(mnx, mxx, mny, mxy, A.p[0], A.p[1]);
    if (noise < 0.0f) return -1.0f;
    const float sgx = (float)acc[0], sgy = (float)acc[1];
    const float sg = sgx > sgy ? sgy : sgx;
    return sg > 0.0f ? (float)sqrt(2.0 * (double)noise * (double)noise / (double)sg) : -1.0f;
  }

  __device__ __forceinline__ void record_failed(double *fields, int nsub, int gid, int status, int iters) {
    if (grp.rank() != 0) return;
    auto f = [&](int s) -> double & { return fields[(size_t)s * nsub + gid]; };
    f(F_SIGMA) = -1.0;
    f(F_MATCH) = -1.0;
    f(F_GAMMA) = -1.0;
    f(F_BETA) = -1.0;
    f(F_NOISE) = -1.0;
    f(F_CONTRAST) = -1.0;
    f(F_ACTIVE_PIXELS) = -1.0;
    f(F_STATUS) = status;
    f(F_ITERATIONS) = iters;
  }

  // initialize_parameters_from_fields of the subset the initializer picked:
  // the subset itself, or its neighbour's values gathered into kp.guess.
  __device__ __forceinline__ void init_params(Affine &A, const double *fields, int nsub, int gid) {
    auto f = [&](int s, int k) {
      return kp.guess ? kp.guess[(size_t)k * nsub + gid] : fields[(size_t)s * nsub + gid];
    };
    A.p[0] = (float)f(F_DISP_X, 0);
    A.p[1] = (float)f(F_DISP_Y, 1);
    A.p[2] = kp.en_rot ? (float)f(F_ROT_Z, 2) : 0.0f;
    A.p[3] = kp.en_normal ? (float)f(F_NS_XX, 3) : 0.0f;
    A.p[4] = kp.en_normal ? (float)f(F_NS_YY, 4) : 0.0f;
    A.p[5] = kp.en_shear ? (float)f(F_SS_XY, 5) : 0.0f;
  }

  // One subset of the differentiation path: gamma at p (forward) or the
  // adjoint; a BATCHED wave whose window does not fit hands the subset over.
  __device__ __forceinline__ void run_gamma(const DSubsets &ss, int gid, const GammaArgs &ga, int *ov_list,
                                            int *ov_count) {
    set_subset(ss, gid);
    float p[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) p[k] = (float)ga.params[(size_t)gid * 6 + k];
    // the register-resident path of square subsets (BATCHED waves), else the general one
    auto square = [&](auto adj, double w, double &g, double (&dp)[6], bool &ok) {
      if constexpr (kIsWave) {
        if (gamma_square_ok()) {
          ok = gamma_square<decltype(adj)::value>(p, w, g, dp, ga.dimage);
          return true;
        }
      }
      return false;
    };
    bool ok;
    if (ga.adjoint) {
      double dp[6], g;
      const double w = ga.dgamma[gid];
      if (!square(std::true_type{}, w, g, dp, ok)) ok = gamma_adjoint(p, w, dp, ga.dimage);
      if (!ok) {
        push(ov_list, ov_count, gid);
        return;
      }
      if (grp.rank() == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) ga.dparams[(size_t)gid * 6 + k] = dp[k];
    } else {
      double g, dp[6];
      if (!square(std::false_type{}, 0.0, g, dp, ok)) ok = gamma_at(p, g);
      if (!ok) {
        push(ov_list, ov_count, gid);
        return;
      }
      if (grp.rank() == 0) ga.gamma[gid] = g;
    }
  }

  __device__ __forceinline__ void push(int *list, int *count, int gid) {
    if (grp.rank() == 0) list[atomicAdd(count, 1)] = gid;
  }

  // generic_correlation_routine (DICe_Schema.cpp:2025-2368) for one subset,
  // split in two passes so that each kernel holds one optimizer:
  //   pass 0: FOV / initial-guess checks, the primary optimizer, and either
  //           the epilogue (sigma, gamma, record) or a fallback request;
  //   pass 1: the fallback optimizer of GRADIENT_BASED_THEN_SIMPLEX /
  //           SIMPLEX_THEN_GRADIENT_BASED for the subsets that asked for it.
  // SIMPLEX_OPT selects which optimizer this instantiation runs.  A BATCHED
  // wave whose window overflows its LDS slot pushes the subset to ov_list and
  // leaves it to the PLAIN kernel (nothing it wrote is an input of the solve).
  // nm: per-subset Nelder-Mead results of simplex_groups_kernel for this pass
  // (SIMPLEX_OPT only), or null to run the simplex here.  NM_ONLY: the
  // instantiation holds no simplex (and so needs far fewer registers); a
  // subset without a result goes to ov_list, and the PLAIN kernel runs its
  // simplex.
  template <bool SIMPLEX_OPT, bool NM_ONLY = false>
  __device__ __forceinline__ void run(const DSubsets &ss, int gid, double *fields, int pass, int *fb_list,
                                      int *fb_count, int *ov_list, int *ov_count, const float *nm = nullptr) {
    const int nsub = ss.n;
    set_subset(ss, gid);
    last_fast = false;
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
      // Field_Value_Initializer (DICe_Initializer.cpp:398-434): the sigma of the source subset
      if ((kp.guess ? kp.guess[(size_t)6 * nsub + gid] : fv(F_SIGMA)) == -1.0) {
        record_failed(fields, nsub, gid, INITIALIZE_FAILED, -1);
        return;
      }
    }
    Affine A;
    init_params(A, fields, nsub, gid);
    const int init_status = kp.guess ? (int)kp.guess[(size_t)7 * nsub + gid] : INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL;
    int iters = -1;
    if constexpr (!SIMPLEX_OPT && kFast) {
      precompute_phi();  // and meanF
    } else if (!SIMPLEX_OPT) {
      // mean of the reference subset over defined-active pixels (DICe_Objective.cpp:464)
      double acc[2] = {0.0, 0.0};
      for_pixels([&](int x, int y, bool a) {
        if (!a) return;
        acc[0] += (double)ref_value(x - ref.ox, y - ref.oy);
        acc[1] += 1.0;
      });
      grp.sum(acc);
      meanF = acc[1] > 0.0 ? (float)(acc[0] / acc[1]) : 0.0f;
    }
    if (pass == 0 && kp.init_gamma_thr != -1.0) {
      int act;
      float con;
      A.prepare();
      const float g0 = gamma_eval(A, act, con);
      if (overflow) { push(ov_list, ov_count, gid); return; }
      if (g0 > kp.init_gamma_thr || g0 < 0.0f) {
        record_failed(fields, nsub, gid, INITIALIZE_FAILED, iters);
        return;
      }
    }
    if constexpr (!SIMPLEX_OPT) {
      // a forced subset goes to the Nelder-Mead pass, which starts it from its initial guess
      if (pass == 0 && forced) {
        push(fb_list, fb_count, gid);
        return;
      }
    }
    const float prev_u = (float)fv(F_DISP_X), prev_v = (float)fv(F_DISP_Y), prev_t = (float)fv(F_ROT_Z);
    float p_state[6];
    int corr;
    if (SIMPLEX_OPT) {
      const float *r = nm ? nm + (size_t)gid * kNmStride : nullptr;
      if (r && __float_as_int(r[7]) != kNmNoResult) {
#pragma unroll
        for (int k = 0; k < 6; ++k) A.p[k] = r[k];
        iters = __float_as_int(r[6]);
        corr = __float_as_int(r[7]);
      } else if constexpr (NM_ONLY) {
        corr = kOverflow;
      } else {
        corr = simplex(A, iters);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) p_state[k] = A.p[k];
    } else {
      corr = gauss_newton(A, iters, p_state, fields, nsub, gid);
    }
    if (corr == kOverflow || overflow) { push(ov_list, ov_count, gid); return; }
    const float du = A.p[0] - prev_u, dv = A.p[1] - prev_v, dt = (kp.en_rot ? A.p[2] : 0.0f) - prev_t;
    const bool jump = fabsf(du) > kp.disp_jump_tol || fabsf(dv) > kp.disp_jump_tol || fabsf(dt) > kp.theta_jump_tol;
    if (pass == 0) {
      if (corr != CORRELATION_SUCCESSFUL || jump) {
        if ((kp.opt == 3 || kp.opt == 4) && !forced) {
          // the fallback starts again from the initial guess (:2238, :2266)
          push(fb_list, fb_count, gid);
          return;
        }
        if (corr != CORRELATION_SUCCESSFUL) {
          record_failed(fields, nsub, gid, corr, iters);
          return;
        }
        // a successful solve that jumped goes on and fails the later jump test
      }
    } else if (corr != CORRELATION_SUCCESSFUL) {
      record_failed(fields, nsub, gid, corr, iters);
      return;
    }
    float noise = 0.0f;
    const float sig = sigma(p_state, A, noise);
    if (sig < 0.0f) {
      record_failed(fields, nsub, gid, FRAME_FAILED_DUE_TO_NEGATIVE_SIGMA, iters);
      return;
    }
    A.prepare();
    int active;
    float contrast;
    const float gam = gamma_eval(A, active, contrast);
    if (overflow) { push(ov_list, ov_count, gid); return; }
    if ((kp.final_gamma_thr != -1.0 && gam > kp.final_gamma_thr) || gam < 0.0f) {
      record_failed(fields, nsub, gid, FRAME_FAILED_DUE_TO_HIGH_GAMMA, iters);
      return;
    }
    if (jump) {
      record_failed(fields, nsub, gid, JUMP_TOLERANCE_EXCEEDED, iters);
      return;
    }
    if (grp.rank() == 0) {
      auto f = [&](int s) -> double & { return fields[(size_t)s * nsub + gid]; };
      f(F_DISP_X) = A.p[0];
      f(F_DISP_Y) = A.p[1];
      if (kp.en_rot) f(F_ROT_Z) = A.p[2];
      if (kp.en_normal) { f(F_NS_XX) = A.p[3]; f(F_NS_YY) = A.p[4]; }
      if (kp.en_shear) f(F_SS_XY) = A.p[5];
      f(F_SIGMA) = sig;
      f(F_MATCH) = 0.0;
      f(F_GAMMA) = gam;
      f(F_BETA) = 0.0;
      f(F_NOISE) = noise;
      f(F_CONTRAST) = contrast;
      f(F_ACTIVE_PIXELS) = active;
      f(F_STATUS) = init_status;
      f(F_ITERATIONS) = iters;
    }
  }

  // Pass 0 of a Gauss-Newton correlation in the lean kernel: run() for the
  // common case only -- a square subset inside the reference image, no blocked
  // pixels, not forced to Nelder-Mead, every warped pixel kept at every step
  // and at the final parameters, a safely positive definite S (the host
  // launches it only for six parameters without an initial-gamma threshold,
  // LM term, momentum or deformed-image gradients).  Any other subset goes to
  // ov_list for the general kernel, which starts it again from its initial
  // guess: the lean pass has then written nothing but CONDITION and FIELD_1..7,
  // which that run writes again.
  __device__ __forceinline__ void run_lean(const DSubsets &ss, int gid, double *fields, int *fb_list,
                                           int *fb_count, int *ov_list, int *ov_count) {
    const int nsub = ss.n;
    set_subset(ss, gid);
    if (S <= 0 || !ref_in || cx - hw < 0 || cy - hw < 0 || mw > 0 || forced) {
      push(ov_list, ov_count, gid);
      return;
    }
    auto fv = [&](int s) { return fields[(size_t)s * nsub + gid]; };
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
    // Field_Value_Initializer (DICe_Initializer.cpp:398-434): the sigma of the source subset
    if ((kp.guess ? kp.guess[(size_t)6 * nsub + gid] : fv(F_SIGMA)) == -1.0) {
      record_failed(fields, nsub, gid, INITIALIZE_FAILED, -1);
      return;
    }
    Affine A;
    init_params(A, fields, nsub, gid);
    const int init_status = kp.guess ? (int)kp.guess[(size_t)7 * nsub + gid] : INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL;
    if (!lean_setup(gid)) {
      push(ov_list, ov_count, gid);
      return;
    }
    const float prev_u = (float)fv(F_DISP_X), prev_v = (float)fv(F_DISP_Y), prev_t = (float)fv(F_ROT_Z);
    float p_state[6];
    int iters = -1;
    const int corr = gauss_newton_lean(A, iters, p_state, fields, nsub, gid);
    if (corr == kOverflow) {
      push(ov_list, ov_count, gid);
      return;
    }
    const float du = A.p[0] - prev_u, dv = A.p[1] - prev_v, dt = (kp.en_rot ? A.p[2] : 0.0f) - prev_t;
    const bool jump = fabsf(du) > kp.disp_jump_tol || fabsf(dv) > kp.disp_jump_tol || fabsf(dt) > kp.theta_jump_tol;
    if (corr != CORRELATION_SUCCESSFUL || jump) {
      if (kp.opt == 3 || kp.opt == 4) {
        // the fallback starts again from the initial guess (:2238, :2266)
        push(fb_list, fb_count, gid);
        return;
      }
      if (corr != CORRELATION_SUCCESSFUL) {
        record_failed(fields, nsub, gid, corr, iters);
        return;
      }
    }
    // sigma with the last warp keeping every pixel: noise over the subset bbox
    // shifted by (u, v), gradient sums sum(gx^2), sum(gy^2) from S
    // (Objective::sigma, DICe_Objective.cpp:150-184)
    const float noise = noise_std_dev<true>(bx0, bx1, by0, by1, A.p[0], A.p[1]);
    float sig = -1.0f;
    if (noise >= 0.0f) {
      const float sgx = (float)uniform(pers[kLeanS00]), sgy = (float)uniform(pers[kLeanS11]);
      const float sg = sgx > sgy ? sgy : sgx;
      sig = sg > 0.0f ? (float)sqrt(2.0 * (double)noise * (double)noise / (double)sg) : -1.0f;
    }
    if (sig < 0.0f) {
      record_failed(fields, nsub, gid, FRAME_FAILED_DUE_TO_NEGATIVE_SIGMA, iters);
      return;
    }
    A.prepare();
    if (!stage_affine(A) || !all_kept(A)) {
      push(ov_list, ov_count, gid);
      return;
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    gamma_sums_every<true>(A, acc);
    grp.sum(acc);
    acc[0] = (double)n_def_active;
    int active;
    float contrast;
    const float gam = gamma_of(acc, active, contrast);
    if ((kp.final_gamma_thr != -1.0 && gam > kp.final_gamma_thr) || gam < 0.0f) {
      record_failed(fields, nsub, gid, FRAME_FAILED_DUE_TO_HIGH_GAMMA, iters);
      return;
    }
    if (jump) {
      record_failed(fields, nsub, gid, JUMP_TOLERANCE_EXCEEDED, iters);
      return;
    }
    if (grp.rank() == 0) {
      auto f = [&](int s) -> double & { return fields[(size_t)s * nsub + gid]; };
      f(F_DISP_X) = A.p[0];
      f(F_DISP_Y) = A.p[1];
      if (kp.en_rot) f(F_ROT_Z) = A.p[2];
      if (kp.en_normal) { f(F_NS_XX) = A.p[3]; f(F_NS_YY) = A.p[4]; }
      if (kp.en_shear) f(F_SS_XY) = A.p[5];
      f(F_SIGMA) = sig;
      f(F_MATCH) = 0.0;
      f(F_GAMMA) = gam;
      f(F_BETA) = 0.0;
      f(F_NOISE) = noise;
      f(F_CONTRAST) = contrast;
      f(F_ACTIVE_PIXELS) = active;
      f(F_STATUS) = init_status;
      f(F_ITERATIONS) = iters;
    }
  }
};

// A PLAIN pass over a device list launches min(n, kListGroups) workgroups.
constexpr int kListGroups = 2048;

// Work items of a pass: all subsets (list == nullptr) or a device list.
struct Items {
  const int *list;
  const int *count;
  const int *order;  // all subsets: the work order (DSubsets::order), or null
  __device__ __forceinline__ int total(int n) const { return list ? *count : n; }
  __device__ __forceinline__ int at(int k) const { return list ? list[k] : (order ? order[k] : k); }
};

// --------------------------------- 8 or 16 lanes per subset: simplex ---
// The register budget of simplex_groups_kernel as waves per SIMD (0: the
// compiler's choice).
#ifndef DICE_LANE_WAVES_PER_EU
#define DICE_LANE_WAVES_PER_EU 4
#endif
#if DICE_LANE_WAVES_PER_EU > 0
#define DICE_LANE_BOUNDS __attribute__((amdgpu_waves_per_eu(DICE_LANE_WAVES_PER_EU)))
#else
#define DICE_LANE_BOUNDS
#endif

// One lane as a solver group: a Solver on it runs every pixel loop serially in
// the lane and samples the images from global memory (no LDS window).
struct LaneGroup {
  static constexpr int kSize = 1;
  __device__ __forceinline__ int rank() const { return 0; }
  __device__ __forceinline__ void sync() const {}
  template <int K>
  __device__ __forceinline__ void sum(double (&v)[K]) const { (void)v; }
  __device__ __forceinline__ int imin(int v) const { return v; }
  __device__ __forceinline__ int imax(int v) const { return v; }
};

// Simplex::minimize (DICe_Simplex.cpp:60-277) of one subset per row of 8 or 16
// lanes, written as a state machine whose every step ends in exactly one
// objective evaluation.  The rows of a wave are each at their own stage of
// their own subset (vertex set-up, reflection, expansion, contraction,
// shrink), and all of them run the evaluation -- the expensive part --
// together; only the cheap bookkeeping between evaluations diverges.
// Comparisons, update order and arithmetic follow Solver::simplex.  Every lane
// of a row holds the same scalars and computes the same values; the vertex
// table, the vertices' objective values and the initial point live in the
// row's LDS (nm, kNmFloats floats), written by the row's lane 0.
constexpr int kNmFloats = 56;  // pts 7 x 6, gv 7, init 6, one spare
template <typename T, int INTERP, bool SQUARE>
struct LaneSimplex {
  enum : int { kVertex = 0, kReflect = 1, kExpand = 2, kContract = 3, kShrink = 4 };
  Solver<T, LaneGroup, INTERP, false> s;
  Affine A;
  lds_float *nm;
  bool writer;
  float sums[6], phi[6];  // phi: the high vertex; the low one while shrinking
  float g_lo, g_hi, g_nhi, ysave, old_rtol, gamma_new, gamma_old;
  int ilo, ihi, inhi, vk, iteration, repeats, phase;
  int nd, mpts;

  __device__ __forceinline__ LaneSimplex(const KParams &kp, const DImage &r, const DImage &d, lds_float *table,
                                         bool w)
      : s(kp, r, d, LaneGroup{}, nullptr, 0), nm(table), writer(w) {
    nd = 2 + (kp.en_rot ? 1 : 0) + (kp.en_normal ? 2 : 0) + (kp.en_shear ? 1 : 0);
    mpts = nd + 1;
  }
  __device__ __forceinline__ float pt(int i, int j) const { return nm[i * 6 + j]; }
  __device__ __forceinline__ void set_pt(int i, int j, float v) const {
    if (writer) nm[i * 6 + j] = v;
  }
  __device__ __forceinline__ float gval(int i) const { return nm[42 + i]; }
  __device__ __forceinline__ void set_gval(int i, float v) const {
    if (writer) nm[42 + i] = v;
  }
  __device__ __forceinline__ float init(int j) const { return nm[49 + j]; }

  // Start subset gid: vertices p0 and p0 + delta_i e_i; the first step
  // evaluates vertex 0.
  __device__ __forceinline__ void start(const DSubsets &ss, int gid, const double *fields) {
    s.set_subset(ss, gid);
    s.init_params(A, fields, ss.n, gid);
    if (writer) {
#pragma unroll
      for (int j = 0; j < 6; ++j) nm[49 + j] = A.p[j];
#pragma unroll
      for (int i = 0; i < 7; ++i) nm[42 + i] = 0.0f;
      // vertex i moves the (i-1)-th enabled parameter (reference order u, v,
      // theta, ex, ey, gxy) by its delta; rank: enabled parameters before slot j
      int rank = 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const bool on = s.en[j] != 0.0f;
        const float delta = (j < 2) ? s.kp.delta_disp : s.kp.delta_theta;
#pragma unroll
        for (int i = 0; i < 7; ++i) nm[i * 6 + j] = (on && i - 1 == rank) ? A.p[j] + delta : A.p[j];
        if (on) ++rank;
      }
    }
    phase = kVertex;
    vk = 0;
  }

  __device__ __forceinline__ void column_sums() {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      float acc = 0.0f;
#pragma unroll
      for (int i = 0; i < 7; ++i) if (i < mpts) acc += pt(i, j);
      sums[j] = acc;
    }
  }

  // Head of an iteration (no evaluation): stop tests, then the high vertex
  // into phi for the reflection.  Returns the optimizer status when the
  // subset is done, else -1.
  __device__ __forceinline__ int head(int &iters) {
    if (iteration >= s.kp.max_robust) { iters = iteration; return CORRELATION_SUCCESSFUL; }
    if (iteration >= s.kp.max_robust - 1) { iters = iteration; return MAX_ITERATIONS_REACHED; }
    ilo = 0;
    const float g0 = gval(0), g1 = gval(1);
    if (g0 > g1) { inhi = 1; ihi = 0; } else { inhi = 0; ihi = 1; }
    g_lo = g0;
    g_hi = ihi == 0 ? g0 : g1;
    g_nhi = inhi == 0 ? g0 : g1;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (i < mpts) {
        const float gi = gval(i);
        if (gi <= g_lo) { ilo = i; g_lo = gi; }
        if (gi > g_hi) { inhi = ihi; g_nhi = g_hi; ihi = i; g_hi = gi; }
        else if (gi > g_nhi && i != ihi) { inhi = i; g_nhi = gi; }
      }
    }
    const float tiny = 1.0e-10f;
    const float rtol = (float)(2.0 * fabsf(g_hi - g_lo) / (double)(fabsf(g_hi) + fabsf(g_lo) + tiny));
    const float rtol_diff = fabsf(rtol - old_rtol);
    old_rtol = rtol;
    if (fabsf(gamma_new - gamma_old) < 1.0e-8f) repeats++;
    gamma_old = gamma_new;
    if (rtol < s.kp.robust_tol || rtol_diff < s.kp.robust_tol || repeats > 15) {
      iters = iteration;
      return CORRELATION_SUCCESSFUL;
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) phi[j] = pt(ihi, j);
    phase = kReflect;
    return -1;
  }

  // Consume the objective value g at A.p; set A.p to the next point to
  // evaluate (return -1) or finish (return the status; iters, and A.p = the
  // last evaluated point, are the result).  Each piece of bookkeeping has one
  // site, so rows in different phases pay for each once.  The extra
  // evaluation at the end of an iteration (DICe_Simplex.cpp:264) is at the
  // last evaluated point, so its value is g.
  __device__ __forceinline__ int step(float g, int &iters) {
    float fac = 0.0f;       // next: the trial point sums * fac1 - phi * fac2
    bool shrink = false;    // next: the following vertex of a shrink
    bool next_iter = false; // the iteration ended with evaluation g
    bool first = false;     // the vertices are done: the first iteration head
    if (phase == kVertex) {
      set_gval(vk, g);
      if (vk == 0 && g < s.kp.skip_thr && g >= 0.0f) { iters = 0; return CORRELATION_SUCCESSFUL; }
      ++vk;
      if (vk < mpts) {
#pragma unroll
        for (int j = 0; j < 6; ++j) A.p[j] = pt(vk, j);
        return -1;
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) A.p[j] = init(j);
      column_sums();
      old_rtol = 0.0f;
      gamma_new = 0.0f;
      gamma_old = -1.0f;
      repeats = 0;
      iteration = 0;
      first = true;
    } else if (phase == kShrink) {
      set_gval(vk, g);
      shrink = true;
    } else {
      // a trial point: it replaces the high vertex if it is better
      if (g < g_hi) {
        set_gval(ihi, g);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          sums[j] += A.p[j] - pt(ihi, j);
          set_pt(ihi, j, A.p[j]);
          phi[j] = A.p[j];
        }
        g_hi = g;
      }
      if (phase == kReflect && g <= g_lo) {
        phase = kExpand;
        fac = 2.0f;
      } else if (phase == kReflect && g >= g_nhi) {
        ysave = g_hi;
        phase = kContract;
        fac = 0.5f;
      } else if (phase == kContract && g >= ysave) {
        // shrink every vertex but the low one towards it (held in phi)
#pragma unroll
        for (int j = 0; j < 6; ++j) phi[j] = pt(ilo, j);
        phase = kShrink;
        vk = -1;
        shrink = true;
      } else {
        next_iter = true;
      }
    }
    if (shrink) {
      int next = 7;
#pragma unroll
      for (int i = 6; i >= 0; --i)
        if (i > vk && i < mpts && i != ilo) next = i;
      if (next < 7) {
        vk = next;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          const float v = (float)(0.5 * ((double)pt(vk, j) + (double)phi[j]));
          set_pt(vk, j, v);
          sums[j] = v;
          A.p[j] = v;
        }
        return -1;
      }
      column_sums();
      next_iter = true;
    }
    if (next_iter) {
      gamma_new = g;
      ++iteration;
    }
    if (next_iter || first) {
      const int st = head(iters);
      if (st >= 0) return st;
      fac = -1.0f;
    }
    // trial point; disabled parameters stay at their initial values
    const float fac1 = (float)((1.0 - fac) / nd);
    const float fac2 = fac1 - fac;
#pragma unroll
    for (int j = 0; j < 6; ++j) A.p[j] = fmaf(sums[j], fac1, -(phi[j] * fac2));
#pragma unroll
    for (int j = 2; j < 6; ++j) if (s.en[j] == 0.0f) A.p[j] = init(j);
    return -1;
  }
};

// BATCHED Nelder-Mead iterations for small subsets, 64 / W per wave: each row
// of W lanes (16, or 8 when no subset has more than kNarrowSimplexMaxPixels
// pixels) owns one subset, runs LaneSimplex's state machine in all of its
// lanes (the same values in each) and evaluates the objective together --
// lane l takes the pixels l, l + W, ... (Solver::small_partial), the row adds
// the partials (row_sum<W>).  The wave pays for the state machine's
// bookkeeping once per step whatever the number of rows, so narrower rows
// spend less of it per subset.  The row keeps its lanes' reference values in
// registers and stages the deformed window of the current point, with a
// margin of up to 3 px, into its LDS slot of `slot` floats (the last kNmFloats
// of them hold LaneSimplex's vertex table); a point whose footprint leaves the
// window restages it.  A point whose window does not fit the slot reads its
// taps from the image (bilinear) or ends the row's work on the subset with
// status kNmNoResult (Keys): the epilogue kernel hands such a subset to the
// PLAIN kernel, which runs its simplex (in the same summation order, so with
// the same result).  Rows that finish a subset write its result to
// nm_out and take the next items: one atomic per wave and refill, the rows
// numbered by their rank in the ballot.  A row whose item index is past the
// end stays idle, and the wave exits when all of its rows are idle: every
// simplex ends within max_robust iterations, so every wave reaches that exit.
template <typename T, int INTERP, bool SQUARE, int W, int KPER>
__global__ __launch_bounds__(256) DICE_LANE_BOUNDS void simplex_groups_kernel(KParams kp, DImage ref, DImage def,
                                                                              DSubsets ss, const double *fields,
                                                                              PassArgs pa, float *nm_out,
                                                                              int *counter, int slot) {
  extern __shared__ float lds[];
  const Items items{pa.items, pa.item_count, ss.order};
  const int total = items.total(ss.n);
  const int lane = (int)(threadIdx.x % kWave), l = lane & (W - 1);
  // the row's slot: the window, then the vertex table
  lds_float *const win = (lds_float *)lds + (size_t)(threadIdx.x / W) * slot;
  const int cap = slot - kNmFloats;
  LaneSimplex<T, INTERP, SQUARE> m(kp, ref, def, win + cap, l == 0);
  auto &s = m.s;
  constexpr int kPer = KPER;
  constexpr int lead = INTERP == 0 ? 0 : 2, trail = INTERP == 0 ? 1 : 3;
  float rv[kPer];  // reference values of the lane's pixels
#pragma unroll
  for (int k = 0; k < kPer; ++k) rv[k] = 0.0f;
  int wx0 = 0, wy0 = 0, wp = 1, wr = 0;  // the row's window (local deformed coordinates); wr = 0: none
  int gid = 0;
  bool need = true, live = false;
#pragma unroll 1
  while (true) {
    const unsigned long long want = __ballot(need && l == 0);
    if (want) {
      const int leader = __ffsll((unsigned long long)want) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(counter, __popcll(want));
      base = __builtin_amdgcn_readlane(base, leader);
      if (need) {
        const int k = base + __popcll(want & ((1ull << (lane & ~(W - 1))) - 1ull));
        live = k < total;
        if (live) {
          gid = items.at(k);
          m.start(ss, gid, fields);
          s.template small_walk<W, KPER>(l, [&](int q, int x, int y, bool) { rv[q] = s.ref_value(x - ref.ox, y - ref.oy); });
          wr = 0;
        }
        need = false;
      }
    }
    if (__ballot(live) == 0) break;
    if (live) {
      m.A.prepare();
      // the deformed pixels under this point's footprint (as Solver::stage_affine)
      float lo_x, hi_x, lo_y, hi_y;
      s.warp_bounds(m.A, lo_x, hi_x, lo_y, hi_y);
      const float mg = 1e-3f;
      const int nx0 = max(s.cx - def.ox + (int)floorf(lo_x - mg) - lead, 0);
      const int nx1 = min(s.cx - def.ox + (int)floorf(hi_x + mg) + trail, def.w - 1);
      const int ny0 = max(s.cy - def.oy + (int)floorf(lo_y - mg) - lead, 0);
      const int ny1 = min(s.cy - def.oy + (int)floorf(hi_y + mg) + trail, def.h - 1);
      const bool inwin = wr > 0 && nx0 >= wx0 && nx1 < wx0 + wp && ny0 >= wy0 && ny1 < wy0 + wr;
      bool fits = true;
      if (!inwin && nx1 >= nx0 && ny1 >= ny0) {
        const int bw = nx1 - nx0 + 1, bh = ny1 - ny0 + 1;
        int e = 3;
        while (e > 0 && (bw + 2 * e) * (bh + 2 * e) > cap) --e;
        fits = bw * bh <= cap;
        if (fits) {
          const int x0 = max(nx0 - e, 0), y0 = max(ny0 - e, 0);
          wx0 = x0;
          wy0 = y0;
          wp = min(nx1 + e, def.w - 1) - x0 + 1;
          wr = min(ny1 + e, def.h - 1) - y0 + 1;
          const auto *P = Pix<T>::ptr(def.pix) + y0 * def.w + x0;
          const float inv = 1.0f / (float)wp;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
#pragma unroll 4
          for (int c = l; c < wp * wr; c += W) {
            const int r = (int)(((float)c + 0.5f) * inv);  // exact: c < 2^12
            win[c] = (float)P[r * def.w + (c - r * wp)];
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        }
      }
      // bilinear: a point whose window does not fit reads its taps from the
      // image itself (bilinear_sum's arithmetic does not depend on where the
      // taps come from); Keys: the subset goes back to the epilogue kernel
      const bool glob = INTERP == 0 && !fits;
      if (!fits && INTERP != 0) {
        if (l == 0) nm_out[(size_t)gid * kNmStride + 7] = __int_as_float(kNmNoResult);
        need = true;
        continue;
      }
      // the interpolant from the window (the arithmetic of Solver::sample)
      auto smp = [&](float ox, float oy) -> float {
        const float flx = floorf(ox), fly = floorf(oy);
        const float tx = ox - flx, ty = oy - fly;
        if (INTERP == 0 && glob) {
          const int gx = s.cx - def.ox + (int)flx, gy = s.cy - def.oy + (int)fly;
          return bilinear_sum(Pix<T>::ptr(def.pix) + (gy * def.w + gx), def.w, tx, ty);
        }
        const int ix = s.cx - def.ox + (int)flx - wx0, iy = s.cy - def.oy + (int)fly - wy0;
        if (INTERP == 0) return bilinear_sum(win + (iy * wp + ix), wp, tx, ty);
        float wx[6], wy[6];
        keys_weights(tx, wx);
        keys_weights(ty, wy);
        return keys_sum(win + ((iy - 2) * wp + ix - 2), wp, wx, wy);
      };
      double acc[6] = {0, 0, 0, 0, 0, 0};
      s.template small_partial<W, KPER>(m.A, l, acc, smp, [&](int, int, int k) { return rv[k]; });
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k] = row_sum<W>(acc[k]);
      int active;
      float contrast;
      const float g = s.gamma_of(acc, active, contrast);
      int iters = 0;
      const int st = m.step(g, iters);
      if (st >= 0) {
        if (l == 0) {
          float *o = nm_out + (size_t)gid * kNmStride;
#pragma unroll
          for (int j = 0; j < 6; ++j) o[j] = m.A.p[j];
          o[6] = __int_as_float(iters);
          o[7] = __int_as_float(st);
        }
        need = true;
      }
    }
  }
}

// LDS floats of one row's slot in simplex_groups_kernel, rows of W lanes: 16
// rows of 16 lanes (600 floats each) or 32 rows of 8 lanes (312 floats: a 16 x 16
// window and the vertex table) per 256-thread workgroup; four workgroups per
// CU: 150 KB or 156 KB.
template <int W>
constexpr int group_slot_floats() {
  return W == 8 ? 312 : 600;
}

template <typename T, int INTERP, bool SQUARE, int W, int KPER>
hipError_t launch_simplex_lanes_variant(const KParams &kp, const DImage &ref, const DImage &def, const DSubsets &subs,
                                        const double *fields, const PassArgs &pa, float *nm_out, int *counter,
                                        int groups, hipStream_t stream) {
  hipLaunchKernelGGL((simplex_groups_kernel<T, INTERP, SQUARE, W, KPER>), dim3(groups), dim3(256),
                     (size_t)(256 / W) * group_slot_floats<W>() * sizeof(float), stream, kp, ref, def, subs, fields,
                     pa, nm_out, counter, group_slot_floats<W>());
  return hipGetLastError();
}

// PLAIN: workgroup k handles items k, k + gridDim.x, ...
template <typename T, int INTERP, bool DEFGRAD, bool SIMPLEX_OPT>
__global__ __launch_bounds__(kPlainThreads) void correlate_plain_kernel(KParams kp, DImage ref, DImage def,
                                                                        DSubsets ss, double *fields, PassArgs pa) {
  __shared__ double red[4 * kMaxRed];
  __shared__ float small_tab[3 * kLaneSimplexMaxPixels];
  const Items items{pa.items, pa.item_count, ss.order};
  const int total = items.total(ss.n);
  BlockGroup g{(lds_double *)red};
  Solver<T, BlockGroup, INTERP, DEFGRAD> s(kp, ref, def, g, (lds_float *)red, 0);
  s.small_tab = (lds_float *)small_tab;
#pragma unroll 1
  for (int k = blockIdx.x; k < total; k += gridDim.x)
    s.template run<SIMPLEX_OPT>(ss, items.at(k), fields, pa.pass, pa.fb_list, pa.fb_count, nullptr, nullptr,
                                pa.nm_result);
}

// BATCHED: persistent waves pull items from work counters until every range
// is exhausted (every wave reaches that exit).  The items are split into
// kWorkRanges contiguous ranges with a counter each.  Workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup b starts on range b mod 8: the
// subsets that one XCD works on at a time are neighbours in the grid, whose
// windows overlap, and its L2 serves them.  An exhausted range sends the wave
// on to the next one (work stealing).  The register budget is set for WPE
// waves per SIMD: DICE_BATCHED_WAVES_PER_EU (4: 128 VGPRs), or 2 (256 VGPRs,
// no spills) for slots too large for four waves per SIMD in the LDS.
#ifndef DICE_BATCHED_WAVES_PER_EU
#define DICE_BATCHED_WAVES_PER_EU 4
#endif
template <typename T, int INTERP, bool DEFGRAD, bool SIMPLEX_OPT, int WPE = DICE_BATCHED_WAVES_PER_EU,
          bool NM_ONLY = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void correlate_batched_kernel(KParams kp, DImage ref, DImage def, DSubsets ss,
                                                                double *fields, PassArgs pa, int lds_per_wave) {
  extern __shared__ float lds[];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const Items items{pa.items, pa.item_count, ss.order};
  WaveGroup g{lane};
  Solver<T, WaveGroup, INTERP, DEFGRAD> s(kp, ref, def, g, (lds_float *)(lds + (size_t)wave * lds_per_wave),
                                          lds_per_wave);
  const int total = items.total(ss.n);
  // more waves than items (a short list): the first `total` waves take every item between them, and
  // the others leave without touching the counters
  if ((int)blockIdx.x * (int)(blockDim.x / kWave) + wave >= total) return;
  int range = (int)(blockIdx.x % kWorkRanges);
#pragma unroll 1
  for (int visited = 0; visited < kWorkRanges;) {
    const int lo = (int)((long long)total * range / kWorkRanges);
    const int hi = (int)((long long)total * (range + 1) / kWorkRanges);
    int k = 0;
    if (lane == 0) k = atomicAdd(pa.counter + range * kCounterStride, 1);
    k = __builtin_amdgcn_readfirstlane(k) + lo;
    if (k >= hi) {
      range = range + 1 == kWorkRanges ? 0 : range + 1;
      ++visited;
      continue;
    }
    s.template run<SIMPLEX_OPT, NM_ONLY>(ss, items.at(k), fields, pa.pass, pa.fb_list, pa.fb_count, pa.ov_list,
                                         pa.ov_count, pa.nm_result);
  }
}

// BATCHED lean Gauss-Newton pass: the work loop of correlate_batched_kernel
// over Solver::run_lean; the subsets it hands over go to pa.ov_list.
template <typename T, int INTERP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DICE_BATCHED_WAVES_PER_EU))) void correlate_lean_kernel(
    KParams kp, DImage ref, DImage def, DSubsets ss, double *fields, PassArgs pa, int lds_per_wave) {
  extern __shared__ float lds[];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const Items items{pa.items, pa.item_count, ss.order};
  WaveGroup g{lane};
  Solver<T, WaveGroup, INTERP, false> s(kp, ref, def, g, (lds_float *)(lds + (size_t)wave * lds_per_wave),
                                        lds_per_wave);
  const int total = items.total(ss.n);
  if ((int)blockIdx.x * (int)(blockDim.x / kWave) + wave >= total) return;  // as in correlate_batched_kernel
  int range = (int)(blockIdx.x % kWorkRanges);
#pragma unroll 1
  for (int visited = 0; visited < kWorkRanges;) {
    const int lo = (int)((long long)total * range / kWorkRanges);
    const int hi = (int)((long long)total * (range + 1) / kWorkRanges);
    int k = 0;
    if (lane == 0) k = atomicAdd(pa.counter + range * kCounterStride, 1);
    k = __builtin_amdgcn_readfirstlane(k) + lo;
    if (k >= hi) {
      range = range + 1 == kWorkRanges ? 0 : range + 1;
      ++visited;
      continue;
    }
    s.run_lean(ss, items.at(k), fields, pa.fb_list, pa.fb_count, pa.ov_list, pa.ov_count);
  }
}

template <typename T, int INTERP>
hipError_t launch_lean_variant(const Launch &L, const KParams &kp, const DImage &ref, const DImage &def,
                               const DSubsets &subs, double *fields, const PassArgs &pa, hipStream_t stream) {
  const size_t shm = (size_t)L.waves_per_group * L.lds_floats_per_wave * sizeof(float);
  hipLaunchKernelGGL((correlate_lean_kernel<T, INTERP>), dim3(L.groups), dim3(L.waves_per_group * kWave), shm,
                     stream, kp, ref, def, subs, fields, pa, L.lds_floats_per_wave);
  return hipGetLastError();
}

// Differentiation path: forward gamma or adjoint for every item.
template <typename T, int INTERP>
__global__ __launch_bounds__(kPlainThreads) void gamma_plain_kernel(KParams kp, DImage ref, DImage def, DSubsets ss,
                                                                    GammaArgs ga, PassArgs pa) {
  __shared__ double red[4 * kMaxRed];
  const Items items{pa.items, pa.item_count, ss.order};
  const int total = items.total(ss.n);
  BlockGroup g{(lds_double *)red};
  Solver<T, BlockGroup, INTERP, false> s(kp, ref, def, g, (lds_float *)red, 0);
#pragma unroll 1
  for (int k = blockIdx.x; k < total; k += gridDim.x) s.run_gamma(ss, items.at(k), ga, nullptr, nullptr);
}

template <typename T, int INTERP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DICE_BATCHED_WAVES_PER_EU))) void gamma_batched_kernel(KParams kp, DImage ref, DImage def, DSubsets ss,
                                                            GammaArgs ga, PassArgs pa, int lds_per_wave) {
  extern __shared__ float lds[];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const Items items{pa.items, pa.item_count, ss.order};
  WaveGroup g{lane};
  Solver<T, WaveGroup, INTERP, false> s(kp, ref, def, g, (lds_float *)(lds + (size_t)wave * lds_per_wave),
                                        lds_per_wave);
  const int total = items.total(ss.n);
  const int waves = (int)gridDim.x * (int)(blockDim.x / kWave);
  if ((int)blockIdx.x * (int)(blockDim.x / kWave) + wave >= total) return;  // as in correlate_batched_kernel
  // One evaluation per item, so items are short and alike: a wave claims a
  // chunk of up to 4 (about a quarter of its share) per atomic, and claims its
  // next chunk before it works on the current one, so that the atomic's round
  // trip overlaps the work.
  const int chunk = max(1, min(4, total / (4 * waves)));
  int range = (int)(blockIdx.x % kWorkRanges);
  auto claim = [&](int r) {
    int k = 0;
    if (lane == 0) k = atomicAdd(pa.counter + r * kCounterStride, chunk);
    return k;
  };
  int next = claim(range);
#pragma unroll 1
  for (int visited = 0; visited < kWorkRanges;) {
    const int lo = (int)((long long)total * range / kWorkRanges);
    const int hi = (int)((long long)total * (range + 1) / kWorkRanges);
    const int k0 = __builtin_amdgcn_readfirstlane(next) + lo;
    if (k0 >= hi) {
      range = range + 1 == kWorkRanges ? 0 : range + 1;
      if (++visited < kWorkRanges) next = claim(range);
      continue;
    }
    next = claim(range);
    const int k1 = min(k0 + chunk, hi);
#pragma unroll 1
    for (int k = k0; k < k1; ++k) s.run_gamma(ss, items.at(k), ga, pa.ov_list, pa.ov_count);
  }
}

template <typename T, int INTERP, int MODE>
hipError_t launch_gamma_variant(const Launch &L, const KParams &kp, const DImage &ref, const DImage &def,
                                const DSubsets &subs, const GammaArgs &ga, const PassArgs &pa, hipStream_t stream) {
  if (MODE == 0) {
    const int groups = pa.items ? (subs.n < kListGroups ? subs.n : kListGroups) : subs.n;
    hipLaunchKernelGGL((gamma_plain_kernel<T, INTERP>), dim3(groups), dim3(kPlainThreads), 0, stream, kp, ref, def,
                       subs, ga, pa);
    return hipGetLastError();
  }
  const size_t shm = (size_t)L.waves_per_group * L.lds_floats_per_wave * sizeof(float);
  hipLaunchKernelGGL((gamma_batched_kernel<T, INTERP>), dim3(L.groups), dim3(L.waves_per_group * kWave), shm, stream,
                     kp, ref, def, subs, ga, pa, L.lds_floats_per_wave);
  return hipGetLastError();
}

// Launch one pass of one kernel variant.  MODE 0: PLAIN, MODE 1: BATCHED.
// A PLAIN pass over a device list launches min(n, list_groups) workgroups.
template <typename T, int INTERP, bool DEFGRAD, bool SX, int MODE>
hipError_t launch_variant(const Launch &L, const KParams &kp, const DImage &ref, const DImage &def,
                          const DSubsets &subs, double *fields, const PassArgs &pa, hipStream_t stream) {
  if (MODE == 0) {
    const int groups = pa.items ? (subs.n < kListGroups ? subs.n : kListGroups) : subs.n;
    hipLaunchKernelGGL((correlate_plain_kernel<T, INTERP, DEFGRAD, SX>), dim3(groups), dim3(kPlainThreads), 0,
                       stream, kp, ref, def, subs, fields, pa);
    return hipGetLastError();
  }
  const size_t shm = (size_t)L.waves_per_group * L.lds_floats_per_wave * sizeof(float);
  // slots larger than a sixteenth of the CU's 160 KB leave fewer than four waves
  // per SIMD: the variant with twice the registers then costs no occupancy at
  // two waves per SIMD and spills nothing (built for the Gauss-Newton kernel
  // with reference gradients, the large-conformal-subset case)
  // the epilogue of simplex_groups_kernel's results: the variant without a
  // simplex (the subsets it has no result for go to the PLAIN kernel)
  if constexpr (SX && !DEFGRAD) {
    if (pa.nm_result && pa.ov_list) {
      hipLaunchKernelGGL((correlate_batched_kernel<T, INTERP, DEFGRAD, SX, DICE_BATCHED_WAVES_PER_EU, true>),
                         dim3(L.groups), dim3(L.waves_per_group * kWave), shm, stream, kp, ref, def, subs, fields,
                         pa, L.lds_floats_per_wave);
      return hipGetLastError();
    }
  }
  if constexpr (!SX && !DEFGRAD) {
    if (L.lds_floats_per_wave > kWideSlotFloats) {
      hipLaunchKernelGGL((correlate_batched_kernel<T, INTERP, DEFGRAD, SX, 2>), dim3(L.groups),
                         dim3(L.waves_per_group * kWave), shm, stream, kp, ref, def, subs, fields, pa,
                         L.lds_floats_per_wave);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((correlate_batched_kernel<T, INTERP, DEFGRAD, SX>), dim3(L.groups),
                     dim3(L.waves_per_group * kWave), shm, stream, kp, ref, def, subs, fields, pa,
                     L.lds_floats_per_wave);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace dice_hip
