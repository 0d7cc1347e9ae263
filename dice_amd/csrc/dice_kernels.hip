// Image-level kernels of the subset-correlation library and the dispatch of
// the correlation passes (the solver kernels live in dice_k_*.hip).
#include "dice_device.h"

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

namespace dice_hip {

typedef hipError_t (*PassLauncher)(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                                   const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_gn_plain(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                           const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_gn_batched(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                             const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_simplex_plain(int, int, int, const Launch &, const KParams &, const DImage &, const DImage &,
                                const DSubsets &, double *, const PassArgs &, hipStream_t);
hipError_t launch_gn_lean(int, int, const Launch &, const KParams &, const DImage &, const DImage &, const DSubsets &,
                          double *, const PassArgs &, hipStream_t);
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

// Gauss filter masks of Image_::gauss_filter (DICe_Image.cpp:54-126, 1194-1240).
// The 2-D masks are the outer products c_i c_j of these 1-D factors (listed
// from the edge to the centre; the 5-point one sums to 0.9458, so its mask is
// not normalised).  gauss_coefficients forms each product in float from the
// float factors: that reproduces the reference's filtered-image goldens
// (tests/component/images/outFilter*.rawi) bit for bit, where its source
// tables, the same products written to nine digits, move a few pixels by an
// ulp.
static const double kGauss5[3] = {0.0014, 0.1574, 0.62825};
static const double kGauss7[4] = {0.006, 0.0606, 0.2418, 0.3831};
static const double kGauss9[5] = {0.0007, 0.0108, 0.0748, 0.2384, 0.3505};
static const double kGauss11[6] = {0.0001, 0.0017, 0.0168, 0.087, 0.2328, 0.3231};
static const double kGauss13[7] = {0.0001, 0.0012, 0.0085, 0.038, 0.1109, 0.2108, 0.2611};

constexpr int kGfMaxMask = 13;
struct GaussCoef {
  float c[kGfMaxMask * kGfMaxMask];  // c[i * M + j]: i the x offset, j the y offset
};

static bool gauss_coefficients(int M, GaussCoef &g) {
  const double *f = M == 5 ? kGauss5 : M == 7 ? kGauss7 : M == 9 ? kGauss9 : M == 11 ? kGauss11 : M == 13 ? kGauss13 : nullptr;
  if (!f) return false;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      g.c[i * M + j] = (float)f[i <= M / 2 ? i : M - 1 - i] * (float)f[j <= M / 2 ? j : M - 1 - j];
  return true;
}

// One 32 x 32 output tile per 32 x 8 block, each thread four rows; the tile
// and its halo are staged in LDS (reads clamped to the image: clamped values
// only reach the copied border band).  Pixels at least M / 2 + 1 from every
// edge get sum_i sum_j c_ij I(x + i - M/2, y + j - M/2), accumulated in float
// with a separate multiply and add per tap (no fused multiply-add) in the
// reference's order, x offset outer; the others are copied.
constexpr int kGfTile = 32;
template <typename T, int M>
__global__ __launch_bounds__(256) void gauss_filter_kernel(const void *in, void *out, int w, int h, GaussCoef g) {
#pragma clang fp contract(off)
  constexpr int R = M / 2, P = kGfTile + 2 * R;
  __shared__ float tile[P * P];
  const T *src = (const T *)in;
  T *dst = (T *)out;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 32 + tx;
  const int x0 = blockIdx.x * kGfTile, y0 = blockIdx.y * kGfTile;
  for (int k = tid; k < P * P; k += 256) {
    const int ly = k / P, lx = k - ly * P;
    const int gx = min(max(x0 + lx - R, 0), w - 1), gy = min(max(y0 + ly - R, 0), h - 1);
    tile[k] = (float)src[(size_t)gy * w + gx];
  }
  __syncthreads();
  const int hm = R + 1;  // gauss_filter_half_mask_
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int lx = tx, ly = ty + 8 * r;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) continue;
    const size_t o = (size_t)y * w + x;
    if (x >= hm && x < w - hm && y >= hm && y < h - hm) {
      float value = 0.0f;
#pragma unroll
      for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) value += g.c[i * M + j] * tile[(ly + j) * P + lx + i];
      // cast to the storage type as the reference does (truncating for uint16; the
      // 13-point mask sums to 1.0002, so a saturated 16-bit pixel is clamped)
      if constexpr (std::is_same<T, float>::value) dst[o] = value;
      else dst[o] = (T)fminf(value, 65535.0f);
    } else {
      dst[o] = src[o];
    }
  }
}

template <typename T>
hipError_t launch_gauss_t(const void *in, void *out, int w, int h, int mask, hipStream_t stream) {
  GaussCoef g;
  if (!gauss_coefficients(mask, g)) return hipErrorInvalidValue;
  const dim3 block(32, 8), grid((w + kGfTile - 1) / kGfTile, (h + kGfTile - 1) / kGfTile);
  switch (mask) {
    case 5: hipLaunchKernelGGL((gauss_filter_kernel<T, 5>), grid, block, 0, stream, in, out, w, h, g); break;
    case 7: hipLaunchKernelGGL((gauss_filter_kernel<T, 7>), grid, block, 0, stream, in, out, w, h, g); break;
    case 9: hipLaunchKernelGGL((gauss_filter_kernel<T, 9>), grid, block, 0, stream, in, out, w, h, g); break;
    case 11: hipLaunchKernelGGL((gauss_filter_kernel<T, 11>), grid, block, 0, stream, in, out, w, h, g); break;
    case 13: hipLaunchKernelGGL((gauss_filter_kernel<T, 13>), grid, block, 0, stream, in, out, w, h, g); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

__global__ void zero_ints_kernel(int *p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

__global__ void copy_doubles_kernel(double *dst, const double *src, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// One subset per thread: gather the initial guess of subset i from its
// neighbour's fields (see launch_neighbor_guess).
__global__ void neighbor_guess_kernel(const double *start, const double *cur, double *guess, const double *prev,
                                      double tol, int *list, int *count, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nb = (int)start[(size_t)F_NEIGHBOR_ID * n + i];
  const bool self = nb < 0 || nb == i;  // the host has checked nb < n
  const int j = self ? i : nb;
  const double *src = (!self && nb < i) ? cur : start;
  const int slots[7] = {F_DISP_X, F_DISP_Y, F_ROT_Z, F_NS_XX, F_NS_YY, F_SS_XY, F_SIGMA};
  double g[kGuessSlots];
#pragma unroll
  for (int k = 0; k < 7; ++k) g[k] = src[(size_t)slots[k] * n + j];
  g[7] = self ? 1.0 : 3.0;  // INITIALIZE_USING_PREVIOUS_FRAME / _NEIGHBOR_VALUE_SUCCESSFUL
  bool redo = prev == nullptr;
  if (!redo) {
#pragma unroll
    for (int k = 0; k < 6; ++k) redo |= tol > 0.0 ? fabs(g[k] - prev[(size_t)k * n + i]) > tol : g[k] != prev[(size_t)k * n + i];
    const double ps = prev[(size_t)6 * n + i];
    redo |= tol > 0.0 ? (g[6] == -1.0) != (ps == -1.0) : g[6] != ps;
    redo |= g[7] != prev[(size_t)7 * n + i];
  }
#pragma unroll
  for (int k = 0; k < kGuessSlots; ++k) guess[(size_t)k * n + i] = redo ? g[k] : prev[(size_t)k * n + i];
  if (redo) list[atomicAdd(count, 1)] = i;
}

__global__ void reset_listed_kernel(const int *list, const int *count, const double *start, double *fields, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= *count) return;
  const int i = list[k];
#pragma unroll
  for (int s = 0; s < kNumFields; ++s) fields[(size_t)s * n + i] = start[(size_t)s * n + i];
}

// 3 x 3 LU with partial pivoting and the inverse from its factors, in double
// (GETRF + GETRI).  Returns false if a pivot is zero.
__device__ bool inverse3(const double (&a_in)[3][3], double (&inv)[3][3]) {
  double a[3][3];
  int piv[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = a_in[i][j];
  for (int k = 0; k < 3; ++k) {
    int p = k;
    for (int i = k + 1; i < 3; ++i)
      if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
    if (a[p][k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < 3; ++j) {
        const double t = a[k][j];
        a[k][j] = a[p][j];
        a[p][j] = t;
      }
      const int t = piv[k];
      piv[k] = piv[p];
      piv[p] = t;
    }
    for (int i = k + 1; i < 3; ++i) {
      a[i][k] /= a[k][k];
      for (int j = k + 1; j < 3; ++j) a[i][j] -= a[i][k] * a[k][j];
    }
  }
  // solve L U x = P e_c for each column c
  for (int c = 0; c < 3; ++c) {
    double x[3];
    for (int i = 0; i < 3; ++i) {
      double v = piv[i] == c ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) v -= a[i][j] * x[j];
      x[i] = v;
    }
    for (int i = 2; i >= 0; --i) {
      double v = x[i];
      for (int j = i + 1; j < 3; ++j) v -= a[i][j] * x[j];
      x[i] = v / a[i][i];
    }
    for (int i = 0; i < 3; ++i) inv[i][c] = x[i];
  }
  return true;
}

__device__ double norm1_3(const double (&a)[3][3]) {
  double m = 0.0;
  for (int j = 0; j < 3; ++j) {
    const double c = fabs(a[0][j]) + fabs(a[1][j]) + fabs(a[2][j]);
    m = c > m ? c : m;
  }
  return m;
}

// Bounds of the subset centres for the VSG grid, in two launches: each of up
// to kVsgBoundBlocks workgroups reduces a grid-stride share into 5 partials
// (min x, min y, max x, max y over the finite centres, count of non-finite
// ones), then one workgroup folds the partials into out[0..4].
constexpr int kVsgBoundBlocks = 256;

__device__ void vsg_bounds_fold(double (&red)[5][256], double x0, double y0, double x1, double y1, double bad,
                                double *out) {
  const int t = threadIdx.x;
  red[0][t] = x0;
  red[1][t] = y0;
  red[2][t] = x1;
  red[3][t] = y1;
  red[4][t] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = fmin(red[0][t], red[0][t + s]);
      red[1][t] = fmin(red[1][t], red[1][t + s]);
      red[2][t] = fmax(red[2][t], red[2][t + s]);
      red[3][t] = fmax(red[3][t], red[3][t + s]);
      red[4][t] += red[4][t + s];
    }
    __syncthreads();
  }
  if (t < 5) out[t] = red[t][0];
}

__global__ __launch_bounds__(256) void vsg_bounds_partial_kernel(const double *fields, int n, double *partial) {
  __shared__ double red[5][256];
  const double *X = fields + (size_t)F_COORD_X * n, *Y = fields + (size_t)F_COORD_Y * n;
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, bad = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double x = X[i], y = Y[i];
    if (!isfinite(x) || !isfinite(y)) {
      bad += 1.0;
      continue;
    }
    x0 = fmin(x0, x);
    y0 = fmin(y0, y);
    x1 = fmax(x1, x);
    y1 = fmax(y1, y);
  }
  vsg_bounds_fold(red, x0, y0, x1, y1, bad, partial + 5 * blockIdx.x);
}

__global__ __launch_bounds__(256) void vsg_bounds_final_kernel(const double *partial, int blocks, double *out) {
  __shared__ double red[5][256];
  const int t = threadIdx.x;
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, bad = 0.0;
  if (t < blocks) {
    const double *q = partial + 5 * t;
    x0 = q[0];
    y0 = q[1];
    x1 = q[2];
    y1 = q[3];
    bad = q[4];
  }
  vsg_bounds_fold(red, x0, y0, x1, y1, bad, out);
}

// Cell of every subset centre (the same arithmetic as vsg_strain_kernel's own
// cell) and its id, the keys of the sort.
__global__ void vsg_cells_kernel(const double *fields, int n, VsgGrid g, int *keys, int *ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = fields[(size_t)F_COORD_X * n + i], yi = fields[(size_t)F_COORD_Y * n + i];
  const int ci = min(max((int)floor((xi - g.x0) / g.cell), 0), g.ncx - 1);
  const int cj = min(max((int)floor((yi - g.y0) / g.cell), 0), g.ncy - 1);
  keys[i] = cj * g.ncx + ci;
  ids[i] = i;
}

// Cell ranges from the sorted keys: entry k opens its cell when the previous
// key differs and closes it when the next one does (cells without entries keep
// the empty range [0, 0) of the zero fill).
__global__ void vsg_ranges_kernel(const int *keys, int n, int *begin, int *end) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int c = keys[k];
  if (k == 0 || keys[k - 1] != c) begin[c] = k;
  if (k == n - 1 || keys[k + 1] != c) end[c] = k + 1;
}

// Entry k of the cell-ordered list: the centre and displacement of subset
// ids[k] as float (the precision the strain uses), and ids[k] replaced by ~id
// when the subset failed (SIGMA < 0), so that the neighbour loop reads one
// 16-byte record and one word per candidate.
__global__ void vsg_pack_kernel(const double *fields, int n, int *ids, float4 *rec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int j = ids[k];
  auto F = [&](int s) { return fields[(size_t)s * n + j]; };
  rec[k] = make_float4((float)F(F_COORD_X), (float)F(F_COORD_Y), (float)F(F_DISP_X), (float)F(F_DISP_Y));
  if (!(F(F_SIGMA) >= 0.0)) ids[k] = ~j;
}

// One subset per thread (see launch_vsg_strain).
// Candidate records of the radius scan: straight from the cell-ordered list,
// or from a workgroup's LDS copy of the three cell-row ranges it needs.
struct VsgGlobalReader {
  const float4 *rec;
  const int *ids;
  __device__ float4 at(int, int k) const { return rec[k]; }
  __device__ int id(int, int k) const { return ids[k]; }
};
struct VsgTileReader {
  const float4 *rec;  // LDS
  const int *ids;     // LDS
  int row0;           // cell row of range 0
  int glo[3], loff[3];
  __device__ int slot(int cy, int k) const {
    const int r = cy - row0;
    return loff[r] + (k - glo[r]);
  }
  __device__ float4 at(int cy, int k) const { return rec[slot(cy, k)]; }
  __device__ int id(int cy, int k) const { return ids[slot(cy, k)]; }
};

// Strain of subset i (see launch_vsg_strain); radius candidates through R.
// L lanes of one wave per subset (L = 1, 2, 4, ...; lane `part` of them): the
// radius scan deals the candidates of each cell out round-robin and the lanes'
// sums are folded by butterfly shuffles; part 0 finishes.  Every lane of a
// group must reach the fold.
template <int L, class Reader>
__device__ void vsg_subset(const VsgGrid &g, float radius, int knn, double *fields, double *vsg, int n, int i,
                           int part, const Reader &R) {
  auto F = [&](int s, int k) -> double & { return fields[(size_t)s * n + k]; };
  const double xi = F(F_COORD_X, i), yi = F(F_COORD_Y, i);
  const float xf = (float)xi, yf = (float)yi;
  const float r2 = radius * radius + 1.0e-5f;  // the reference's rad^2 + tiny in scalar_t
  double s1 = 0, sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0, su = 0, sxu = 0, syu = 0, sv = 0, sxv = 0, syv = 0;
  int valid = 0;
  // the reference keeps the neighbour offsets and the displacements as scalar_t (float)
  auto add = [&](float uf, float vf, float ddx, float ddy) {
    const double dx = ddx, dy = ddy;
    const double u = uf, v = vf;
    ++valid;
    s1 += 1.0; sx += dx; sy += dy;
    sxx += dx * dx; sxy += dx * dy; syy += dy * dy;
    su += u; sxu += dx * u; syu += dy * u;
    sv += v; sxv += dx * v; syv += dy * v;
  };
  const int ci = min(max((int)floor((xi - g.x0) / g.cell), 0), g.ncx - 1);
  const int cj = min(max((int)floor((yi - g.y0) / g.cell), 0), g.ncy - 1);
  if (F(F_SIGMA, i) >= 0.0 && knn == 0) {
    for (int cy = max(cj - 1, 0); cy <= min(cj + 1, g.ncy - 1); ++cy)
      for (int cx = max(ci - 1, 0); cx <= min(ci + 1, g.ncx - 1); ++cx) {
        const int c = cy * g.ncx + cx;
        for (int k = g.cell_begin[c] + part; k < g.cell_end[c]; k += L) {
          const float4 q = R.at(cy, k);
          const float ddx = q.x - xf, ddy = q.y - yf;
          if (ddx * ddx + ddy * ddy < r2 && R.id(cy, k) >= 0) add(q.z, q.w, ddx, ddy);
        }
      }
  } else if (F(F_SIGMA, i) >= 0.0) {
    // the knn nearest subsets (this one included) by squared float distance, then id: rings of
    // cells around this one until no unvisited cell can hold a closer subset
    float bd[kVsgMaxK];
    int bj[kVsgMaxK];
    int have = 0;
    const int rings = max(g.ncx, g.ncy);
    for (int r = 0; r <= rings; ++r) {
      for (int cy = cj - r; cy <= cj + r; ++cy) {
        if (cy < 0 || cy >= g.ncy) continue;
        const int step = (cy == cj - r || cy == cj + r) ? 1 : 2 * r;
        for (int cx = ci - r; cx <= ci + r; cx += max(step, 1)) {
          if (cx < 0 || cx >= g.ncx) continue;
          const int c = cy * g.ncx + cx;
          for (int k = g.cell_begin[c]; k < g.cell_end[c]; ++k) {
            const int j = g.cell_ids[k] < 0 ? ~g.cell_ids[k] : g.cell_ids[k];
            const float4 q = g.rec[k];
            const float ddx = q.x - xf, ddy = q.y - yf;
            const float d2 = ddx * ddx + ddy * ddy;
            if (have == knn && !(d2 < bd[knn - 1] || (d2 == bd[knn - 1] && j < bj[knn - 1]))) continue;
            int pos = have < knn ? have++ : knn - 1;
            while (pos > 0 && (d2 < bd[pos - 1] || (d2 == bd[pos - 1] && j < bj[pos - 1]))) {
              bd[pos] = bd[pos - 1];
              bj[pos] = bj[pos - 1];
              --pos;
            }
            bd[pos] = d2;
            bj[pos] = j;
          }
        }
      }
      const float reach = (float)(r * g.cell);
      if (have == knn && bd[knn - 1] < reach * reach) break;
    }
    if (have > 0 && !(F(F_SIGMA, bj[0]) >= 0.0)) have = 0;  // the nearest must be valid
    for (int k = 0; k < have; ++k) {
      const int j = bj[k];
      if (F(F_SIGMA, j) >= 0.0)
        add((float)F(F_DISP_X, j), (float)F(F_DISP_Y, j), (float)F(F_COORD_X, j) - xf, (float)F(F_COORD_Y, j) - yf);
    }
  }
  if constexpr (L > 1) {
    auto fold = [](double &v) {
#pragma unroll
      for (int o = L / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    };
    fold(s1); fold(sx); fold(sy); fold(sxx); fold(sxy); fold(syy);
    fold(su); fold(sxu); fold(syu); fold(sv); fold(sxv); fold(syv);
#pragma unroll
    for (int o = L / 2; o >= 1; o >>= 1) valid += __shfl_xor(valid, o);
    if (part != 0) return;
  }
  double out[kVsgFields] = {0, 0, 0, 0, 0, 0, 0};
  bool ok = valid >= 3;
  if (ok) {
    const double a[3][3] = {{s1, sx, sy}, {sx, sxx, sxy}, {sy, sxy, syy}};
    double inv[3][3];
    ok = inverse3(a, inv);
    // the reciprocal 1-norm condition number (the reference estimates it with GECON)
    if (ok) ok = 1.0 / (norm1_3(a) * norm1_3(inv)) >= 1.0e-12;
    if (ok) {
      const double bu[3] = {su, sxu, syu}, bv[3] = {sv, sxv, syv};
      double cu[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          cu[r] += inv[r][c] * bu[c];
          cv[r] += inv[r][c] * bv[c];
        }
      const double dudx = cu[1], dudy = cu[2], dvdx = cv[1], dvdy = cv[2];
      out[0] = (float)(0.5 * (2.0 * dudx + dudx * dudx + dvdx * dvdx));
      out[1] = (float)(0.5 * (2.0 * dvdy + dudy * dudy + dvdy * dvdy));
      out[2] = (float)(0.5 * (dudy + dvdx + dudx * dudy + dvdx * dvdy));
      out[3] = (float)dudx;  // stored in scalar_t fields
      out[4] = (float)dudy;
      out[5] = (float)dvdx;
      out[6] = (float)dvdy;
    }
  }
  for (int k = 0; k < kVsgFields; ++k) vsg[(size_t)k * n + i] = out[k];
  if (!ok) F(F_MATCH, i) = -1.0;
}

// k-nearest windows: one subset per thread, in subset order.
__global__ void vsg_strain_kernel(VsgGrid g, float radius, int knn, double *fields, double *vsg, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vsg_subset<1>(g, radius, knn, fields, vsg, n, i, 0, VsgGlobalReader{g.rec, g.cell_ids});
}

__device__ int lower_bound_int(const int *a, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (a[m] < v) lo = m + 1;
    else hi = m;
  }
  return lo;
}

// Radius windows: kVsgLanes lanes per subset; workgroup b takes entries
// [64 b, 64 b + 64) of the cell-ordered list.  When they lie in one row of cells, every candidate they
// need lies in three contiguous ranges of the list (the cell rows above, at and
// below, from one column left of the first cell to one right of the last),
// which the workgroup copies to LDS once; otherwise, or when the ranges exceed
// kVsgTileEntries, its threads read the list directly.  Both give the same
// sums in the same order.
constexpr int kVsgTileEntries = 2048;
#ifndef DICE_VSG_LANES
#define DICE_VSG_LANES 4
#endif
constexpr int kVsgLanes = DICE_VSG_LANES;
constexpr int kVsgGroupSubsets = 256 / kVsgLanes;
__global__ __launch_bounds__(256) void vsg_strain_tiled_kernel(VsgGrid g, float radius, double *fields, double *vsg,
                                                               int n) {
  __shared__ float4 srec[kVsgTileEntries];
  __shared__ int sid[kVsgTileEntries];
  __shared__ int sh[6];  // start and length of the three ranges
  const int k0 = blockIdx.x * kVsgGroupSubsets;
  const int k1 = min(k0 + kVsgGroupSubsets, n);
  const int cfirst = g.cell_keys[k0], clast = g.cell_keys[k1 - 1];
  const int cj = cfirst / g.ncx;
  bool tiled = false;
  if (clast / g.ncx == cj) {  // uniform over the workgroup
    if (threadIdx.x < 3) {
      const int cy = cj - 1 + (int)threadIdx.x;
      int lo = 0, hi = 0;
      if (cy >= 0 && cy < g.ncy) {
        const int a = cy * g.ncx + max(cfirst - cj * g.ncx - 1, 0);
        const int b = cy * g.ncx + min(clast - cj * g.ncx + 1, g.ncx - 1) + 1;
        lo = lower_bound_int(g.cell_keys, n, a);
        hi = lower_bound_int(g.cell_keys, n, b);
      }
      sh[threadIdx.x] = lo;
      sh[3 + threadIdx.x] = hi - lo;
    }
    __syncthreads();
    tiled = sh[3] + sh[4] + sh[5] <= kVsgTileEntries;
    if (tiled) {
      int off = 0;
      for (int r = 0; r < 3; ++r) {
        for (int t = threadIdx.x; t < sh[3 + r]; t += blockDim.x) {
          srec[off + t] = g.rec[sh[r] + t];
          sid[off + t] = g.cell_ids[sh[r] + t];
        }
        off += sh[3 + r];
      }
    }
    __syncthreads();
  }
  const int k = k0 + (int)threadIdx.x / kVsgLanes, part = (int)threadIdx.x % kVsgLanes;
  if (k >= n) return;  // whole groups of lanes
  const int e = g.cell_ids[k];
  const int i = e < 0 ? ~e : e;
  if (tiled) {
    const VsgTileReader R{srec, sid, cj - 1, {sh[0], sh[1], sh[2]}, {0, sh[3], sh[3] + sh[4]}};
    vsg_subset<kVsgLanes>(g, radius, 0, fields, vsg, n, i, part, R);
  } else {
    vsg_subset<kVsgLanes>(g, radius, 0, fields, vsg, n, i, part, VsgGlobalReader{g.rec, g.cell_ids});
  }
}

__global__ void uncertainty_kernel(const double *fields, double *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float angle = (float)fields[(size_t)F_FIELD_1 * n + i];
  const float sig = (float)fields[(size_t)F_SIGMA * n + i];
  // 1.0 / angle * sig with the double literal of the reference, stored as scalar_t
  out[i] = sig < 0.0f ? 0.0f : angle == 0.0f ? 0.0f : (float)(1.0 / (double)angle * (double)sig);
  out[(size_t)n + i] = angle;
}

}  // namespace

hipError_t launch_uncertainty(const double *fields, double *out, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(uncertainty_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, fields, out, n);
  return hipGetLastError();
}

hipError_t launch_vsg_bounds(const double *fields, int n, double *bounds, hipStream_t stream) {
  if (n <= 0) return hipErrorInvalidValue;
  // bounds holds 5 results, then 5 partials per workgroup
  const int blocks = std::min(kVsgBoundBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(vsg_bounds_partial_kernel, dim3(blocks), dim3(256), 0, stream, fields, n, bounds + 5);
  hipLaunchKernelGGL(vsg_bounds_final_kernel, dim3(1), dim3(256), 0, stream, bounds + 5, blocks, bounds);
  return hipGetLastError();
}

static int vsg_key_bits(long long ncells) {
  int bits = 1;
  while ((1LL << bits) < ncells) ++bits;
  return bits;
}

hipError_t vsg_grid_temp_bytes(int n, long long ncells, size_t *bytes) {
  return rocprim::radix_sort_pairs(nullptr, *bytes, (const int *)nullptr, (int *)nullptr, (const int *)nullptr,
                                   (int *)nullptr, (size_t)n, 0, vsg_key_bits(ncells));
}

hipError_t launch_vsg_grid(const double *fields, int n, VsgGrid &g, const VsgWork &w, hipStream_t stream) {
  if (n <= 0) return hipErrorInvalidValue;
  const long long ncells = (long long)g.ncx * g.ncy;
  hipLaunchKernelGGL(zero_ints_kernel, dim3((int)((2 * ncells + 255) / 256)), dim3(256), 0, stream, w.cell_range,
                     (int)(2 * ncells));
  hipLaunchKernelGGL(vsg_cells_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, fields, n, g, w.keys, w.ids);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // ids ordered by cell; the radix sort is stable, so ids ascend within a cell
  size_t bytes = w.temp_bytes;
  e = rocprim::radix_sort_pairs(w.temp, bytes, (const int *)w.keys, w.keys_sorted, (const int *)w.ids, w.ids_sorted,
                                (size_t)n, 0, vsg_key_bits(ncells), stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(vsg_ranges_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, w.keys_sorted, n, w.cell_range,
                     w.cell_range + ncells);
  hipLaunchKernelGGL(vsg_pack_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, fields, n, w.ids_sorted, w.rec);
  g.cell_begin = w.cell_range;
  g.cell_end = w.cell_range + ncells;
  g.cell_ids = w.ids_sorted;
  g.cell_keys = w.keys_sorted;
  g.rec = w.rec;
  return hipGetLastError();
}

hipError_t launch_vsg_strain(const VsgGrid &g, float radius, int knn, double *fields, double *vsg, int n,
                             hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (knn < 0 || knn > kVsgMaxK) return hipErrorInvalidValue;
  if (knn == 0)
    hipLaunchKernelGGL(vsg_strain_tiled_kernel, dim3((n + kVsgGroupSubsets - 1) / kVsgGroupSubsets), dim3(256), 0,
                       stream, g, radius, fields, vsg, n);
  else
    hipLaunchKernelGGL(vsg_strain_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, g, radius, knn, fields, vsg,
                       n);
  return hipGetLastError();
}

hipError_t launch_neighbor_guess(const double *start, const double *cur, double *guess, const double *prev_guess,
                                 double tol, int *list, int *count, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(neighbor_guess_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, start, cur, guess, prev_guess,
                     tol, list, count, n);
  return hipGetLastError();
}

hipError_t launch_reset_listed(const int *list, const int *count, const double *start, double *fields, int n,
                               hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(reset_listed_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, list, count, start, fields, n);
  return hipGetLastError();
}

// One optimizer pass over `items` (nullptr: all subsets).  BATCHED mode runs the
// persistent kernel, then the PLAIN kernel over the subsets it handed over.
// A Nelder-Mead pass with L.lane_simplex first runs the iterations in rows of 8
// or 16 lanes per subset (results in nm); the persistent kernel then does the
// checks and the epilogue of each subset around them, and hands any subset
// that kernel left with kNmNoResult to the PLAIN kernel, which runs its simplex.
static hipError_t launch_pass(bool simplex, int storage, const Launch &L, const KParams &kp, const DImage &ref,
                              const DImage &def, const DSubsets &subs, double *fields, PassArgs pa, int *ov_list,
                              int *ov_count, float *nm, hipStream_t stream, int *lean_list = nullptr) {
  const int defgrad = simplex ? 0 : kp.def_grads;
  PassLauncher plain = simplex ? launch_simplex_plain : launch_gn_plain;
  pa.nm_result = nullptr;
  if (L.mode == 0) {
    pa.ov_list = pa.ov_count = nullptr;
    return plain(storage, kp.interp, defgrad, L, kp, ref, def, subs, fields, pa, stream);
  }
  PassLauncher batched = simplex ? launch_simplex_batched : launch_gn_batched;
  hipError_t e;
  // Gauss-Newton pass 0 of square subsets with none of the options the lean
  // kernel leaves out: the lean kernel takes every subset first and hands the
  // others (and those that leave its common case) to the general kernel
  if (!simplex && pa.pass == 0 && subs.S > 0 && !subs.mbox && !kp.def_grads &&
      kp.momentum == 0.0f && kp.init_gamma_thr == -1.0 && !kp.has_forced && kp.en_rot && kp.en_normal &&
      kp.en_shear && kp.lm_alpha <= 0.0f && lean_list) {
    e = launch_zero_ints(pa.counter, kWorkCounterInts, stream);
    if (e != hipSuccess) return e;
    PassArgs pl = pa;
    pl.ov_list = lean_list;
    pl.ov_count = pa.counter + kWorkLeanCount;
    e = launch_gn_lean(storage, kp.interp, L, kp, ref, def, subs, fields, pl, stream);
    if (e != hipSuccess) return e;
    pa.items = lean_list;
    pa.item_count = pa.counter + kWorkLeanCount;
  }
  if (simplex && L.lane_simplex) {
    int *lane_counter = pa.counter + kWorkLaneCount;
    e = launch_zero_ints(lane_counter, 1, stream);
    if (e != hipSuccess) return e;
    e = launch_simplex_lanes(storage, kp.interp, L.lane_simplex, kp, ref, def, subs, fields, pa, nm, lane_counter,
                             stream);
    if (e != hipSuccess) return e;
    pa.nm_result = nm;
  }
  e = launch_zero_ints(pa.counter, kWorkCounterInts, stream);
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
                            const DSubsets &subs, double *fields, int *work, hipStream_t stream, const int *items,
                            const int *item_count) {
  if (subs.n <= 0) return hipSuccess;
  // work: the header (range counters, fallback and overflow counts), then the
  // fallback list and the overflow lists of pass 0 and pass 1 (n each)
  const size_t n = (size_t)subs.n;
  int *fb_list = work + kWorkHeader, *ov0 = fb_list + n, *ov1 = fb_list + 2 * n;
  float *nm = reinterpret_cast<float *>(fb_list + 3 * n);
  hipError_t e = launch_zero_ints(work, kWorkHeader, stream);
  if (e != hipSuccess) return e;
  const bool simplex_first = kp.opt == 0 || kp.opt == 3;
  PassArgs pa{0, items, items ? item_count : nullptr, fb_list, work + kWorkFbCount, nullptr, nullptr, work, nullptr};
  e = launch_pass(simplex_first, storage, L, kp, ref, def, subs, fields, pa, ov0, work + kWorkOv0Count, nm, stream,
                  ov1);
  // the second pass: the fallback optimizer, or Nelder-Mead for the forced subsets of a Gauss-Newton pass
  if (e != hipSuccess || !(kp.opt == 3 || kp.opt == 4 || (kp.has_forced && !simplex_first))) return e;
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

hipError_t launch_gauss_filter(int storage, const void *in, void *out, int w, int h, int mask_size,
                               hipStream_t stream) {
  if (w < mask_size || h < mask_size) return hipErrorInvalidValue;
  return storage == 0 ? launch_gauss_t<float>(in, out, w, h, mask_size, stream)
                      : launch_gauss_t<uint16_t>(in, out, w, h, mask_size, stream);
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
