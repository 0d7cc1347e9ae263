// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dice_hip {

constexpr int kWave = 64;
constexpr int kPlainThreads = 256;
constexpr int kNumFields = 26;  // DICE_F_COUNT

// Status_Flag values, DICe.h:609-677.
enum Status : int {
  CORRELATION_SUCCESSFUL = 0,
  INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL = 1,
  INITIALIZE_FAILED = 5,
  CORRELATION_FAILED = 8,
  MAX_ITERATIONS_REACHED = 11,
  INITIALIZE_FAILED_BY_EXCEPTION = 12,
  JUMP_TOLERANCE_EXCEEDED = 16,
  HESSIAN_SINGULAR = 21,
  FRAME_FAILED_DUE_TO_HIGH_GAMMA = 23,
  FRAME_FAILED_DUE_TO_NEGATIVE_SIGMA = 24,
};

// Field slots (enum dice_field).
enum Field : int {
  F_COORD_X = 0, F_COORD_Y, F_DISP_X, F_DISP_Y, F_ROT_Z, F_NS_XX, F_NS_YY, F_SS_XY,
  F_SIGMA, F_GAMMA, F_BETA, F_MATCH, F_NOISE, F_CONTRAST, F_ACTIVE_PIXELS,
  F_STATUS, F_ITERATIONS, F_CONDITION, F_NEIGHBOR_ID,
  F_FIELD_1, F_FIELD_2, F_FIELD_3, F_FIELD_4, F_FIELD_5, F_FIELD_6, F_FIELD_7,
};

// Solver options resolved for the device (all wave-uniform).
struct KParams {
  int interp;            // 0 bilinear, 2 keys
  int opt;               // optimization_method
  int en_rot, en_normal, en_shear;
  int max_fast, max_robust;
  float fast_tol;
  double robust_tol;
  float delta_disp, delta_theta;
  float skip_thr;
  double init_gamma_thr, final_gamma_thr;
  double disp_jump_tol, theta_jump_tol;
  float lm_alpha;        // > 0 enables the Levenberg-Marquardt diagonal
  float momentum;        // != 0 enables momentum
  int def_grads;         // deformed image has gradients: GN uses interpolated def gradients
  int normalize_gamma;
  int uncertainty;
  int first_frame;
  int has_forced;        // some subset has force[i] set: a GN primary pass hands them to a Nelder-Mead pass
  int small_w;           // lanes of the fixed summation order of small subsets' objective (8 or 16)
  // lean kernel: per-subset cache of kLeanCacheDoubles doubles (or null), and the
  // generation its entries must carry to be current (entries of S that is not
  // positive definite carry -generation)
  double *lean_cache;
  int lean_gen;
  // USE_NEIGHBOR_VALUES: per-subset initial guesses, slot-major n x kGuessSlots
  // (u, v, theta, ex, ey, gxy, sigma of the source subset, init status), or null
  // to start every subset from its own fields (USE_FIELD_VALUES)
  const double *guess;
};
constexpr int kGuessSlots = 8;

// Device image: pixels (float or uint16) + optional float gradients.
struct DImage {
  const void *pix;
  const float *gx, *gy;
  int w, h, ox, oy;
};

// Subset set.  square: S > 0, centre (cx, cy).  conformal: S == 0 and a CSR pixel list.
struct DSubsets {
  int n;
  int S;                       // square side (0: conformal)
  int fov;                     // subset_dim for the first-frame field-of-view test (0: off)
  const int *cx, *cy;
  const int *off;              // n + 1 (conformal)
  const int *px, *py;          // pixel coordinates (conformal)
  const uint8_t *act;          // per-pixel active flag (conformal, may be null)
  const int4 *bbox;            // per-subset pixel-coordinate bbox (min_x, min_y, max_x, max_y)
  const int *n_active_def;     // per-subset count of defined-active pixels (conformal)
  const uint32_t *pxy;         // conformal, may be null: pixel i packed as x | y << 15 | inactive << 30
                               // (coordinates clamped at 0, below 2^15)
  const int4 *abox;            // conformal: bbox of the active pixels (of all pixels if none is active)
  // Per-subset options (either may be null).  force[i] != 0: Nelder-Mead only (Schema::set_force_simplex
  // and the blocking subset ids, DICe_Schema.cpp:2160-2185).  mbox[i] = (x0, y0, w, h) and the row-major
  // bits at mbits + moff[i]: image pixels that deactivate a warped pixel of subset i rounding onto them
  // (Subset::obstructed_coords_, pixels_blocked_by_other_subsets_; w = 0: none).
  const uint8_t *force;
  const int4 *mbox;
  const int *moff;
  const uint32_t *mbits;
  // Work order of a pass over all subsets (may be null: subset order): item k is
  // subset order[k].  Conformal subsets go in Morton order of their centres, so
  // that the contiguous item range each XCD takes is one region of the images
  // and its L2 holds that region's pixels.
  const int *order;
};

// BATCHED Gauss-Newton reduction: kGnSums per-lane partials (q 6, H upper
// triangle 21, sum r 6, sum G, count) transposed through a table of row stride
// kGnRedStride floats, totals (52 doubles) at float offset kGnTotOffset.
constexpr int kGnSums = 35;
constexpr int kGnRedStride = kWave + 1;
constexpr int kGnTotOffset = kGnSums * kGnRedStride + 1;
constexpr int kGnTotQ = 0, kGnTotH = 8, kGnTotR = 44;
// Per-subset constants of the BATCHED fast Gauss-Newton pass at the end of a
// wave's slot: the 6 x 6 matrix sum(phi phi^T), sum(phi), sum(F) and the pixel
// count (44 doubles), then the sums of the current pass, sum(gmf phi) and
// sum(G) (7 doubles), and one double of padding.
constexpr int kPersistFloats = 104;
constexpr int kPersistPass = 44;  // double offset of the current pass's sums
// The lean kernel's layout of the same area (double offsets): the L D L^T
// factors of S (L strictly lower, packed by rows, 15; then 1/d, 6), sum(phi)
// (6), S00, S01, S11, the mean reference intensity, the cache tag, then the
// current pass's sums (7).  The first kLeanCacheDoubles are also the per-subset
// cache entry in global memory.
constexpr int kLeanL = 0, kLeanDinv = 15, kLeanSphi = 21, kLeanS00 = 27, kLeanS01 = 28, kLeanS11 = 29;
constexpr int kLeanMeanF = 30, kLeanTag = 31, kLeanPass = 32;
constexpr int kLeanCacheDoubles = 32;
constexpr int kMinSlotFloats = kGnTotOffset + 2 * 52 + kPersistFloats;  // smallest LDS slot per wave
// the largest slot with which four waves per SIMD (16 per CU) fit the CU's 160 KB of LDS
constexpr int kWideSlotFloats = 160 * 1024 / (16 * 4);

struct Launch {
  int mode;                    // 0 plain, 1 batched
  int waves_per_group;         // batched
  int groups;                  // batched: persistent grid size
  int lds_floats_per_wave;     // batched: LDS window budget
  int lane_simplex;            // batched: 0, or the most pixels of a subset (<= kLaneSimplexMaxPixels):
                               // Nelder-Mead passes run rows of 8 or 16 lanes per subset first
                               // (simplex_groups_kernel)
};

// Subsets of at most this many pixels run the Nelder-Mead iterations of the
// BATCHED mode in rows of 16 lanes per subset (simplex_groups_kernel): a wave
// then works on four subsets instead of spreading a few dozen pixels over 64
// lanes.
constexpr int kLaneSimplexMaxPixels = 256;
// When no subset has more than this many pixels, rows of 8 lanes (eight subsets
// per wave) instead of 16.
constexpr int kNarrowSimplexMaxPixels = 88;
// Per-subset result of simplex_groups_kernel: the last evaluated point (6
// floats), then the iteration count and the optimizer status as int bits.
// Status kNmNoResult: the kernel left the subset (its window did not fit the
// LDS slot), and the pass runs the simplex itself.
constexpr int kNmStride = 8;
constexpr int kNmNoResult = -1;

// Device work lists of one correlation pass.  items == nullptr: every subset;
// otherwise the *item_count subsets listed in items.  Subsets that ask for the
// fallback optimizer are appended to fb_list (pass 0 only); BATCHED waves
// append subsets whose deformed window does not fit their LDS slot to ov_list.
struct PassArgs {
  int pass;
  const int *items;
  const int *item_count;
  int *fb_list, *fb_count;
  int *ov_list, *ov_count;
  int *counter;                // BATCHED work counters, kWorkRanges of them kCounterStride ints apart (zeroed before the launch)
  const float *nm_result;      // Nelder-Mead passes: per-subset results of simplex_groups_kernel, or null
};

// Differentiation path (forward gamma / adjoint) of one launch.
struct GammaArgs {
  const double *params;        // n x 6: u, v, theta, ex, ey, gxy
  const double *dgamma;        // adjoint: dL/dgamma, n
  double *gamma;               // forward: gamma, n
  double *dparams;             // adjoint: dL/dp, n x 6
  float *dimage;               // adjoint: dL/dI_def, W x H (added to; may be null)
  int adjoint;
};

// BATCHED work ranges (one per XCD of the MI355X).  Each range counter sits on
// a 128-byte line of its own: device-scope atomics on one line are served one
// after another, so eight counters on one line would serialize every work
// fetch of the chip (one per subset).
constexpr int kWorkRanges = 8;
constexpr int kCounterStride = 32;  // ints
constexpr int kWorkCounterInts = kWorkRanges * kCounterStride;
// Header of the work area: the range counters, then (on the next line) the
// fallback count and the overflow counts of pass 0 and pass 1.
constexpr int kWorkHeader = kWorkCounterInts + kCounterStride;
constexpr int kWorkFbCount = kWorkCounterInts, kWorkOv0Count = kWorkCounterInts + 1,
              kWorkOv1Count = kWorkCounterInts + 2;
constexpr int kWorkLaneCount = kWorkCounterInts + 3;  // work counter of simplex_groups_kernel
// count of the lean pass's hand-over list, which lives in the storage of the pass-1
// overflow list (pass 1 starts after pass 0 has finished with it)
constexpr int kWorkLeanCount = kWorkCounterInts + 4;

// Device ints of the work area of launch_correlate: the header, the fallback
// and overflow lists (n each), the Nelder-Mead results (kNmStride per subset).
inline size_t work_ints(int n) { return kWorkHeader + (3 + kNmStride) * (size_t)(n > 0 ? n : 0); }

// Enqueue the correlation of all subsets (primary pass, then the fallback pass
// when the optimization method has one).  storage: 0 f32, 1 u16.  work: device
// ints, at least work_ints(n) (counters, fallback and overflow lists).
// Field_Value_Initializer with USE_NEIGHBOR_VALUES (DICe_Initializer.cpp:398-434)
// for one round of the neighbour iteration: the guess of subset i from the
// fields of i's neighbour (NEIGHBOR_ID of start; itself when -1): from cur when
// the neighbour comes earlier in subset order (the reference has already
// correlated it in this frame), else from start (its frame-start values).
// prev_guess null (the first round): every subset goes to list.  Otherwise a
// subset whose guess moved by more than tol (any parameter), or whose sigma = -1
// flag or init status changed, takes the new guess and goes to list; the others
// keep prev_guess.  list / count (zeroed by the caller) receive the subsets to
// correlate again; fields of the listed subsets are then reset to start by
// launch_reset_listed.
hipError_t launch_neighbor_guess(const double *start, const double *cur, double *guess, const double *prev_guess,
                                 double tol, int *list, int *count, int n, hipStream_t stream);
hipError_t launch_reset_listed(const int *list, const int *count, const double *start, double *fields, int n,
                               hipStream_t stream);

// VSG_Strain_Post_Processor::execute (DICe_PostProcessor.cpp:931-1136) for
// every subset: the neighbours within the window radius r (squared distance
// below r^2 + 1e-5, formed in float as the reference's float kd-tree does),
// found through a uniform grid of cell side cell >= r (subsets sorted by cell,
// each cell a range of the sorted list), a least-squares plane of u and
// v over the valid ones (sigma >= 0), its gradient and the Green-Lagrange
// strains.  vsg: 7 x n doubles (strain xx, yy, xy, du/dx, du/dy, dv/dx, dv/dy);
// subsets whose own sigma is < 0, that have fewer than 3 valid neighbours, or
// whose normal matrix has rcond < 1e-12 get zeros and MATCH = -1 in fields.
struct VsgGrid {
  double x0, y0, cell;
  int ncx, ncy;
  const int *cell_begin, *cell_end;  // per cell: its range of the cell-ordered list
  const int *cell_ids;    // subset id per entry, ~id for failed subsets (SIGMA < 0)
  const float4 *rec;      // (x, y, u, v) of the entry's subset as float
  const int *cell_keys;   // the entry's cell (ascending)
};
// Scratch of the device-side grid build (n entries each unless noted).
struct VsgWork {
  int *keys, *ids, *keys_sorted, *ids_sorted;
  int *cell_range;  // 2 * ncells: begin, then end, of each cell
  float4 *rec;
  void *temp;       // rocPRIM scratch of vsg_grid_temp_bytes
  size_t temp_bytes;
};
// Bounds of the finite centres: bounds[0..4] = min x, min y, max x, max y,
// count of non-finite centres; bounds holds kVsgBoundsDoubles device doubles
// (the rest are partials).
constexpr int kVsgBoundsDoubles = 5 + 5 * 256;
hipError_t launch_vsg_bounds(const double *fields, int n, double *bounds, hipStream_t stream);
hipError_t vsg_grid_temp_bytes(int n, long long ncells, size_t *bytes);
// Bin the centres into g's cells (g.x0, y0, cell, ncx, ncy set) on the device:
// keys, a stable radix sort of (cell, id), cell ranges, packed records; sets
// g.cell_begin, g.cell_end, g.cell_ids, g.rec and g.cell_keys to w's buffers.
hipError_t launch_vsg_grid(const double *fields, int n, VsgGrid &g, const VsgWork &w, hipStream_t stream);
// knn > 0 (at most kVsgMaxK): the knn nearest subsets instead (this one included, ties in
// squared float distance broken by subset id), the grid of any cell side.
hipError_t launch_vsg_strain(const VsgGrid &g, float radius, int knn, double *fields, double *vsg, int n,
                             hipStream_t stream);
constexpr int kVsgFields = 7;
constexpr int kVsgMaxK = 64;

// Uncertainty_Post_Processor::execute (DICe_PostProcessor.cpp:1507-1537) for a
// 2-D analysis: out[0][i] = UNCERTAINTY = SIGMA / FIELD_1 (0 when FIELD_1 is 0
// or SIGMA < 0), out[1][i] = UNCERTAINTY_ANGLE = FIELD_1, in scalar_t (float).
hipError_t launch_uncertainty(const double *fields, double *out, int n, hipStream_t stream);

// One Objective_ZNSSD call for subset gid (call 0 gamma, 1 sigma, 2
// computeUpdateFast, 3 computeUpdateRobust) on one workgroup; io: 8 device
// doubles (parameters in / out, then two results).  computeUpdateFast writes the
// subset's CONDITION_NUMBER and uncertainty fields into fields, as the reference.
hipError_t launch_objective(int storage, const KParams &kp, const DImage &ref, const DImage &def, const DSubsets &ss,
                            double *fields, int gid, int call, double *io, hipStream_t stream);

// Enqueue the correlation of all subsets, or of the device list items (count
// *item_count) when items is not null.
hipError_t launch_correlate(const Launch &L, int storage, const KParams &kp, const DImage &ref,
                            const DImage &def, const DSubsets &subs, double *fields, int *work,
                            hipStream_t stream, const int *items = nullptr, const int *item_count = nullptr);

// Enqueue the forward gamma or the adjoint of every subset (BATCHED, then the
// PLAIN kernel over the subsets whose window did not fit).
hipError_t launch_gamma(const Launch &L, int storage, const KParams &kp, const DImage &ref, const DImage &def,
                        const DSubsets &subs, const GammaArgs &ga, int *work, hipStream_t stream);

// Enqueue the Nelder-Mead iterations of the items of pa, rows of 8 lanes per
// subset when max_pixels <= kNarrowSimplexMaxPixels, else 16 (subsets of at
// most max_pixels <= kLaneSimplexMaxPixels pixels): results to
// nm_out (kNmStride floats per subset); counter: one int, zeroed before the
// launch.
hipError_t launch_simplex_lanes(int storage, int interp, int max_pixels, const KParams &kp, const DImage &ref,
                                const DImage &def, const DSubsets &subs, const double *fields, const PassArgs &pa,
                                float *nm_out, int *counter, hipStream_t stream);

// Finite-difference gradients (+ optional 5x5 smoothing) of an image.
hipError_t launch_gradients(int storage, const void *pix, int w, int h, float *gx, float *gy, int smooth,
                            float *tmpx, float *tmpy, hipStream_t stream);

// Image_::gauss_filter: the mask_size Gauss mask (5, 7, 9, 11, 13) from in to out
// (different buffers), border band of mask_size / 2 + 1 pixels copied.
hipError_t launch_gauss_filter(int storage, const void *in, void *out, int w, int h, int mask_size,
                               hipStream_t stream);

// Zero a small int buffer with a kernel (graph-safe, no memset node).
hipError_t launch_zero_ints(int *p, int count, hipStream_t stream);

// Copy a double buffer with a kernel (for the per-frame history inside graphs).
hipError_t launch_copy_doubles(double *dst, const double *src, size_t count, hipStream_t stream);

}  // namespace dice_hip
