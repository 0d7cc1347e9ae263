/*
 * CPU oracle for the DICe local subset-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker: only the
 * test suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product library (libdice_hip.so) never links, loads or calls it.
 *
 * It is a from-scratch restatement of the reference algorithm (DICe, NTESS,
 * BSD-3) written from the reference's mathematics and control flow; every
 * function cites the reference file:line it follows.  It is templated on the
 * scalar type so that the same source gives the reference's default build
 * (scalar_t = float, FP64 Hessian accumulation) and its DICE_USE_DOUBLE build.
 *
 * What pins each part (tests/test_oracle.py):
 *   fd_gradients ........ analytic 255 cos(x/4pi) sin(y/4pi) within 1e-3 and the one-sided border
 *                         differences (DICe_TestImage.cpp:166-232)
 *   interp .............. baboon apply_transform golden, L2 < 0.1 in f32 and f64 (DICe_TestImage.cpp:236-266),
 *                         a planted sign flip fails it; bilinear exact on bilinear polynomials, Keys exact on
 *                         cubics; the border fall-through of keys_fourth_all (DICe_Image.cpp:851-857)
 *   subset_gamma ........ stripe closed form gamma = 0.4 k (DICe_TestSubsetGamma.cpp:76-117)
 *   correlate (GN) ...... parallel_subsets gold (49 subsets: u, v, gamma, iterations, status) and
 *                         single_speckled_shift gold (u, v, sigma) from tests/regression; the 6-parameter
 *                         known answers of DICe_TestCorrelateParams.cpp:78-140; sub-pixel shifts
 *                         (DICe_TestInterpShift.cpp:70-128); CONDITION against ||H||_F^2 / |det H| of the u,v
 *                         block built from the pinned gradients
 *   correlate (simplex) . DICe_TestCorrelateParams (SIMPLEX, SIMPLEX_THEN_GRADIENT_BASED) and the
 *                         dice_correlate integer shifts of DICe_TestAPI.cpp:70-150
 *   sigma / noise ....... SIGMA of the single_speckled_shift gold (2e-6)
 *   statuses ............ flat subsets, field of view, failed initial guess, max iterations + fallback
 *   polygon_pixels ...... conformal(_d).rawi golden raster (DICe_TestSubset.cpp:248-309)
 *   lu_inverse .......... numpy inverse; exactly singular input leaves the LU factors
 * and in the other test files:
 *   gauss_filter ........ the ten filtered-image goldens outFilter{5..13}(_d).rawi bit for bit
 *                         (tests/test_image_io.py)
 *   correlate (neighbour  parallel_subsets_seed and parallel_seed_first_step_only golds
 *   initialization) ..... (tests/test_host_cpp.py)
 *   vsg_strain .......... dic_challenge_12_sorted gold (VSG_STRAIN_XX/YY), closed forms on linear fields
 *                         (tests/test_host_cpp.py)
 *   objective, FIELD_1,   bitmap_star gold: u, v, sigma, gamma and UNCERTAINTY = sigma / FIELD_1
 *   uncertainty ......... (tests/test_uncertainty_gold.py)
 * parity unpinned: the uncertainty fields FIELD_2..7 (computeUncertaintyFields), Nelder-Mead
 * iteration counts and which stop rule fired, and the noise level beyond what sigma pins; they
 * follow the cited lines by review only.
 */
#ifndef DICE_ORACLE_H
#define DICE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as dice_params in include/dice_hip.h (checked by the tests). */
typedef struct oracle_params {
  int32_t interpolation_method;          /* BILINEAR=0 BICUBIC=1 KEYS_FOURTH=2 */
  int32_t optimization_method;           /* SIMPLEX=0 GRADIENT_BASED=1 GRADIENT_THEN_SEARCH=2
                                            SIMPLEX_THEN_GRADIENT_BASED=3 GRADIENT_BASED_THEN_SIMPLEX=4 */
  int32_t correlation_routine;           /* GENERIC_ROUTINE=0 TRACKING_ROUTINE=1 */
  int32_t initialization_method;         /* USE_FIELD_VALUES=0 USE_NEIGHBOR_VALUES=1 */
  int32_t enable_translation;
  int32_t enable_rotation;
  int32_t enable_normal_strain;
  int32_t enable_shear_strain;
  int32_t max_solver_iterations_fast;
  int32_t max_solver_iterations_robust;
  double fast_solver_tolerance;
  double robust_solver_tolerance;
  double robust_delta_disp;
  double robust_delta_theta;
  double skip_solve_gamma_threshold;
  double initial_gamma_threshold;
  double final_gamma_threshold;
  double disp_jump_tol;
  double theta_jump_tol;
  double levenberg_marquardt_regularization_factor;
  double momentum_factor;
  int32_t compute_def_gradients;
  int32_t normalize_gamma_with_active_pixels;
  int32_t compute_uncertainty_fields;
  int32_t gradient_method;               /* FINITE_DIFFERENCE=0 CONVOLUTION_5_POINT=1 */
  int32_t gauss_filter_images;           /* Gauss-filter both images first (DICe_Image.cpp:215-230) */
  int32_t gauss_filter_mask_size;        /* 5, 7, 9, 11, 13 */
  double neighbor_settle_tolerance;      /* not used: the oracle correlates neighbour chains in order */
} oracle_params;

/* Field slots of the per-subset SoA double store: fields[slot*n + i]. */
enum oracle_field {
  OF_COORD_X = 0, OF_COORD_Y, OF_DISP_X, OF_DISP_Y, OF_ROT_Z, OF_NS_XX, OF_NS_YY, OF_SS_XY,
  OF_SIGMA, OF_GAMMA, OF_BETA, OF_MATCH, OF_NOISE, OF_CONTRAST, OF_ACTIVE_PIXELS,
  OF_STATUS, OF_ITERATIONS, OF_CONDITION, OF_NEIGHBOR_ID,
  OF_FIELD_1, OF_FIELD_2, OF_FIELD_3, OF_FIELD_4, OF_FIELD_5, OF_FIELD_6, OF_FIELD_7,
  OF_COUNT
};

void oracle_default_params(oracle_params *p);          /* dice_default_params */
void oracle_tracking_default_params(oracle_params *p); /* tracking_default_params */

/* Image-level pieces (S = float for *_f, double for *_d). */
void oracle_fd_gradients_f(int w, int h, const float *img, float *gx, float *gy);
void oracle_fd_gradients_d(int w, int h, const double *img, double *gx, double *gy);
void oracle_smooth_gradients_f(int w, int h, float *gx, float *gy);
void oracle_interp_f(int method, int w, int h, const float *img, int n, const float *x, const float *y, float *out);
void oracle_interp_d(int method, int w, int h, const double *img, int n, const double *x, const double *y, double *out);
void oracle_apply_transform_d(int w, int h, const double *in, double *out, int cx, int cy,
                              double u, double v, double theta);
void oracle_apply_transform_f(int w, int h, const float *in, float *out, int cx, int cy,
                              float u, float v, float theta);
/* Un-warped subset gamma (Subset::initialize without a map, then Subset::gamma). */
double oracle_subset_gamma_f(int w, int h, const float *ref, const float *def, int cx, int cy, int size);
/* Polygon::get_owned_pixels: writes up to cap (x,y) pairs in (y,x) order; returns the count.
 * The winding angles in double (the reference's DICE_USE_DOUBLE build) or, _f, in float (its
 * default build); they differ on pixels that lie on an edge. */
int oracle_polygon_pixels(int nv, const int *vx, const int *vy, int cap, int *out_x, int *out_y);
int oracle_polygon_pixels_f(int nv, const int *vx, const int *vy, int cap, int *out_x, int *out_y);
/* Polygon::get_owned_pixels(shape_function, cx, cy, skin_factor) (DICe_Shape.cpp:136-219), the pixels
 * a blocking subset's polygon covers at its current shape function: the vertices mapped by the affine
 * parameters params[6] (u, v, theta, ex, ey, gxy) about (cx, cy) and rounded, stretched by
 * skin_factor about their integer centroid, then the float winding rule. */
int oracle_deformed_polygon_pixels(int nv, const int *vx, const int *vy, const double *params, int cx, int cy,
                                   double skin_factor, int cap, int *out_x, int *out_y);

/*
 * One Schema::execute_correlation (GENERIC_ROUTINE) over n subsets.
 *   storage: 0 = float32 pixels, 1 = uint16 pixels, 2 = float64 pixels (double oracle only).
 *   square subsets when conf_offsets == NULL (size = subset_size), otherwise conformal pixel lists
 *   (pixels of subset i are conf_x/conf_y/conf_active[conf_offsets[i] .. conf_offsets[i+1])).
 *   fields: OF_COUNT*n doubles, read (coordinates, initial guess, sigma) and written.
 *   first_frame: 1 when frame_id == first_frame_id (enables the field-of-view check).
 *   threads: OpenMP threads (>=1) over an even contiguous split of the subsets.
 * Returns 0, or -1 on a bad argument.
 */
int oracle_correlate_f(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int n, int subset_size, const int *conf_offsets, const int *conf_x,
                       const int *conf_y, const unsigned char *conf_active, double *fields,
                       int first_frame, int threads);
int oracle_correlate_d(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int n, int subset_size, const int *conf_offsets, const int *conf_x,
                       const int *conf_y, const unsigned char *conf_active, double *fields,
                       int first_frame, int threads);

/* The same with per-subset options: force[i] != 0 runs Nelder-Mead only on subset i (Schema::set_force_simplex,
 * and the subsets in BLOCKING_SUBSETS lists, DICe_Schema.cpp:2160-2185); mask_x/mask_y[mask_off[i] ..
 * mask_off[i+1]) are image pixels that deactivate a warped pixel of subset i rounding onto them
 * (Subset::obstructed_coords_ and pixels_blocked_by_other_subsets_, DICe_Subset.cpp:162-167,412-419,447-460).
 * Either may be null. */
int oracle_correlate_ex_f(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                          int n, int subset_size, const int *conf_offsets, const int *conf_x, const int *conf_y,
                          const unsigned char *conf_active, double *fields, int first_frame, int threads,
                          const unsigned char *force, const int *mask_off, const int *mask_x, const int *mask_y);
int oracle_correlate_ex_d(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                          int n, int subset_size, const int *conf_offsets, const int *conf_x, const int *conf_y,
                          const unsigned char *conf_active, double *fields, int first_frame, int threads,
                          const unsigned char *force, const int *mask_off, const int *mask_x, const int *mask_y);

/* Image_::gauss_filter in place (mask 5, 7, 9, 11 or 13): float arithmetic over float32
 * (storage 0) or uint16 (storage 1) pixels, or double arithmetic over double pixels.
 * Returns 0, or -1 on a bad mask or an image smaller than the mask. */
int oracle_gauss_filter_f(int w, int h, int storage, void *img, int mask_size);
int oracle_gauss_filter_d(int w, int h, double *img, int mask_size);

/* VSG_Strain_Post_Processor over n subsets: fields as for oracle_correlate (COORD, DISP, SIGMA
 * read; MATCH set to -1 where the strain fails); vsg: 7 * n doubles (strain xx, yy, xy, du/dx,
 * du/dy, dv/dx, dv/dy).  window > 0: radius window / 2; window < 0: the -window nearest subsets
 * (ties by id).  Returns 0 or -1. */
int oracle_vsg_strain(int n, double *fields, int window, double *vsg);

/* One Objective_ZNSSD call on the square subset of side subset_size centred at (cx, cy):
 * call 0 gamma, 1 sigma, 2 computeUpdateFast, 3 computeUpdateRobust.  params: 6 doubles
 * (u, v, theta, ex, ey, gxy), updated by the updates.  result: 10 doubles (gamma | sigma |
 * status, noise | iterations, then CONDITION_NUMBER and FIELD_1..7 after computeUpdateFast).
 * _f: float arithmetic on float32 (storage 0) or uint16 (1) pixels; _d: double arithmetic on
 * float64 (storage 2) or uint16 pixels.  Returns 0 or -1. */
int oracle_objective_f(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int subset_size, int cx, int cy, int call, double *params, double *result);
int oracle_objective_d(const oracle_params *p, int w, int h, int storage, const void *ref, const void *def,
                       int subset_size, int cx, int cy, int call, double *params, double *result);

/* Uncertainty_Post_Processor (2-D): out[i] = SIGMA / FIELD_1 (0 if FIELD_1 == 0 or SIGMA < 0),
 * out[n + i] = FIELD_1, in float.  Returns 0 or -1. */
int oracle_uncertainty(int n, const double *fields, double *out);

/* Simplex::minimize (DICe_Simplex.cpp:60-277) on a test objective of nd <= 6 variables:
 * kind 0: sum_i coef[i] (p_i - coef[nd + i])^2 (a quadratic bowl), kind 1: sum_i (p_i^2 - 1)^2
 * (a double well).  p: start point in, the last evaluated point out.  trace (may be null): up to
 * cap evaluations as (point, value) rows of nd + 1 doubles, in evaluation order; n_evals: all
 * evaluations; num_iterations; stop: bits of the rules that ended it (1 gamma(p0) below the
 * threshold, 2 rtol, 4 rtol difference, 8 repeated gamma, 16 iteration cap).  Returns the
 * Status_Flag (0 success, 11 max iterations) or -1 on a bad argument. */
int oracle_nelder_mead(int kind, int nd, double *p, const double *deltas, const double *coef, int max_iterations,
                       double tolerance, double threshold, int cap, double *trace, int *n_evals, int *num_iterations,
                       int *stop);

/* 6x6 LU with partial pivoting + inverse (GETRF/GETRI semantics); returns INFO. */
int oracle_lu_inverse(int n, double *a_rowmajor);

#ifdef __cplusplus
}
#endif
#endif
