/*
 * dice_hip.h -- C ABI of the MI355X (gfx950) subset-correlation library.
 *
 * This is the drop-in boundary for DICe's local subset-correlation path.
 * Each entry point names the reference interface it replaces.  Plain C types
 * only: pointers, sizes, a POD parameter struct.  No exceptions cross it.
 *
 * Conventions
 *  - Return codes: DICE_HIP_OK (0) or a negative DICE_HIP_E* code; the text of
 *    the last error is available from dice_hip_last_error().
 *  - Host arrays are copied; the caller keeps ownership.  *_device entry points
 *    take device pointers that must stay valid until the work is synchronized.
 *  - Status values written to the STATUS field are the reference's Status_Flag
 *    numbers (DICe.h:609-677); enum values of dice_params match DICe.h.
 *  - One context per host thread; a context is not thread-safe.  All work of a
 *    context runs in order on its own HIP stream.
 */
#ifndef DICE_HIP_H
#define DICE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DICE_HIP_OK 0
#define DICE_HIP_EINVAL (-1)      /* bad argument / unsupported option */
#define DICE_HIP_EHIP (-2)        /* HIP runtime error */
#define DICE_HIP_ESTATE (-3)      /* call order: e.g. no image or subsets yet */

/* Storage of image pixels (DICe.h:105-112 storage_t). */
#define DICE_STORAGE_F32 0
#define DICE_STORAGE_U16 1

/* Execution modes of dice_hip_correlate*. */
#define DICE_MODE_PLAIN 0    /* one subset per 256-thread workgroup, global-memory gathers */
#define DICE_MODE_BATCHED 1  /* persistent work-stealing waves, LDS-staged deformed windows */

/* Correlation parameters: the subset of DICe's correlation parameter list
 * (DICe.h:144-397, defaults DICe_ParameterUtilities.cpp:297-420) that the
 * local subset-correlation path reads.  Enum-valued members use DICe.h values. */
typedef struct dice_params {
  int32_t interpolation_method;          /* BILINEAR=0, KEYS_FOURTH=2 (BICUBIC=1 rejected) */
  int32_t optimization_method;           /* SIMPLEX=0, GRADIENT_BASED=1, SIMPLEX_THEN_GRADIENT_BASED=3,
                                            GRADIENT_BASED_THEN_SIMPLEX=4 (GRADIENT_THEN_SEARCH=2 rejected) */
  int32_t correlation_routine;           /* GENERIC_ROUTINE=0, TRACKING_ROUTINE=1 */
  int32_t initialization_method;         /* USE_FIELD_VALUES=0 USE_NEIGHBOR_VALUES=1
                                            USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY=2 */
  int32_t enable_translation;            /* must be 1 */
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
  double initial_gamma_threshold;        /* -1 = off */
  double final_gamma_threshold;          /* -1 = off */
  double disp_jump_tol;
  double theta_jump_tol;
  double levenberg_marquardt_regularization_factor; /* <= 0 = off */
  double momentum_factor;                /* 0 = off */
  int32_t compute_def_gradients;         /* deformed image carries FD gradients: GN uses them */
  int32_t normalize_gamma_with_active_pixels;
  int32_t compute_uncertainty_fields;    /* FIELD_1..7 on converged Gauss-Newton subsets */
  int32_t gradient_method;               /* FINITE_DIFFERENCE=0, CONVOLUTION_5_POINT=1 */
  int32_t gauss_filter_images;           /* Gauss-filter both images on upload (DICe_Image.cpp:215-230) */
  int32_t gauss_filter_mask_size;        /* 5, 7, 9, 11 or 13 (default 7) */
  double neighbor_settle_tolerance;      /* USE_NEIGHBOR_VALUES on this library: 0 (default) gives the
                                            reference's subset-by-subset results; > 0 stops the rounds of
                                            dice_hip_correlate once no initial guess moves by more than this
                                            (px for u, v; rad / strain for the others).  No counterpart in
                                            the reference. */
} dice_params;

/* Per-subset field store: SoA of doubles, fields[slot * n + i]
 * (the Schema's MultiField per Field_Spec, DICe_Schema.cpp:1363-1440). */
enum dice_field {
  DICE_F_COORD_X = 0, DICE_F_COORD_Y, DICE_F_DISP_X, DICE_F_DISP_Y, DICE_F_ROT_Z,
  DICE_F_NS_XX, DICE_F_NS_YY, DICE_F_SS_XY, DICE_F_SIGMA, DICE_F_GAMMA, DICE_F_BETA,
  DICE_F_MATCH, DICE_F_NOISE, DICE_F_CONTRAST, DICE_F_ACTIVE_PIXELS, DICE_F_STATUS,
  DICE_F_ITERATIONS, DICE_F_CONDITION, DICE_F_NEIGHBOR_ID,
  DICE_F_FIELD_1, DICE_F_FIELD_2, DICE_F_FIELD_3, DICE_F_FIELD_4, DICE_F_FIELD_5,
  DICE_F_FIELD_6, DICE_F_FIELD_7,
  DICE_F_COUNT
};

typedef struct dice_hip_context dice_hip_context;

/* Context lifetime.  Replaces the state the reference keeps in a static
 * DICe::Schema (DICe_api.cpp:97-115): device buffers, stream, graphs. */
int dice_hip_create(dice_hip_context **ctx, int device);
int dice_hip_destroy(dice_hip_context *ctx);
const char *dice_hip_last_error(const dice_hip_context *ctx);

/* Default parameter sets: dice_default_params / tracking_default_params
 * (DICe_ParameterUtilities.cpp:360-420 / :297-358). */
void dice_hip_default_params(dice_params *p);
void dice_hip_tracking_default_params(dice_params *p);

/* Schema::set_params (DICe_Schema.cpp:508-883) for the options above. */
int dice_hip_set_params(dice_hip_context *ctx, const dice_params *p);

/* Schema::set_ref_image(w, h, ArrayRCP) (DICe_Schema.cpp:395-445): uploads the
 * pixels and computes the reference gradients on the device
 * (Image_::compute_gradients, DICe_Image.cpp:1013-1094). */
int dice_hip_set_ref_image(dice_hip_context *ctx, int width, int height, int storage,
                           const void *host_pixels);

/* Schema::set_def_image(w, h, ArrayRCP) (DICe_Schema.cpp:327-348): uploads the
 * deformed pixels; gradients only when compute_def_gradients is set. */
int dice_hip_set_def_image(dice_hip_context *ctx, int width, int height, int storage,
                           const void *host_pixels);

/* Same, for pixels already in device memory (no copy; the pointer is used by
 * later correlate calls and must stay valid).  Gradients, if requested, go to
 * context-owned buffers; with gauss_filter_images the filtered copy does too. */
int dice_hip_set_def_image_device(dice_hip_context *ctx, int width, int height, int storage,
                                  const void *device_pixels);

/* Image_::gauss_filter (DICe_Image.cpp:1194-1240) on device `device`: the
 * mask_size x mask_size Gauss mask (5, 7, 9, 11 or 13) over the pixels at least
 * mask_size / 2 + 1 from every edge, the others copied; the result is cast to
 * the storage type (truncated for uint16).  host_in / host_out: width x height
 * pixels of `storage` (may be the same buffer).  Synchronous.  The contexts
 * apply the same filter to both images when gauss_filter_images is set. */
int dice_hip_gauss_filter(int device, int width, int height, int storage, const void *host_in, void *host_out,
                          int mask_size);

/* The subset set of the Schema (DICe_Schema.cpp:141-150 + Subset ctors,
 * DICe_Subset.cpp:78-168).  Square subsets of side subset_size centred on
 * (coord_x[i], coord_y[i]) when conf_offsets is NULL.  Otherwise conformal
 * subsets: pixels of subset i are conf_x/conf_y[conf_offsets[i] ..
 * conf_offsets[i+1]) in the reference's (y, x)-sorted order, with per-pixel
 * active flags (excluded areas = 0; NULL = all active).  Centroids are the
 * coordinates.  The coordinates are also written to the COORD fields. */
int dice_hip_set_subsets(dice_hip_context *ctx, int n, const double *coord_x, const double *coord_y,
                         int subset_size, const int32_t *conf_offsets, const int32_t *conf_x,
                         const int32_t *conf_y, const uint8_t *conf_active);

/* Per-subset Nelder-Mead override of the current subset set: generic_correlation_routine forces
 * computeUpdateRobust, with no fallback optimizer, on the subsets of Schema::set_force_simplex
 * (FORCE_SIMPLEX in a subset file) and on every subset that blocks or is blocked by another
 * (BLOCKING_SUBSETS; DICe_Schema.cpp:2160-2185, 2197-2204, 2226-2229).  force: n flags, 1 = forced;
 * NULL clears.  Cleared by dice_hip_set_subsets. */
int dice_hip_set_force_simplex(dice_hip_context *ctx, const uint8_t *force);

/* Image pixels that block the current subsets: the union of Subset::obstructed_coords_ (the
 * pixels of a conformal subset's OBSTRUCTED shapes, DICe_Subset.cpp:162-167) and
 * pixels_blocked_by_other_subsets_ (Schema::check_for_blocking_subsets, DICe_Schema.cpp:4529-4556).
 * A warped subset pixel whose rounded position is one of subset i's blocked pixels is deactivated
 * for that evaluation (Subset::initialize, DICe_Subset.cpp:412-419; is_obstructed_pixel :447-460).
 * Pixels of subset i: (mask_x, mask_y)[mask_offsets[i] .. mask_offsets[i+1]) in global image
 * coordinates; mask_offsets NULL clears.  Cleared by dice_hip_set_subsets. */
int dice_hip_set_subset_masks(dice_hip_context *ctx, const int32_t *mask_offsets, const int32_t *mask_x,
                              const int32_t *mask_y);

/* Image_::offset_x / offset_y (DICe_Image.h:168-176): the global coordinates of the (0, 0) pixel
 * of the reference and deformed images when they hold a sub-region of the frame (the sub-image
 * extents of Schema::update_extents, DICe_Schema.cpp:886-985).  Subset coordinates and fields
 * stay global; every gather subtracts the offset, and the 4-px guard, the field-of-view test and
 * the noise window use the sub-image bounds, as in Subset::initialize (DICe_Subset.cpp:380-411).
 * Default 0, 0, 0, 0. */
int dice_hip_set_image_offsets(dice_hip_context *ctx, int ref_offset_x, int ref_offset_y, int def_offset_x,
                               int def_offset_y);

/* Schema::execute_correlation (DICe_Schema.cpp:1622-1792, GENERIC_ROUTINE with
 * generic_correlation_routine :2025-2368 per subset).  fields: host SoA,
 * DICE_F_COUNT * n doubles; read (displacements, rotation, strains, SIGMA as
 * the initial guess of Field_Value_Initializer, DICe_Initializer.cpp:398-434;
 * NEIGHBOR_ID with USE_NEIGHBOR_VALUES) and written (record_step /
 * record_failed_step, :1958-2009).  first_frame = 1 enables the field-of-view
 * test that the reference applies when frame_id == first_frame_id.
 * USE_NEIGHBOR_VALUES (and USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY on a first
 * frame) gives the results of the reference's subset-by-subset order, in which
 * a subset starts from its neighbour's result of this frame when the neighbour
 * has a lower id (else from the neighbour's frame-start values, or its own when
 * NEIGHBOR_ID is -1); NEIGHBOR_ID must be -1 or a subset id.  Synchronous. */
int dice_hip_correlate(dice_hip_context *ctx, int mode, double *fields, int first_frame);

/* Same on a device-resident field store (DICE_F_COUNT * n doubles); enqueued on
 * the context stream and returns without waiting, except with neighbour
 * initialization, whose rounds wait on the stream between launches. */
int dice_hip_correlate_device(dice_hip_context *ctx, int mode, double *device_fields, int first_frame);

/* dice_hip_correlate_device without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory at this call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The correlation passes' work area (work-stealing counters,
 * fallback and overflow lists, Nelder-Mead results) is a static device array
 * of the library, one per GPU, sized for at most 262,144 subsets (11.5 MB, held
 * until the process ends); the call makes no hipMalloc / hipFree.  The counters
 * are cleared by a kernel at each call.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (several contexts or streams) race.
 * More subsets, or USE_NEIGHBOR_VALUES initialization (which needs the
 * context's own buffers), give DICE_HIP_EINVAL.  Enqueued on the context
 * stream; results are those of dice_hip_correlate_device. */
int dice_hip_correlate_device_noalloc(dice_hip_context *ctx, int mode, double *device_fields, int first_frame);

/* VSG_Strain_Post_Processor::execute (DICe_PostProcessor.cpp:877-1136): for each
 * subset, a least-squares plane of u and v over the subsets within
 * window_size / 2 px (window_size > 0), or over its -window_size nearest subsets
 * (window_size < 0, at most 64; ties in distance go to the lower subset id,
 * where the reference's kd-tree has an order of its own), whose SIGMA >= 0, and
 * from its gradient the Green-Lagrange strains.  vsg: DICE_VSG_COUNT * n doubles, slot-major in the
 * order of enum dice_vsg_field.  Subsets that fail (own SIGMA < 0, fewer than 3
 * valid neighbours, a singular normal matrix) get zeros and MATCH = -1 in
 * fields, as in the reference.  dice_hip_vsg_strain takes host arrays and is
 * synchronous; the _device form takes device arrays on the context stream and
 * waits for the kernel.  The neighbour grid is built on the device; a non-finite
 * centre coordinate is an error (DICE_HIP_EINVAL). */
enum dice_vsg_field {
  DICE_VSG_STRAIN_XX = 0, DICE_VSG_STRAIN_YY, DICE_VSG_STRAIN_XY,
  DICE_VSG_DUDX, DICE_VSG_DUDY, DICE_VSG_DVDX, DICE_VSG_DVDY, DICE_VSG_COUNT
};
int dice_hip_vsg_strain(dice_hip_context *ctx, int window_size, double *fields, double *vsg);
int dice_hip_vsg_strain_device(dice_hip_context *ctx, int window_size, double *device_fields, double *device_vsg);

/* Objective_ZNSSD(schema, subset_id) calls for one subset (DICe_Objective.h:82-242) on
 * the context's images and subsets:
 *   DICE_OBJ_GAMMA          Objective::gamma(sf) (DICe_Objective.cpp:58-76): result[0]
 *   DICE_OBJ_SIGMA          Objective::sigma(sf, noise) (:150-190): result[0] sigma,
 *                           result[1] noise
 *   DICE_OBJ_UPDATE_FAST    computeUpdateFast(sf, iters) (:430-657): params updated,
 *                           result[0] Status_Flag, result[1] iterations, result[2]
 *                           CONDITION_NUMBER and result[3..9] FIELD_1..7 of the subset
 *                           (the fields it writes in the reference)
 *   DICE_OBJ_UPDATE_ROBUST  computeUpdateRobust(sf, iters) (:369-388): params updated,
 *                           result[0] Status_Flag, result[1] iterations
 * params: 6 doubles (u, v, theta, ex, ey, gxy; disabled ones read as 0).  result: 10
 * doubles.  Synchronous; one workgroup, for callers that drive subsets one by one as
 * the reference's routines do (dice_hip_correlate runs whole frames). */
enum dice_objective_call { DICE_OBJ_GAMMA = 0, DICE_OBJ_SIGMA, DICE_OBJ_UPDATE_FAST, DICE_OBJ_UPDATE_ROBUST };
int dice_hip_objective(dice_hip_context *ctx, int subset_id, int call, double *params, double *result);

/* Uncertainty_Post_Processor::execute (DICe_PostProcessor.cpp:1494-1537), 2-D
 * analysis: out[i] = UNCERTAINTY = SIGMA / FIELD_1 (FIELD_1 is the cosine of the
 * angle of computeUncertaintyFields; 0 when it is 0 or when SIGMA < 0) and
 * out[n + i] = UNCERTAINTY_ANGLE = FIELD_1, as scalar_t (float) values.
 * dice_hip_uncertainty takes host arrays and is synchronous; the _device form is
 * enqueued on the context stream. */
int dice_hip_uncertainty(dice_hip_context *ctx, const double *fields, double *out);
int dice_hip_uncertainty_device(dice_hip_context *ctx, const double *device_fields, double *device_out);

/* Correlation rounds the last neighbour-initialized frame took (0 after a
 * USE_FIELD_VALUES frame).  No counterpart in the reference, which correlates
 * the subsets in order, once each: a diagnostic of this library's rounds. */
int dice_hip_neighbor_rounds(dice_hip_context *ctx);

/* Work-list sizes of the last correlation launch of the context (waits for the
 * stream): counts[0] subsets handed to the fallback optimizer pass
 * (GRADIENT_BASED_THEN_SIMPLEX / SIMPLEX_THEN_GRADIENT_BASED), counts[1] and
 * counts[2] subsets whose deformed window outgrew the BATCHED wave's LDS slot in
 * the primary and the fallback pass and were solved by the PLAIN kernel.  No
 * counterpart in the reference: a diagnostic of this library's launches. */
int dice_hip_pass_counts(dice_hip_context *ctx, int *counts);

/* A multi-frame sequence: the CLI frame loop of execute_correlation calls
 * (DICe_Main.cpp:324-374) for n_frames deformed frames already in device
 * memory, each initialised from the previous frame's fields (USE_FIELD_VALUES
 * only).  When
 * use_graph != 0 the whole sequence is captured once into a hipGraph and
 * replayed.  device_history (optional, may be NULL): n_frames * DICE_F_COUNT * n
 * doubles receiving each frame's fields.  Asynchronous. */
int dice_hip_correlate_sequence(dice_hip_context *ctx, int mode, int n_frames, const void *const *device_frames,
                                double *device_fields, double *device_history, int use_graph);

/* Differentiation path.  No counterpart entry point in the reference: the
 * objective is the one Objective::gamma evaluates (DICe_Objective.cpp:58-76,
 * Subset::initialize + Subset::gamma, DICe_Subset.cpp:323-440) for the current
 * images and subsets, at given affine parameters, without a solve:
 * gamma_i = 2 - 2 rho_i, the ZNSSD of the kept pixels (no clamp, no
 * normalisation; -1 for a flat subset).  device_params: n x 6 doubles
 * (u, v, theta, ex, ey, gxy of subset i at [6 i, 6 i + 6)).  mode: PLAIN or
 * BATCHED.  compute_def_gradients must be 0.  Enqueued on the context stream.
 *
 * dice_hip_gamma writes gamma_i (n doubles). */
int dice_hip_gamma(dice_hip_context *ctx, int mode, const double *device_params, double *device_gamma);

/* Adjoint of dice_hip_gamma for L = sum_i w_i gamma_i with w = device_dgamma
 * (n doubles): writes dL/dp (n x 6 doubles) and ADDS dL/dI_def to
 * device_dimage (width x height floats of the deformed image, row-major; NULL
 * to skip).  Interpolation weights and their derivatives are recomputed, not
 * stored.  Flat subsets contribute 0. */
int dice_hip_gamma_adjoint(dice_hip_context *ctx, int mode, const double *device_params,
                           const double *device_dgamma, double *device_dparams, float *device_dimage);

/* Waits for all work of the context. */
int dice_hip_synchronize(dice_hip_context *ctx);

/* The context's HIP stream (hipStream_t), for callers that time or order work. */
void *dice_hip_stream(dice_hip_context *ctx);

/* Average duration (ms) of the correlation launches (primary + fallback pass)
 * enqueued eagerly since the last call, measured with HIP events recorded on
 * the context stream around each launch (launches replayed from a graph are
 * not instrumented); resets the counters.  Returns the number measured. */
int dice_hip_kernel_timing(dice_hip_context *ctx, double *avg_ms);
int dice_hip_set_kernel_timing(dice_hip_context *ctx, int enable);

/* Autotuner for DICE_MODE_BATCHED: times candidate launch shapes (waves per
 * workgroup, workgroups per CU, LDS window budget) on the current images,
 * subsets and fields (without modifying fields) and keeps the fastest for
 * this problem signature.  No counterpart in the reference (CPU only). */
int dice_hip_autotune(dice_hip_context *ctx, const double *device_fields, int first_frame);
int dice_hip_get_launch_shape(dice_hip_context *ctx, int *waves_per_group, int *groups_per_cu,
                              int *lds_floats_per_wave);
int dice_hip_set_launch_shape(dice_hip_context *ctx, int waves_per_group, int groups_per_cu,
                              int lds_floats_per_wave);

/* dice_correlate (include/dice_api.h) with an int flag: (DICe_api.h:129-133, DICe_api.cpp:51-153): stride-9
 * points [x, y, u, v, theta, sigma, gamma, beta, status], float images of equal
 * size, a static context created on the first call with the given parameters
 * (NULL = tracking defaults) and kept between calls; the parameters are
 * applied again only when update_params != 0.  Returns 0, or -1 on a size
 * mismatch or bad argument, as the reference does. */
int dice_hip_correlate_points(float *points, int n_points, int subset_size, const float *ref_img, int ref_w,
                              int ref_h, const float *def_img, int def_w, int def_h, const dice_params *params,
                              int update_params);

#ifdef __cplusplus
}
#endif
#endif
