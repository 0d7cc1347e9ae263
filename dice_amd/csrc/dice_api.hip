// Host side of the C ABI declared in include/dice_hip.h.
//
// A context owns the device copies of the reference and deformed images, their
// gradients, the subset set, a device field store, the work buffers of the
// correlation passes, one HIP stream and (for multi-frame sequences) a
// hipGraph.  All entry points return a code and never throw.
#include "dice_hip.h"
#include "dice_api.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "dice_kernels.h"

using dice_hip::DImage;
using dice_hip::DSubsets;
using dice_hip::KParams;
using dice_hip::Launch;

// Work area of dice_hip_correlate_device_noalloc: a file-scope device array of
// work_ints(kNoAllocMaxSubsets) ints (11.5 MB), held by every GPU the library
// runs on until the process ends.  Host code never uses the array's name as an
// address: dice_hip_create resolves each device's copy once with
// hipGetSymbolAddress (that device current) and the entry point passes that
// pointer to the kernels.
static constexpr int kNoAllocMaxSubsets = 1 << 18;
static constexpr size_t kNoAllocWorkInts = dice_hip::kWorkHeader + (3 + dice_hip::kNmStride) * (size_t)kNoAllocMaxSubsets;
static __device__ __align__(256) int scratch[kNoAllocWorkInts];

namespace {

// dynamic LDS per workgroup of the batched kernel (default launch limit)
constexpr size_t kMaxLdsBytes = 64 * 1024;
// LDS slot per wave: deformed window, reused for the Gauss-Newton reduction
// (35 x 64 floats + totals); 10 KB lets 4 four-wave workgroups share a CU
constexpr int kDefaultSlotFloats = 2560;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), resolved once per device.
hipError_t scratch_address(int device, int **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<int *>(ptr);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

struct GraphKey {
  int mode, n_frames, n;
  const void *frames0;
  double *fields, *history;
  bool operator==(const GraphKey &o) const {
    return mode == o.mode && n_frames == o.n_frames && n == o.n && frames0 == o.frames0 && fields == o.fields &&
           history == o.history;
  }
};

}  // namespace

struct dice_hip_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  dice_params params{};
  bool have_params = false;
  int num_cus = 256;
  // reference image
  int rw = 0, rh = 0, rstorage = 0;
  DevBuf ref_pix, ref_gx, ref_gy;
  // deformed image
  int dw = 0, dh = 0, dstorage = 0;
  DevBuf def_pix;
  const void *def_ptr = nullptr;
  DevBuf def_gx, def_gy, tmp_x, tmp_y;
  DevBuf ref_raw, def_filt;  // upload target of a filtered reference image; the filtered deformed image
  DevBuf nb_start, nb_guess[2], nb_list, nb_counts;  // USE_NEIGHBOR_VALUES: frame-start fields, guesses,
                                                     // subsets to correlate again, their count per round
  int nb_rounds = 0;                         // correlation rounds of the last neighbour-initialized frame
  DevBuf vsg_cells, vsg_ids, vsg_out;        // VSG strain: subset grid, host-call output
  DevBuf vsg_keys, vsg_keys2, vsg_ids2, vsg_rec, vsg_temp, vsg_bounds;  // VSG grid build
  DevBuf obj_io;                             // dice_hip_objective: parameters and results
  bool def_has_grads = false;
  // subsets
  int n = 0, S = 0, fov = 0;
  // largest and 90th-percentile bounding-box side of the subsets (S for square subsets): the
  // autotuner sizes its LDS window candidates from them
  int max_side = 0, p90_side = 0;
  int max_count = 0;  // most pixels in one subset
  DevBuf cx, cy, off, px, py, act, bbox, nact, pxy, abox;
  DevBuf order;         // conformal: the work order (Morton order of the centres)
  bool packed = false;  // conformal pixels also packed into pxy
  // per-subset options: force simplex flags, blocked-pixel masks (box, word offset, bits)
  DevBuf force, mbox, moff, mbits;
  bool has_force = false, has_masks = false;
  // global coordinates of the images' (0, 0) pixels (Image_::offset_x / offset_y)
  int rox = 0, roy = 0, dox = 0, doy = 0;
  // fields + work
  DevBuf fields, work;
  // lean kernel: per-subset constants of the reference (kLeanCacheDoubles each),
  // current while their tag equals lean_gen; every change of parameters,
  // images, subsets or offsets starts a new generation (drop_graph)
  DevBuf lean_cache;
  int lean_gen = 1;
  int *scratch_work = nullptr;  // this device's copy of the static work area (no-allocation entry point)
  int *last_work = nullptr;     // work area of the last correlation launch
  // launch shape of the batched mode
  Launch batched{1, 4, 0, kDefaultSlotFloats};
  std::map<std::tuple<int, int, int, int, int>, Launch> tuned;
  // timing of correlation launches
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t events_used = 0;
  // graph
  hipGraphExec_t graph_exec = nullptr;
  GraphKey graph_key{};
  std::vector<const void *> graph_frames;

  int fail(int code, const char *what, hipError_t e = hipSuccess) {
    err = what;
    if (e != hipSuccess) {
      err += ": ";
      err += hipGetErrorString(e);
    }
    return code;
  }
};

namespace {

#define HIP_TRY(ctx, expr, what)                                    \
  do {                                                              \
    hipError_t _e = (expr);                                         \
    if (_e != hipSuccess) return (ctx)->fail(DICE_HIP_EHIP, what, _e); \
  } while (0)

bool params_valid(const dice_params *p, std::string &why) {
  if (p->interpolation_method != 0 && p->interpolation_method != 2) {
    why = "interpolation_method must be BILINEAR (0) or KEYS_FOURTH (2)";
    return false;
  }
  if (p->optimization_method < 0 || p->optimization_method > 4 || p->optimization_method == 2) {
    why = "optimization_method must be SIMPLEX, GRADIENT_BASED, SIMPLEX_THEN_GRADIENT_BASED or "
          "GRADIENT_BASED_THEN_SIMPLEX";
    return false;
  }
  if (!(p->neighbor_settle_tolerance >= 0.0)) {
    why = "neighbor_settle_tolerance must be >= 0";
    return false;
  }
  if (p->initialization_method < 0 || p->initialization_method > 2) {
    why = "initialization_method must be USE_FIELD_VALUES, USE_NEIGHBOR_VALUES or USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY";
    return false;
  }
  if (!p->enable_translation) {
    why = "enable_translation is required by the affine shape function";
    return false;
  }
  if (p->max_solver_iterations_fast < 0 || p->max_solver_iterations_robust < 1) {
    why = "iteration limits out of range";
    return false;
  }
  if (p->gauss_filter_images && p->gauss_filter_mask_size != 5 && p->gauss_filter_mask_size != 7 &&
      p->gauss_filter_mask_size != 9 && p->gauss_filter_mask_size != 11 && p->gauss_filter_mask_size != 13) {
    why = "the Gauss filter mask size is invalid (options include 5,7,9,11,13)";
    return false;
  }
  if (p->gradient_method != 0 && p->gradient_method != 1) {
    why = "gradient_method must be FINITE_DIFFERENCE or CONVOLUTION_5_POINT";
    return false;
  }
  return true;
}

KParams make_kparams(const dice_params &p, int first_frame) {
  KParams k{};
  k.interp = p.interpolation_method;
  k.opt = p.optimization_method;
  k.en_rot = p.enable_rotation != 0;
  k.en_normal = p.enable_normal_strain != 0;
  k.en_shear = p.enable_shear_strain != 0;
  k.max_fast = p.max_solver_iterations_fast;
  k.max_robust = p.max_solver_iterations_robust;
  k.fast_tol = (float)p.fast_solver_tolerance;
  k.robust_tol = p.robust_solver_tolerance;
  k.delta_disp = (float)p.robust_delta_disp;
  k.delta_theta = (float)p.robust_delta_theta;
  k.skip_thr = (float)p.skip_solve_gamma_threshold;
  k.init_gamma_thr = p.initial_gamma_threshold;
  k.final_gamma_thr = p.final_gamma_threshold;
  k.disp_jump_tol = p.disp_jump_tol;
  k.theta_jump_tol = p.theta_jump_tol;
  k.lm_alpha = p.levenberg_marquardt_regularization_factor > 0.0 ? (float)p.levenberg_marquardt_regularization_factor : 0.0f;
  k.momentum = (float)p.momentum_factor;
  k.def_grads = p.compute_def_gradients != 0;
  k.normalize_gamma = p.normalize_gamma_with_active_pixels != 0;
  k.uncertainty = p.compute_uncertainty_fields != 0;
  k.first_frame = first_frame != 0;
  k.small_w = 16;
  return k;
}

// The lanes of the small-subset objective order for the current subset set: the
// width that simplex_groups_kernel runs these subsets at.
int small_width(const dice_hip_context *ctx) {
  return ctx->max_count <= dice_hip::kNarrowSimplexMaxPixels ? 8 : 16;
}

size_t pix_bytes(int storage) { return storage == DICE_STORAGE_U16 ? 2 : 4; }

int check_ready(dice_hip_context *ctx) {
  if (!ctx->have_params) return ctx->fail(DICE_HIP_ESTATE, "parameters not set");
  if (!ctx->ref_pix.p) return ctx->fail(DICE_HIP_ESTATE, "reference image not set");
  if (!ctx->def_ptr) return ctx->fail(DICE_HIP_ESTATE, "deformed image not set");
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_ESTATE, "subsets not set");
  // with sub-images (dice_hip_set_image_offsets) the two images may cover different parts of the frame
  if (ctx->rstorage != ctx->dstorage) return ctx->fail(DICE_HIP_EINVAL, "reference and deformed images differ in storage");
  if ((ctx->rw != ctx->dw || ctx->rh != ctx->dh) && ctx->rox == 0 && ctx->roy == 0 && ctx->dox == 0 && ctx->doy == 0)
    return ctx->fail(DICE_HIP_EINVAL, "reference and deformed images differ in size");
  if (ctx->params.compute_def_gradients && !ctx->def_has_grads)
    return ctx->fail(DICE_HIP_ESTATE, "compute_def_gradients set after the deformed image; set the image again");
  return DICE_HIP_OK;
}

void fill_views(dice_hip_context *ctx, DImage &ref, DImage &def, DSubsets &ss) {
  ref = DImage{ctx->ref_pix.p, ctx->ref_gx.as<float>(), ctx->ref_gy.as<float>(), ctx->rw, ctx->rh, ctx->rox, ctx->roy};
  def = DImage{ctx->def_ptr, ctx->def_has_grads ? ctx->def_gx.as<float>() : nullptr,
               ctx->def_has_grads ? ctx->def_gy.as<float>() : nullptr, ctx->dw, ctx->dh, ctx->dox, ctx->doy};
  ss = DSubsets{ctx->n,
                ctx->S,
                ctx->fov,
                ctx->cx.as<int>(),
                ctx->cy.as<int>(),
                ctx->S > 0 ? nullptr : ctx->off.as<int>(),
                ctx->S > 0 ? nullptr : ctx->px.as<int>(),
                ctx->S > 0 ? nullptr : ctx->py.as<int>(),
                ctx->S > 0 ? nullptr : ctx->act.as<uint8_t>(),
                ctx->S > 0 ? nullptr : ctx->bbox.as<int4>(),
                ctx->S > 0 ? nullptr : ctx->nact.as<int>(),
                ctx->S > 0 || !ctx->packed ? nullptr : ctx->pxy.as<uint32_t>(),
                ctx->S > 0 ? nullptr : ctx->abox.as<int4>(),
                ctx->has_force ? ctx->force.as<uint8_t>() : nullptr,
                ctx->has_masks ? ctx->mbox.as<int4>() : nullptr,
                ctx->has_masks ? ctx->moff.as<int>() : nullptr,
                ctx->has_masks ? ctx->mbits.as<uint32_t>() : nullptr,
                ctx->S > 0 ? nullptr : ctx->order.as<int>()};
}

Launch current_launch(dice_hip_context *ctx, int mode) {
  if (mode == DICE_MODE_PLAIN) return Launch{0, 0, 0, 0};
  Launch L = ctx->batched;
  L.mode = 1;
  if (L.groups <= 0) {
    // persistent grid: enough workgroups to fill every CU at the occupancy the
    // kernel allows, capped by the work available
    int per_cu = 4;
    L.groups = ctx->num_cus * per_cu;
  }
  const int max_groups = (ctx->n + L.waves_per_group - 1) / L.waves_per_group;
  if (L.groups > max_groups) L.groups = max_groups > 0 ? max_groups : 1;
  L.lane_simplex = ctx->max_count <= dice_hip::kLaneSimplexMaxPixels ? std::max(ctx->max_count, 1) : 0;
  return L;
}

hipEvent_t *next_event_pair(dice_hip_context *ctx) {
  if (ctx->events_used >= ctx->events.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return nullptr;
    ctx->events.emplace_back(a, b);
  }
  return &ctx->events[ctx->events_used++].first;
}

int enqueue_correlate(dice_hip_context *ctx, int mode, double *d_fields, int first_frame, bool capturing,
                      const double *guess = nullptr, const int *items = nullptr, const int *item_count = nullptr,
                      int *work = nullptr) {
  DImage ref, def;
  DSubsets ss;
  fill_views(ctx, ref, def, ss);
  KParams kp = make_kparams(ctx->params, first_frame);
  kp.small_w = small_width(ctx);
  kp.guess = guess;
  kp.has_forced = ctx->has_force ? 1 : 0;
  // the cache holds one entry per subset of the current set (square subsets only)
  kp.lean_cache = ctx->S > 0 && ctx->lean_cache.p ? ctx->lean_cache.as<double>() : nullptr;
  kp.lean_gen = ctx->lean_gen;
  const Launch L = current_launch(ctx, mode);
  // per-launch timing events on the context stream (eager launches only; a captured
  // graph is timed as a whole by the caller)
  const bool timed = ctx->timing && !capturing && !work;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (timed) {
    hipEvent_t *p = next_event_pair(ctx);
    if (!p) return ctx->fail(DICE_HIP_EHIP, "hipEventCreate failed");
    ev = ctx->events[ctx->events_used - 1];
    HIP_TRY(ctx, hipEventRecord(ev.first, ctx->stream), "hipEventRecord");
  }
  ctx->last_work = work ? work : ctx->work.as<int>();
  HIP_TRY(ctx, dice_hip::launch_correlate(L, ctx->rstorage, kp, ref, def, ss, d_fields, ctx->last_work, ctx->stream,
                                          items, item_count),
          "correlation launch");
  if (timed) HIP_TRY(ctx, hipEventRecord(ev.second, ctx->stream), "hipEventRecord");
  return DICE_HIP_OK;
}

void drop_graph(dice_hip_context *ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  ctx->graph_exec = nullptr;
  ctx->graph_key = GraphKey{};
  // the configuration changes: the cached per-subset constants are stale
  ctx->lean_gen = ctx->lean_gen >= (1 << 30) ? 1 : ctx->lean_gen + 1;
}

int upload_image(dice_hip_context *ctx, DevBuf &buf, int w, int h, int storage, const void *host) {
  const size_t bytes = (size_t)w * h * pix_bytes(storage);
  HIP_TRY(ctx, buf.ensure(bytes), "hipMalloc image");
  HIP_TRY(ctx, hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy image");
  return DICE_HIP_OK;
}

int gradients(dice_hip_context *ctx, const void *pix, int w, int h, int storage, DevBuf &gx, DevBuf &gy) {
  const size_t n = (size_t)w * h;
  HIP_TRY(ctx, gx.ensure(n * 4), "hipMalloc gradients");
  HIP_TRY(ctx, gy.ensure(n * 4), "hipMalloc gradients");
  const int smooth = ctx->params.gradient_method == 1;
  if (smooth) {
    HIP_TRY(ctx, ctx->tmp_x.ensure(n * 4), "hipMalloc gradients");
    HIP_TRY(ctx, ctx->tmp_y.ensure(n * 4), "hipMalloc gradients");
  }
  HIP_TRY(ctx, dice_hip::launch_gradients(storage, pix, w, h, gx.as<float>(), gy.as<float>(), smooth,
                                           ctx->tmp_x.as<float>(), ctx->tmp_y.as<float>(), ctx->stream),
          "gradient kernel");
  return DICE_HIP_OK;
}

}  // namespace

extern "C" {

int dice_hip_create(dice_hip_context **out, int device) {
  if (!out) return DICE_HIP_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DICE_HIP_EHIP;
  if (device < 0 || device >= count) return DICE_HIP_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return DICE_HIP_EHIP;
  dice_hip_context *ctx = new dice_hip_context();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return DICE_HIP_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cus = prop.multiProcessorCount;
  if (scratch_address(device, &ctx->scratch_work) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DICE_HIP_EHIP;
  }
  dice_hip_default_params(&ctx->params);
  ctx->have_params = true;
  *out = ctx;
  return DICE_HIP_OK;
}

int dice_hip_destroy(dice_hip_context *ctx) {
  if (!ctx) return DICE_HIP_EINVAL;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  for (auto &e : ctx->events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (DevBuf *b : {&ctx->ref_pix, &ctx->ref_gx, &ctx->ref_gy, &ctx->def_pix, &ctx->def_gx, &ctx->def_gy, &ctx->tmp_x,
                    &ctx->tmp_y, &ctx->cx, &ctx->cy, &ctx->off, &ctx->px, &ctx->py, &ctx->act, &ctx->bbox, &ctx->nact,
                    &ctx->pxy, &ctx->abox, &ctx->ref_raw, &ctx->def_filt, &ctx->nb_start, &ctx->nb_guess[0],
                    &ctx->nb_guess[1], &ctx->nb_list, &ctx->nb_counts, &ctx->vsg_cells, &ctx->vsg_ids, &ctx->vsg_keys, &ctx->vsg_keys2, &ctx->vsg_ids2,
                    &ctx->vsg_rec, &ctx->vsg_temp, &ctx->vsg_bounds,
                    &ctx->vsg_out, &ctx->obj_io, &ctx->fields, &ctx->work, &ctx->force, &ctx->mbox, &ctx->moff,
                    &ctx->mbits, &ctx->lean_cache, &ctx->order})
    b->release();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return DICE_HIP_OK;
}

const char *dice_hip_last_error(const dice_hip_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void dice_hip_default_params(dice_params *p) {
  // dice_default_params, DICe_ParameterUtilities.cpp:360-420
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->interpolation_method = 2;
  p->optimization_method = 4;
  p->correlation_routine = 0;
  p->initialization_method = 0;
  p->enable_translation = 1;
  p->max_solver_iterations_fast = 25;
  p->max_solver_iterations_robust = 1000;
  p->fast_solver_tolerance = 1.0e-4;
  p->robust_solver_tolerance = 1.0e-6;
  p->robust_delta_disp = 1.0;
  p->robust_delta_theta = 0.1;
  p->skip_solve_gamma_threshold = 1.0e-10;
  p->initial_gamma_threshold = -1.0;
  p->final_gamma_threshold = -1.0;
  p->disp_jump_tol = 10000.0;
  p->theta_jump_tol = 100.0;
  p->levenberg_marquardt_regularization_factor = -1.0;
  p->momentum_factor = 0.0;
  p->compute_uncertainty_fields = 1;
  p->gauss_filter_images = 0;
  p->gauss_filter_mask_size = 7;
}

void dice_hip_tracking_default_params(dice_params *p) {
  // tracking_default_params, DICe_ParameterUtilities.cpp:297-358
  if (!p) return;
  dice_hip_default_params(p);
  p->correlation_routine = 1;
  p->optimization_method = 0;
  p->enable_rotation = 1;
  p->disp_jump_tol = 8.0;
  p->theta_jump_tol = 0.1;
}

int dice_hip_set_params(dice_hip_context *ctx, const dice_params *p) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!p) return ctx->fail(DICE_HIP_EINVAL, "null parameters");
  std::string why;
  if (!params_valid(p, why)) return ctx->fail(DICE_HIP_EINVAL, why.c_str());
  const bool regrad = ctx->params.gradient_method != p->gradient_method;
  ctx->params = *p;
  ctx->have_params = true;
  drop_graph(ctx);
  (void)hipSetDevice(ctx->device);
  if (regrad && ctx->ref_pix.p) {
    int rc = gradients(ctx, ctx->ref_pix.p, ctx->rw, ctx->rh, ctx->rstorage, ctx->ref_gx, ctx->ref_gy);
    if (rc) return rc;
  }
  return DICE_HIP_OK;
}

int dice_hip_set_ref_image(dice_hip_context *ctx, int width, int height, int storage, const void *host_pixels) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (width < 8 || height < 8 || !host_pixels || (storage != DICE_STORAGE_F32 && storage != DICE_STORAGE_U16))
    return ctx->fail(DICE_HIP_EINVAL, "bad reference image");
  (void)hipSetDevice(ctx->device);
  drop_graph(ctx);
  const bool filter = ctx->params.gauss_filter_images != 0;
  if (filter && (width < ctx->params.gauss_filter_mask_size || height < ctx->params.gauss_filter_mask_size))
    return ctx->fail(DICE_HIP_EINVAL, "image too small for gauss filtering with this mask size");
  int rc = upload_image(ctx, filter ? ctx->ref_raw : ctx->ref_pix, width, height, storage, host_pixels);
  if (rc) return rc;
  if (filter) {
    // images are filtered when they are built (DICe_Image.cpp:215-230), before their gradients
    HIP_TRY(ctx, ctx->ref_pix.ensure((size_t)width * height * pix_bytes(storage)), "hipMalloc image");
    HIP_TRY(ctx, dice_hip::launch_gauss_filter(storage, ctx->ref_raw.p, ctx->ref_pix.p, width, height,
                                               ctx->params.gauss_filter_mask_size, ctx->stream),
            "gauss filter");
  }
  ctx->rw = width;
  ctx->rh = height;
  ctx->rstorage = storage;
  // reference gradients are always computed (compute_ref_gradients, DICe_Schema.cpp:628,759-761)
  rc = gradients(ctx, ctx->ref_pix.p, width, height, storage, ctx->ref_gx, ctx->ref_gy);
  if (rc) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "set_ref_image");
  return DICE_HIP_OK;
}

static int set_def_common(dice_hip_context *ctx, int width, int height, int storage, const void *dptr) {
  if (ctx->params.gauss_filter_images) {
    const int m = ctx->params.gauss_filter_mask_size;
    if (width < m || height < m) return ctx->fail(DICE_HIP_EINVAL, "image too small for gauss filtering with this mask size");
    HIP_TRY(ctx, ctx->def_filt.ensure((size_t)width * height * pix_bytes(storage)), "hipMalloc image");
    HIP_TRY(ctx, dice_hip::launch_gauss_filter(storage, dptr, ctx->def_filt.p, width, height, m, ctx->stream),
            "gauss filter");
    dptr = ctx->def_filt.p;
  }
  ctx->dw = width;
  ctx->dh = height;
  ctx->dstorage = storage;
  ctx->def_ptr = dptr;
  ctx->def_has_grads = false;
  if (ctx->params.compute_def_gradients) {
    int rc = gradients(ctx, dptr, width, height, storage, ctx->def_gx, ctx->def_gy);
    if (rc) return rc;
    ctx->def_has_grads = true;
  }
  return DICE_HIP_OK;
}

int dice_hip_set_def_image(dice_hip_context *ctx, int width, int height, int storage, const void *host_pixels) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (width < 8 || height < 8 || !host_pixels || (storage != DICE_STORAGE_F32 && storage != DICE_STORAGE_U16))
    return ctx->fail(DICE_HIP_EINVAL, "bad deformed image");
  (void)hipSetDevice(ctx->device);
  int rc = upload_image(ctx, ctx->def_pix, width, height, storage, host_pixels);
  if (rc) return rc;
  rc = set_def_common(ctx, width, height, storage, ctx->def_pix.p);
  if (rc) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "set_def_image");
  return DICE_HIP_OK;
}

int dice_hip_set_def_image_device(dice_hip_context *ctx, int width, int height, int storage, const void *device_pixels) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (width < 8 || height < 8 || !device_pixels || (storage != DICE_STORAGE_F32 && storage != DICE_STORAGE_U16))
    return ctx->fail(DICE_HIP_EINVAL, "bad deformed image");
  (void)hipSetDevice(ctx->device);
  return set_def_common(ctx, width, height, storage, device_pixels);
}

int dice_hip_set_subsets(dice_hip_context *ctx, int n, const double *coord_x, const double *coord_y, int subset_size,
                         const int32_t *conf_offsets, const int32_t *conf_x, const int32_t *conf_y,
                         const uint8_t *conf_active) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (n < 1 || !coord_x || !coord_y) return ctx->fail(DICE_HIP_EINVAL, "bad subset coordinates");
  if (!conf_offsets && subset_size < 1) return ctx->fail(DICE_HIP_EINVAL, "subset_size must be >= 1");
  (void)hipSetDevice(ctx->device);
  drop_graph(ctx);
  std::vector<int> hx(n), hy(n);
  for (int i = 0; i < n; ++i) {
    // Objective ctor: static_cast<int_t> of the coordinate fields (DICe_Objective.h:88-92)
    hx[i] = (int)coord_x[i];
    hy[i] = (int)coord_y[i];
  }
  HIP_TRY(ctx, ctx->cx.ensure(n * sizeof(int)), "hipMalloc subsets");
  HIP_TRY(ctx, ctx->cy.ensure(n * sizeof(int)), "hipMalloc subsets");
  HIP_TRY(ctx, hipMemcpy(ctx->cx.p, hx.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy subsets");
  HIP_TRY(ctx, hipMemcpy(ctx->cy.p, hy.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy subsets");
  ctx->fov = subset_size > 0 ? subset_size : 0;
  if (!conf_offsets) {
    ctx->S = subset_size;
    ctx->max_side = ctx->p90_side = subset_size;
    ctx->max_count = (2 * (subset_size / 2) + 1) * (2 * (subset_size / 2) + 1);
  } else {
    if (!conf_x || !conf_y) return ctx->fail(DICE_HIP_EINVAL, "conformal subsets need pixel lists");
    const int total = conf_offsets[n];
    if (conf_offsets[0] != 0 || total < 1) return ctx->fail(DICE_HIP_EINVAL, "bad conformal offsets");
    std::vector<int4> bb(n), ab(n);
    std::vector<int> na(n), sides(n);
    // pixels packed into one word each (one load per pixel per pass) when every
    // coordinate fits 15 bits
    std::vector<uint32_t> pk(total);
    bool packed = true;
    for (int i = 0; i < n; ++i) {
      const int b = conf_offsets[i], e = conf_offsets[i + 1];
      if (e <= b) return ctx->fail(DICE_HIP_EINVAL, "empty conformal subset");
      int4 box{1 << 30, 1 << 30, -(1 << 30), -(1 << 30)};
      int4 abx = box;
      int cnt = 0;
      for (int k = b; k < e; ++k) {
        const int x = conf_x[k] < 0 ? 0 : conf_x[k], y = conf_y[k] < 0 ? 0 : conf_y[k];
        const bool on = conf_active ? (conf_active[k] != 0) : true;
        box.x = std::min(box.x, x);
        box.y = std::min(box.y, y);
        box.z = std::max(box.z, x);
        box.w = std::max(box.w, y);
        if (on) {
          abx.x = std::min(abx.x, x);
          abx.y = std::min(abx.y, y);
          abx.z = std::max(abx.z, x);
          abx.w = std::max(abx.w, y);
        }
        cnt += on ? 1 : 0;
        if (x >= (1 << 15) || y >= (1 << 15)) packed = false;
        pk[k] = (uint32_t)(x & 0x7fff) | ((uint32_t)(y & 0x7fff) << 15) | (on ? 0u : 1u << 30);
      }
      bb[i] = box;
      ab[i] = cnt > 0 ? abx : box;
      na[i] = cnt;
      sides[i] = std::max(box.z - box.x, box.w - box.y) + 1;
    }
    auto p90 = sides.begin() + (size_t)(0.9 * (n - 1));
    std::nth_element(sides.begin(), p90, sides.end());
    ctx->p90_side = *p90;
    ctx->max_side = *std::max_element(sides.begin(), sides.end());
    ctx->max_count = 0;
    for (int i = 0; i < n; ++i) ctx->max_count = std::max(ctx->max_count, conf_offsets[i + 1] - conf_offsets[i]);
    HIP_TRY(ctx, ctx->off.ensure((n + 1) * sizeof(int)), "hipMalloc subsets");
    HIP_TRY(ctx, ctx->px.ensure(total * sizeof(int)), "hipMalloc subsets");
    HIP_TRY(ctx, ctx->py.ensure(total * sizeof(int)), "hipMalloc subsets");
    HIP_TRY(ctx, ctx->act.ensure(total), "hipMalloc subsets");
    HIP_TRY(ctx, ctx->bbox.ensure(n * sizeof(int4)), "hipMalloc subsets");
    HIP_TRY(ctx, ctx->nact.ensure(n * sizeof(int)), "hipMalloc subsets");
    std::vector<uint8_t> a(total, 1);
    if (conf_active) std::memcpy(a.data(), conf_active, total);
    HIP_TRY(ctx, hipMemcpy(ctx->off.p, conf_offsets, (n + 1) * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(ctx, hipMemcpy(ctx->px.p, conf_x, total * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(ctx, hipMemcpy(ctx->py.p, conf_y, total * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(ctx, hipMemcpy(ctx->act.p, a.data(), total, hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(ctx, hipMemcpy(ctx->bbox.p, bb.data(), n * sizeof(int4), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(ctx, hipMemcpy(ctx->nact.p, na.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(ctx, ctx->abox.ensure(n * sizeof(int4)), "hipMalloc subsets");
    HIP_TRY(ctx, hipMemcpy(ctx->abox.p, ab.data(), n * sizeof(int4), hipMemcpyHostToDevice), "hipMemcpy");
    // work order: Morton order of the centres (ties by subset id)
    {
      auto spread = [](uint32_t v) {
        v &= 0xffffu;
        v = (v | (v << 8)) & 0x00ff00ffu;
        v = (v | (v << 4)) & 0x0f0f0f0fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
      };
      std::vector<std::pair<uint32_t, int>> key(n);
      for (int i = 0; i < n; ++i) {
        const uint32_t x = (uint32_t)std::min(std::max(hx[i], 0), 0xffff);
        const uint32_t y = (uint32_t)std::min(std::max(hy[i], 0), 0xffff);
        key[i] = {spread(x) | (spread(y) << 1), i};
      }
      std::sort(key.begin(), key.end());
      std::vector<int> ord(n);
      for (int i = 0; i < n; ++i) ord[i] = key[i].second;
      HIP_TRY(ctx, ctx->order.ensure(n * sizeof(int)), "hipMalloc subsets");
      HIP_TRY(ctx, hipMemcpy(ctx->order.p, ord.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    }
    ctx->packed = packed;
    if (packed) {
      HIP_TRY(ctx, ctx->pxy.ensure((size_t)total * sizeof(uint32_t)), "hipMalloc subsets");
      HIP_TRY(ctx, hipMemcpy(ctx->pxy.p, pk.data(), (size_t)total * sizeof(uint32_t), hipMemcpyHostToDevice),
              "hipMemcpy");
    }
    ctx->S = 0;
  }
  ctx->n = n;
  ctx->has_force = ctx->has_masks = false;  // per-subset options belong to the old subset set
  ctx->last_work = nullptr;  // the work area may move
  HIP_TRY(ctx, ctx->work.ensure(dice_hip::work_ints(n) * sizeof(int)), "hipMalloc work");
  HIP_TRY(ctx, ctx->fields.ensure((size_t)DICE_F_COUNT * n * sizeof(double)), "hipMalloc fields");
  // square subsets: the lean kernel's cache, zeroed (no entry carries a generation >= 1)
  const size_t cache_ints = (size_t)n * dice_hip::kLeanCacheDoubles * 2;
  if (ctx->S > 0 && cache_ints <= (size_t)INT_MAX) {
    HIP_TRY(ctx, ctx->lean_cache.ensure(cache_ints * sizeof(int)), "hipMalloc lean cache");
    HIP_TRY(ctx, dice_hip::launch_zero_ints(ctx->lean_cache.as<int>(), (int)cache_ints, ctx->stream), "zero lean cache");
  } else {
    ctx->lean_cache.release();
  }
  return DICE_HIP_OK;
}

int dice_hip_set_force_simplex(dice_hip_context *ctx, const uint8_t *force) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_ESTATE, "subsets not set");
  (void)hipSetDevice(ctx->device);
  ctx->has_force = false;
  if (!force) return DICE_HIP_OK;
  bool any = false;
  for (int i = 0; i < ctx->n && !any; ++i) any = force[i] != 0;
  if (!any) return DICE_HIP_OK;
  HIP_TRY(ctx, ctx->force.ensure((size_t)ctx->n), "hipMalloc force");
  HIP_TRY(ctx, hipMemcpy(ctx->force.p, force, (size_t)ctx->n, hipMemcpyHostToDevice), "hipMemcpy force");
  ctx->has_force = true;
  drop_graph(ctx);
  return DICE_HIP_OK;
}

int dice_hip_set_subset_masks(dice_hip_context *ctx, const int32_t *mask_offsets, const int32_t *mask_x,
                              const int32_t *mask_y) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_ESTATE, "subsets not set");
  (void)hipSetDevice(ctx->device);
  ctx->has_masks = false;
  drop_graph(ctx);
  if (!mask_offsets) return DICE_HIP_OK;
  if (!mask_x || !mask_y) return ctx->fail(DICE_HIP_EINVAL, "null mask coordinates");
  const int n = ctx->n;
  if (mask_offsets[0] != 0) return ctx->fail(DICE_HIP_EINVAL, "mask offsets must start at 0");
  for (int i = 0; i < n; ++i)
    if (mask_offsets[i + 1] < mask_offsets[i]) return ctx->fail(DICE_HIP_EINVAL, "mask offsets must not decrease");
  // per subset: the bbox of its blocked pixels and one bit per pixel of the box, row-major
  std::vector<int4> box(n);
  std::vector<int> woff(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const int b = mask_offsets[i], e = mask_offsets[i + 1];
    int x0 = 0, y0 = 0, x1 = -1, y1 = -1;
    for (int k = b; k < e; ++k) {
      if (k == b) { x0 = x1 = mask_x[k]; y0 = y1 = mask_y[k]; }
      x0 = std::min(x0, (int)mask_x[k]);
      x1 = std::max(x1, (int)mask_x[k]);
      y0 = std::min(y0, (int)mask_y[k]);
      y1 = std::max(y1, (int)mask_y[k]);
    }
    const long long w = e > b ? (long long)x1 - x0 + 1 : 0, h = e > b ? (long long)y1 - y0 + 1 : 0;
    if (w * h > (1ll << 28)) return ctx->fail(DICE_HIP_EINVAL, "a subset's mask spans too large a box");
    box[i] = make_int4(x0, y0, (int)w, (int)h);
    const long long words = (w * h + 31) / 32;
    if ((long long)woff[i] + words > (1ll << 30)) return ctx->fail(DICE_HIP_EINVAL, "masks too large");
    woff[i + 1] = woff[i] + (int)words;
  }
  std::vector<uint32_t> bits((size_t)std::max(woff[n], 1), 0u);
  for (int i = 0; i < n; ++i)
    for (int k = mask_offsets[i]; k < mask_offsets[i + 1]; ++k) {
      const unsigned bit = (unsigned)(mask_y[k] - box[i].y) * (unsigned)box[i].z + (unsigned)(mask_x[k] - box[i].x);
      bits[(size_t)woff[i] + (bit >> 5)] |= 1u << (bit & 31u);
    }
  HIP_TRY(ctx, ctx->mbox.ensure((size_t)n * sizeof(int4)), "hipMalloc masks");
  HIP_TRY(ctx, ctx->moff.ensure((size_t)(n + 1) * sizeof(int)), "hipMalloc masks");
  HIP_TRY(ctx, ctx->mbits.ensure(bits.size() * sizeof(uint32_t)), "hipMalloc masks");
  HIP_TRY(ctx, hipMemcpy(ctx->mbox.p, box.data(), (size_t)n * sizeof(int4), hipMemcpyHostToDevice), "hipMemcpy masks");
  HIP_TRY(ctx, hipMemcpy(ctx->moff.p, woff.data(), (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice),
          "hipMemcpy masks");
  HIP_TRY(ctx, hipMemcpy(ctx->mbits.p, bits.data(), bits.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
          "hipMemcpy masks");
  ctx->has_masks = mask_offsets[n] > 0;
  return DICE_HIP_OK;
}

int dice_hip_set_image_offsets(dice_hip_context *ctx, int ref_offset_x, int ref_offset_y, int def_offset_x,
                               int def_offset_y) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (ref_offset_x < 0 || ref_offset_y < 0 || def_offset_x < 0 || def_offset_y < 0)
    return ctx->fail(DICE_HIP_EINVAL, "image offsets must be >= 0");
  ctx->rox = ref_offset_x;
  ctx->roy = ref_offset_y;
  ctx->dox = def_offset_x;
  ctx->doy = def_offset_y;
  drop_graph(ctx);
  return DICE_HIP_OK;
}

static bool neighbor_mode(const dice_hip_context *ctx, int first_frame) {
  return ctx->params.initialization_method == 1 || (ctx->params.initialization_method == 2 && first_frame);
}

// USE_NEIGHBOR_VALUES.  The reference correlates the subsets one after another,
// each starting from the values of its neighbour (NEIGHBOR_ID_FS): this frame's
// result when the neighbour comes earlier in subset order, else its frame-start
// values.  Here round 0 correlates every subset at once from the guesses
// gathered from the frame-start fields; each later round gathers the guesses
// again from the latest results and correlates again only the subsets whose
// guess changed, each from its own frame-start fields.  A subset's result is a
// fixed function of its guess, so when a round finds no guess changed, every
// result is the one its current guess gives, and these are the sequential
// results: the one solution, found subset by subset in order.  Round r has
// every subset right whose chain of earlier neighbours is shorter than r, so
// the rounds end after at most n + 1.  With neighbor_settle_tolerance > 0 a
// guess counts as changed only when it moves by more than the tolerance.
// Rounds are enqueued kNeighborBatch at a time between two waits on the host.
static constexpr int kNeighborBatch = 16;

static int correlate_neighbors(dice_hip_context *ctx, int mode, double *d_fields, int first_frame) {
  const int n = ctx->n;
  const size_t per = (size_t)DICE_F_COUNT * n;
  std::vector<double> ids(n);
  HIP_TRY(ctx, hipMemcpyAsync(ids.data(), d_fields + (size_t)DICE_F_NEIGHBOR_ID * n, n * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy neighbor ids");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "neighbor ids");
  for (int i = 0; i < n; ++i)
    if (ids[i] != (double)(int)ids[i] || ids[i] < -1.0 || ids[i] >= n)
      return ctx->fail(DICE_HIP_EINVAL, "NEIGHBOR_ID must be -1 or the id of a subset");
  HIP_TRY(ctx, ctx->nb_start.ensure(per * sizeof(double)), "hipMalloc neighbor fields");
  HIP_TRY(ctx, ctx->nb_guess[0].ensure((size_t)dice_hip::kGuessSlots * n * sizeof(double)), "hipMalloc guesses");
  HIP_TRY(ctx, ctx->nb_guess[1].ensure((size_t)dice_hip::kGuessSlots * n * sizeof(double)), "hipMalloc guesses");
  HIP_TRY(ctx, ctx->nb_list.ensure((size_t)n * sizeof(int)), "hipMalloc neighbor list");
  HIP_TRY(ctx, ctx->nb_counts.ensure(kNeighborBatch * sizeof(int)), "hipMalloc neighbor counts");
  double *start = ctx->nb_start.as<double>();
  int *list = ctx->nb_list.as<int>(), *counts = ctx->nb_counts.as<int>();
  const double tol = ctx->params.neighbor_settle_tolerance;
  HIP_TRY(ctx, dice_hip::launch_copy_doubles(start, d_fields, per, ctx->stream), "copy fields");
  const double *prev = nullptr;
  int which = 0, round = 0;
  while (true) {
    HIP_TRY(ctx, dice_hip::launch_zero_ints(counts, kNeighborBatch, ctx->stream), "zero counts");
    for (int b = 0; b < kNeighborBatch; ++b) {
      double *guess = ctx->nb_guess[which].as<double>();
      HIP_TRY(ctx, dice_hip::launch_neighbor_guess(start, d_fields, guess, prev, tol, list, counts + b, n, ctx->stream),
              "neighbor guess");
      HIP_TRY(ctx, dice_hip::launch_reset_listed(list, counts + b, start, d_fields, n, ctx->stream), "reset fields");
      int rc = enqueue_correlate(ctx, mode, d_fields, first_frame, false, guess, list, counts + b);
      if (rc) return rc;
      prev = guess;
      which ^= 1;
    }
    int host_counts[kNeighborBatch];
    HIP_TRY(ctx, hipMemcpyAsync(host_counts, counts, sizeof host_counts, hipMemcpyDeviceToHost, ctx->stream),
            "hipMemcpy counts");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "neighbor rounds");
    for (int b = 0; b < kNeighborBatch; ++b)
      if (host_counts[b] == 0) {
        ctx->nb_rounds = round + b;  // rounds that correlated something
        return DICE_HIP_OK;
      }
    round += kNeighborBatch;
    if (round > n + 1 + kNeighborBatch) return ctx->fail(DICE_HIP_EHIP, "neighbor initialization did not settle");
  }
}

int dice_hip_correlate_device(dice_hip_context *ctx, int mode, double *device_fields, int first_frame) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (mode != DICE_MODE_PLAIN && mode != DICE_MODE_BATCHED) return ctx->fail(DICE_HIP_EINVAL, "unknown mode");
  if (!device_fields) return ctx->fail(DICE_HIP_EINVAL, "null fields");
  int rc = check_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  if (neighbor_mode(ctx, first_frame)) return correlate_neighbors(ctx, mode, device_fields, first_frame);
  ctx->nb_rounds = 0;
  return enqueue_correlate(ctx, mode, device_fields, first_frame, false);
}

int dice_hip_correlate_device_noalloc(dice_hip_context *ctx, int mode, double *device_fields, int first_frame) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (mode != DICE_MODE_PLAIN && mode != DICE_MODE_BATCHED) return ctx->fail(DICE_HIP_EINVAL, "unknown mode");
  if (!device_fields) return ctx->fail(DICE_HIP_EINVAL, "null fields");
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->n > kNoAllocMaxSubsets)
    return ctx->fail(DICE_HIP_EINVAL, "dice_hip_correlate_device_noalloc: more subsets than its static work area "
                                      "holds (262144)");
  if (neighbor_mode(ctx, first_frame))
    return ctx->fail(DICE_HIP_EINVAL, "dice_hip_correlate_device_noalloc: USE_NEIGHBOR_VALUES needs the buffers of "
                                      "dice_hip_correlate_device");
  if (!ctx->scratch_work) return ctx->fail(DICE_HIP_ESTATE, "static work area not resolved");
  (void)hipSetDevice(ctx->device);
  ctx->nb_rounds = 0;
  return enqueue_correlate(ctx, mode, device_fields, first_frame, false, nullptr, nullptr, nullptr, ctx->scratch_work);
}

int dice_hip_neighbor_rounds(dice_hip_context *ctx) { return ctx ? ctx->nb_rounds : DICE_HIP_EINVAL; }

int dice_hip_pass_counts(dice_hip_context *ctx, int *counts) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!counts) return ctx->fail(DICE_HIP_EINVAL, "null counts");
  counts[0] = counts[1] = counts[2] = 0;
  if (!ctx->last_work) return DICE_HIP_OK;
  (void)hipSetDevice(ctx->device);
  int h[dice_hip::kWorkHeader];
  HIP_TRY(ctx, hipMemcpyAsync(h, ctx->last_work, sizeof h, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy counts");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "pass counts");
  counts[0] = h[dice_hip::kWorkFbCount];
  counts[1] = h[dice_hip::kWorkOv0Count];
  counts[2] = h[dice_hip::kWorkOv1Count];
  return DICE_HIP_OK;
}

// VSG strains: the subsets binned on the device into a uniform grid (cells of
// side max(r, 1) px for radius windows, about k subsets each for k-nearest):
// the bounds of the centres (read back, 40 B, to size the grid), a cell key per
// subset, a stable radix sort of (cell, id), each cell's range and one packed
// record per subset, then one thread per subset on the context stream.  The
// bounds read-back and the final wait make this call synchronous.
static int vsg_enqueue(dice_hip_context *ctx, int window_size, double *d_fields, double *d_vsg) {
  if (window_size == 0) return ctx->fail(DICE_HIP_EINVAL, "strain_window_size_in_pixels must not be 0");
  if (window_size < -dice_hip::kVsgMaxK)
    return ctx->fail(DICE_HIP_EINVAL, "k-nearest VSG windows are limited to 64 neighbours");
  const int knn = window_size < 0 ? -window_size : 0;
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_EINVAL, "no subsets");
  const int n = ctx->n;
  // bounds of the centres on the device; only these five numbers come to the host
  HIP_TRY(ctx, ctx->vsg_bounds.ensure(dice_hip::kVsgBoundsDoubles * sizeof(double)), "hipMalloc VSG grid");
  HIP_TRY(ctx, dice_hip::launch_vsg_bounds(d_fields, n, ctx->vsg_bounds.as<double>(), ctx->stream), "VSG bounds");
  double b[5];
  HIP_TRY(ctx, hipMemcpyAsync(b, ctx->vsg_bounds.p, sizeof b, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy VSG bounds");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "VSG bounds");
  if (b[4] != 0.0) return ctx->fail(DICE_HIP_EINVAL, "non-finite coordinates");
  const double x0 = b[0], y0 = b[1], x1 = b[2], y1 = b[3];
  const float radius = knn ? 0.0f : (float)window_size / 2.0f;
  dice_hip::VsgGrid g{};
  g.x0 = x0;
  g.y0 = y0;
  // radius windows: cells of side >= r, so that every neighbour lies in the 3 x 3 cells around a
  // subset; k-nearest: cells holding about k subsets; at most 4000 cells along a side
  const double side = knn ? std::sqrt((x1 - x0 + 1.0) * (y1 - y0 + 1.0) * knn / n) : (double)radius;
  g.cell = std::max({side, 1.0, (x1 - x0) / 4000.0, (y1 - y0) / 4000.0});
  g.ncx = (int)std::floor((x1 - x0) / g.cell) + 1;
  g.ncy = (int)std::floor((y1 - y0) / g.cell) + 1;
  const long long ncells = (long long)g.ncx * g.ncy;
  size_t temp_bytes = 0;
  HIP_TRY(ctx, dice_hip::vsg_grid_temp_bytes(n, ncells, &temp_bytes), "VSG grid scratch size");
  const size_t nb = (size_t)n * sizeof(int);
  HIP_TRY(ctx, ctx->vsg_cells.ensure(2 * (size_t)ncells * sizeof(int)), "hipMalloc VSG grid");
  HIP_TRY(ctx, ctx->vsg_ids.ensure(nb), "hipMalloc VSG grid");
  HIP_TRY(ctx, ctx->vsg_ids2.ensure(nb), "hipMalloc VSG grid");
  HIP_TRY(ctx, ctx->vsg_keys.ensure(nb), "hipMalloc VSG grid");
  HIP_TRY(ctx, ctx->vsg_keys2.ensure(nb), "hipMalloc VSG grid");
  HIP_TRY(ctx, ctx->vsg_rec.ensure((size_t)n * sizeof(float4)), "hipMalloc VSG grid");
  HIP_TRY(ctx, ctx->vsg_temp.ensure(temp_bytes), "hipMalloc VSG grid");
  const dice_hip::VsgWork w{ctx->vsg_keys.as<int>(), ctx->vsg_ids.as<int>(), ctx->vsg_keys2.as<int>(),
                            ctx->vsg_ids2.as<int>(),  ctx->vsg_cells.as<int>(),   ctx->vsg_rec.as<float4>(),
                            ctx->vsg_temp.p,          temp_bytes};
  HIP_TRY(ctx, dice_hip::launch_vsg_grid(d_fields, n, g, w, ctx->stream), "VSG grid");
  HIP_TRY(ctx, dice_hip::launch_vsg_strain(g, radius, knn, d_fields, d_vsg, n, ctx->stream), "VSG strain kernel");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "VSG strain");
  return DICE_HIP_OK;
}

int dice_hip_vsg_strain_device(dice_hip_context *ctx, int window_size, double *device_fields, double *device_vsg) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!device_fields || !device_vsg) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  (void)hipSetDevice(ctx->device);
  return vsg_enqueue(ctx, window_size, device_fields, device_vsg);
}

int dice_hip_objective(dice_hip_context *ctx, int subset_id, int call, double *params, double *result) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!params || !result) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  if (call < DICE_OBJ_GAMMA || call > DICE_OBJ_UPDATE_ROBUST) return ctx->fail(DICE_HIP_EINVAL, "unknown call");
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (subset_id < 0 || subset_id >= ctx->n) return ctx->fail(DICE_HIP_EINVAL, "subset id out of range");
  (void)hipSetDevice(ctx->device);
  DImage ref, def;
  DSubsets ss;
  fill_views(ctx, ref, def, ss);
  KParams kp = make_kparams(ctx->params, 0);
  kp.small_w = small_width(ctx);
  HIP_TRY(ctx, ctx->obj_io.ensure(8 * sizeof(double)), "hipMalloc objective");
  double io[8] = {params[0], params[1], params[2], params[3], params[4], params[5], 0.0, 0.0};
  HIP_TRY(ctx, hipMemcpyAsync(ctx->obj_io.p, io, sizeof io, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy objective");
  HIP_TRY(ctx, dice_hip::launch_objective(ctx->rstorage, kp, ref, def, ss, ctx->fields.as<double>(), subset_id, call,
                                          ctx->obj_io.as<double>(), ctx->stream),
          "objective kernel");
  HIP_TRY(ctx, hipMemcpyAsync(io, ctx->obj_io.p, sizeof io, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy objective");
  double extra[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // CONDITION_NUMBER, FIELD_1..7 of the subset
  if (call == DICE_OBJ_UPDATE_FAST) {
    const int slots[8] = {DICE_F_CONDITION, DICE_F_FIELD_1, DICE_F_FIELD_2, DICE_F_FIELD_3,
                          DICE_F_FIELD_4,   DICE_F_FIELD_5, DICE_F_FIELD_6, DICE_F_FIELD_7};
    for (int k = 0; k < 8; ++k)
      HIP_TRY(ctx, hipMemcpyAsync(&extra[k], ctx->fields.as<double>() + (size_t)slots[k] * ctx->n + subset_id,
                                  sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
              "hipMemcpy objective fields");
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "objective");
  if (call >= DICE_OBJ_UPDATE_FAST)
    for (int k = 0; k < 6; ++k) params[k] = io[k];
  result[0] = io[6];
  result[1] = io[7];
  for (int k = 0; k < 8; ++k) result[2 + k] = extra[k];
  return DICE_HIP_OK;
}

int dice_hip_uncertainty_device(dice_hip_context *ctx, const double *device_fields, double *device_out) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!device_fields || !device_out) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_EINVAL, "no subsets");
  (void)hipSetDevice(ctx->device);
  HIP_TRY(ctx, dice_hip::launch_uncertainty(device_fields, device_out, ctx->n, ctx->stream), "uncertainty kernel");
  return DICE_HIP_OK;
}

int dice_hip_uncertainty(dice_hip_context *ctx, const double *fields, double *out) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!fields || !out) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_EINVAL, "no subsets");
  (void)hipSetDevice(ctx->device);
  const size_t bytes = (size_t)DICE_F_COUNT * ctx->n * sizeof(double), obytes = 2 * (size_t)ctx->n * sizeof(double);
  HIP_TRY(ctx, ctx->vsg_out.ensure(std::max(obytes, (size_t)dice_hip::kVsgFields * ctx->n * sizeof(double))),
          "hipMalloc output");
  HIP_TRY(ctx, hipMemcpyAsync(ctx->fields.p, fields, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy fields");
  HIP_TRY(ctx, dice_hip::launch_uncertainty(ctx->fields.as<double>(), ctx->vsg_out.as<double>(), ctx->n, ctx->stream),
          "uncertainty kernel");
  HIP_TRY(ctx, hipMemcpyAsync(out, ctx->vsg_out.p, obytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy output");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "uncertainty");
  return DICE_HIP_OK;
}

int dice_hip_vsg_strain(dice_hip_context *ctx, int window_size, double *fields, double *vsg) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!fields || !vsg) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  if (ctx->n <= 0) return ctx->fail(DICE_HIP_EINVAL, "no subsets");
  (void)hipSetDevice(ctx->device);
  const size_t bytes = (size_t)DICE_F_COUNT * ctx->n * sizeof(double);
  const size_t vbytes = (size_t)dice_hip::kVsgFields * ctx->n * sizeof(double);
  HIP_TRY(ctx, ctx->vsg_out.ensure(vbytes), "hipMalloc VSG fields");
  HIP_TRY(ctx, hipMemcpyAsync(ctx->fields.p, fields, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy fields");
  int rc = vsg_enqueue(ctx, window_size, ctx->fields.as<double>(), ctx->vsg_out.as<double>());
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(fields, ctx->fields.p, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy fields");
  HIP_TRY(ctx, hipMemcpyAsync(vsg, ctx->vsg_out.p, vbytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy VSG");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "VSG strain");
  return DICE_HIP_OK;
}

static int enqueue_gamma(dice_hip_context *ctx, int mode, const dice_hip::GammaArgs &ga) {
  if (mode != DICE_MODE_PLAIN && mode != DICE_MODE_BATCHED) return ctx->fail(DICE_HIP_EINVAL, "unknown mode");
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->params.compute_def_gradients)
    return ctx->fail(DICE_HIP_EINVAL, "the differentiation path interpolates the deformed image itself; "
                                      "compute_def_gradients must be 0");
  (void)hipSetDevice(ctx->device);
  DImage ref, def;
  DSubsets ss;
  fill_views(ctx, ref, def, ss);
  KParams kp = make_kparams(ctx->params, 0);
  kp.small_w = small_width(ctx);
  const Launch L = current_launch(ctx, mode);
  HIP_TRY(ctx, dice_hip::launch_gamma(L, ctx->rstorage, kp, ref, def, ss, ga, ctx->work.as<int>(), ctx->stream),
          "gamma launch");
  return DICE_HIP_OK;
}

int dice_hip_gamma(dice_hip_context *ctx, int mode, const double *device_params, double *device_gamma) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!device_params || !device_gamma) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  return enqueue_gamma(ctx, mode, dice_hip::GammaArgs{device_params, nullptr, device_gamma, nullptr, nullptr, 0});
}

int dice_hip_gamma_adjoint(dice_hip_context *ctx, int mode, const double *device_params,
                           const double *device_dgamma, double *device_dparams, float *device_dimage) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!device_params || !device_dgamma || !device_dparams) return ctx->fail(DICE_HIP_EINVAL, "null pointer");
  return enqueue_gamma(ctx, mode,
                       dice_hip::GammaArgs{device_params, device_dgamma, nullptr, device_dparams, device_dimage, 1});
}

int dice_hip_correlate(dice_hip_context *ctx, int mode, double *fields, int first_frame) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!fields) return ctx->fail(DICE_HIP_EINVAL, "null fields");
  int rc = check_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const size_t bytes = (size_t)DICE_F_COUNT * ctx->n * sizeof(double);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->fields.p, fields, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy fields");
  rc = dice_hip_correlate_device(ctx, mode, ctx->fields.as<double>(), first_frame);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(fields, ctx->fields.p, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy fields");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "correlate");
  return DICE_HIP_OK;
}

int dice_hip_correlate_sequence(dice_hip_context *ctx, int mode, int n_frames, const void *const *device_frames,
                                double *device_fields, double *device_history, int use_graph) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (n_frames < 1 || !device_frames || !device_fields) return ctx->fail(DICE_HIP_EINVAL, "bad sequence arguments");
  if (mode != DICE_MODE_PLAIN && mode != DICE_MODE_BATCHED) return ctx->fail(DICE_HIP_EINVAL, "unknown mode");
  (void)hipSetDevice(ctx->device);
  if (ctx->params.compute_def_gradients)
    return ctx->fail(DICE_HIP_EINVAL, "sequences need compute_def_gradients off (gradients are per frame)");
  if (ctx->params.initialization_method != 0)
    return ctx->fail(DICE_HIP_EINVAL, "sequences need USE_FIELD_VALUES (the neighbour rounds of USE_NEIGHBOR_VALUES "
                                      "wait on the host between launches)");
  if (!ctx->def_ptr) {
    int rc = set_def_common(ctx, ctx->rw, ctx->rh, ctx->rstorage, device_frames[0]);
    if (rc) return rc;
  }
  int rc = check_ready(ctx);
  if (rc) return rc;
  const size_t per_frame = (size_t)DICE_F_COUNT * ctx->n;
  auto enqueue_all = [&](bool capturing) -> int {
    for (int f = 0; f < n_frames; ++f) {
      ctx->def_ptr = device_frames[f];
      int r = enqueue_correlate(ctx, mode, device_fields, f == 0 ? 1 : 0, capturing);
      if (r) return r;
      if (device_history) {
        HIP_TRY(ctx, dice_hip::launch_copy_doubles(device_history + f * per_frame, device_fields, per_frame, ctx->stream),
                "history copy");
      }
    }
    return DICE_HIP_OK;
  };
  if (!use_graph) return enqueue_all(false);
  GraphKey key{mode, n_frames, ctx->n, device_frames[0], device_fields, device_history};
  std::vector<const void *> frames(device_frames, device_frames + n_frames);
  if (!(ctx->graph_exec && key == ctx->graph_key && frames == ctx->graph_frames)) {
    drop_graph(ctx);
    hipGraph_t graph = nullptr;
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    int r = enqueue_all(true);
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (r) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    HIP_TRY(ctx, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(ctx, e, "hipGraphInstantiate");
    ctx->graph_key = key;
    ctx->graph_frames = frames;
  }
  HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream), "hipGraphLaunch");
  return DICE_HIP_OK;
}

int dice_hip_synchronize(dice_hip_context *ctx) {
  if (!ctx) return DICE_HIP_EINVAL;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return DICE_HIP_OK;
}

void *dice_hip_stream(dice_hip_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int dice_hip_set_kernel_timing(dice_hip_context *ctx, int enable) {
  if (!ctx) return DICE_HIP_EINVAL;
  ctx->timing = enable != 0;
  ctx->events_used = 0;
  drop_graph(ctx);
  return DICE_HIP_OK;
}

int dice_hip_kernel_timing(dice_hip_context *ctx, double *avg_ms) {
  if (!ctx || !avg_ms) return DICE_HIP_EINVAL;
  (void)hipStreamSynchronize(ctx->stream);
  double sum = 0.0;
  int cnt = 0;
  for (size_t i = 0; i < ctx->events_used && i < ctx->events.size(); ++i) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->events[i].first, ctx->events[i].second) == hipSuccess) {
      sum += ms;
      ++cnt;
    }
  }
  *avg_ms = cnt ? sum / cnt : 0.0;
  ctx->events_used = 0;
  return cnt;
}

int dice_hip_get_launch_shape(dice_hip_context *ctx, int *waves_per_group, int *groups_per_cu, int *lds_floats) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (waves_per_group) *waves_per_group = ctx->batched.waves_per_group;
  if (groups_per_cu) *groups_per_cu = ctx->batched.groups > 0 ? ctx->batched.groups / ctx->num_cus : 0;
  if (lds_floats) *lds_floats = ctx->batched.lds_floats_per_wave;
  return DICE_HIP_OK;
}

int dice_hip_set_launch_shape(dice_hip_context *ctx, int waves_per_group, int groups_per_cu, int lds_floats) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (waves_per_group < 1 || waves_per_group > 4 || groups_per_cu < 0 || groups_per_cu > 64 || lds_floats < 0 ||
      (lds_floats & 3) != 0 || lds_floats < dice_hip::kMinSlotFloats ||
      (size_t)lds_floats * waves_per_group * 4 > kMaxLdsBytes)
    return ctx->fail(DICE_HIP_EINVAL, "launch shape out of range");
  ctx->batched = Launch{1, waves_per_group, groups_per_cu * ctx->num_cus, lds_floats};
  drop_graph(ctx);
  return DICE_HIP_OK;
}

int dice_hip_autotune(dice_hip_context *ctx, const double *device_fields, int first_frame) {
  if (!ctx) return DICE_HIP_EINVAL;
  if (!device_fields) return ctx->fail(DICE_HIP_EINVAL, "null fields");
  int rc = check_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const auto key = std::make_tuple(ctx->n, ctx->S > 0 ? ctx->S : -ctx->max_side, ctx->params.interpolation_method,
                                   ctx->params.optimization_method, ctx->rw * 65536 + ctx->rh);
  auto it = ctx->tuned.find(key);
  if (it != ctx->tuned.end()) {
    ctx->batched = it->second;
    drop_graph(ctx);
    return DICE_HIP_OK;
  }
  // scratch copy of the fields: the candidates run on it, the caller's store is untouched
  const size_t bytes = (size_t)DICE_F_COUNT * ctx->n * sizeof(double);
  double *scratch = nullptr;
  HIP_TRY(ctx, hipMalloc(&scratch, bytes), "hipMalloc autotune");
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const bool timing = ctx->timing;
  ctx->timing = false;
  Launch best = ctx->batched;
  float best_ms = 1e30f;
  // candidate LDS window budgets (floats per wave) cover the warped subset with margin: the
  // largest subset, and for uneven subsets also the 90th percentile (the rest overflow to
  // the plain kernel)
  std::vector<int> windows;
  for (int side : {ctx->max_side, ctx->p90_side}) {
    for (int extra : {8, 16, 28}) {
      const int w = std::max(kDefaultSlotFloats, ((side + extra) * (side + extra) + 3) & ~3);
      if ((size_t)w * 4 <= kMaxLdsBytes && std::find(windows.begin(), windows.end(), w) == windows.end())
        windows.push_back(w);
    }
    if (ctx->p90_side == ctx->max_side) break;
  }
  if (windows.empty()) windows.push_back(kDefaultSlotFloats);
  for (int waves : {1, 2, 4})
    for (int per_cu : {1, 2, 4, 8})
      for (int win : windows) {
        if ((size_t)waves * win * 4 > kMaxLdsBytes) continue;
        ctx->batched = Launch{1, waves, per_cu * ctx->num_cus, win};
        if (hipMemcpyAsync(scratch, device_fields, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) continue;
        (void)hipEventRecord(a, ctx->stream);
        rc = enqueue_correlate(ctx, DICE_MODE_BATCHED, scratch, first_frame, false);
        (void)hipEventRecord(b, ctx->stream);
        if (rc) break;
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best_ms) {
          best_ms = ms;
          best = ctx->batched;
        }
      }
  ctx->timing = timing;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(scratch);
  if (rc) return rc;
  ctx->batched = best;
  ctx->tuned[key] = best;
  drop_graph(ctx);
  return DICE_HIP_OK;
}

int dice_hip_gauss_filter(int device, int width, int height, int storage, const void *host_in, void *host_out,
                          int mask_size) {
  if (!host_in || !host_out || (storage != DICE_STORAGE_F32 && storage != DICE_STORAGE_U16)) return DICE_HIP_EINVAL;
  if (mask_size != 5 && mask_size != 7 && mask_size != 9 && mask_size != 11 && mask_size != 13)
    return DICE_HIP_EINVAL;
  if (width < mask_size || height < mask_size) return DICE_HIP_EINVAL;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return DICE_HIP_EHIP;
  if (hipSetDevice(device) != hipSuccess) return DICE_HIP_EHIP;
  const size_t bytes = (size_t)width * height * pix_bytes(storage);
  void *din = nullptr, *dout = nullptr;
  int rc = DICE_HIP_OK;
  if (hipMalloc(&din, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
      hipMemcpy(din, host_in, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      dice_hip::launch_gauss_filter(storage, din, dout, width, height, mask_size, nullptr) != hipSuccess ||
      hipMemcpy(host_out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    rc = DICE_HIP_EHIP;
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  return rc;
}

// dice_correlate (DICe_api.cpp:51-153): one static context for the process, created with the
// subsets and the reference image of the first call; later calls set the deformed image and correlate
// the same points.  The state is committed only once a call has gone through: a first call that
// fails leaves nothing behind, so the next call starts again.
static int correlate_points_impl(float *points, int n_points, int subset_size, const float *ref_img, int ref_w,
                                 int ref_h, const float *def_img, int def_w, int def_h, const dice_params *params,
                                 bool update_params) {
  static std::mutex m;
  static dice_hip_context *ctx = nullptr;
  static int n_init = 0;
  std::lock_guard<std::mutex> lock(m);
  if (ref_w != def_w || ref_h != def_h) return -1;
  if (subset_size < 1 || n_points < 1 || !points || !ref_img || !def_img) return -1;
  const int stride = 9;
  const bool first = ctx == nullptr;
  dice_hip_context *c = ctx;
  auto give_up = [&]() {
    if (first && c) dice_hip_destroy(c);
    return -1;
  };
  if (first) {
    if (dice_hip_create(&c, 0) != DICE_HIP_OK) return -1;
    std::vector<double> cx(n_points), cy(n_points);
    for (int i = 0; i < n_points; ++i) {
      cx[i] = points[i * stride + 0];
      cy[i] = points[i * stride + 1];
    }
    if (dice_hip_set_subsets(c, n_points, cx.data(), cy.data(), subset_size, nullptr, nullptr, nullptr, nullptr))
      return give_up();
  } else if (n_points != n_init) {
    return -1;
  }
  if (first || update_params) {
    dice_params p;
    if (params) p = *params;
    else dice_hip_tracking_default_params(&p);
    // set_def_image(w, h, array) builds the deformed image with compute_ref_gradients, which is on by
    // default (DICe_Schema.cpp:339, :628), so the API path always has deformed-image gradients
    p.compute_def_gradients = 1;
    if (dice_hip_set_params(c, &p) != DICE_HIP_OK) return give_up();
  }
  if (first && dice_hip_set_ref_image(c, ref_w, ref_h, DICE_STORAGE_F32, ref_img) != DICE_HIP_OK) return give_up();
  if (dice_hip_set_def_image(c, def_w, def_h, DICE_STORAGE_F32, def_img) != DICE_HIP_OK) return give_up();
  std::vector<double> f((size_t)DICE_F_COUNT * n_points, 0.0);
  auto F = [&](int slot, int i) -> double & { return f[(size_t)slot * n_points + i]; };
  for (int i = 0; i < n_points; ++i) {
    F(DICE_F_COORD_X, i) = points[i * stride + 0];
    F(DICE_F_COORD_Y, i) = points[i * stride + 1];
    F(DICE_F_DISP_X, i) = points[i * stride + 2];
    F(DICE_F_DISP_Y, i) = points[i * stride + 3];
    F(DICE_F_ROT_Z, i) = points[i * stride + 4];
    F(DICE_F_SIGMA, i) = points[i * stride + 5];
    F(DICE_F_GAMMA, i) = points[i * stride + 6];
    F(DICE_F_BETA, i) = points[i * stride + 7];
    F(DICE_F_STATUS, i) = points[i * stride + 8];
  }
  if (dice_hip_correlate(c, DICE_MODE_BATCHED, f.data(), first ? 1 : 0) != DICE_HIP_OK) return give_up();
  ctx = c;
  n_init = n_points;
  for (int i = 0; i < n_points; ++i) {
    points[i * stride + 2] = (float)F(DICE_F_DISP_X, i);
    points[i * stride + 3] = (float)F(DICE_F_DISP_Y, i);
    points[i * stride + 4] = (float)F(DICE_F_ROT_Z, i);
    points[i * stride + 5] = (float)F(DICE_F_SIGMA, i);
    points[i * stride + 6] = (float)F(DICE_F_GAMMA, i);
    points[i * stride + 7] = (float)F(DICE_F_BETA, i);
    points[i * stride + 8] = (float)F(DICE_F_STATUS, i);
  }
  return 0;
}

int dice_hip_correlate_points(float *points, int n_points, int subset_size, const float *ref_img, int ref_w, int ref_h,
                              const float *def_img, int def_w, int def_h, const dice_params *params,
                              int update_params) {
  return correlate_points_impl(points, n_points, subset_size, ref_img, ref_w, ref_h, def_img, def_w, def_h, params,
                               update_params != 0);
}

int dice_correlate(float points[], int n_points, int subset_size, float ref_img[], int ref_w, int ref_h,
                   float def_img[], int def_w, int def_h, const dice_params *input_params, bool update_params) {
  return correlate_points_impl(points, n_points, subset_size, ref_img, ref_w, ref_h, def_img, def_w, def_h,
                               input_params, update_params);
}

}  // extern "C"
