// C++ host mirror of the reference's Schema interface (see dice_schema.hpp).
#include "dice_schema.hpp"

#include <algorithm>
#include <cmath>
#include <set>

namespace dice_hip {

namespace {

// Parameter names of DICe.h:155-285 that this path reads, with their slot in
// dice_params; plus the two switches of this library.
enum Kind { I32, F64 };
struct Slot {
  Kind kind;
  size_t offset;
};
const std::map<std::string, Slot> &slots() {
  static const std::map<std::string, Slot> m = {
      {"interpolation_method", {I32, offsetof(dice_params, interpolation_method)}},
      {"optimization_method", {I32, offsetof(dice_params, optimization_method)}},
      {"correlation_routine", {I32, offsetof(dice_params, correlation_routine)}},
      {"initialization_method", {I32, offsetof(dice_params, initialization_method)}},
      {"enable_translation", {I32, offsetof(dice_params, enable_translation)}},
      {"enable_rotation", {I32, offsetof(dice_params, enable_rotation)}},
      {"enable_normal_strain", {I32, offsetof(dice_params, enable_normal_strain)}},
      {"enable_shear_strain", {I32, offsetof(dice_params, enable_shear_strain)}},
      {"max_solver_iterations_fast", {I32, offsetof(dice_params, max_solver_iterations_fast)}},
      {"max_solver_iterations_robust", {I32, offsetof(dice_params, max_solver_iterations_robust)}},
      {"fast_solver_tolerance", {F64, offsetof(dice_params, fast_solver_tolerance)}},
      {"robust_solver_tolerance", {F64, offsetof(dice_params, robust_solver_tolerance)}},
      {"robust_delta_disp", {F64, offsetof(dice_params, robust_delta_disp)}},
      {"robust_delta_theta", {F64, offsetof(dice_params, robust_delta_theta)}},
      {"skip_solve_gamma_threshold", {F64, offsetof(dice_params, skip_solve_gamma_threshold)}},
      {"initial_gamma_threshold", {F64, offsetof(dice_params, initial_gamma_threshold)}},
      {"final_gamma_threshold", {F64, offsetof(dice_params, final_gamma_threshold)}},
      {"disp_jump_tol", {F64, offsetof(dice_params, disp_jump_tol)}},
      {"theta_jump_tol", {F64, offsetof(dice_params, theta_jump_tol)}},
      {"levenberg_marquardt_regularization_factor",
       {F64, offsetof(dice_params, levenberg_marquardt_regularization_factor)}},
      {"momentum_factor", {F64, offsetof(dice_params, momentum_factor)}},
      {"compute_def_gradients", {I32, offsetof(dice_params, compute_def_gradients)}},
      {"normalize_gamma_with_active_pixels", {I32, offsetof(dice_params, normalize_gamma_with_active_pixels)}},
      {"compute_uncertainty_fields", {I32, offsetof(dice_params, compute_uncertainty_fields)}},
      {"gradient_method", {I32, offsetof(dice_params, gradient_method)}},
  };
  return m;
}
const char *kBackend = "correlation_backend";

void to_list(const dice_params &p, ParameterList *out) {
  for (const auto &kv : slots()) {
    const char *base = reinterpret_cast<const char *>(&p) + kv.second.offset;
    const double v = kv.second.kind == I32 ? (double)*reinterpret_cast<const int32_t *>(base)
                                           : *reinterpret_cast<const double *>(base);
    out->set(kv.first, v);
  }
  out->set(kBackend, HIP_BATCHED);
}

// atan2 difference wrapped to [-pi, pi] (angle_2d, DICe_Shape.cpp:88-103).
double angle_2d(double x1, double y1, double x2, double y2) {
  const double pi = 3.14159265358979323846;
  double d = std::atan2(y2, x2) - std::atan2(y1, x1);
  while (d > pi) d -= 2.0 * pi;
  while (d < -pi) d += 2.0 * pi;
  return d;
}

}  // namespace

ParameterList &ParameterList::set(const std::string &name, double value) {
  if (name != kBackend && slots().count(name) == 0)
    throw std::invalid_argument("Error, unknown correlation parameter: " + name);
  values_[name] = value;
  return *this;
}

double ParameterList::get(const std::string &name) const {
  auto it = values_.find(name);
  if (it == values_.end()) throw std::invalid_argument("Error, parameter not set: " + name);
  return it->second;
}

void dice_default_params(ParameterList *params) {
  dice_params p;
  dice_hip_default_params(&p);
  to_list(p, params);
}

void tracking_default_params(ParameterList *params) {
  dice_params p;
  dice_hip_tracking_default_params(&p);
  to_list(p, params);
}

Polygon::Polygon(const std::vector<int> &coords_x, const std::vector<int> &coords_y) : vx_(coords_x), vy_(coords_y) {
  if (vx_.size() != vy_.size()) throw std::invalid_argument("Error, polygon x and y vertex counts differ");
  if (vx_.size() < 3) throw std::invalid_argument("Error, a polygon needs at least three vertices");
  min_x_ = max_x_ = vx_[0];
  min_y_ = max_y_ = vy_[0];
  for (size_t i = 0; i < vx_.size(); ++i) {
    min_x_ = std::min(min_x_, vx_[i]);
    max_x_ = std::max(max_x_, vx_[i]);
    min_y_ = std::min(min_y_, vy_[i]);
    max_y_ = std::max(max_y_, vy_[i]);
  }
  vx_.push_back(vx_[0]);
  vy_.push_back(vy_[0]);
}

bool Polygon::inside(int x, int y) const {
  double angle = 0.0;
  for (size_t i = 0; i + 1 < vx_.size(); ++i)
    angle += angle_2d(vx_[i] - x, vy_[i] - y, vx_[i + 1] - x, vy_[i + 1] - y);
  return std::abs(angle) >= 3.14159265358979323846;
}

std::vector<std::pair<int, int>> Polygon::get_owned_pixels() const {
  std::vector<std::pair<int, int>> out;
  for (int y = min_y_; y <= max_y_; ++y)
    for (int x = min_x_; x <= max_x_; ++x)
      if (inside(x, y)) out.emplace_back(y, x);
  return out;
}

void Polygon::deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const {
  for (int j = 0; j < size; ++j)
    if (inside(x[j], y[j])) flags[j] = 0;
}

Schema::Schema(const ParameterList *params, int device) {
  create(device);
  if (params) set_params(*params);
}

Schema::Schema(int roi_width, int roi_height, int step_size_x, int step_size_y, int subset_size,
               const ParameterList *params, int device) {
  if (subset_size <= 0) throw std::runtime_error("Error: width cannot be equal to or less than zero.");
  if (step_size_x <= 0) throw std::runtime_error("Error, step size x is <= 0");
  if (step_size_y <= 0) throw std::runtime_error("Error, step size y is <= 0");
  // a border of one subset along every edge, first centre at subset_size - 1
  const int nx = (roi_width - 2 * subset_size) / step_size_x + 1;
  const int ny = (roi_height - 2 * subset_size) / step_size_y + 1;
  if (nx <= 0) throw std::runtime_error("Error, numPointsX <= 0.");
  if (ny <= 0) throw std::runtime_error("Error, numPointsY <= 0.");
  std::vector<double> cx((size_t)nx * ny), cy((size_t)nx * ny);
  for (int i = 0; i < nx * ny; ++i) {
    const int yi = i / nx, xi = i - yi * nx;
    cx[i] = subset_size + xi * step_size_x - 1;
    cy[i] = subset_size + yi * step_size_y - 1;
  }
  create(device);
  if (params) set_params(*params);
  init_subsets(cx, cy, subset_size, nullptr);
}

Schema::Schema(const std::vector<double> &coords_x, const std::vector<double> &coords_y, int subset_size,
               const std::map<int, Conformal_Area_Def> *conformal_subset_defs, const ParameterList *params,
               int device) {
  if (coords_x.size() != coords_y.size()) throw std::runtime_error("Error, coordinate arrays differ in size");
  create(device);
  if (params) set_params(*params);
  init_subsets(coords_x, coords_y, subset_size, conformal_subset_defs);
}

Schema::~Schema() {
  if (ctx_) dice_hip_destroy(ctx_);
}

void Schema::create(int device) {
  dice_hip_default_params(&p_);
  const int rc = dice_hip_create(&ctx_, device);
  if (rc != DICE_HIP_OK) {
    ctx_ = nullptr;
    throw std::runtime_error("Error, cannot create the HIP correlation context on device " + std::to_string(device));
  }
  check(dice_hip_set_params(ctx_, &p_), "set_params");
}

void Schema::check(int rc, const char *what) const {
  if (rc == DICE_HIP_OK) return;
  const std::string msg = std::string("Error in ") + what + ": " + (ctx_ ? dice_hip_last_error(ctx_) : "no context");
  if (rc == DICE_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

void Schema::set_params(const ParameterList &params) {
  // parameters that are not in the list keep their defaults (dice_default_params)
  dice_params p;
  dice_hip_default_params(&p);
  for (const auto &kv : params.values()) {
    if (kv.first == kBackend) continue;
    const Slot &s = slots().at(kv.first);
    char *base = reinterpret_cast<char *>(&p) + s.offset;
    if (s.kind == I32) *reinterpret_cast<int32_t *>(base) = (int32_t)kv.second;
    else *reinterpret_cast<double *>(base) = kv.second;
  }
  Correlation_Backend backend = HIP_BATCHED;
  if (params.isParameter(kBackend)) {
    const int b = (int)params.get(kBackend);
    if (b != HIP_BATCHED && b != HIP_PLAIN) throw std::invalid_argument("Error, unknown correlation_backend");
    backend = (Correlation_Backend)b;
  }
  check(dice_hip_set_params(ctx_, &p), "set_params");
  p_ = p;
  backend_ = backend;
}

void Schema::init_subsets(const std::vector<double> &cx, const std::vector<double> &cy, int subset_size,
                          const std::map<int, Conformal_Area_Def> *defs) {
  const int n = (int)cx.size();
  if (n <= 0) throw std::runtime_error("Error, no subsets");
  n_ = n;
  subset_dim_ = subset_size;
  fields_.assign((size_t)DICE_F_COUNT * n, 0.0);
  if (!defs || defs->empty()) {
    if (subset_size <= 0) throw std::runtime_error("Error, square subsets need a subset size");
    check(dice_hip_set_subsets(ctx_, n, cx.data(), cy.data(), subset_size, nullptr, nullptr, nullptr, nullptr),
          "set_subsets");
  } else {
    // every subset as a pixel list: conformal ones from their polygons
    // (DICe_Subset.cpp:115-168), square ones row by row (:78-113)
    std::vector<int32_t> off(1, 0), px, py;
    std::vector<uint8_t> act;
    for (int i = 0; i < n; ++i) {
      auto it = defs->find(i);
      const size_t start = px.size();
      if (it != defs->end()) {
        if (it->second.boundary.empty()) throw std::runtime_error("Error, conformal subset without a boundary");
        std::set<std::pair<int, int>> coords;  // (y, x): y-major order
        for (const Polygon &poly : it->second.boundary)
          for (const auto &yx : poly.get_owned_pixels()) coords.insert(yx);
        for (const auto &yx : coords) {
          px.push_back(yx.second < 0 ? 0 : yx.second);
          py.push_back(yx.first < 0 ? 0 : yx.first);
        }
        const int cnt = (int)(px.size() - start);
        std::vector<uint8_t> flags(cnt, 1);
        for (const Polygon &poly : it->second.excluded_area)
          poly.deactivate_pixels(cnt, flags, px.data() + start, py.data() + start);
        act.insert(act.end(), flags.begin(), flags.end());
      } else {
        if (subset_size <= 0) throw std::runtime_error("Error, square subsets need a subset size");
        const int h = subset_size / 2;
        const int x0 = (int)cx[i] - h, y0 = (int)cy[i] - h;
        for (int y = y0; y < y0 + 2 * h + 1; ++y)
          for (int x = x0; x < x0 + 2 * h + 1; ++x) {
            px.push_back(x < 0 ? 0 : x);
            py.push_back(y < 0 ? 0 : y);
            act.push_back(1);
          }
      }
      off.push_back((int32_t)px.size());
    }
    check(dice_hip_set_subsets(ctx_, n, cx.data(), cy.data(), subset_size, off.data(), px.data(), py.data(),
                               act.data()),
          "set_subsets");
  }
  for (int i = 0; i < n; ++i) {
    fields_[(size_t)DICE_F_COORD_X * n + i] = cx[i];
    fields_[(size_t)DICE_F_COORD_Y * n + i] = cy[i];
  }
}

template <typename T>
void Schema::set_image(bool ref, int width, int height, const T *pix, int storage) {
  if (!pix) throw std::invalid_argument("Error, null image");
  if (ref) {
    check(dice_hip_set_ref_image(ctx_, width, height, storage, pix), "set_ref_image");
    ref_w_ = width;
    ref_h_ = height;
  } else {
    check(dice_hip_set_def_image(ctx_, width, height, storage, pix), "set_def_image");
  }
}

void Schema::set_ref_image(int width, int height, const float *intensities) {
  set_image(true, width, height, intensities, DICE_STORAGE_F32);
}
void Schema::set_ref_image(int width, int height, const uint16_t *intensities) {
  set_image(true, width, height, intensities, DICE_STORAGE_U16);
}
void Schema::set_def_image(int width, int height, const float *intensities) {
  set_image(false, width, height, intensities, DICE_STORAGE_F32);
}
void Schema::set_def_image(int width, int height, const uint16_t *intensities) {
  set_image(false, width, height, intensities, DICE_STORAGE_U16);
}

int Schema::execute_correlation() {
  if (n_ <= 0) throw std::runtime_error("Error, the schema has no subsets");
  const int mode = backend_ == HIP_PLAIN ? DICE_MODE_PLAIN : DICE_MODE_BATCHED;
  check(dice_hip_correlate(ctx_, mode, fields_.data(), frame_id_ == first_frame_id_ ? 1 : 0),
        "execute_correlation");
  ++frame_id_;  // update_frame_id
  return 0;
}

double &Schema::local_field_value(int local_id, Field_Spec fs) {
  if (local_id < 0 || local_id >= n_) throw std::runtime_error("Error, subset id out of range");
  return fields_[(size_t)fs * n_ + local_id];
}

}  // namespace dice_hip

extern "C" {

int dice_polygon_pixel_counts(int n_polygons, const int *vert_offsets, const int *vx, const int *vy,
                              int *pixel_counts) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
  for (int i = 0; i < n_polygons; ++i) {
    const int b = vert_offsets[i], e = vert_offsets[i + 1];
    if (e - b < 3) {
      ++bad;
      pixel_counts[i] = 0;
      continue;
    }
    const dice_hip::Polygon p(std::vector<int>(vx + b, vx + e), std::vector<int>(vy + b, vy + e));
    pixel_counts[i] = (int)p.get_owned_pixels().size();
  }
  return bad ? -1 : 0;
}

int dice_polygon_pixels(int n_polygons, const int *vert_offsets, const int *vx, const int *vy,
                        const int64_t *pixel_offsets, int *px, int *py) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
  for (int i = 0; i < n_polygons; ++i) {
    const int b = vert_offsets[i], e = vert_offsets[i + 1];
    if (e - b < 3) {
      ++bad;
      continue;
    }
    const dice_hip::Polygon p(std::vector<int>(vx + b, vx + e), std::vector<int>(vy + b, vy + e));
    int64_t k = pixel_offsets[i];
    for (const auto &yx : p.get_owned_pixels()) {
      px[k] = yx.second;
      py[k] = yx.first;
      ++k;
    }
  }
  return bad ? -1 : 0;
}

}  // extern "C"
