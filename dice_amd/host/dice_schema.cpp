// C++ host mirror of the reference's Schema interface (see dice_schema.hpp).
#include "dice_schema.hpp"
#include "dice_api.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <set>
#include <sstream>

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
      {"gauss_filter_images", {I32, offsetof(dice_params, gauss_filter_images)}},
      {"gauss_filter_mask_size", {I32, offsetof(dice_params, gauss_filter_mask_size)}},
      {"neighbor_settle_tolerance", {F64, offsetof(dice_params, neighbor_settle_tolerance)}},
  };
  return m;
}
const char *kBackend = "correlation_backend";
// Schema-level parameters of the obstruction and blocking logic (DICe.h:273, 277), not in dice_params
const char *kOverrideForceSimplex = "override_force_simplex";
const char *kSkinFactor = "obstruction_skin_factor";
const char *kReadFullImages = "read_full_images";
bool schema_level(const std::string &name) {
  return name == kBackend || name == kOverrideForceSimplex || name == kSkinFactor || name == kReadFullImages;
}

void to_list(const dice_params &p, ParameterList *out) {
  for (const auto &kv : slots()) {
    const char *base = reinterpret_cast<const char *>(&p) + kv.second.offset;
    const double v = kv.second.kind == I32 ? (double)*reinterpret_cast<const int32_t *>(base)
                                           : *reinterpret_cast<const double *>(base);
    out->set(kv.first, v);
  }
  out->set(kBackend, HIP_BATCHED);
  // DICe_ParameterUtilities.cpp:323,333 (tracking) and :392,396 (full field)
  const bool tracking = p.correlation_routine == TRACKING_ROUTINE;
  out->set(kOverrideForceSimplex, tracking ? 1.0 : 0.0);
  out->set(kSkinFactor, tracking ? 1.8 : 1.0);
  out->set(kReadFullImages, 0.0);  // :326, :377
}

// atan2 difference wrapped to [-pi, pi] (angle_2d, DICe_Shape.cpp:88-103), in
// scalar_t = float as the reference's default build computes it: the float atan2,
// the wrap formed against the double constants and rounded back to float.  A
// pixel on a polygon edge sums to +-pi up to rounding, so the arithmetic decides
// whether it belongs to the polygon (the float build's conformal subset golden,
// tests/component/images/conformal.rawi, has one such pixel fewer than the double
// build's conformal_d.rawi).
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
float angle_2d(float x1, float y1, float x2, float y2) {
  float d = std::atan2(y2, x2) - std::atan2(y1, x1);
  while (d > kPi) d -= kTwoPi;
  while (d < -kPi) d += kTwoPi;
  return d;
}

}  // namespace

ParameterList &ParameterList::set(const std::string &name, double value) {
  if (!schema_level(name) && slots().count(name) == 0)
    throw std::invalid_argument("Error, unknown correlation parameter: " + name);
  values_[name] = value;
  return *this;
}

double ParameterList::get(const std::string &name) const {
  auto it = values_.find(name);
  if (it == values_.end()) throw std::invalid_argument("Error, parameter not set: " + name);
  return it->second;
}

ParameterList read_correlation_params(const std::string &file_name, bool *use_tracking_defaults) {
  std::ifstream in(file_name);
  if (!in) throw std::runtime_error("Error, cannot read the parameter file " + file_name);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();
  ParameterList out;
  if (use_tracking_defaults) *use_tracking_defaults = false;
  // enum-valued parameters and their strings, in enum order (DICe.h:440-600)
  static const std::map<std::string, std::vector<std::string>> enums = {
      {"correlation_routine", {"GENERIC_ROUTINE", "TRACKING_ROUTINE"}},
      {"interpolation_method", {"BILINEAR", "BICUBIC", "KEYS_FOURTH"}},
      {"optimization_method",
       {"SIMPLEX", "GRADIENT_BASED", "GRADIENT_THEN_SEARCH", "SIMPLEX_THEN_GRADIENT_BASED",
        "GRADIENT_BASED_THEN_SIMPLEX"}},
      {"initialization_method", {"USE_FIELD_VALUES", "USE_NEIGHBOR_VALUES", "USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY"}},
      {"gradient_method", {"FINITE_DIFFERENCE", "CONVOLUTION_5_POINT"}}};
  static const std::set<std::string> output_only = {"output_deformed_subset_images",
                                                    "output_deformed_subset_intensity_images",
                                                    "output_evolved_subset_images"};
  auto attr = [](const std::string &tag, const char *key) -> std::string {
    const std::string k = std::string(key) + "=\"";
    const size_t a = tag.find(k);
    if (a == std::string::npos) return "";
    const size_t b = tag.find('"', a + k.size());
    return b == std::string::npos ? "" : tag.substr(a + k.size(), b - a - k.size());
  };
  size_t pos = 0;
  int depth = 0;
  while ((pos = text.find('<', pos)) != std::string::npos) {
    const size_t end = text.find('>', pos);
    if (end == std::string::npos) break;
    const std::string tag = text.substr(pos, end - pos + 1);
    pos = end + 1;
    if (tag.compare(0, 4, "<!--") == 0 || tag.compare(0, 2, "<?") == 0) continue;
    if (tag.compare(0, 14, "<ParameterList") == 0) {
      if (++depth > 1) throw std::invalid_argument("Error, nested parameter lists are not supported: " + attr(tag, "name"));
      continue;
    }
    if (tag.compare(0, 15, "</ParameterList") == 0) {
      --depth;
      continue;
    }
    if (tag.compare(0, 10, "<Parameter") != 0) continue;
    std::string name = attr(tag, "name"), type = attr(tag, "type"), value = attr(tag, "value");
    std::transform(name.begin(), name.end(), name.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    if (name == "use_tracking_default_params") {
      if (use_tracking_defaults) *use_tracking_defaults = value == "true" || value == "1";
      continue;
    }
    if (output_only.count(name)) continue;
    if (name == "shape_function_type") {
      if (value != "AFFINE") throw std::invalid_argument("Error, only the AFFINE shape function is supported");
      continue;
    }
    double v = 0.0;
    auto e = enums.find(name);
    if (e != enums.end()) {
      auto it = std::find(e->second.begin(), e->second.end(), value);
      if (it == e->second.end()) throw std::invalid_argument("Error, unknown value " + value + " of " + name);
      v = (double)(it - e->second.begin());
    } else if (type == "bool") {
      v = (value == "true" || value == "1") ? 1.0 : 0.0;
    } else {
      char *stop = nullptr;
      v = std::strtod(value.c_str(), &stop);
      if (value.empty() || (stop && *stop)) throw std::invalid_argument("Error, bad value of " + name + ": " + value);
    }
    out.set(name, v);  // throws for a name this path does not read
  }
  return out;
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
  float angle = 0.0f;
  for (size_t i = 0; i + 1 < vx_.size(); ++i)
    angle += angle_2d((float)(vx_[i] - x), (float)(vy_[i] - y), (float)(vx_[i + 1] - x), (float)(vy_[i + 1] - y));
  return std::abs(angle) >= kPi;
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

Circle::Circle(int centroid_x, int centroid_y, double radius) : cx_(centroid_x), cy_(centroid_y) {
  if (centroid_x <= 0 || centroid_y <= 0 || !(radius > 0.0))
    throw std::invalid_argument("Error, a circle needs a positive centroid and radius");
  radius2_ = radius * radius;
  // integer members initialised from the floating-point extents
  min_x_ = (int)(cx_ - radius - 1);
  max_x_ = (int)(cx_ + radius + 1);
  min_y_ = (int)(cy_ - radius - 1);
  max_y_ = (int)(cy_ + radius + 1);
}

std::vector<std::pair<int, int>> Circle::get_owned_pixels() const {
  std::vector<std::pair<int, int>> out;
  for (int y = min_y_; y <= max_y_; ++y)
    for (int x = min_x_; x <= max_x_; ++x) {
      const double dx = (double)(x - cx_) * (x - cx_), dy = (double)(y - cy_) * (y - cy_);
      if (dx + dy <= radius2_) out.emplace_back(y, x);
    }
  return out;
}

void Circle::deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const {
  for (int j = 0; j < size; ++j) {
    const double dx = (double)(x[j] - cx_) * (x[j] - cx_), dy = (double)(y[j] - cy_) * (y[j] - cy_);
    if (dx + dy <= radius2_) flags[j] = 0;
  }
}

Rectangle::Rectangle(int centroid_x, int centroid_y, int width, int height)
    : origin_x_(centroid_x - width / 2), origin_y_(centroid_y - height / 2), width_(width), height_(height) {
  if (centroid_x <= 0 || centroid_y <= 0 || width <= 0 || height <= 0 || origin_x_ < 0 || origin_y_ < 0)
    throw std::invalid_argument("Error, a rectangle needs a positive size and must lie at non-negative coordinates");
}

std::vector<std::pair<int, int>> Rectangle::get_owned_pixels() const {
  std::vector<std::pair<int, int>> out;
  out.reserve((size_t)width_ * height_);
  for (int y = 0; y < height_; ++y)
    for (int x = 0; x < width_; ++x) out.emplace_back(origin_y_ + y, origin_x_ + x);
  return out;
}

void Rectangle::deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const {
  for (int j = 0; j < size; ++j)
    if (x[j] >= origin_x_ && x[j] < origin_x_ + width_ && y[j] >= origin_y_ && y[j] < origin_y_ + height_) flags[j] = 0;
}

namespace {

// Affine_Shape_Function::map (DICe_LocalShapeFunction.cpp:220-248) in scalar_t = float: the
// disabled parameters read as 0.
void map_float(const Affine_Shape_Function &sf, float x, float y, float cx, float cy, float &out_x, float &out_y) {
  const double *q = sf.parameters();
  const float dispx = (float)q[0], dispy = (float)q[1], theta = (float)q[2], dudx = (float)q[3],
              dvdy = (float)q[4], gxy = (float)q[5];
  const float cost = std::cos(theta), sint = std::sin(theta);
  const float dx = x - cx, dy = y - cy;
  const float Dx = (float)((1.0 + dudx) * dx + gxy * dy);
  const float Dy = (float)((1.0 + dvdy) * dy + gxy * dx);
  out_x = cost * Dx - sint * Dy + dispx + cx;
  out_y = sint * Dx + cost * Dy + dispy + cy;
}

// The moved, stretched vertices of a closed vertex list (vertex 0 repeated at the end) and their
// bbox, as Polygon/Rectangle::get_owned_pixels(shape_function, ...) form them (DICe_Shape.cpp:
// 147-199, 347-395): each vertex mapped and rounded (c = (int)X, +1 for a fraction >= 0.5), the
// integer centroid of the open list, then x = (int)(skin (x - c) + c).
void moved_vertices(const Affine_Shape_Function &sf, int cx, int cy, double skin, const std::vector<int> &vx0,
                    const std::vector<int> &vy0, std::vector<int> &vx, std::vector<int> &vy, int &min_x, int &max_x,
                    int &min_y, int &max_y) {
  const size_t m = vx0.size();
  vx.assign(m, 0);
  vy.assign(m, 0);
  for (size_t i = 0; i < m; ++i) {
    float X = 0.0f, Y = 0.0f;
    map_float(sf, (float)vx0[i], (float)vy0[i], (float)cx, (float)cy, X, Y);
    int nx = (int)X, ny = (int)Y;
    if (X - (int)X >= 0.5f) nx++;
    if (Y - (int)Y >= 0.5f) ny++;
    vx[i] = nx;
    vy[i] = ny;
  }
  int ccx = 0, ccy = 0;
  for (size_t i = 0; i + 1 < m; ++i) {
    ccx += vx[i];
    ccy += vy[i];
  }
  const double cnt = m - 1 == 0 ? 1.0 : (double)(m - 1);
  ccx = (int)(ccx / cnt);
  ccy = (int)(ccy / cnt);
  const float sk = (float)skin;
  for (size_t i = 0; i < m; ++i) {
    const int nx = (int)(sk * (vx[i] - ccx) + ccx), ny = (int)(sk * (vy[i] - ccy) + ccy);
    vx[i] = nx;
    vy[i] = ny;
    if (i == 0) {
      min_x = max_x = nx;
      min_y = max_y = ny;
    } else {
      min_x = std::min(min_x, nx);
      max_x = std::max(max_x, nx);
      min_y = std::min(min_y, ny);
      max_y = std::max(max_y, ny);
    }
  }
}

// pixels of the bbox whose winding angle around the closed vertex list is at least pi
std::vector<std::pair<int, int>> winding_pixels(const std::vector<int> &vx, const std::vector<int> &vy, int min_x,
                                                int max_x, int min_y, int max_y) {
  std::vector<std::pair<int, int>> out;
  for (int y = min_y; y <= max_y; ++y)
    for (int x = min_x; x <= max_x; ++x) {
      float angle = 0.0f;
      for (size_t i = 0; i + 1 < vx.size(); ++i)
        angle += angle_2d((float)(vx[i] - x), (float)(vy[i] - y), (float)(vx[i + 1] - x), (float)(vy[i + 1] - y));
      if (std::abs(angle) >= kPi) out.emplace_back(y, x);
    }
  return out;
}

}  // namespace

std::vector<std::pair<int, int>> Polygon::get_owned_pixels(const Affine_Shape_Function &sf, int cx, int cy,
                                                           double skin_factor) const {
  std::vector<int> vx, vy;
  int min_x, max_x, min_y, max_y;
  moved_vertices(sf, cx, cy, skin_factor, vx_, vy_, vx, vy, min_x, max_x, min_y, max_y);
  return winding_pixels(vx, vy, min_x, max_x, min_y, max_y);
}

std::vector<std::pair<int, int>> Circle::get_owned_pixels(const Affine_Shape_Function &, int, int, double) const {
  throw std::runtime_error("Error, circle deformation has not been implemented yet");
}

std::vector<std::pair<int, int>> Rectangle::get_owned_pixels(const Affine_Shape_Function &sf, int cx, int cy,
                                                             double skin_factor) const {
  const std::vector<int> vx0 = {origin_x_, origin_x_ + width_, origin_x_ + width_, origin_x_, origin_x_};
  const std::vector<int> vy0 = {origin_y_, origin_y_, origin_y_ + height_, origin_y_ + height_, origin_y_};
  std::vector<int> vx, vy;
  int min_x, max_x, min_y, max_y;
  moved_vertices(sf, cx, cy, skin_factor, vx0, vy0, vx, vy, min_x, max_x, min_y, max_y);
  return winding_pixels(vx, vy, min_x, max_x, min_y, max_y);
}

namespace {

// tokenize_line (DICe_Parser.cpp:1046-1084): blank-delimited, capitalised
// tokens; a line whose first token starts with '#' is empty.
std::vector<std::string> tokenize(const std::string &line, const char *delim = " \t\r") {
  std::vector<std::string> out;
  std::string buf = line;
  for (char *t = std::strtok(&buf[0], delim); t; t = std::strtok(nullptr, delim)) {
    std::string tok(t);
    for (char &c : tok) c = (char)std::toupper((unsigned char)c);
    out.push_back(tok);
  }
  if (!out.empty() && out[0][0] == '#') out.clear();
  return out;
}

// is_number (DICe_Parser.cpp:1088-1094)
bool is_number(const std::string &s) {
  if (s.empty()) return false;
  for (char c : s)
    if (!(std::isdigit((unsigned char)c) || c == '+' || c == '-' || c == 'e' || c == 'E' || c == '.')) return false;
  return true;
}

struct LineReader {
  std::ifstream in;
  std::string name;
  bool next(std::vector<std::string> &tok) {
    std::string line;
    while (std::getline(in, line)) {
      tok = tokenize(line);
      if (!tok.empty()) return true;
    }
    return false;
  }
  [[noreturn]] void fail(const std::string &what) const {
    throw std::runtime_error("Error parsing subset file " + name + ": " + what);
  }
};

int to_int(const LineReader &r, const std::string &s) {
  if (!is_number(s)) r.fail("expected a number, found " + s);
  return (int)std::strtol(s.c_str(), nullptr, 0);
}

std::shared_ptr<Shape> read_polygon(LineReader &r) {
  std::vector<int> vx, vy;
  std::vector<std::string> t;
  while (r.next(t)) {
    if (t[0] == "END") break;
    if (t.size() < 2 || t[0] != "BEGIN" || t[1] != "VERTICES") r.fail("expected BEGIN VERTICES in a polygon");
    while (r.next(t)) {
      if (t[0] == "END") break;
      if (t.size() < 2) r.fail("a vertex needs x and y");
      vx.push_back(to_int(r, t[0]));
      vy.push_back(to_int(r, t[1]));
    }
  }
  if (vx.empty()) r.fail("polygon without vertices");
  return std::make_shared<Polygon>(vx, vy);
}

std::shared_ptr<Shape> read_circle(LineReader &r) {
  int cx = -1, cy = -1;
  double radius = -1.0;
  std::vector<std::string> t;
  while (r.next(t)) {
    if (t[0] == "END") break;
    if (t[0] == "CENTER") {
      if (t.size() < 3) r.fail("not enough values to define circle center point");
      cx = to_int(r, t[1]);
      cy = to_int(r, t[2]);
    } else if (t[0] == "RADIUS") {
      if (t.size() < 2 || !is_number(t[1])) r.fail("not enough values to define circle radius");
      radius = std::strtod(t[1].c_str(), nullptr);
    } else {
      r.fail("invalid token in circle definition " + t[0]);
    }
  }
  if (radius == -1.0 || cx == -1 || cy == -1) r.fail("incomplete circle");
  return std::make_shared<Circle>(cx, cy, radius);
}

std::shared_ptr<Shape> read_rectangle(LineReader &r) {
  int cx = -1, cy = -1, width = -1, height = -1, ulx = -1, uly = -1, lrx = -1, lry = -1;
  bool corners = false;
  std::vector<std::string> t;
  while (r.next(t)) {
    if (t[0] == "END") break;
    if (t[0] == "CENTER") {
      if (t.size() < 3) r.fail("not enough values to define rectangle center point");
      cx = to_int(r, t[1]);
      cy = to_int(r, t[2]);
    } else if (t[0] == "WIDTH" || t[0] == "HEIGHT") {
      if (t.size() < 2 || !is_number(t[1])) r.fail("not enough values to define the " + t[0]);
      (t[0] == "WIDTH" ? width : height) = (int)std::strtod(t[1].c_str(), nullptr);
    } else if (t[0] == "UPPER_LEFT" || t[0] == "LOWER_RIGHT") {
      if (t.size() < 3) r.fail("not enough values to define rectangle " + t[0]);
      corners = true;
      (t[0] == "UPPER_LEFT" ? ulx : lrx) = to_int(r, t[1]);
      (t[0] == "UPPER_LEFT" ? uly : lry) = to_int(r, t[2]);
    } else {
      r.fail("invalid token in rectangle definition " + t[0]);
    }
  }
  if (corners) {
    if (ulx < 0 || uly < 0 || lrx < 0 || lry < 0) r.fail("rectangle corners must be non-negative");
    if (lrx < ulx || lry < uly) r.fail("Rectangle inverted or zero width");
    width = lrx - ulx;
    height = lry - uly;
    cx = width / 2 + ulx;
    cy = height / 2 + uly;
  }
  if (width <= 0 || height <= 0 || cx < 0 || cy < 0) r.fail("incomplete rectangle");
  return std::make_shared<Rectangle>(cx, cy, width, height);
}

// read_shapes (DICe_Parser.cpp:1832-1870): BEGIN POLYGON / CIRCLE / RECTANGLE blocks until END
multi_shape read_shapes(LineReader &r) {
  multi_shape shapes;
  std::vector<std::string> t;
  while (r.next(t)) {
    if (t[0] == "END") break;
    if (t.size() < 2 || t[0] != "BEGIN") r.fail("expected BEGIN <shape>");
    if (t[1] == "POLYGON") shapes.push_back(read_polygon(r));
    else if (t[1] == "CIRCLE") shapes.push_back(read_circle(r));
    else if (t[1] == "RECTANGLE") shapes.push_back(read_rectangle(r));
    else r.fail("unknown shape " + t[1]);
  }
  return shapes;
}

}  // namespace

Subset_File_Info read_subset_file(const std::string &file_name, int width, int height) {
  LineReader r;
  r.name = file_name;
  r.in.open(file_name);
  if (!r.in.good()) throw std::runtime_error("Error, the subset file does not exist: " + file_name);
  Subset_File_Info info;
  std::vector<std::string> t;
  while (r.next(t)) {
    if (t.size() < 2 || t[0] != "BEGIN") r.fail("invalid entry " + t[0]);
    if (t[1] == "SUBSET_COORDINATES") {
      while (r.next(t)) {
        if (t[0] == "END") break;
        if (!is_number(t[0])) r.fail("error parsing subset coordinates: " + t[0]);
        if (t.size() < 2 || !is_number(t[1])) r.fail("invalid coordinate (not enough values)");
        const double x = std::strtod(t[0].c_str(), nullptr), y = std::strtod(t[1].c_str(), nullptr);
        if (x < 0 || (width != -1 && x >= width)) r.fail("invalid subset coordinate x: " + t[0]);
        if (y < 0 || (height != -1 && y >= height)) r.fail("invalid subset coordinate y: " + t[1]);
        info.coordinates_x.push_back(x);
        info.coordinates_y.push_back(y);
      }
    } else if (t[1] == "CONFORMAL_SUBSET") {
      int id = -1;
      Conformal_Area_Def def;
      std::vector<int> blocking;
      bool force_simplex = false;
      while (r.next(t)) {
        if (t[0] == "END") break;
        if (t[0] == "SUBSET_ID") {
          if (t.size() < 2) r.fail("SUBSET_ID needs a value");
          id = to_int(r, t[1]);
        } else if (t[0] == "FORCE_SIMPLEX") {
          force_simplex = true;
        } else if (t[0] == "BEGIN" && t.size() >= 2) {
          if (t[1] == "BOUNDARY") {
            def.boundary = read_shapes(r);
            if (def.boundary.empty()) r.fail("cannot define a boundary with zero shapes");
          } else if (t[1] == "EXCLUDED") {
            def.excluded_area = read_shapes(r);
            if (def.excluded_area.empty()) r.fail("cannot define an exclusion with zero shapes");
          } else if (t[1] == "OBSTRUCTED") {
            def.obstructed_area = read_shapes(r);
            if (def.obstructed_area.empty()) r.fail("cannot define an obstruction with zero shapes");
          } else if (t[1] == "BLOCKING_SUBSETS") {
            while (r.next(t)) {
              if (t[0] == "END") break;
              blocking.push_back(to_int(r, t[0]));
            }
          } else {
            r.fail(t[1] + " blocks are not supported by this path");
          }
        } else {
          r.fail(t[0] + " is not supported by this path");
        }
      }
      if (id < 0) r.fail("invalid subset id for conformal subset (SUBSET_ID <id> missing)");
      if (def.boundary.empty()) r.fail("conformal subset without a boundary");
      info.conformal_area_defs[id] = def;
      info.blocking_subset_ids[id] = blocking;
      if (force_simplex) info.force_simplex.push_back(id);
    } else if (t[1] == "REGION_OF_INTEREST") {
      // DICe_Parser.cpp:1173-1367
      const int roi = (int)info.conformal_area_defs.size();
      info.region_of_interest = true;
      Conformal_Area_Def def;
      while (r.next(t)) {
        if (t[0] == "END") break;
        if (t[0] == "BEGIN" && t.size() >= 2) {
          if (t[1] == "BOUNDARY") {
            def.boundary = read_shapes(r);
            if (def.boundary.empty()) r.fail("cannot define a boundary with zero shapes");
          } else if (t[1] == "EXCLUDED") {
            def.excluded_area = read_shapes(r);
            if (def.excluded_area.empty()) r.fail("cannot define an exclusion with zero shapes");
          } else if (t[1] == "SEED") {
            bool has_location = false, has_disp = false;
            while (r.next(t)) {
              if (t[0] == "END") break;
              auto num = [&](size_t k) {
                if (t.size() <= k || !is_number(t[k])) r.fail("not enough values specified for seed " + t[0]);
                return std::strtod(t[k].c_str(), nullptr);
              };
              if (t[0] == "LOCATION") {
                info.size_map[roi] = {(int)std::strtol(t.size() > 1 ? t[1].c_str() : "", nullptr, 0),
                                      (int)std::strtol(t.size() > 2 ? t[2].c_str() : "", nullptr, 0)};
                num(1);
                num(2);
                has_location = true;
              } else if (t[0] == "DISPLACEMENT") {
                info.displacement_map[roi] = {num(1), num(2)};
                has_disp = true;
              } else if (t[0] == "NORMAL_STRAIN") {
                info.normal_strain_map[roi] = {num(1), num(2)};
              } else if (t[0] == "SHEAR_STRAIN") {
                info.shear_strain_map[roi] = num(1);
              } else if (t[0] == "ROTATION") {
                info.rotation_map[roi] = num(1);
              } else {
                r.fail("unrecognized command in seed block " + t[0]);
              }
            }
            if (!has_location) r.fail("seed must have location specified");
            if (!has_disp) r.fail("seed must have displacement guess specified");
          } else {
            r.fail(t[1] + " blocks are not supported by this path");
          }
        } else {
          r.fail(t[0] + " is not supported by this path");
        }
      }
      info.conformal_area_defs[roi] = def;
    } else {
      r.fail(t[1] + " blocks are not supported by this path");
    }
  }
  return info;
}

namespace {
const std::vector<std::pair<Field_Spec, const char *>> &field_names() {
  static const std::vector<std::pair<Field_Spec, const char *>> v = {
      {SUBSET_COORDINATES_X_FS, "COORDINATE_X"}, {SUBSET_COORDINATES_Y_FS, "COORDINATE_Y"},
      {SUBSET_DISPLACEMENT_X_FS, "DISPLACEMENT_X"}, {SUBSET_DISPLACEMENT_Y_FS, "DISPLACEMENT_Y"},
      {ROTATION_Z_FS, "ROTATION_Z"}, {NORMAL_STRETCH_XX_FS, "NORMAL_STRAIN_X"},
      {NORMAL_STRETCH_YY_FS, "NORMAL_STRAIN_Y"}, {SHEAR_STRETCH_XY_FS, "SHEAR_STRAIN_XY"},
      {SIGMA_FS, "SIGMA"}, {GAMMA_FS, "GAMMA"}, {BETA_FS, "BETA"}, {MATCH_FS, "MATCH"},
      {NOISE_LEVEL_FS, "NOISE_LEVEL"}, {CONTRAST_LEVEL_FS, "CONTRAST_LEVEL"}, {ACTIVE_PIXELS_FS, "ACTIVE_PIXELS"},
      {STATUS_FLAG_FS, "STATUS_FLAG"}, {ITERATIONS_FS, "ITERATIONS"}, {CONDITION_NUMBER_FS, "CONDITION_NUMBER"},
      {NEIGHBOR_ID_FS, "NEIGHBOR_ID"}, {FIELD_1_FS, "FIELD_1"}, {FIELD_2_FS, "FIELD_2"}, {FIELD_3_FS, "FIELD_3"},
      {FIELD_4_FS, "FIELD_4"}, {FIELD_5_FS, "FIELD_5"}, {FIELD_6_FS, "FIELD_6"}, {FIELD_7_FS, "FIELD_7"},
      {VSG_STRAIN_XX_FS, "VSG_STRAIN_XX"}, {VSG_STRAIN_YY_FS, "VSG_STRAIN_YY"}, {VSG_STRAIN_XY_FS, "VSG_STRAIN_XY"},
      {VSG_DUDX_FS, "VSG_DUDX"}, {VSG_DUDY_FS, "VSG_DUDY"}, {VSG_DVDX_FS, "VSG_DVDX"}, {VSG_DVDY_FS, "VSG_DVDY"},
      {UNCERTAINTY_FS, "UNCERTAINTY"}, {UNCERTAINTY_ANGLE_FS, "UNCERTAINTY_ANGLE"}};
  return v;
}
}  // namespace

const char *field_name(Field_Spec fs) {
  for (const auto &e : field_names())
    if (e.first == fs) return e.second;
  return "";
}

Field_Spec field_spec_from_name(const std::string &name) {
  std::string up = name;
  for (char &c : up) c = (char)std::toupper((unsigned char)c);
  for (const auto &e : field_names())
    if (up == e.second) return e.first;
  throw std::invalid_argument("Error, the output field " + name + " is not produced by this path");
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
  step_x_ = step_size_x;
  step_y_ = step_size_y;
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
    if (schema_level(kv.first)) continue;
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
  override_force_simplex_ = params.isParameter(kOverrideForceSimplex) && params.get(kOverrideForceSimplex) != 0.0;
  skin_factor_ = params.isParameter(kSkinFactor) ? params.get(kSkinFactor) : 1.0;
  read_full_images_ = params.isParameter(kReadFullImages) && params.get(kReadFullImages) != 0.0;
}

void create_regular_grid_of_correlation_points(int img_w, int img_h, int subset_size, int step_size,
                                               std::vector<double> &points_x, std::vector<double> &points_y,
                                               std::vector<int> &neighbor_ids,
                                               const std::map<int, Conformal_Area_Def> *roi_defs,
                                               Subset_File_Info *seeds) {
  if (step_size <= 0) throw std::runtime_error("Error: step size is <= 0");
  if (subset_size <= 0) throw std::runtime_error("Error, the subset size was not specified");
  points_x.clear();
  points_y.clear();
  neighbor_ids.clear();
  std::map<int, Conformal_Area_Def> whole;
  if (!roi_defs || roi_defs->empty()) {
    whole[0].boundary.push_back(std::make_shared<Rectangle>(img_w / 2, img_h / 2, img_w, img_h));
    roi_defs = &whole;
  }
  const bool seeded = seeds && !seeds->size_map.empty();
  if (seeded && seeds->size_map.size() != seeds->displacement_map.size())
    throw std::runtime_error("Error the number of displacement guesses and seed locations must be the same");
  for (const auto &roi : *roi_defs) {
    std::set<std::pair<int, int>> coords, excluded;
    for (const auto &shape : roi.second.boundary)
      for (const auto &yx : shape->get_owned_pixels()) coords.insert(yx);
    for (const auto &shape : roi.second.excluded_area)
      for (const auto &yx : shape->get_owned_pixels()) excluded.insert(yx);
    // valid_correlation_point (DICe_Decomp.cpp:1072-1103) without the gradient threshold
    auto valid = [&](int x, int y) {
      if (x < subset_size - 1 || x > img_w - subset_size || y < subset_size - 1 || y > img_h - subset_size)
        return false;
      if (!coords.count({y, x})) return false;
      const int x0 = x - subset_size / 2, y0 = y - subset_size / 2;
      const int cxs[4] = {x0, x0 + subset_size, x0 + subset_size, x0};
      const int cys[4] = {y0, y0, y0 + subset_size, y0 + subset_size};
      for (int i = 0; i < 4; ++i)
        if (excluded.count({cys[i], cxs[i]})) return false;
      return true;
    };
    // the seed's grid row and column: the nearest grid point to its location
    const bool has_seed = seeded && seeds->size_map.count(roi.first);
    const int seed_x = has_seed ? seeds->size_map.at(roi.first).first : 0;
    const int seed_y = has_seed ? seeds->size_map.at(roi.first).second : 0;
    int num_rows = 0, num_cols = 0, seed_row = 0, seed_col = 0;
    for (int x = subset_size - 1; x < img_w - subset_size; x += step_size, ++num_cols)
      if (x + step_size / 2 < seed_x) ++seed_col;
    for (int y = subset_size - 1; y < img_h - subset_size; y += step_size, ++num_rows)
      if (y + step_size / 2 < seed_y) ++seed_row;
    int seed_id = -1;
    auto add = [&](int col, int row, int neighbor) {
      const int x = subset_size - 1 + col * step_size, y = subset_size - 1 + row * step_size;
      if (!valid(x, y)) return false;
      points_x.push_back(x);
      points_y.push_back(y);
      neighbor_ids.push_back(neighbor);
      return true;
    };
    if (add(seed_col, seed_row, -1)) {
      if (has_seed) {
        seed_id = (int)points_x.size() - 1;
        seeds->seed_subset_ids[seed_id] = roi.first;
      }
    } else if (!(seed_row == 0 && seed_col == 0)) {
      throw std::runtime_error("Error, seed specified does not meet the SSSIG criteria for suffient image gradients");
    }
    // the serpentine to the right of the seed, then the one to the left
    for (int side = 1; side >= -1; side -= 2) {
      const int start = (int)points_x.size();
      int direction = side, row = seed_row, col = seed_col;
      while (row >= 0 && row < num_rows) {
        if ((direction == 1 && row + 1 >= num_rows) || (direction == -1 && row - 1 < 0)) {
          direction = -direction;
          col += side;
        } else {
          row += direction;
        }
        if (col >= num_cols || col < 0) break;
        const int id = (int)points_x.size();
        add(col, row, id == start ? seed_id : id - 1);
      }
    }
  }
}

void conformal_subset_pixels(const Conformal_Area_Def &def, std::vector<int> &x, std::vector<int> &y,
                             std::vector<uint8_t> &active) {
  if (def.boundary.empty()) throw std::runtime_error("Error, conformal subset without a boundary");
  std::set<std::pair<int, int>> coords;  // (y, x): y-major order
  for (const auto &shape : def.boundary)
    for (const auto &yx : shape->get_owned_pixels()) coords.insert(yx);
  x.clear();
  y.clear();
  for (const auto &yx : coords) {
    x.push_back(yx.second < 0 ? 0 : yx.second);
    y.push_back(yx.first < 0 ? 0 : yx.first);
  }
  active.assign(x.size(), 1);
  for (const auto &shape : def.excluded_area) shape->deactivate_pixels((int)x.size(), active, x.data(), y.data());
}

void Schema::init_subsets(const std::vector<double> &cx, const std::vector<double> &cy, int subset_size,
                          const std::map<int, Conformal_Area_Def> *defs) {
  const int n = (int)cx.size();
  if (n <= 0) throw std::runtime_error("Error, no subsets");
  n_ = n;
  subset_dim_ = subset_size;
  fields_.assign((size_t)DICE_F_COUNT * n, 0.0);
  defs_.clear();
  obstructed_.assign(n, {});
  obstructing_ids_.clear();
  force_simplex_ids_.clear();
  if (defs) {
    defs_ = *defs;
    // Subset::obstructed_coords_ (DICe_Subset.cpp:162-167): the pixels of the obstructed shapes
    for (const auto &kv : *defs) {
      if (kv.first < 0 || kv.first >= n) continue;
      std::set<std::pair<int, int>> yx;
      for (const auto &shape : kv.second.obstructed_area) {
        const auto px = shape->get_owned_pixels();
        yx.insert(px.begin(), px.end());
      }
      for (const auto &p : yx) obstructed_[kv.first].emplace_back(p.second, p.first);
    }
  }
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
      if (it != defs->end()) {
        std::vector<int> sx, sy;
        std::vector<uint8_t> flags;
        conformal_subset_pixels(it->second, sx, sy, flags);
        px.insert(px.end(), sx.begin(), sx.end());
        py.insert(py.end(), sy.begin(), sy.end());
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
  upload_masks({});
}

void Schema::set_obstructing_subset_ids(const std::map<int, std::vector<int>> &ids) {
  for (const auto &kv : ids) {
    if (kv.first < 0 || kv.first >= n_) throw std::runtime_error("Error, blocked subset id out of range");
    for (int b : kv.second)
      if (b < 0 || b >= n_) throw std::runtime_error("Error, blocking subset id out of range");
  }
  obstructing_ids_ = ids;
}

void Schema::set_force_simplex(const std::vector<int> &ids) {
  for (int id : ids)
    if (id < 0 || id >= n_) throw std::runtime_error("Error, force simplex subset id out of range");
  force_simplex_ids_ = ids;
}

// generic_correlation_routine's force_simplex (DICe_Schema.cpp:2160-2185): the FORCE_SIMPLEX subsets,
// the subsets with blockers and the blockers, unless override_force_simplex
void Schema::upload_force_flags() {
  std::vector<uint8_t> force(n_, 0);
  bool any = false;
  if (!override_force_simplex_) {
    for (int id : force_simplex_ids_) force[id] = 1;
    for (const auto &kv : obstructing_ids_) {
      if (kv.second.empty()) continue;
      force[kv.first] = 1;
      for (int b : kv.second) force[b] = 1;
    }
    for (uint8_t f : force) any = any || f;
  }
  check(dice_hip_set_force_simplex(ctx_, any ? force.data() : nullptr), "set_force_simplex");
}

// The obstructed pixels of every subset, with the pixels blocked by other subsets (blocked[i], may be
// empty) added.
void Schema::upload_masks(const std::vector<std::vector<std::pair<int, int>>> &blocked) {
  std::vector<int32_t> off(1, 0), mx, my;
  for (int i = 0; i < n_; ++i) {
    for (const auto &p : obstructed_[i]) {
      mx.push_back(p.first);
      my.push_back(p.second);
    }
    if (i < (int)blocked.size())
      for (const auto &p : blocked[i]) {
        mx.push_back(p.first);
        my.push_back(p.second);
      }
    off.push_back((int32_t)mx.size());
  }
  check(dice_hip_set_subset_masks(ctx_, mx.empty() ? nullptr : off.data(), mx.data(), my.data()), "set_subset_masks");
}

// check_for_blocking_subsets (DICe_Schema.cpp:4529-4556): the pixels of the conformal blockers of
// `subset`, each moved by the blocker's current shape function (its fields) about its centroid and
// stretched by the skin factor, as (x, y); square blockers block nothing (Subset::deformed_shapes,
// DICe_Subset.cpp:462-474).
std::vector<std::pair<int, int>> Schema::blocking_pixels(int subset) const {
  std::vector<std::pair<int, int>> out;
  auto it = obstructing_ids_.find(subset);
  if (it == obstructing_ids_.end()) return out;
  std::set<std::pair<int, int>> yx;
  for (int b : it->second) {
    auto d = defs_.find(b);
    if (d == defs_.end()) continue;
    Affine_Shape_Function sf(rotation_enabled(), normal_strain_enabled(), shear_strain_enabled());
    sf.initialize_parameters_from_fields(this, b);
    const int cx = (int)local_field_value(b, SUBSET_COORDINATES_X_FS);
    const int cy = (int)local_field_value(b, SUBSET_COORDINATES_Y_FS);
    for (const auto &shape : d->second.boundary) {
      const auto px = shape->get_owned_pixels(sf, cx, cy, skin_factor_);
      yx.insert(px.begin(), px.end());
    }
  }
  for (const auto &p : yx) out.emplace_back(p.second, p.first);
  return out;
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
Affine_Shape_Function::Affine_Shape_Function(bool enable_rotation, bool enable_normal_strain,
                                             bool enable_shear_strain) {
  en_[0] = en_[1] = true;
  en_[2] = enable_rotation;
  en_[3] = en_[4] = enable_normal_strain;
  en_[5] = enable_shear_strain;
}

void Affine_Shape_Function::set(int index, double value) {
  if (index < 0 || index > 5) throw std::invalid_argument("Error, shape function parameter index out of range");
  p_[index] = en_[index] ? value : 0.0;
}

void Affine_Shape_Function::initialize_parameters_from_fields(const Schema *schema, int subset_gid) {
  const Field_Spec fs[6] = {SUBSET_DISPLACEMENT_X_FS, SUBSET_DISPLACEMENT_Y_FS, ROTATION_Z_FS,
                            NORMAL_STRETCH_XX_FS,     NORMAL_STRETCH_YY_FS,     SHEAR_STRETCH_XY_FS};
  for (int k = 0; k < 6; ++k) set(k, schema->local_field_value(subset_gid, fs[k]));
}

void Affine_Shape_Function::save_fields(Schema *schema, int subset_gid) const {
  const Field_Spec fs[6] = {SUBSET_DISPLACEMENT_X_FS, SUBSET_DISPLACEMENT_Y_FS, ROTATION_Z_FS,
                            NORMAL_STRETCH_XX_FS,     NORMAL_STRETCH_YY_FS,     SHEAR_STRETCH_XY_FS};
  for (int k = 0; k < 6; ++k)
    if (en_[k]) schema->local_field_value(subset_gid, fs[k]) = p_[k];
}

Objective_ZNSSD::Objective_ZNSSD(Schema *schema, int correlation_point_global_id)
    : schema_(schema), gid_(correlation_point_global_id) {
  if (!schema) throw std::invalid_argument("Error, null schema");
  if (gid_ < 0 || gid_ >= schema->local_num_subsets()) throw std::runtime_error("Error, subset id out of range");
}

double Objective_ZNSSD::call(int which, const Affine_Shape_Function &sf, double *params, double *result) const {
  for (int k = 0; k < 6; ++k) params[k] = sf.p_[k];
  const int rc = dice_hip_objective(schema_->context(), gid_, which, params, result);
  if (rc != DICE_HIP_OK)
    throw std::runtime_error(std::string("Error, objective call failed: ") + dice_hip_last_error(schema_->context()));
  return result[0];
}

Status_Flag Objective_ZNSSD::computeUpdateFast(Affine_Shape_Function &sf, int &num_iterations) {
  double params[6], result[10];
  const double status = call(DICE_OBJ_UPDATE_FAST, sf, params, result);
  for (int k = 0; k < 6; ++k) sf.set(k, params[k]);
  num_iterations = (int)result[1];
  schema_->local_field_value(gid_, CONDITION_NUMBER_FS) = result[2];
  const Field_Spec unc[7] = {FIELD_1_FS, FIELD_2_FS, FIELD_3_FS, FIELD_4_FS, FIELD_5_FS, FIELD_6_FS, FIELD_7_FS};
  if (status == CORRELATION_SUCCESSFUL)
    for (int k = 0; k < 7; ++k) schema_->local_field_value(gid_, unc[k]) = result[3 + k];
  return (Status_Flag)(int)status;
}

Status_Flag Objective_ZNSSD::computeUpdateRobust(Affine_Shape_Function &sf, int &num_iterations) {
  double params[6], result[10];
  const double status = call(DICE_OBJ_UPDATE_ROBUST, sf, params, result);
  for (int k = 0; k < 6; ++k) sf.set(k, params[k]);
  num_iterations = (int)result[1];
  return (Status_Flag)(int)status;
}

double Objective_ZNSSD::gamma(const Affine_Shape_Function &sf) const {
  double params[6], result[10];
  return call(DICE_OBJ_GAMMA, sf, params, result);
}

double Objective_ZNSSD::sigma(const Affine_Shape_Function &sf, double &noise_level) const {
  double params[6], result[10];
  const double s = call(DICE_OBJ_SIGMA, sf, params, result);
  noise_level = result[1];
  return s;
}

void Schema::set_neighbor_ids(const std::vector<int> &neighbor_ids) {
  if ((int)neighbor_ids.size() != n_) throw std::invalid_argument("Error, one neighbor id per subset is needed");
  for (int i = 0; i < n_; ++i) {
    if (neighbor_ids[i] < -1 || neighbor_ids[i] >= n_) throw std::invalid_argument("Error, invalid neighbor id");
    fields_[(size_t)DICE_F_NEIGHBOR_ID * n_ + i] = neighbor_ids[i];
  }
}

void Schema::set_seed_values(const Subset_File_Info &info) {
  for (const auto &seed : info.seed_subset_ids) {
    const int id = seed.first, roi = seed.second;
    if (id < 0 || id >= n_) throw std::runtime_error("Error, seed subset id out of range");
    auto d = info.displacement_map.find(roi);
    if (d == info.displacement_map.end()) throw std::runtime_error("Error, seed without a displacement guess");
    local_field_value(id, SUBSET_DISPLACEMENT_X_FS) = d->second.first;
    local_field_value(id, SUBSET_DISPLACEMENT_Y_FS) = d->second.second;
    auto e = info.normal_strain_map.find(roi);
    if (e != info.normal_strain_map.end()) {
      local_field_value(id, NORMAL_STRETCH_XX_FS) = e->second.first;
      local_field_value(id, NORMAL_STRETCH_YY_FS) = e->second.second;
    }
    auto g = info.shear_strain_map.find(roi);
    if (g != info.shear_strain_map.end()) local_field_value(id, SHEAR_STRETCH_XY_FS) = g->second;
    auto t = info.rotation_map.find(roi);
    if (t != info.rotation_map.end()) local_field_value(id, ROTATION_Z_FS) = t->second;
  }
}

void Schema::update_extents() {
  if (!defs_.empty() || read_full_images_) {
    has_extents_ = false;
    return;
  }
  // in scalar_t = float from INT_MAX / 0, as the reference
  float mnx_r = (float)std::numeric_limits<int>::max(), mxx_r = 0.0f, mny_r = mnx_r, mxy_r = 0.0f;
  float mnx_d = mnx_r, mxx_d = 0.0f, mny_d = mnx_r, mxy_d = 0.0f;
  for (int i = 0; i < n_; ++i) {
    const float x = (float)local_field_value(i, SUBSET_COORDINATES_X_FS);
    const float y = (float)local_field_value(i, SUBSET_COORDINATES_Y_FS);
    const float xd = (float)(local_field_value(i, SUBSET_COORDINATES_X_FS) + local_field_value(i, SUBSET_DISPLACEMENT_X_FS));
    const float yd = (float)(local_field_value(i, SUBSET_COORDINATES_Y_FS) + local_field_value(i, SUBSET_DISPLACEMENT_Y_FS));
    mnx_r = std::min(mnx_r, x); mxx_r = std::max(mxx_r, x);
    mny_r = std::min(mny_r, y); mxy_r = std::max(mxy_r, y);
    mnx_d = std::min(mnx_d, xd); mxx_d = std::max(mxx_d, xd);
    mny_d = std::min(mny_d, yd); mxy_d = std::max(mxy_d, yd);
  }
  const int buffer = 100;
  auto lower = [&](float v) { v -= buffer; return v < buffer ? 0.0f : std::round(v); };
  // the scalar_t extents stored in the int_t vectors
  ref_ext_[0] = (int)lower(mnx_r);
  ref_ext_[1] = (int)(mxx_r + buffer);
  ref_ext_[2] = (int)lower(mny_r);
  ref_ext_[3] = (int)(mxy_r + buffer);
  def_ext_[0] = (int)lower(mnx_d);
  def_ext_[1] = (int)(mxx_d + buffer);
  def_ext_[2] = (int)lower(mny_d);
  def_ext_[3] = (int)(mxy_d + buffer);
  has_extents_ = true;
}

// The sub-image of the frame that the extents cover (DICe_Schema.cpp:200-210, 366-373): a side whose
// extent lies within 100 px of the image edge takes the whole image on that side.
void Schema::read_with_extents(const std::string &file_name, const int *ext, std::vector<float> &pix, int &w,
                               int &h, int &ox, int &oy) const {
  ox = oy = 0;
  if (!has_extents_) {
    read_image(file_name, pix, w, h);
    return;
  }
  int fw = 0, fh = 0;
  read_image_dimensions(file_name, fw, fh);
  const int buffer = 100;
  ox = ext[0] > buffer && ext[0] < fw - buffer ? ext[0] : 0;
  oy = ext[2] > buffer && ext[2] < fh - buffer ? ext[2] : 0;
  const int end_x = ext[1] > buffer && ext[1] < fw - buffer ? ext[1] : fw;
  const int end_y = ext[3] > buffer && ext[3] < fh - buffer ? ext[3] : fh;
  read_image(file_name, pix, w, h, ox, oy, end_x - ox, end_y - oy);
}

void Schema::set_ref_image(const std::string &file_name) {
  std::vector<float> pix;
  int w = 0, h = 0;
  read_with_extents(file_name, ref_ext_, pix, w, h, rox_, roy_);
  check(dice_hip_set_image_offsets(ctx_, rox_, roy_, dox_, doy_), "set_image_offsets");
  set_image(true, w, h, pix.data(), DICE_STORAGE_F32);
  ref_name_ = file_name;
}
void Schema::set_def_image(const std::string &file_name) {
  std::vector<float> pix;
  int w = 0, h = 0;
  read_with_extents(file_name, def_ext_, pix, w, h, dox_, doy_);
  check(dice_hip_set_image_offsets(ctx_, rox_, roy_, dox_, doy_), "set_image_offsets");
  set_image(false, w, h, pix.data(), DICE_STORAGE_F32);
  def_name_ = file_name;
}

int Schema::execute_correlation() {
  if (n_ <= 0) throw std::runtime_error("Error, the schema has no subsets");
  const int mode = backend_ == HIP_PLAIN ? DICE_MODE_PLAIN : DICE_MODE_BATCHED;
  const int first = frame_id_ == first_frame_id_ ? 1 : 0;
  upload_force_flags();
  // TRACKING_ROUTINE with conformal blockers: the subsets in id order, each blocked by its blockers'
  // current positions (a lower-id blocker's result of this frame).  The subsets go in runs in which no
  // subset has a conformal blocker earlier in the same run; each run is correlated with the masks its
  // blockers give at the start of the run, and only the run's results are kept.
  bool rounds = false;
  if (correlation_routine() == TRACKING_ROUTINE)
    for (const auto &kv : obstructing_ids_)
      for (int b : kv.second) rounds = rounds || defs_.count(b) != 0;
  if (!rounds) {
    check(dice_hip_correlate(ctx_, mode, fields_.data(), first), "execute_correlation");
  } else {
    std::vector<double> work;
    int begin = 0;
    while (begin < n_) {
      int end = begin + 1;
      for (; end < n_; ++end) {
        auto it = obstructing_ids_.find(end);
        bool dep = false;
        if (it != obstructing_ids_.end())
          for (int b : it->second) dep = dep || (b >= begin && b < end && defs_.count(b) != 0);
        if (dep) break;
      }
      std::vector<std::vector<std::pair<int, int>>> blocked(n_);
      for (int i = begin; i < end; ++i) blocked[i] = blocking_pixels(i);
      upload_masks(blocked);
      work = fields_;
      check(dice_hip_correlate(ctx_, mode, work.data(), first), "execute_correlation");
      for (int s = 0; s < DICE_F_COUNT; ++s)
        for (int i = begin; i < end; ++i) fields_[(size_t)s * n_ + i] = work[(size_t)s * n_ + i];
      begin = end;
    }
    upload_masks({});
  }
  // post_execution_tasks: the post-processors of the frame
  if (vsg_window_ != 0) {
    vsg_.assign((size_t)DICE_VSG_COUNT * n_, 0.0);
    check(dice_hip_vsg_strain(ctx_, vsg_window_, fields_.data(), vsg_.data()), "VSG strain");
  }
  if (uncertainty_pp_) {
    unc_.assign(2 * (size_t)n_, 0.0);
    check(dice_hip_uncertainty(ctx_, fields_.data(), unc_.data()), "uncertainty");
  }
  ++frame_id_;  // update_frame_id
  return 0;
}

void Schema::set_vsg_strain_window(int window_size_in_pixels) {
  if (window_size_in_pixels < -64)
    throw std::invalid_argument("Error, k-nearest VSG windows are limited to 64 neighbours");
  vsg_window_ = window_size_in_pixels;
}

double &Schema::local_field_value(int local_id, Field_Spec fs) {
  if (local_id < 0 || local_id >= n_) throw std::runtime_error("Error, subset id out of range");
  if ((int)fs >= 110) {
    if (unc_.empty()) unc_.assign(2 * (size_t)n_, 0.0);
    return unc_[(size_t)((int)fs - 110) * n_ + local_id];
  }
  if ((int)fs >= 100) {
    if (vsg_.empty()) vsg_.assign((size_t)DICE_VSG_COUNT * n_, 0.0);
    return vsg_[(size_t)((int)fs - 100) * n_ + local_id];
  }
  return fields_[(size_t)fs * n_ + local_id];
}

double Schema::local_field_value(int local_id, Field_Spec fs) const {
  if (local_id < 0 || local_id >= n_) throw std::runtime_error("Error, subset id out of range");
  if ((int)fs >= 110) return unc_.empty() ? 0.0 : unc_[(size_t)((int)fs - 110) * n_ + local_id];
  if ((int)fs >= 100) return vsg_.empty() ? 0.0 : vsg_[(size_t)((int)fs - 100) * n_ + local_id];
  return fields_[(size_t)fs * n_ + local_id];
}

namespace {
const char *interp_name(int m) { return m == 0 ? "BILINEAR" : m == 1 ? "BICUBIC" : m == 2 ? "KEYS_FOURTH" : "UNKNOWN"; }
const char *opt_name(int m) {
  static const char *n[] = {"SIMPLEX", "GRADIENT_BASED", "GRADIENT_THEN_SEARCH", "SIMPLEX_THEN_GRADIENT_BASED",
                            "GRADIENT_BASED_THEN_SIMPLEX"};
  return m >= 0 && m < 5 ? n[m] : "UNKNOWN";
}
int digits(int v) {
  int d = 0;
  if (v == 0) return 1;
  while (v) { v /= 10; ++d; }
  return d;
}
}  // namespace

Output_Spec::Output_Spec(const Schema *schema, bool omit_row_id, const std::vector<std::string> &fields,
                         const std::string &delimiter)
    : schema_(schema), omit_row_id_(omit_row_id), delimiter_(delimiter) {
  if (fields.empty()) {
    // DICe_Schema.cpp:4683-4696 with the affine shape function
    names_ = {"COORDINATE_X", "COORDINATE_Y", "DISPLACEMENT_X", "DISPLACEMENT_Y", "ROTATION_Z",
              "NORMAL_STRAIN_X", "NORMAL_STRAIN_Y", "SHEAR_STRAIN_XY", "SIGMA", "STATUS_FLAG"};
  } else {
    names_ = fields;
  }
  for (std::string &n : names_) {
    for (char &c : n) c = (char)std::toupper((unsigned char)c);
    specs_.push_back(field_spec_from_name(n));
  }
}

void Output_Spec::write_info(std::FILE *file) const {
  // Output_Spec::write_info (DICe_Schema.cpp:4795-4856) for a local analysis;
  // the first line is the masthead, which the comparison skips
  const Schema &s = *schema_;
  std::fprintf(file, "***\n");
  std::fprintf(file, "*** Digital Image Correlation Engine (DICe) format, written by the MI355X subset-correlation library\n");
  std::fprintf(file, "***\n");
  std::fprintf(file, "*** Reference image: %s \n", s.ref_image_file_name().c_str());
  std::fprintf(file, "*** Deformed image: %s \n", s.def_image_file_name().c_str());
  std::fprintf(file, "*** DIC method : local \n");
  std::fprintf(file, "*** Correlation method: ZNSSD\n");
  std::fprintf(file, "*** Interpolation method: %s\n", interp_name(s.interpolation_method()));
  std::fprintf(file, "*** Image gradient method: %s\n",
               s.gradient_method() == CONVOLUTION_5_POINT ? "CONVOLUTION_5_POINT" : "FINITE_DIFFERENCE");
  std::fprintf(file, "*** Optimization method: %s\n", opt_name(s.optimization_method()));
  std::fprintf(file, "*** Projection method: DISPLACEMENT_BASED\n");
  static const char *const init_names[] = {"USE_FIELD_VALUES", "USE_NEIGHBOR_VALUES",
                                           "USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY"};
  const int im = s.initialization_method();
  std::fprintf(file, "*** Guess initialization method: %s\n", im >= 0 && im <= 2 ? init_names[im] : "UNKNOWN");
  std::fprintf(file, "*** Seed location: N/A\n");
  std::fprintf(file, "*** Shape functions: ");
  if (s.translation_enabled()) std::fprintf(file, "Translation (u,v) ");
  if (s.rotation_enabled()) std::fprintf(file, "Rotation (theta) ");
  if (s.normal_strain_enabled()) std::fprintf(file, "Normal Strain (ex,ey) ");
  if (s.shear_strain_enabled()) std::fprintf(file, "Shear Strain (gamma_xy) ");
  std::fprintf(file, "\n");
  std::fprintf(file, "*** Incremental correlation: false\n");
  std::fprintf(file, "*** Subset size: %i\n", s.subset_dim());
  std::fprintf(file, "*** Step size: x %i y %i (-1 implies not regular grid)\n", s.step_size_x(), s.step_size_y());
  // the first strain post-processor's window (DICe_Schema.cpp:4843-4847)
  if (s.strain_window_size() != 0 && s.strain_window_size() != -1)
    std::fprintf(file, "*** Strain window size in pixels: %i (only first strain post-processor is reported)\n",
                 s.strain_window_size());
  else
    std::fprintf(file, "*** Strain window: N/A\n");
  std::fprintf(file, "*** Coordinates given with (0,0) as upper left corner of image, x positive right, y positive down\n");
  std::fprintf(file, "***\n");
}

void Output_Spec::write_header(std::FILE *file, const std::string &row_id) const {
  if (!omit_row_id_) std::fprintf(file, "%s%s", row_id.c_str(), delimiter_.c_str());
  for (size_t i = 0; i < names_.size(); ++i)
    std::fprintf(file, "%s%s", i == 0 ? "" : delimiter_.c_str(), names_[i].c_str());
  std::fprintf(file, "\n");
}

void Output_Spec::write_frame(std::FILE *file, int row_index, int subset) const {
  // values in the reference's scalar_t (float) with %4.4E
  if (!omit_row_id_) std::fprintf(file, "%i%s", row_index, delimiter_.c_str());
  for (size_t i = 0; i < specs_.size(); ++i) {
    const float v = (float)schema_->local_field_value(subset, specs_[i]);
    std::fprintf(file, "%s%4.4E", i == 0 ? "" : delimiter_.c_str(), (double)v);
  }
  std::fprintf(file, "\n");
}

void Schema::write_output(const std::string &output_folder, const std::string &prefix,
                          bool separate_files_per_subset, const std::vector<std::string> &fields,
                          const std::string &delimiter, bool omit_row_id, int num_frames) const {
  if (n_ <= 0) throw std::runtime_error("Error, the schema has no subsets");
  const Output_Spec spec(this, omit_row_id, fields, delimiter);
  const int frame = frame_id_ - 1;  // written after execute_correlation advanced the frame
  auto open = [](const std::string &name, const char *mode) {
    std::FILE *f = std::fopen(name.c_str(), mode);
    if (!f) throw std::runtime_error("Error, cannot write " + name);
    return f;
  };
  if (separate_files_per_subset) {
    const int total = digits(n_);
    for (int i = 0; i < n_; ++i) {
      std::string name = output_folder + prefix + "_" + std::string(total - digits(i), '0') + std::to_string(i) +
                         ".txt";
      if (frame == first_frame_id_) {
        std::FILE *f = open(name, "w");
        spec.write_info(f);
        spec.write_header(f, "FRAME");
        std::fclose(f);
      }
      std::FILE *f = open(name, "a");
      spec.write_frame(f, frame, i);
      std::fclose(f);
    }
  } else {
    int zeros = 0;
    if (num_frames > 0) zeros = std::max(0, digits(first_frame_id_ + num_frames) - digits(frame));
    const std::string name = output_folder + prefix + "_" + std::string(zeros, '0') + std::to_string(frame) + ".txt";
    std::FILE *f = open(name, "w");
    spec.write_info(f);
    spec.write_header(f, "SUBSET_ID");
    std::vector<int> order(n_);
    for (int i = 0; i < n_; ++i) order[i] = i;
    if (sort_txt_output_)
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double xa = fields_[(size_t)DICE_F_COORD_X * n_ + a], xb = fields_[(size_t)DICE_F_COORD_X * n_ + b];
        const double ya = fields_[(size_t)DICE_F_COORD_Y * n_ + a], yb = fields_[(size_t)DICE_F_COORD_Y * n_ + b];
        return xa < xb || (xa == xb && ya < yb);
      });
    for (int i : order) spec.write_frame(f, i, i);
    std::fclose(f);
  }
}

int diff_files(const std::string &file_a, const std::string &file_b, double rel_tol, double floor, bool use_floor,
               bool numerical_only, std::string *report) {
  std::ifstream a(file_a), b(file_b);
  std::ostringstream rep;
  int errors = 0;
  if (!a.good() || !b.good()) {
    if (report) *report = "Error, cannot read " + (a.good() ? file_b : file_a);
    return 1;
  }
  std::string la, lb;
  int line = 0;
  while (std::getline(a, la)) {
    ++line;
    if (!std::getline(b, lb)) {
      rep << "Error, File A has more lines than FileB\n";
      ++errors;
      break;
    }
    const std::vector<std::string> ta = tokenize(la, " ,\r"), tb = tokenize(lb, " ,\r");
    if (ta.size() >= 2 && ta[1].find("DIGITAL") != std::string::npos) continue;  // the masthead
    if (ta.size() != tb.size()) {
      rep << "Error, Different number of tokens per line " << line << " A:" << ta.size() << " B: " << tb.size() << "\n";
      ++errors;
      break;
    }
    bool line_diff = false;
    for (size_t i = 0; i < ta.size(); ++i) {
      if (is_number(ta[i])) {
        if (!is_number(tb[i])) { line_diff = true; continue; }
        const double va = std::strtod(ta[i].c_str(), nullptr), vb = std::strtod(tb[i].c_str(), nullptr);
        const double diff = std::abs((va - vb) / va);
        const bool tiny = std::abs(va) + std::abs(vb) < 1.0e-8;
        const bool below_floor = std::abs(va) < floor;
        if ((!below_floor || !use_floor) && !tiny && diff > rel_tol) line_diff = true;
      } else if (ta[i] != tb[i] && !numerical_only) {
        line_diff = true;
      }
    }
    if (line_diff) {
      ++errors;
      rep << "< " << line << ": " << la << "\n> " << line << ": " << lb << "\n";
    }
  }
  if (report) *report = rep.str();
  return errors;
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

extern "C" int dice_correlate_conformal(float points[], float ref_img[], int ref_w, int ref_h, float def_img[],
                                        int def_w, int def_h, const char *subset_file, const char *param_file,
                                        bool write_output) {
  // DICe_api.cpp:155-282: the schema, its subsets and reference image persist from the first call
  using namespace dice_hip;
  static std::mutex m;
  static std::unique_ptr<Schema> schema;
  std::lock_guard<std::mutex> lock(m);
  if (ref_w != def_w || ref_h != def_h) return -1;
  if (!points || !ref_img || !def_img || !subset_file) return -1;
  const int stride = 9;
  try {
    if (!schema) {
      ParameterList params;
      tracking_default_params(&params);
      if (param_file) {
        // the user's parameters over the tracking defaults (use_tracking_default_params set, :196-199)
        const ParameterList user = read_correlation_params(param_file);
        for (const auto &kv : user.values()) params.set(kv.first, kv.second);
      }
      const Subset_File_Info info = read_subset_file(subset_file, ref_w, ref_h);
      const int n = (int)info.coordinates_x.size();
      if ((int)info.conformal_area_defs.size() != n)
        throw std::runtime_error("Error there is a mismatch between the number of conformal subsets defined and the "
                                 "number of coordinate sets");
      std::unique_ptr<Schema> s(new Schema(info.coordinates_x, info.coordinates_y, -1, &info.conformal_area_defs,
                                           &params));
      s->set_obstructing_subset_ids(info.blocking_subset_ids);
      s->set_force_simplex(info.force_simplex);
      s->set_ref_image(ref_w, ref_h, ref_img);
      schema = std::move(s);
    }
    const int n = schema->local_num_subsets();
    // the centroids from the subset file
    for (int i = 0; i < n; ++i) {
      points[i * stride + 0] = (float)schema->local_field_value(i, SUBSET_COORDINATES_X_FS);
      points[i * stride + 1] = (float)schema->local_field_value(i, SUBSET_COORDINATES_Y_FS);
    }
    schema->set_def_image(ref_w, ref_h, def_img);
    const Field_Spec slots[7] = {SUBSET_DISPLACEMENT_X_FS, SUBSET_DISPLACEMENT_Y_FS, ROTATION_Z_FS, SIGMA_FS,
                                 GAMMA_FS, BETA_FS, STATUS_FLAG_FS};
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 7; ++k) schema->local_field_value(i, slots[k]) = points[i * stride + 2 + k];
    schema->execute_correlation();
    if (write_output) schema->write_output("./");
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 7; ++k) points[i * stride + 2 + k] = (float)schema->local_field_value(i, slots[k]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "dice_correlate_conformal: %s\n", e.what());
    return -1;
  }
  return 0;
}
