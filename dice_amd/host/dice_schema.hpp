// C++ host mirror of the reference's Schema interface for the local
// subset-correlation path, on top of the C ABI of include/dice_hip.h.
//
// Names, argument meaning and error behaviour follow DICe::Schema
// (src/core/DICe_Schema.h:200-600): configuration errors throw
// std::runtime_error / std::invalid_argument, execute_correlation() returns 0
// and per-subset failures become STATUS_FLAG values; results are read through
// local_field_value(lid, Field_Spec).  Parameters are set by the reference's
// string names (DICe.h:155-285) in a ParameterList.  One process, one GPU:
// local ids equal global ids.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dice_hip.h"

namespace dice_hip {

// Enum values of DICe.h:473-600 (the subset this path uses).
enum Interpolation_Method { BILINEAR = 0, BICUBIC = 1, KEYS_FOURTH = 2 };
enum Optimization_Method {
  SIMPLEX = 0,
  GRADIENT_BASED = 1,
  GRADIENT_THEN_SEARCH = 2,
  SIMPLEX_THEN_GRADIENT_BASED = 3,
  GRADIENT_BASED_THEN_SIMPLEX = 4
};
enum Correlation_Routine { GENERIC_ROUTINE = 0, TRACKING_ROUTINE = 1 };
enum Initialization_Method { USE_FIELD_VALUES = 0 };
enum Gradient_Method { FINITE_DIFFERENCE = 0, CONVOLUTION_5_POINT = 1 };
// Where execute_correlation runs (parameter "correlation_backend"; no
// counterpart in the reference, which has only its CPU loop).
enum Correlation_Backend { HIP_BATCHED = 0, HIP_PLAIN = 1 };

// Field_Spec names of DICe_FieldEnums.h for the fields this path reads and writes.
enum Field_Spec {
  SUBSET_COORDINATES_X_FS = DICE_F_COORD_X,
  SUBSET_COORDINATES_Y_FS = DICE_F_COORD_Y,
  SUBSET_DISPLACEMENT_X_FS = DICE_F_DISP_X,
  SUBSET_DISPLACEMENT_Y_FS = DICE_F_DISP_Y,
  ROTATION_Z_FS = DICE_F_ROT_Z,
  NORMAL_STRETCH_XX_FS = DICE_F_NS_XX,
  NORMAL_STRETCH_YY_FS = DICE_F_NS_YY,
  SHEAR_STRETCH_XY_FS = DICE_F_SS_XY,
  SIGMA_FS = DICE_F_SIGMA,
  GAMMA_FS = DICE_F_GAMMA,
  BETA_FS = DICE_F_BETA,
  MATCH_FS = DICE_F_MATCH,
  NOISE_LEVEL_FS = DICE_F_NOISE,
  CONTRAST_LEVEL_FS = DICE_F_CONTRAST,
  ACTIVE_PIXELS_FS = DICE_F_ACTIVE_PIXELS,
  STATUS_FLAG_FS = DICE_F_STATUS,
  ITERATIONS_FS = DICE_F_ITERATIONS,
  CONDITION_NUMBER_FS = DICE_F_CONDITION,
  NEIGHBOR_ID_FS = DICE_F_NEIGHBOR_ID,
  FIELD_1_FS = DICE_F_FIELD_1,
  FIELD_2_FS = DICE_F_FIELD_2,
  FIELD_3_FS = DICE_F_FIELD_3,
  FIELD_4_FS = DICE_F_FIELD_4,
  FIELD_5_FS = DICE_F_FIELD_5,
  FIELD_6_FS = DICE_F_FIELD_6,
  FIELD_7_FS = DICE_F_FIELD_7
};

// Parameter list keyed by the reference's parameter names.  Values are
// numbers (enums and bools as their integer values).  Unknown names throw
// std::invalid_argument when set.
class ParameterList {
 public:
  ParameterList &set(const std::string &name, double value);
  bool isParameter(const std::string &name) const { return values_.count(name) != 0; }
  double get(const std::string &name) const;
  const std::map<std::string, double> &values() const { return values_; }

 private:
  std::map<std::string, double> values_;
};

// dice_default_params / tracking_default_params (DICe_ParameterUtilities.cpp:297-420).
void dice_default_params(ParameterList *params);
void tracking_default_params(ParameterList *params);

// Polygon with integer vertices (DICe_Shape.cpp:50-219): the pixels it owns
// are those whose winding angle around the vertices is at least pi.
class Polygon {
 public:
  Polygon(const std::vector<int> &coords_x, const std::vector<int> &coords_y);
  // (y, x) pairs in (y, x) order, as std::set<std::pair<int,int>> in the reference
  std::vector<std::pair<int, int>> get_owned_pixels() const;
  // clears flags[j] for the pixels (x[j], y[j]) inside the polygon
  void deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const;

 private:
  bool inside(int x, int y) const;
  std::vector<int> vx_, vy_;  // closed: vertex 0 repeated at the end
  int min_x_, min_y_, max_x_, max_y_;
};

// Conformal_Area_Def (DICe_Shape.h): union of boundary polygons, with excluded
// areas whose pixels stay in the subset but are inactive.
struct Conformal_Area_Def {
  std::vector<Polygon> boundary;
  std::vector<Polygon> excluded_area;
};

class Schema {
 public:
  // Schema(correlation_params) with no subsets yet.
  explicit Schema(const ParameterList *params = nullptr, int device = 0);
  // Regular grid of square subsets (DICe_Schema.cpp:101-139).
  Schema(int roi_width, int roi_height, int step_size_x, int step_size_y, int subset_size,
         const ParameterList *params = nullptr, int device = 0);
  // Explicit centroids (DICe_Schema.cpp:141-150); conformal_subset_defs maps a
  // subset id to its definition, the others are square of side subset_size.
  Schema(const std::vector<double> &coords_x, const std::vector<double> &coords_y, int subset_size,
         const std::map<int, Conformal_Area_Def> *conformal_subset_defs = nullptr,
         const ParameterList *params = nullptr, int device = 0);
  ~Schema();
  Schema(const Schema &) = delete;
  Schema &operator=(const Schema &) = delete;

  void set_params(const ParameterList &params);
  void set_ref_image(int width, int height, const float *intensities);
  void set_ref_image(int width, int height, const uint16_t *intensities);
  void set_def_image(int width, int height, const float *intensities);
  void set_def_image(int width, int height, const uint16_t *intensities);

  // Schema::execute_correlation (DICe_Schema.cpp:1622-1792): all subsets of the
  // frame, then the frame id advances.  Returns 0.
  int execute_correlation();

  double &local_field_value(int local_id, Field_Spec fs);
  double &global_field_value(int global_id, Field_Spec fs) { return local_field_value(global_id, fs); }
  int local_num_subsets() const { return n_; }
  int global_num_subsets() const { return n_; }
  int subset_dim() const { return subset_dim_; }
  int frame_id() const { return frame_id_; }
  int first_frame_id() const { return first_frame_id_; }
  Interpolation_Method interpolation_method() const { return (Interpolation_Method)p_.interpolation_method; }
  Optimization_Method optimization_method() const { return (Optimization_Method)p_.optimization_method; }
  Correlation_Routine correlation_routine() const { return (Correlation_Routine)p_.correlation_routine; }
  Initialization_Method initialization_method() const { return (Initialization_Method)p_.initialization_method; }
  Correlation_Backend correlation_backend() const { return backend_; }
  dice_hip_context *context() { return ctx_; }

 private:
  void create(int device);
  void init_subsets(const std::vector<double> &cx, const std::vector<double> &cy, int subset_size,
                    const std::map<int, Conformal_Area_Def> *defs);
  void check(int rc, const char *what) const;
  template <typename T>
  void set_image(bool ref, int width, int height, const T *pix, int storage);

  dice_hip_context *ctx_ = nullptr;
  dice_params p_{};
  Correlation_Backend backend_ = HIP_BATCHED;
  int n_ = 0;
  int subset_dim_ = -1;
  int frame_id_ = 0;
  int first_frame_id_ = 0;
  int ref_w_ = -1, ref_h_ = -1;
  std::vector<double> fields_;
};

}  // namespace dice_hip

extern "C" {
// Batch rasterisation of polygons by Polygon::get_owned_pixels (OpenMP over
// polygons).  Polygon i has the vertices [vert_offsets[i], vert_offsets[i+1]).
// pixel_counts: owned pixels of each polygon.  pixels: fills x / y of polygon i
// at [pixel_offsets[i], pixel_offsets[i+1]) in (y, x) order.  Return 0, or -1 on
// a polygon with fewer than three vertices.
int dice_polygon_pixel_counts(int n_polygons, const int *vert_offsets, const int *vx, const int *vy,
                              int *pixel_counts);
int dice_polygon_pixels(int n_polygons, const int *vert_offsets, const int *vx, const int *vy,
                        const int64_t *pixel_offsets, int *px, int *py);
}
