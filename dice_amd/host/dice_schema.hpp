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
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dice_hip.h"
#include "dice_image_io.hpp"

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
enum Initialization_Method {
  USE_FIELD_VALUES = 0,
  USE_NEIGHBOR_VALUES = 1,
  USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY = 2
};
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
  FIELD_7_FS = DICE_F_FIELD_7,
  // fields of the VSG strain post-processor (DICe_FieldEnums.h), held beside the correlation fields
  VSG_STRAIN_XX_FS = 100 + DICE_VSG_STRAIN_XX,
  VSG_STRAIN_YY_FS = 100 + DICE_VSG_STRAIN_YY,
  VSG_STRAIN_XY_FS = 100 + DICE_VSG_STRAIN_XY,
  VSG_DUDX_FS = 100 + DICE_VSG_DUDX,
  VSG_DUDY_FS = 100 + DICE_VSG_DUDY,
  VSG_DVDX_FS = 100 + DICE_VSG_DVDX,
  VSG_DVDY_FS = 100 + DICE_VSG_DVDY,
  // fields of the uncertainty post-processor
  UNCERTAINTY_FS = 110,
  UNCERTAINTY_ANGLE_FS = 111
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

// read_correlation_params (DICe_Parser.cpp:353-430): the <Parameter name=".." type=".." value=".."/>
// entries of a Teuchos XML parameter file, names lower-cased, enum-valued strings (correlation_routine,
// interpolation_method, optimization_method, initialization_method, gradient_method) translated to
// their values, bools as 0 / 1.  use_tracking_default_params is returned in *use_tracking_defaults
// (may be null).  Parameters that only switch the reference's image output are accepted and ignored;
// shape_function_type must be AFFINE; any other parameter this path does not read, a nested
// ParameterList or an unknown enum string throws std::invalid_argument naming it.
ParameterList read_correlation_params(const std::string &file_name, bool *use_tracking_defaults = nullptr);

// Shapes of conformal subsets (DICe_Shape.h:50-200): the pixels a shape owns,
// as (y, x) pairs in (y, x) order (std::set<std::pair<int,int>> in the
// reference), and clearing flags[j] for the pixels (x[j], y[j]) it covers.
class Affine_Shape_Function;
class Shape {
 public:
  virtual ~Shape() = default;
  virtual std::vector<std::pair<int, int>> get_owned_pixels() const = 0;
  // The pixels of the shape moved by a blocking subset's shape function about its centroid (cx, cy)
  // and stretched by skin_factor about the moved vertices' centroid (get_owned_pixels(shape_function,
  // cx, cy, skin_factor), DICe_Shape.cpp:136-219, 322-420).
  virtual std::vector<std::pair<int, int>> get_owned_pixels(const Affine_Shape_Function &sf, int cx, int cy,
                                                            double skin_factor) const = 0;
  virtual void deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const = 0;
};

// Polygon with integer vertices (DICe_Shape.cpp:50-219): the pixels it owns
// are those whose winding angle around the vertices is at least pi.
class Polygon : public Shape {
 public:
  Polygon(const std::vector<int> &coords_x, const std::vector<int> &coords_y);
  std::vector<std::pair<int, int>> get_owned_pixels() const override;
  std::vector<std::pair<int, int>> get_owned_pixels(const Affine_Shape_Function &sf, int cx, int cy,
                                                    double skin_factor) const override;
  void deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const override;

 private:
  bool inside(int x, int y) const;
  std::vector<int> vx_, vy_;  // closed: vertex 0 repeated at the end
  int min_x_, min_y_, max_x_, max_y_;
};

// Circle (DICe_Shape.cpp:221-279): integer pixels with (x - cx)^2 + (y - cy)^2
// <= radius^2, searched over [c - radius - 1, c + radius + 1].
class Circle : public Shape {
 public:
  Circle(int centroid_x, int centroid_y, double radius);
  std::vector<std::pair<int, int>> get_owned_pixels() const override;
  // the reference has no deformed circle: throws std::runtime_error, as it does
  std::vector<std::pair<int, int>> get_owned_pixels(const Affine_Shape_Function &sf, int cx, int cy,
                                                    double skin_factor) const override;
  void deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const override;

 private:
  int cx_, cy_;
  double radius2_;
  int min_x_, max_x_, min_y_, max_y_;
};

// Rectangle (DICe_Shape.cpp:281-420, undeformed): origin = centroid -
// size / 2 (integer division), pixels [origin, origin + size) on each axis.
class Rectangle : public Shape {
 public:
  Rectangle(int centroid_x, int centroid_y, int width, int height);
  std::vector<std::pair<int, int>> get_owned_pixels() const override;
  std::vector<std::pair<int, int>> get_owned_pixels(const Affine_Shape_Function &sf, int cx, int cy,
                                                    double skin_factor) const override;
  void deactivate_pixels(int size, std::vector<uint8_t> &flags, const int *x, const int *y) const override;

 private:
  int origin_x_, origin_y_, width_, height_;
};

typedef std::vector<std::shared_ptr<Shape>> multi_shape;

// Conformal_Area_Def (DICe_Shape.h): the union of the boundary shapes, with
// excluded areas whose pixels stay in the subset but are inactive
// (DICe_Subset.cpp:115-168), and obstructed areas: static image pixels that
// deactivate any warped subset pixel rounding onto them (:162-167, 412-415).
struct Conformal_Area_Def {
  multi_shape boundary;
  multi_shape excluded_area;
  multi_shape obstructed_area;
};

// The subset file of the reference (read_subset_file, DICe_Parser.cpp:1099-1700):
// SUBSET_COORDINATES and CONFORMAL_SUBSET blocks (SUBSET_ID; BOUNDARY,
// EXCLUDED and OBSTRUCTED blocks of POLYGON / CIRCLE / RECTANGLE shapes;
// BLOCKING_SUBSETS; FORCE_SIMPLEX), and REGION_OF_INTEREST blocks (BOUNDARY,
// EXCLUDED, and a SEED with LOCATION, DISPLACEMENT and optional NORMAL_STRAIN,
// SHEAR_STRAIN, ROTATION).  Keywords are case-insensitive, '#' starts a
// comment.  The blocks this path has no use for (motion windows, paths,
// skip-solve lists, the boundary conditions and IC values of global DIC) throw
// std::runtime_error naming them.  width / height (or -1) bound the
// coordinates as in the reference.
struct Subset_File_Info {
  // REGION_OF_INTEREST_INFO: conformal_area_defs are regions of interest (keyed
  // by region number) for create_regular_grid_of_correlation_points, not subsets
  bool region_of_interest = false;
  std::vector<double> coordinates_x, coordinates_y;
  std::map<int, Conformal_Area_Def> conformal_area_defs;
  std::map<int, std::vector<int>> blocking_subset_ids;
  std::vector<int> force_simplex;
  // seeds of the regions of interest, keyed by region (the reference's size_map
  // holds the seed locations)
  std::map<int, std::pair<int, int>> size_map;
  std::map<int, std::pair<double, double>> displacement_map, normal_strain_map;
  std::map<int, double> shear_strain_map, rotation_map;
  // subset id of each region's seed -> region, filled by create_regular_grid_of_correlation_points
  std::map<int, int> seed_subset_ids;
};
Subset_File_Info read_subset_file(const std::string &file_name, int width = -1, int height = -1);

// create_regular_grid_of_correlation_points (DICe_Decomp.cpp:868-1070) without
// a gradient threshold: in each region of interest (the whole image when
// roi_defs is null or empty), centres at subset_size - 1 + k * step_size that
// lie at least a subset inside the image, inside the region, with no subset
// corner in an excluded area.  Per region: the seed point first (the grid point
// nearest the seed location, or row 0 / column 0 without a seed), then the
// serpentine to the right of it (down the seed's column, up the next, ...),
// then the one to the left (up the seed's column, down the previous, ...).
// neighbor_ids: -1 for the seed point, the seed for the first point of each
// serpentine (-1 without a seed), else the previous point.  seeds (optional):
// the regions' seed locations (size_map); seed_subset_ids is filled.  Throws
// std::runtime_error when a seed's grid point is not a valid correlation point.
void create_regular_grid_of_correlation_points(int img_w, int img_h, int subset_size, int step_size,
                                               std::vector<double> &points_x, std::vector<double> &points_y,
                                               std::vector<int> &neighbor_ids,
                                               const std::map<int, Conformal_Area_Def> *roi_defs = nullptr,
                                               Subset_File_Info *seeds = nullptr);

// The pixel list of a conformal subset (Subset::Subset(cx, cy, Conformal_Area_Def),
// DICe_Subset.cpp:115-168): the union of the boundary shapes' pixels in (y, x)
// order, negative coordinates clamped to 0, active unless an excluded shape
// covers the pixel.
void conformal_subset_pixels(const Conformal_Area_Def &def, std::vector<int> &x, std::vector<int> &y,
                             std::vector<uint8_t> &active);

// Output names of the Field_Specs (DICe_FieldEnums.cpp:94-171), e.g.
// DISPLACEMENT_X for SUBSET_DISPLACEMENT_X_FS.  field_spec_from_name throws
// std::invalid_argument for a name this path does not produce.
const char *field_name(Field_Spec fs);
Field_Spec field_spec_from_name(const std::string &name);

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
  // The post_process_vsg_strain block of the correlation parameters
  // (strain_window_size_in_pixels, DICe_PostProcessor.cpp:877-906): after every
  // execute_correlation the VSG strains of the frame (VSG_*_FS fields).  0: no
  // VSG post-processor; a negative size -k: the k nearest subsets (k <= 64,
  // std::invalid_argument beyond).
  void set_vsg_strain_window(int window_size_in_pixels);
  int strain_window_size() const { return vsg_window_; }
  // The post_process_uncertainty block (Uncertainty_Post_Processor,
  // DICe_PostProcessor.cpp:1494-1537): UNCERTAINTY_FS and UNCERTAINTY_ANGLE_FS
  // after every execute_correlation.
  void set_uncertainty_post_processor(bool on) { uncertainty_pp_ = on; }
  // sort_txt_output (DICe_Schema.cpp:4446-4466): write_output's one-file-per-frame
  // rows in order of x, then y
  void set_sort_txt_output(bool sort) { sort_txt_output_ = sort; }
  // The id of the first frame (the reference's first_frame_id_, from
  // start_image_index): frame ids and output file numbers count from it.
  void set_first_frame_index(int index) { first_frame_id_ = frame_id_ = index; }
  // Schema::set_obstructing_subset_ids (DICe_Schema.h:1076-1082): subset id -> the ids of the subsets
  // that can block it (BLOCKING_SUBSETS of a subset file).  Every subset in such a list, and every
  // subset with a non-empty list, runs Nelder-Mead only unless override_force_simplex is set
  // (generic_correlation_routine, DICe_Schema.cpp:2160-2185); under TRACKING_ROUTINE the pixels of a
  // conformal blocker, moved by its current shape function and stretched by obstruction_skin_factor,
  // are blocked for the subsets it blocks (check_for_blocking_subsets, :4529-4556), and the subsets are
  // correlated in id order (a blocker with a lower id blocks with this frame's result).
  void set_obstructing_subset_ids(const std::map<int, std::vector<int>> &ids);
  // Schema::set_force_simplex (DICe_Schema.h:1083-1086): FORCE_SIMPLEX subsets run Nelder-Mead only.
  void set_force_simplex(const std::vector<int> &ids);
  // NEIGHBOR_ID_FS of every subset (Schema::initialize, DICe_Schema.cpp:1298-1303), the
  // source of the initial guess under USE_NEIGHBOR_VALUES.
  void set_neighbor_ids(const std::vector<int> &neighbor_ids);
  // The seeds' initial values (Schema::initialize, DICe_Schema.cpp:1111-1143): the
  // displacement, and the strains and rotation when given, of each seed subset.
  void set_seed_values(const Subset_File_Info &info);
  // Schema::set_ref_image(const std::string &) / set_def_image(const std::string &)
  // (DICe_Schema.cpp:351-394, 179-304): the whole image read by read_image
  // (.rawi or TIFF), filtered and differentiated as the parameters say; the
  // name goes into the output header.
  void set_ref_image(const std::string &file_name);
  void set_def_image(const std::string &file_name);
  // Schema::update_extents (DICe_Schema.cpp:886-985): unless the subsets are conformal or
  // read_full_images is set, the next file-based set_ref_image / set_def_image read only the part of
  // the frame the subsets reach -- the bbox of the centroids (reference) or of the centroids moved by
  // their displacements (deformed), widened by 100 px, the whole extent on a side within 100 px of the
  // image edge (:200-210, 366-373) -- and hand the sub-image to the device with its offsets
  // (dice_hip_set_image_offsets).  The frame loop calls it before each image, as DICe_Main.cpp does.
  void update_extents();
  bool has_extents() const { return has_extents_; }

  // Schema::execute_correlation (DICe_Schema.cpp:1622-1792): all subsets of the
  // frame, then the frame id advances.  Returns 0.
  int execute_correlation();

  double &local_field_value(int local_id, Field_Spec fs);
  double local_field_value(int local_id, Field_Spec fs) const;
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
  Gradient_Method gradient_method() const { return (Gradient_Method)p_.gradient_method; }
  bool translation_enabled() const { return p_.enable_translation != 0; }
  bool rotation_enabled() const { return p_.enable_rotation != 0; }
  bool normal_strain_enabled() const { return p_.enable_normal_strain != 0; }
  bool shear_strain_enabled() const { return p_.enable_shear_strain != 0; }
  int step_size_x() const { return step_x_; }
  int step_size_y() const { return step_y_; }
  // the step sizes reported in the output header for subsets laid on a grid (-1: not a regular grid)
  void set_step_sizes(int step_x, int step_y) {
    step_x_ = step_x;
    step_y_ = step_y;
  }
  dice_hip_context *context() { return ctx_; }

  // File names written into the header block of the output (the reference
  // takes them from its Image objects).
  void set_image_file_names(const std::string &ref_name, const std::string &def_name) {
    ref_name_ = ref_name;
    def_name_ = def_name;
  }
  const std::string &ref_image_file_name() const { return ref_name_; }
  const std::string &def_image_file_name() const { return def_name_; }

  // Schema::write_output (DICe_Schema.cpp:4327-4470), text files only, called
  // after execute_correlation: with separate_files_per_subset, one file
  // <folder><prefix>_<id>.txt per subset (ids zero-padded to the digits of the
  // subset count), the header block and the field names on the first frame,
  // then one row per frame ("FRAME" column); otherwise one file per frame,
  // <folder><prefix>_<frame>.txt (frame zero-padded to the digits of
  // num_frames when num_frames > 0), with one row per subset ("SUBSET_ID").
  // fields: output names (Output_Spec order); empty = the reference's default
  // list.  Throws std::runtime_error if a file cannot be written.
  void write_output(const std::string &output_folder, const std::string &prefix = "DICe_solution",
                    bool separate_files_per_subset = false, const std::vector<std::string> &fields = {},
                    const std::string &delimiter = " ", bool omit_row_id = false, int num_frames = 0) const;

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
  int step_x_ = -1, step_y_ = -1;
  std::string ref_name_, def_name_;
  std::vector<double> fields_;
  int vsg_window_ = 0;
  bool sort_txt_output_ = false;
  std::vector<double> vsg_;  // DICE_VSG_COUNT x n
  bool uncertainty_pp_ = false;
  std::vector<double> unc_;  // 2 x n: UNCERTAINTY, UNCERTAINTY_ANGLE
  // obstruction and blocking
  std::map<int, Conformal_Area_Def> defs_;                  // conformal subset definitions
  std::vector<std::vector<std::pair<int, int>>> obstructed_;  // per subset: (x, y) of its obstructed pixels
  std::map<int, std::vector<int>> obstructing_ids_;
  std::vector<int> force_simplex_ids_;
  bool override_force_simplex_ = false;
  double skin_factor_ = 1.0;
  // sub-image extents (x min, x max, y min, y max) and the images' offsets
  bool read_full_images_ = false, has_extents_ = false;
  int ref_ext_[4] = {0, 0, 0, 0}, def_ext_[4] = {0, 0, 0, 0};  // std::vector<int_t> in the reference
  int rox_ = 0, roy_ = 0, dox_ = 0, doy_ = 0;
  void read_with_extents(const std::string &file_name, const int *ext, std::vector<float> &pix, int &w, int &h,
                         int &ox, int &oy) const;
  void upload_force_flags();
  void upload_masks(const std::vector<std::vector<std::pair<int, int>>> &blocked);
  std::vector<std::pair<int, int>> blocking_pixels(int subset) const;
};

// Status_Flag values of DICe.h:609-677 that the correlation records.
enum Status_Flag {
  CORRELATION_SUCCESSFUL = 0,
  INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL = 1,
  INITIALIZE_USING_NEIGHBOR_VALUE_SUCCESSFUL = 3,
  INITIALIZE_SUCCESSFUL = 4,
  INITIALIZE_FAILED = 5,
  CORRELATION_FAILED = 8,
  MAX_ITERATIONS_REACHED = 11,
  INITIALIZE_FAILED_BY_EXCEPTION = 12,
  JUMP_TOLERANCE_EXCEEDED = 16
};

// Affine_Shape_Function (DICe_LocalShapeFunction.h:266-398): the six parameters
// u, v, theta, ex, ey, gxy, with the disabled ones held at 0.
class Affine_Shape_Function {
 public:
  Affine_Shape_Function(bool enable_rotation = true, bool enable_normal_strain = true,
                        bool enable_shear_strain = true);
  double u() const { return p_[0]; }
  double v() const { return p_[1]; }
  double theta() const { return p_[2]; }
  double ex() const { return p_[3]; }
  double ey() const { return p_[4]; }
  double gxy() const { return p_[5]; }
  // sets a parameter (index 0..5 in the order above); a disabled one stays 0
  void set(int index, double value);
  const double *parameters() const { return p_; }
  // initialize_parameters_from_fields / save_fields (DICe_LocalShapeFunction.cpp:251-312, 65-74)
  void initialize_parameters_from_fields(const Schema *schema, int subset_gid);
  void save_fields(Schema *schema, int subset_gid) const;

 private:
  friend class Objective_ZNSSD;
  double p_[6] = {0, 0, 0, 0, 0, 0};
  bool en_[6];
};

// Objective_ZNSSD (DICe_Objective.h:82-242) for one subset of a Schema whose
// images and parameters are set: the calls run on the GPU, one workgroup each
// (dice_hip_objective).  computeUpdateFast stores the subset's CONDITION_NUMBER and
// uncertainty fields (FIELD_1..7) in the Schema, as the reference does.
class Objective_ZNSSD {
 public:
  Objective_ZNSSD(Schema *schema, int correlation_point_global_id);
  int correlation_point_global_id() const { return gid_; }
  Status_Flag computeUpdateFast(Affine_Shape_Function &shape_function, int &num_iterations);
  Status_Flag computeUpdateRobust(Affine_Shape_Function &shape_function, int &num_iterations);
  double gamma(const Affine_Shape_Function &shape_function) const;
  double sigma(const Affine_Shape_Function &shape_function, double &noise_level) const;

 private:
  double call(int which, const Affine_Shape_Function &sf, double *params, double *result) const;
  Schema *schema_;
  int gid_;
};

// Output_Spec (DICe_Schema.cpp:4673-4945, text output of a local analysis):
// the header block, the field-name row and the %4.4E rows.
class Output_Spec {
 public:
  // fields empty: COORDINATE_X, COORDINATE_Y, DISPLACEMENT_X, DISPLACEMENT_Y,
  // ROTATION_Z, NORMAL_STRAIN_X, NORMAL_STRAIN_Y, SHEAR_STRAIN_XY, SIGMA,
  // STATUS_FLAG (the affine shape function's default list)
  Output_Spec(const Schema *schema, bool omit_row_id = false, const std::vector<std::string> &fields = {},
              const std::string &delimiter = " ");
  void write_info(std::FILE *file) const;
  void write_header(std::FILE *file, const std::string &row_id) const;
  void write_frame(std::FILE *file, int row_index, int subset) const;
  const std::vector<std::string> &field_names() const { return names_; }

 private:
  const Schema *schema_;
  bool omit_row_id_;
  std::string delimiter_;
  std::vector<std::string> names_;
  std::vector<Field_Spec> specs_;
};

// DICe_Diff (tools/DICe_Diff.cpp:60-360, serial files): line by line, the
// masthead line skipped; numbers differ when |(a - b) / a| > rel_tol, unless
// |a| + |b| < 1e-8 or (use_floor and |a| < floor); strings must match unless
// numerical_only.  Returns the number of differing lines (a line-count or
// token-count mismatch, or a missing file, counts as one and stops); report,
// when set, receives the differing lines.
int diff_files(const std::string &file_a, const std::string &file_b, double rel_tol = 1.0e-6, double floor = 0.0,
               bool use_floor = false, bool numerical_only = false, std::string *report = nullptr);

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
