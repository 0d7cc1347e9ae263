// C++ tests of the Schema mirror, in the shape of the reference's component
// tests: DICe_TestCorrelateParams.cpp:78-140, DICe_TestCorrelateConformal.cpp:75-135,
// the parallel_subsets regression gold, and the conformal raster golden.
//
//   test_schema correlate_params <ref.f32> <def.f32> <w> <h> <backend 0|1>
//   test_schema correlate_conformal <ref.f32> <def.f32> <w> <h> <backend 0|1>
//   test_schema parallel_subsets <ref.f32> <def.f32> <w> <h> <gold.txt> <backend 0|1>
//   test_schema polygon            (prints the owned pixels "y x" of the golden's polygons)
//   test_schema subset_file_info <subset file>   (what read_subset_file makes of it)
//   test_schema deck <config>      (a regression deck end to end, see deck())
//   test_schema read_params <xml>  (read_correlation_params: "name value" lines)
//   test_schema grid <w> <h> <subset size> <step> [subset file]   (prints "x y neighbor" of the
//                                  regular grid, then "seed id u v" for each seed)
//   test_schema errors             (error behaviour that needs no GPU)
//   test_schema objective <ref.f32> <def.f32> <w> <h> <S> <cx> <cy> <u> <v>   (Objective_ZNSSD calls on
//                                  one subset, affine: gamma, sigma, computeUpdateFast, computeUpdateRobust)
//   test_schema read_image <image file> <out.rawi> [ox oy w h]   (read_image, then write_rawi_image)
//   test_schema tiff_samples <tiff file> <out>   (uint32 w, h, bits, then the uint16 samples)
//
// Images are raw row-major float32.  The exit status is the number of failures.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "dice_schema.hpp"

using namespace dice_hip;

namespace {

std::vector<float> read_raw(const char *path, int w, int h) {
  std::vector<float> v((size_t)w * h);
  FILE *f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(float), v.size(), f) != v.size()) {
    std::cerr << "cannot read " << path << "\n";
    std::exit(100);
  }
  std::fclose(f);
  return v;
}

const double u_exact = 9.8, v_exact = -7.62, t_exact = 0.262, ex_exact = 0.121, ey_exact = 0.121;

// One pass over the four optimizers x two interpolants, as the reference's tests loop.
int optimizer_sweep(Schema &schema, ParameterList &params, double errtol, const char *label) {
  const Optimization_Method opts[] = {SIMPLEX, SIMPLEX_THEN_GRADIENT_BASED, GRADIENT_BASED,
                                      GRADIENT_BASED_THEN_SIMPLEX};
  const Interpolation_Method interps[] = {BILINEAR, KEYS_FOURTH};
  int errors = 0;
  for (Optimization_Method opt : opts)
    for (Interpolation_Method interp : interps) {
      params.set("correlation_routine", GENERIC_ROUTINE);
      params.set("interpolation_method", interp);
      params.set("initialization_method", USE_FIELD_VALUES);
      params.set("optimization_method", opt);
      schema.set_params(params);
      schema.local_field_value(0, SUBSET_COORDINATES_X_FS) = 30;
      schema.local_field_value(0, SUBSET_COORDINATES_Y_FS) = 30;
      schema.local_field_value(0, SIGMA_FS) = 0.0;
      schema.local_field_value(0, GAMMA_FS) = 0.0;
      schema.local_field_value(0, STATUS_FLAG_FS) = 0;
      schema.local_field_value(0, SUBSET_DISPLACEMENT_X_FS) = u_exact - 0.25;
      schema.local_field_value(0, SUBSET_DISPLACEMENT_Y_FS) = v_exact - 0.25;
      schema.local_field_value(0, ROTATION_Z_FS) = t_exact - 0.05;
      schema.local_field_value(0, NORMAL_STRETCH_XX_FS) = 0.0;
      schema.local_field_value(0, NORMAL_STRETCH_YY_FS) = 0.0;
      schema.local_field_value(0, SHEAR_STRETCH_XY_FS) = 0.0;
      schema.execute_correlation();
      bool bad = false;
      bad |= std::abs(schema.local_field_value(0, SUBSET_DISPLACEMENT_X_FS) - u_exact) > errtol;
      bad |= std::abs(schema.local_field_value(0, SUBSET_DISPLACEMENT_Y_FS) - v_exact) > errtol;
      bad |= std::abs(schema.local_field_value(0, ROTATION_Z_FS) - t_exact) > errtol;
      bad |= std::abs(schema.local_field_value(0, NORMAL_STRETCH_XX_FS) - ex_exact) > errtol;
      bad |= std::abs(schema.local_field_value(0, NORMAL_STRETCH_YY_FS) - ey_exact) > errtol;
      bad |= std::abs(schema.local_field_value(0, SHEAR_STRETCH_XY_FS)) > errtol;
      bad |= std::abs(schema.local_field_value(0, STATUS_FLAG_FS) - 1.0) > errtol;  // INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL
      bad |= std::abs(schema.local_field_value(0, SIGMA_FS) + 1.0) < errtol;      // sigma != -1
      std::cout << label << (bad ? " FAIL" : " PASS") << " opt " << schema.optimization_method() << " interp "
                << schema.interpolation_method() << " u " << schema.local_field_value(0, SUBSET_DISPLACEMENT_X_FS)
                << " v " << schema.local_field_value(0, SUBSET_DISPLACEMENT_Y_FS) << " t "
                << schema.local_field_value(0, ROTATION_Z_FS) << " ex "
                << schema.local_field_value(0, NORMAL_STRETCH_XX_FS) << " ey "
                << schema.local_field_value(0, NORMAL_STRETCH_YY_FS) << " g "
                << schema.local_field_value(0, SHEAR_STRETCH_XY_FS) << " sigma "
                << schema.local_field_value(0, SIGMA_FS) << " gamma " << schema.local_field_value(0, GAMMA_FS)
                << " flag " << schema.local_field_value(0, STATUS_FLAG_FS) << "\n";
      errors += bad ? 1 : 0;
    }
  return errors;
}

ParameterList all_shape_params(int backend) {
  ParameterList params;
  params.set("enable_translation", 1);
  params.set("enable_rotation", 1);
  params.set("enable_normal_strain", 1);
  params.set("enable_shear_strain", 1);
  params.set("robust_solver_tolerance", 1.0E-4);
  params.set("correlation_backend", backend);
  return params;
}

int correlate_params(const std::vector<float> &ref, const std::vector<float> &def, int w, int h, int backend) {
  ParameterList params = all_shape_params(backend);
  Schema schema(std::vector<double>{30}, std::vector<double>{30}, 21, nullptr, &params);
  schema.set_ref_image(w, h, ref.data());
  schema.set_def_image(w, h, def.data());
  return optimizer_sweep(schema, params, 1.0E-2, "square");
}

int correlate_conformal(const std::vector<float> &ref, const std::vector<float> &def, int w, int h, int backend) {
  // four sided shape at odd angles (DICe_TestCorrelateConformal.cpp:94-103)
  Conformal_Area_Def subset_def;
  subset_def.boundary.push_back(std::make_shared<Polygon>(std::vector<int>{25, 35, 39, 20}, std::vector<int>{22, 20, 40, 35}));
  std::map<int, Conformal_Area_Def> defs{{0, subset_def}};
  ParameterList params = all_shape_params(backend);
  Schema schema(std::vector<double>{30}, std::vector<double>{30}, -1, &defs, &params);
  schema.set_ref_image(w, h, ref.data());
  schema.set_def_image(w, h, def.data());
  return optimizer_sweep(schema, params, 5.0E-2, "conformal");
}

int parallel_subsets(const std::vector<float> &ref, const std::vector<float> &def, int w, int h,
                     const char *gold_path, int backend) {
  // tests/regression/parallel_subsets: default parameters, S = 27
  std::ifstream in(gold_path);
  std::string line;
  std::vector<std::vector<double>> rows;
  while (std::getline(in, line)) {
    if (line.empty() || !std::isdigit((unsigned char)line[0])) continue;
    std::istringstream ss(line);
    std::vector<double> r;
    double v;
    while (ss >> v) r.push_back(v);
    rows.push_back(r);
  }
  std::vector<double> cx, cy;
  for (auto &r : rows) {
    cx.push_back(r[1]);
    cy.push_back(r[2]);
  }
  ParameterList params;
  dice_default_params(&params);
  params.set("correlation_backend", backend);
  Schema schema(cx, cy, 27, nullptr, &params);
  schema.set_ref_image(w, h, ref.data());
  schema.set_def_image(w, h, def.data());
  schema.execute_correlation();
  int errors = 0;
  for (size_t i = 0; i < rows.size(); ++i) {
    const double du = std::abs(schema.local_field_value((int)i, SUBSET_DISPLACEMENT_X_FS) - rows[i][3]);
    const double dv = std::abs(schema.local_field_value((int)i, SUBSET_DISPLACEMENT_Y_FS) - rows[i][4]);
    const bool bad = du > 5e-5 || dv > 5e-5 || schema.local_field_value((int)i, ITERATIONS_FS) != rows[i][9] ||
                     schema.local_field_value((int)i, STATUS_FLAG_FS) != rows[i][10];
    if (bad) {
      std::cout << "subset " << i << " FAIL du " << du << " dv " << dv << "\n";
      ++errors;
    }
  }
  std::cout << "parallel_subsets " << rows.size() << " subsets, " << errors << " failures\n";
  return errors + (rows.size() == 49 ? 0 : 1);
}

int polygon() {
  // the two boundary polygons of the conformal raster golden (DICe_TestSubset.cpp:248-309)
  const Polygon a({940, 951, 964, 980, 959}, {422, 399, 410, 413, 432});
  const Polygon b({982, 1000, 991, 978}, {426, 428, 456, 445});
  for (const Polygon *p : {&a, &b})
    for (const auto &yx : p->get_owned_pixels()) std::cout << yx.first << " " << yx.second << "\n";
  return 0;
}

// Prints what read_subset_file and conformal_subset_pixels make of a subset
// file: "coord x y" per subset, then per conformal subset "conformal id n_pixels
// n_active n_obstructed_shapes blocking..." and its pixels as "y x active".
int subset_file_info(const char *path) {
  const Subset_File_Info info = read_subset_file(path);
  for (size_t i = 0; i < info.coordinates_x.size(); ++i)
    std::cout << "coord " << info.coordinates_x[i] << " " << info.coordinates_y[i] << "\n";
  for (const auto &kv : info.conformal_area_defs) {
    std::vector<int> x, y;
    std::vector<uint8_t> a;
    conformal_subset_pixels(kv.second, x, y, a);
    int na = 0;
    for (uint8_t v : a) na += v;
    std::cout << "conformal " << kv.first << " " << x.size() << " " << na << " " << kv.second.obstructed_area.size();
    for (int b : info.blocking_subset_ids.at(kv.first)) std::cout << " " << b;
    std::cout << "\n";
    for (size_t k = 0; k < x.size(); ++k) std::cout << y[k] << " " << x[k] << " " << (int)a[k] << "\n";
  }
  return 0;
}

// A deck of the reference's regression suite through the Schema mirror.  The
// configuration file has "key value" lines: ref / def (raw float32 images,
// or ref_file / def_file: image files read by Schema::set_ref_image(name)),
// width, height, subset_size, subset_file or step_size, out_dir, prefix,
// separate (0|1), backend (0|1), ref_name, def_name, routine (0|1), and
// optional fields (comma-separated output names), optimization /
// interpolation / initialization (enum values), gauss_filter_mask_size (turns
// the Gauss filter on), frames (the deformed image correlated that many
// times, one output file per frame; def_names: the comma-separated names for
// the header of each frame; def_files: one image file per frame),
// enable_rotation / enable_normal_strain / enable_shear_strain, vsg_window
// (the VSG strain post-processor), sort, omit_row_id, delimiter,
// first_frame_index and num_frames (the frame count of the whole sequence,
// which sets the digits of the output file numbers).  A subset file of regions of interest
// gives the grid (step_size) inside them, with the seeds' values and the
// neighbour ids.  Default parameters otherwise.
int deck(const char *config_path) {
  std::ifstream cfg(config_path);
  std::map<std::string, std::string> c;
  std::string key, value;
  while (cfg >> key && std::getline(cfg >> std::ws, value)) c[key] = value;
  auto num = [&](const char *k) { return std::atoi(c.at(k).c_str()); };
  const int w = num("width"), h = num("height");
  const bool files = c.count("ref_file") != 0;
  std::vector<float> ref, def;
  if (!files) {
    ref = read_raw(c.at("ref").c_str(), w, h);
    def = read_raw(c.at("def").c_str(), w, h);
  }
  ParameterList params;
  dice_default_params(&params);
  params.set("correlation_routine", num("routine"));
  params.set("correlation_backend", num("backend"));
  if (c.count("optimization")) params.set("optimization_method", num("optimization"));
  if (c.count("interpolation")) params.set("interpolation_method", num("interpolation"));
  if (c.count("initialization")) params.set("initialization_method", num("initialization"));
  for (const char *k : {"enable_rotation", "enable_normal_strain", "enable_shear_strain"})
    if (c.count(k)) params.set(k, num(k));
  if (c.count("gauss_filter_mask_size")) {
    params.set("gauss_filter_images", 1);
    params.set("gauss_filter_mask_size", num("gauss_filter_mask_size"));
  }
  std::vector<double> cx, cy;
  std::vector<int> neighbors;
  Subset_File_Info info;
  const bool grid = c.count("step_size") != 0;
  if (c.count("subset_file")) info = read_subset_file(c.at("subset_file"), w, h);
  if (grid) {
    create_regular_grid_of_correlation_points(w, h, num("subset_size"), num("step_size"), cx, cy, neighbors,
                                              &info.conformal_area_defs, &info);
  } else {
    cx = info.coordinates_x;
    cy = info.coordinates_y;
  }
  if (c.count("override_force_simplex")) params.set("override_force_simplex", num("override_force_simplex"));
  if (c.count("read_full_images")) params.set("read_full_images", num("read_full_images"));
  Schema schema(cx, cy, num("subset_size"), grid ? nullptr : &info.conformal_area_defs, &params);
  if (!grid) {
    // Schema(input_params, correlation_params) with a subset file (DICe_Schema.cpp:1052-1068)
    schema.set_obstructing_subset_ids(info.blocking_subset_ids);
    schema.set_force_simplex(info.force_simplex);
  }
  if (grid) {
    schema.set_step_sizes(num("step_size"), num("step_size"));
    schema.set_neighbor_ids(neighbors);
    schema.set_seed_values(info);
  }
  if (c.count("vsg_window")) schema.set_vsg_strain_window(num("vsg_window"));
  if (c.count("sort")) schema.set_sort_txt_output(num("sort") != 0);
  if (c.count("first_frame_index")) schema.set_first_frame_index(num("first_frame_index"));
  const std::string delimiter = c.count("delimiter") ? c.at("delimiter") : " ";
  const bool omit_row_id = c.count("omit_row_id") && num("omit_row_id") != 0;
  std::vector<std::string> def_files;
  if (c.count("def_files")) {
    std::stringstream ss(c.at("def_files"));
    std::string f;
    while (std::getline(ss, f, ',')) def_files.push_back(f);
  }
  // DICe_Main.cpp:314-330: the extents are updated before the reference and each deformed image is read
  if (files) {
    schema.update_extents();
    schema.set_ref_image(c.at("ref_file"));
  } else {
    schema.set_ref_image(w, h, ref.data());
  }
  schema.set_image_file_names(c.at("ref_name"), c.at("def_name"));
  std::vector<std::string> fields;
  if (c.count("fields")) {
    std::stringstream ss(c.at("fields"));
    std::string f;
    while (std::getline(ss, f, ',')) fields.push_back(f);
  }
  const int frames = c.count("frames") ? num("frames") : 1;
  std::vector<std::string> def_names;
  if (c.count("def_names")) {
    std::stringstream ss(c.at("def_names"));
    std::string f;
    while (std::getline(ss, f, ',')) def_names.push_back(f);
  }
  for (int frame = 0; frame < frames; ++frame) {
    if (files) schema.update_extents();
    if (!def_files.empty()) schema.set_def_image(def_files.at(frame));
    else if (files) schema.set_def_image(c.at("def_file"));
    else schema.set_def_image(w, h, def.data());
    schema.set_image_file_names(c.at("ref_name"), frame < (int)def_names.size() ? def_names[frame] : c.at("def_name"));
    schema.execute_correlation();
    schema.write_output(c.at("out_dir") + "/", c.at("prefix"), num("separate") != 0, fields, delimiter, omit_row_id,
                        c.count("num_frames") ? num("num_frames") : frames);
  }
  std::cout << "extents " << (schema.has_extents() ? 1 : 0) << "\n";
  for (int i = 0; i < schema.local_num_subsets(); ++i)
    std::cout << "subset " << i << " x " << cx[i] << " y " << cy[i] << " u "
              << schema.local_field_value(i, SUBSET_DISPLACEMENT_X_FS) << " v "
              << schema.local_field_value(i, SUBSET_DISPLACEMENT_Y_FS) << " status "
              << schema.local_field_value(i, STATUS_FLAG_FS) << "\n";
  return 0;
}

int errors_without_gpu() {
  int errors = 0;
  ParameterList params;
  try {
    params.set("no_such_parameter", 1.0);
    ++errors;
  } catch (const std::invalid_argument &) {
  }
  try {
    Polygon p({1, 2}, {1, 2});
    ++errors;
  } catch (const std::invalid_argument &) {
  }
  try {
    params.get("optimization_method");
    ++errors;
  } catch (const std::invalid_argument &) {
  }
  dice_default_params(&params);
  if (params.get("optimization_method") != GRADIENT_BASED_THEN_SIMPLEX) ++errors;
  if (params.get("interpolation_method") != KEYS_FOURTH) ++errors;
  tracking_default_params(&params);
  if (params.get("optimization_method") != SIMPLEX || params.get("correlation_routine") != TRACKING_ROUTINE) ++errors;
  std::cout << "errors " << (errors ? "FAIL" : "PASS") << "\n";
  return errors;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return 99;
  const std::string cmd = argv[1];
  try {
    if (cmd == "polygon") return polygon();
    if (cmd == "errors") return errors_without_gpu();
    if (cmd == "subset_file_info" && argc >= 3) return subset_file_info(argv[2]);
    if (cmd == "deck" && argc >= 3) return deck(argv[2]);
    if (cmd == "read_params" && argc >= 3) {
      bool tracking = false;
      const ParameterList pl = read_correlation_params(argv[2], &tracking);
      for (const auto &kv : pl.values()) std::cout << kv.first << " " << kv.second << "\n";
      std::cout << "use_tracking_default_params " << (tracking ? 1 : 0) << "\n";
      return 0;
    }
    if (cmd == "objective" && argc >= 11) {
      const int w = std::atoi(argv[4]), h = std::atoi(argv[5]), S = std::atoi(argv[6]);
      const std::vector<float> ref = read_raw(argv[2], w, h), def = read_raw(argv[3], w, h);
      ParameterList params;
      dice_default_params(&params);
      params.set("enable_rotation", 1);
      params.set("enable_normal_strain", 1);
      params.set("enable_shear_strain", 1);
      Schema schema({std::atof(argv[7])}, {std::atof(argv[8])}, S, nullptr, &params);
      schema.set_ref_image(w, h, ref.data());
      schema.set_def_image(w, h, def.data());
      Objective_ZNSSD obj(&schema, 0);
      Affine_Shape_Function sf(true, true, true);
      sf.set(0, std::atof(argv[9]));
      sf.set(1, std::atof(argv[10]));
      double noise = 0.0;
      std::printf("gamma %.9g\n", obj.gamma(sf));
      const double sig = obj.sigma(sf, noise);
      std::printf("sigma %.9g %.9g\n", sig, noise);
      Affine_Shape_Function fast = sf, robust = sf;
      int it_fast = 0, it_robust = 0;
      const int st_fast = obj.computeUpdateFast(fast, it_fast);
      const int st_robust = obj.computeUpdateRobust(robust, it_robust);
      std::printf("fast %d %d", st_fast, it_fast);
      for (int k = 0; k < 6; ++k) std::printf(" %.9g", fast.parameters()[k]);
      std::printf(" %.9g\nrobust %d %d", schema.local_field_value(0, CONDITION_NUMBER_FS), st_robust, it_robust);
      for (int k = 0; k < 6; ++k) std::printf(" %.9g", robust.parameters()[k]);
      std::printf("\n");
      return 0;
    }
    if (cmd == "read_image" && argc >= 4) {
      std::vector<float> pix;
      int w = 0, h = 0;
      if (argc >= 8)
        read_image(argv[2], pix, w, h, std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7]));
      else
        read_image(argv[2], pix, w, h);
      int dw = 0, dh = 0;
      if (argc < 8) {
        read_image_dimensions(argv[2], dw, dh);
        if (dw != w || dh != h) return 1;
      }
      write_rawi_image(argv[3], w, h, pix.data());
      return 0;
    }
    if (cmd == "tiff_samples" && argc >= 4) {
      std::vector<uint16_t> v;
      int w = 0, h = 0, bits = 0;
      read_tiff_samples(argv[2], v, w, h, bits);
      std::FILE *f = std::fopen(argv[3], "wb");
      if (!f) return 1;
      const uint32_t hdr[3] = {(uint32_t)w, (uint32_t)h, (uint32_t)bits};
      const bool ok = std::fwrite(hdr, sizeof hdr, 1, f) == 1 && std::fwrite(v.data(), 2, v.size(), f) == v.size();
      std::fclose(f);
      return ok ? 0 : 1;
    }
    if (cmd == "grid" && argc >= 6) {
      std::vector<double> x, y;
      std::vector<int> nb;
      Subset_File_Info info;
      if (argc >= 7) info = read_subset_file(argv[6], std::atoi(argv[2]), std::atoi(argv[3]));
      create_regular_grid_of_correlation_points(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]),
                                                std::atoi(argv[5]), x, y, nb, &info.conformal_area_defs, &info);
      for (size_t i = 0; i < x.size(); ++i) std::cout << x[i] << " " << y[i] << " " << nb[i] << "\n";
      for (const auto &sd : info.seed_subset_ids) {
        const auto &d = info.displacement_map.at(sd.second);
        std::cout << "seed " << sd.first << " " << d.first << " " << d.second << "\n";
      }
      return 0;
    }
    if (argc < 7) return 99;
    const int w = std::atoi(argv[4]), h = std::atoi(argv[5]);
    const std::vector<float> ref = read_raw(argv[2], w, h), def = read_raw(argv[3], w, h);
    if (cmd == "correlate_params") return correlate_params(ref, def, w, h, std::atoi(argv[6]));
    if (cmd == "correlate_conformal") return correlate_conformal(ref, def, w, h, std::atoi(argv[6]));
    if (cmd == "parallel_subsets" && argc >= 8) return parallel_subsets(ref, def, w, h, argv[6], std::atoi(argv[7]));

  } catch (const std::exception &e) {
    std::cerr << "exception: " << e.what() << "\n";
    return 98;
  }
  return 99;
}
