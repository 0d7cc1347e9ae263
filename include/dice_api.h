/*
 * dice_api.h -- the reference's own C entry points for this path, under their names and
 * argument order (DICe_api.h:129-133, 176-180), so that an existing caller of the reference's
 * library can link or dlsym this one instead.
 *
 *  - dice_correlate lives in libdice_hip.so;
 *  - dice_correlate_conformal lives in libdice_schema.so (it reads a subset file and an XML
 *    parameter file through the C++ Schema mirror, which sits on libdice_hip.so).
 *
 * Types are those of the reference's default build: scalar_t = storage_t = float, int_t = int.
 * The reference's Teuchos::ParameterList * argument of dice_correlate is a pointer to this
 * library's dice_params (include/dice_hip.h) instead: NULL keeps the reference's meaning (the
 * tracking defaults, DICe_api.cpp:84-96).  Points have the stride 9 of DICe_api.cpp:125-150:
 * x, y, u, v, theta, sigma, gamma, beta, status.  Like the reference, both keep static state
 * across calls (the subsets and the reference image of the first call) and return 0, or -1 on
 * a size mismatch or bad argument.  Calls are serialised by a lock; they are not re-entrant.
 */
#ifndef DICE_API_H
#define DICE_API_H

#include <stdbool.h>

#include "dice_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dice_correlate (DICe_api.h:129-133, DICe_api.cpp:51-153): correlates n_points square
 * subsets of side subset_size centred on the points' (x, y) from their (u, v, theta, sigma)
 * on the device; the parameters apply on the first call and again when update_params. */
int dice_correlate(float points[], int n_points, int subset_size, float ref_img[], int ref_w, int ref_h,
                   float def_img[], int def_w, int def_h, const dice_params *input_params, bool update_params);

/* dice_correlate_conformal (DICe_api.h:176-180, DICe_api.cpp:155-282): every subset defined
 * in subset_file (read_subset_file: centroids and conformal definitions, blocking subsets),
 * the parameters of param_file (read_correlation_params; NULL: the tracking defaults) over the
 * tracking defaults, read on the first call; the centroids are written into points[x, y].
 * write_output: the text output of Schema::write_output into "./" after each call. */
int dice_correlate_conformal(float points[], float ref_img[], int ref_w, int ref_h, float def_img[], int def_w,
                             int def_h, const char *subset_file, const char *param_file, bool write_output);

#ifdef __cplusplus
}
#endif
#endif
