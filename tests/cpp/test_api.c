/*
 * A C caller of the reference's own entry points, resolved at run time as an existing binary
 * caller of the reference's library would: dlopen the libraries, dlsym dice_correlate and
 * dice_correlate_conformal, and correlate the integer-shift sequence of DICe_TestAPI.cpp:70-150
 * (defSyntheticSpeckled0..5: the reference image shifted by k px in x and y, k = 0..5).
 *
 *   test_api <libdice_hip.so> <libdice_schema.so> <ref.f32> <def0.f32> ... <def5.f32> <subset file>
 *            <param file>
 *
 * Images are 100 x 100 raw float32.  Prints one line per frame and call, and "API PASSED" when
 * every value is within the reference test's tolerance (1e-4); exits 1 otherwise.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>

#define W 100
#define H 100
#define STRIDE 9

typedef int (*correlate_fn)(float *, int, int, float *, int, int, float *, int, int, const void *, bool);
typedef int (*conformal_fn)(float *, float *, int, int, float *, int, int, const char *, const char *, bool);

static float *read_image(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  float *img = (float *)malloc(sizeof(float) * W * H);
  if (img && fread(img, sizeof(float), W * H, f) != W * H) {
    free(img);
    img = NULL;
  }
  fclose(f);
  return img;
}

int main(int argc, char **argv) {
  if (argc != 12) {
    fprintf(stderr, "usage: test_api lib_hip lib_schema ref def0..def5 subset_file param_file\n");
    return 2;
  }
  void *hip = dlopen(argv[1], RTLD_NOW | RTLD_GLOBAL);
  void *sch = hip ? dlopen(argv[2], RTLD_NOW | RTLD_GLOBAL) : NULL;
  if (!hip || !sch) {
    fprintf(stderr, "dlopen: %s\n", dlerror());
    return 2;
  }
  correlate_fn dice_correlate = (correlate_fn)dlsym(hip, "dice_correlate");
  conformal_fn dice_correlate_conformal = (conformal_fn)dlsym(sch, "dice_correlate_conformal");
  if (!dice_correlate || !dice_correlate_conformal) {
    fprintf(stderr, "dlsym: %s\n", dlerror());
    return 2;
  }
  float *ref = read_image(argv[3]);
  float *defs[6];
  for (int k = 0; k < 6; ++k) defs[k] = read_image(argv[4 + k]);
  for (int k = 0; k < 6; ++k)
    if (!ref || !defs[k]) {
      fprintf(stderr, "cannot read the images\n");
      return 2;
    }
  const double tol = 1.0e-4;
  int errors = 0;

  /* square subsets with the tracking defaults (NULL parameters) */
  const int n = 4;
  const float cx[4] = {47, 51, 51, 47}, cy[4] = {47, 47, 51, 51};
  float points[4 * STRIDE] = {0};
  for (int i = 0; i < n; ++i) {
    points[i * STRIDE + 0] = cx[i];
    points[i * STRIDE + 1] = cy[i];
  }
  for (int k = 0; k < 6; ++k) {
    if (dice_correlate(points, n, 21, ref, W, H, defs[k], W, H, NULL, false) != 0) {
      printf("dice_correlate failed on frame %d\n", k);
      return 1;
    }
    for (int i = 0; i < n; ++i) {
      const float *p = points + i * STRIDE;
      printf("dice_correlate frame %d subset %d x %g y %g u %g v %g theta %g sigma %g gamma %g beta %g status %g\n",
             k, i, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]);
      if (fabs(p[0] - cx[i]) > tol || fabs(p[1] - cy[i]) > tol || fabs(p[2] - k) > tol || fabs(p[3] - k) > tol ||
          fabs(p[4]) > tol || fabs(p[6]) > tol || p[8] != 1.0f)
        ++errors;
    }
  }
  /* a size mismatch is rejected */
  if (dice_correlate(points, n, 21, ref, W, H, defs[1], W, H - 1, NULL, false) != -1) ++errors;

  /* conformal subsets from a subset file and an XML parameter file */
  float cpoints[16 * STRIDE] = {0};
  for (int k = 0; k < 6; ++k) {
    if (dice_correlate_conformal(cpoints, ref, W, H, defs[k], W, H, argv[10], argv[11], false) != 0) {
      printf("dice_correlate_conformal failed on frame %d\n", k);
      return 1;
    }
    for (int i = 0; i < 2; ++i) {
      const float *p = cpoints + i * STRIDE;
      printf("dice_correlate_conformal frame %d subset %d x %g y %g u %g v %g theta %g gamma %g status %g\n", k, i,
             p[0], p[1], p[2], p[3], p[4], p[6], p[8]);
      if (fabs(p[2] - k) > tol || fabs(p[3] - k) > tol || fabs(p[6]) > tol || p[8] != 1.0f) ++errors;
    }
  }
  if (cpoints[0] != 45.0f || cpoints[1] != 44.0f || cpoints[STRIDE] != 52.0f || cpoints[STRIDE + 1] != 53.0f)
    ++errors;
  printf(errors ? "API FAILED (%d)\n" : "API PASSED\n", errors);
  return errors ? 1 : 0;
}
