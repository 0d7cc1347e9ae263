"""Python binding of the MI355X subset-correlation library (ctypes over the C ABI).

This module is plumbing for tests and benchmarks: every computation runs in
libdice_hip.so (hand-written HIP kernels for gfx950).  There is no CPU
fallback: if the library is missing or no GPU is visible, the calls raise.

Field store layout: a (NUM_FIELDS, n) float64 array, one row per Field_Spec
(see FIELDS), as DICe's Schema keeps one MultiField per field
(DICe_Schema.cpp:1363-1440).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libdice_hip.so")
ROOT = os.path.dirname(HERE)

# enum dice_field (include/dice_hip.h)
FIELDS = ["COORD_X", "COORD_Y", "DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY",
          "SIGMA", "GAMMA", "BETA", "MATCH", "NOISE", "CONTRAST", "ACTIVE_PIXELS",
          "STATUS", "ITERATIONS", "CONDITION", "NEIGHBOR_ID",
          "FIELD_1", "FIELD_2", "FIELD_3", "FIELD_4", "FIELD_5", "FIELD_6", "FIELD_7"]
F = {name: i for i, name in enumerate(FIELDS)}
# enum dice_vsg_field
VSG_FIELDS = ["VSG_STRAIN_XX", "VSG_STRAIN_YY", "VSG_STRAIN_XY", "VSG_DUDX", "VSG_DUDY", "VSG_DVDX", "VSG_DVDY"]
NUM_FIELDS = len(FIELDS)

MODE_PLAIN = 0
MODE_BATCHED = 1
STORAGE_F32 = 0
STORAGE_U16 = 1

# Status_Flag values (DICe.h:609-677) used by the path
INITIALIZE_USING_PREVIOUS_FRAME_SUCCESSFUL = 1
INITIALIZE_USING_NEIGHBOR_VALUE_SUCCESSFUL = 3
CONVERGED_STATUSES = (1, 2, 3, 4)

# enum values of DICe.h
BILINEAR, BICUBIC, KEYS_FOURTH = 0, 1, 2
SIMPLEX, GRADIENT_BASED, GRADIENT_THEN_SEARCH, SIMPLEX_THEN_GRADIENT_BASED, GRADIENT_BASED_THEN_SIMPLEX = 0, 1, 2, 3, 4
USE_FIELD_VALUES, USE_NEIGHBOR_VALUES, USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY = 0, 1, 2
# enum dice_objective_call
OBJ_GAMMA, OBJ_SIGMA, OBJ_UPDATE_FAST, OBJ_UPDATE_ROBUST = 0, 1, 2, 3


class Params(ctypes.Structure):
    """struct dice_params (include/dice_hip.h)."""
    _fields_ = [
        ("interpolation_method", ctypes.c_int32),
        ("optimization_method", ctypes.c_int32),
        ("correlation_routine", ctypes.c_int32),
        ("initialization_method", ctypes.c_int32),
        ("enable_translation", ctypes.c_int32),
        ("enable_rotation", ctypes.c_int32),
        ("enable_normal_strain", ctypes.c_int32),
        ("enable_shear_strain", ctypes.c_int32),
        ("max_solver_iterations_fast", ctypes.c_int32),
        ("max_solver_iterations_robust", ctypes.c_int32),
        ("fast_solver_tolerance", ctypes.c_double),
        ("robust_solver_tolerance", ctypes.c_double),
        ("robust_delta_disp", ctypes.c_double),
        ("robust_delta_theta", ctypes.c_double),
        ("skip_solve_gamma_threshold", ctypes.c_double),
        ("initial_gamma_threshold", ctypes.c_double),
        ("final_gamma_threshold", ctypes.c_double),
        ("disp_jump_tol", ctypes.c_double),
        ("theta_jump_tol", ctypes.c_double),
        ("levenberg_marquardt_regularization_factor", ctypes.c_double),
        ("momentum_factor", ctypes.c_double),
        ("compute_def_gradients", ctypes.c_int32),
        ("normalize_gamma_with_active_pixels", ctypes.c_int32),
        ("compute_uncertainty_fields", ctypes.c_int32),
        ("gradient_method", ctypes.c_int32),
        ("gauss_filter_images", ctypes.c_int32),
        ("gauss_filter_mask_size", ctypes.c_int32),
        ("neighbor_settle_tolerance", ctypes.c_double),
    ]


# every symbol declared in include/dice_hip.h
EXPORTS = [
    "dice_hip_create", "dice_hip_destroy", "dice_hip_last_error", "dice_hip_default_params",
    "dice_hip_tracking_default_params", "dice_hip_set_params", "dice_hip_set_ref_image",
    "dice_hip_set_def_image", "dice_hip_set_def_image_device", "dice_hip_set_subsets",
    "dice_hip_correlate", "dice_hip_correlate_device", "dice_hip_correlate_sequence",
    "dice_hip_synchronize", "dice_hip_stream", "dice_hip_kernel_timing", "dice_hip_set_kernel_timing",
    "dice_hip_autotune", "dice_hip_get_launch_shape", "dice_hip_set_launch_shape",
    "dice_hip_correlate_points", "dice_hip_gamma", "dice_hip_gamma_adjoint", "dice_hip_gauss_filter",
    "dice_hip_neighbor_rounds", "dice_hip_vsg_strain", "dice_hip_vsg_strain_device",
    "dice_hip_uncertainty", "dice_hip_uncertainty_device", "dice_hip_objective", "dice_hip_pass_counts",
    "dice_hip_correlate_device_noalloc", "dice_hip_set_force_simplex", "dice_hip_set_subset_masks",
    "dice_hip_set_image_offsets",
]

_lib = None


class DiceHipError(RuntimeError):
    pass


def build():
    """Compile the HIP library in-tree (hipcc, gfx950)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(HERE, "csrc")])


def lib():
    """Load libdice_hip.so (raises if it was not built).

    torch (used by callers for device buffers) bundles a HIP runtime whose soname
    equals the system one; importing it first makes the library bind to that
    same runtime instead of loading a second one into the process."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DiceHipError("libdice_hip.so not built (run __graft_entry__.build())")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.dice_hip_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), I]
        L.dice_hip_destroy.argtypes = [P]
        L.dice_hip_last_error.argtypes = [P]
        L.dice_hip_last_error.restype = ctypes.c_char_p
        L.dice_hip_default_params.argtypes = [P]
        L.dice_hip_tracking_default_params.argtypes = [P]
        L.dice_hip_set_params.argtypes = [P, P]
        L.dice_hip_set_ref_image.argtypes = [P, I, I, I, P]
        L.dice_hip_set_def_image.argtypes = [P, I, I, I, P]
        L.dice_hip_set_def_image_device.argtypes = [P, I, I, I, P]
        L.dice_hip_set_subsets.argtypes = [P, I, P, P, I, P, P, P, P]
        L.dice_hip_correlate.argtypes = [P, I, P, I]
        L.dice_hip_correlate_device.argtypes = [P, I, P, I]
        L.dice_hip_correlate_device_noalloc.argtypes = [P, I, P, I]
        L.dice_hip_correlate_sequence.argtypes = [P, I, I, P, P, P, I]
        L.dice_hip_synchronize.argtypes = [P]
        L.dice_hip_stream.argtypes = [P]
        L.dice_hip_stream.restype = P
        L.dice_hip_kernel_timing.argtypes = [P, ctypes.POINTER(D)]
        L.dice_hip_set_kernel_timing.argtypes = [P, I]
        L.dice_hip_autotune.argtypes = [P, P, I]
        L.dice_hip_get_launch_shape.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]
        L.dice_hip_set_launch_shape.argtypes = [P, I, I, I]
        L.dice_hip_correlate_points.argtypes = [P, I, I, P, I, I, P, I, I, P, I]
        L.dice_hip_gamma.argtypes = [P, I, P, P]
        L.dice_hip_gamma_adjoint.argtypes = [P, I, P, P, P, P]
        L.dice_hip_gauss_filter.argtypes = [I, I, I, I, P, P, I]
        L.dice_hip_neighbor_rounds.argtypes = [P]
        L.dice_hip_pass_counts.argtypes = [P, P]
        L.dice_hip_set_force_simplex.argtypes = [P, P]
        L.dice_hip_set_subset_masks.argtypes = [P, P, P, P]
        L.dice_hip_set_image_offsets.argtypes = [P, I, I, I, I]
        L.dice_hip_vsg_strain.argtypes = [P, I, P, P]
        L.dice_hip_vsg_strain_device.argtypes = [P, I, P, P]
        L.dice_hip_uncertainty.argtypes = [P, P, P]
        L.dice_hip_uncertainty_device.argtypes = [P, P, P]
        L.dice_hip_objective.argtypes = [P, I, I, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def default_params():
    p = Params()
    lib().dice_hip_default_params(ctypes.byref(p))
    return p


def tracking_default_params():
    p = Params()
    lib().dice_hip_tracking_default_params(ctypes.byref(p))
    return p


def new_fields(cx, cy):
    """A fresh field store with coordinates set (all other fields 0)."""
    n = len(cx)
    f = np.zeros((NUM_FIELDS, n), dtype=np.float64)
    f[F["COORD_X"]] = cx
    f[F["COORD_Y"]] = cy
    return f


def regular_grid(width, height, subset_size, step):
    """Centroids of Schema(roi_w, roi_h, step_x, step_y, S) (DICe_Schema.cpp:101-139), row-major."""
    nx = (width - 2 * subset_size) // step + 1
    ny = (height - 2 * subset_size) // step + 1
    xs = subset_size + np.arange(nx) * step - 1
    ys = subset_size + np.arange(ny) * step - 1
    gx, gy = np.meshgrid(xs, ys)
    return gx.ravel().astype(np.float64), gy.ravel().astype(np.float64)


def serpentine_grid(width, height, subset_size, step):
    """create_regular_grid_of_correlation_points (DICe_Decomp.cpp:868-1070) over the whole image with no
    seed: centres subset_size - 1 + k * step down the first column, up the next, ...; neighbour ids -1 for
    the first point and for the first point of the serpentine after it (the reference gives that one the
    seed's id, -1 without a seed), else the previous point.  Returns (cx, cy, neighbor_ids)."""
    xs = list(range(subset_size - 1, width - subset_size, step))
    ys = list(range(subset_size - 1, height - subset_size, step))
    cx, cy = [], []
    for c, x in enumerate(xs):
        for y in (ys if c % 2 == 0 else ys[::-1]):
            if x <= width - subset_size and y <= height - subset_size:
                cx.append(x)
                cy.append(y)
    nb = np.arange(-1, len(cx) - 1, dtype=np.float64)
    nb[1:2] = -1.0
    return np.array(cx, np.float64), np.array(cy, np.float64), nb


class Context:
    """One device context of the C ABI (dice_hip_context)."""

    def __init__(self, device=0):
        self._lib = lib()
        h = ctypes.c_void_p()
        rc = self._lib.dice_hip_create(ctypes.byref(h), device)
        if rc != 0:
            raise DiceHipError("dice_hip_create failed (%d): no usable GPU" % rc)
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            self._lib.dice_hip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.dice_hip_last_error(self.h)
            raise DiceHipError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def set_params(self, p):
        self._check(self._lib.dice_hip_set_params(self.h, ctypes.byref(p)), "set_params")

    @staticmethod
    def _img(img):
        img = np.ascontiguousarray(img)
        if img.dtype == np.uint16:
            return img, STORAGE_U16
        return np.ascontiguousarray(img, dtype=np.float32), STORAGE_F32

    def set_ref_image(self, img):
        img, st = self._img(img)
        self._check(self._lib.dice_hip_set_ref_image(self.h, img.shape[1], img.shape[0], st, _p(img)), "set_ref_image")

    def set_def_image(self, img):
        img, st = self._img(img)
        self._check(self._lib.dice_hip_set_def_image(self.h, img.shape[1], img.shape[0], st, _p(img)), "set_def_image")

    def set_def_image_device(self, ptr, width, height, storage):
        self._check(self._lib.dice_hip_set_def_image_device(self.h, width, height, storage, ctypes.c_void_p(ptr)),
                    "set_def_image_device")

    def set_subsets(self, cx, cy, subset_size=0, conformal=None):
        cx = np.ascontiguousarray(cx, dtype=np.float64)
        cy = np.ascontiguousarray(cy, dtype=np.float64)
        co = px = py = pa = None
        if conformal is not None:
            co, px, py, pa = conformal
            co = np.ascontiguousarray(co, dtype=np.int32)
            px = np.ascontiguousarray(px, dtype=np.int32)
            py = np.ascontiguousarray(py, dtype=np.int32)
            pa = None if pa is None else np.ascontiguousarray(pa, dtype=np.uint8)
        self._check(self._lib.dice_hip_set_subsets(self.h, len(cx), _p(cx), _p(cy), subset_size, _p(co), _p(px),
                                                   _p(py), _p(pa)), "set_subsets")
        self.n = len(cx)

    def set_force_simplex(self, flags):
        """dice_hip_set_force_simplex: per-subset Nelder-Mead override (None clears)."""
        f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
        assert f is None or f.shape == (self.n,)
        self._check(self._lib.dice_hip_set_force_simplex(self.h, _p(f)), "set_force_simplex")

    def set_subset_masks(self, masks):
        """dice_hip_set_subset_masks: masks = (offsets int32[n+1], x, y) of blocked image pixels, or None."""
        if masks is None:
            self._check(self._lib.dice_hip_set_subset_masks(self.h, None, None, None), "set_subset_masks")
            return
        mo, mx, my = [np.ascontiguousarray(a, dtype=np.int32) for a in masks]
        self._check(self._lib.dice_hip_set_subset_masks(self.h, _p(mo), _p(mx), _p(my)), "set_subset_masks")

    def set_image_offsets(self, ref_ox, ref_oy, def_ox, def_oy):
        """dice_hip_set_image_offsets: global coordinates of the images' (0, 0) pixels."""
        self._check(self._lib.dice_hip_set_image_offsets(self.h, ref_ox, ref_oy, def_ox, def_oy),
                    "set_image_offsets")

    def correlate(self, fields, mode=MODE_BATCHED, first_frame=True):
        assert fields.dtype == np.float64 and fields.flags.c_contiguous and fields.shape == (NUM_FIELDS, self.n)
        self._check(self._lib.dice_hip_correlate(self.h, mode, _p(fields), 1 if first_frame else 0), "correlate")
        return fields

    def correlate_device(self, fields_ptr, mode=MODE_BATCHED, first_frame=True):
        self._check(self._lib.dice_hip_correlate_device(self.h, mode, ctypes.c_void_p(fields_ptr),
                                                        1 if first_frame else 0), "correlate_device")

    def correlate_device_noalloc(self, fields_ptr, mode=MODE_BATCHED, first_frame=True):
        """dice_hip_correlate_device_noalloc: as correlate_device, on the library's static work area."""
        self._check(self._lib.dice_hip_correlate_device_noalloc(self.h, mode, ctypes.c_void_p(fields_ptr),
                                                                1 if first_frame else 0), "correlate_device_noalloc")

    def correlate_sequence(self, frame_ptrs, fields_ptr, history_ptr=0, mode=MODE_BATCHED, use_graph=True):
        arr = (ctypes.c_void_p * len(frame_ptrs))(*frame_ptrs)
        self._check(self._lib.dice_hip_correlate_sequence(self.h, mode, len(frame_ptrs), arr,
                                                          ctypes.c_void_p(fields_ptr),
                                                          ctypes.c_void_p(history_ptr) if history_ptr else None,
                                                          1 if use_graph else 0), "correlate_sequence")

    def vsg_strain(self, fields, window_size):
        """dice_hip_vsg_strain: VSG strains of the current subsets from fields (MATCH set to -1 where the
        strain fails, in place); returns a (7, n) array in VSG_FIELDS order."""
        assert fields.dtype == np.float64 and fields.flags.c_contiguous and fields.shape == (NUM_FIELDS, self.n)
        out = np.zeros((len(VSG_FIELDS), self.n))
        self._check(self._lib.dice_hip_vsg_strain(self.h, window_size, _p(fields), _p(out)), "vsg_strain")
        return out

    def uncertainty(self, fields):
        """dice_hip_uncertainty: (2, n) array of UNCERTAINTY and UNCERTAINTY_ANGLE from fields."""
        assert fields.dtype == np.float64 and fields.flags.c_contiguous and fields.shape == (NUM_FIELDS, self.n)
        out = np.zeros((2, self.n))
        self._check(self._lib.dice_hip_uncertainty(self.h, _p(fields), _p(out)), "uncertainty")
        return out

    def objective(self, subset_id, call, params):
        """dice_hip_objective: one Objective_ZNSSD call (OBJ_GAMMA, OBJ_SIGMA, OBJ_UPDATE_FAST,
        OBJ_UPDATE_ROBUST) on subset_id; returns (params after the call, result[10])."""
        p = np.ascontiguousarray(params, dtype=np.float64).copy()
        r = np.zeros(10)
        self._check(self._lib.dice_hip_objective(self.h, subset_id, call, _p(p), _p(r)), "objective")
        return p, r

    def neighbor_rounds(self):
        """Correlation rounds of the last frame with neighbour initialization (0 otherwise)."""
        return self._lib.dice_hip_neighbor_rounds(self.h)

    def pass_counts(self):
        """dice_hip_pass_counts: (fallback list, overflow of the primary pass, overflow of the fallback pass)
        of the last correlation launch."""
        c = np.zeros(3, dtype=np.int32)
        self._check(self._lib.dice_hip_pass_counts(self.h, _p(c)), "pass_counts")
        return tuple(int(v) for v in c)

    def synchronize(self):
        self._check(self._lib.dice_hip_synchronize(self.h), "synchronize")

    def stream(self):
        return self._lib.dice_hip_stream(self.h)

    def set_kernel_timing(self, on):
        self._check(self._lib.dice_hip_set_kernel_timing(self.h, 1 if on else 0), "set_kernel_timing")

    def kernel_timing(self):
        ms = ctypes.c_double()
        cnt = self._lib.dice_hip_kernel_timing(self.h, ctypes.byref(ms))
        return cnt, ms.value

    def autotune(self, fields_ptr, first_frame=True):
        self._check(self._lib.dice_hip_autotune(self.h, ctypes.c_void_p(fields_ptr), 1 if first_frame else 0),
                    "autotune")

    def gamma(self, params_ptr, gamma_ptr, mode=MODE_BATCHED):
        """dice_hip_gamma: gamma of every subset at device params (n x 6 doubles) into gamma (n doubles)."""
        self._check(self._lib.dice_hip_gamma(self.h, mode, ctypes.c_void_p(params_ptr), ctypes.c_void_p(gamma_ptr)),
                    "gamma")

    def gamma_adjoint(self, params_ptr, dgamma_ptr, dparams_ptr, dimage_ptr=0, mode=MODE_BATCHED):
        """dice_hip_gamma_adjoint: dL/dp (n x 6) and, if dimage_ptr, dL/dI_def added to a W x H float image."""
        self._check(self._lib.dice_hip_gamma_adjoint(self.h, mode, ctypes.c_void_p(params_ptr),
                                                     ctypes.c_void_p(dgamma_ptr), ctypes.c_void_p(dparams_ptr),
                                                     ctypes.c_void_p(dimage_ptr) if dimage_ptr else None),
                    "gamma_adjoint")

    def launch_shape(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._lib.dice_hip_get_launch_shape(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def set_launch_shape(self, waves_per_group, groups_per_cu, lds_floats_per_wave):
        self._check(self._lib.dice_hip_set_launch_shape(self.h, waves_per_group, groups_per_cu, lds_floats_per_wave),
                    "set_launch_shape")


def gauss_filter(img, mask_size, device=0):
    """dice_hip_gauss_filter: Image_::gauss_filter (DICe_Image.cpp:1194-1240) of a float32 or uint16 image
    on the GPU; returns a new array of the same type."""
    img, st = Context._img(img)
    out = np.empty_like(img)
    rc = lib().dice_hip_gauss_filter(device, img.shape[1], img.shape[0], st, _p(img), _p(out), mask_size)
    if rc != 0:
        raise DiceHipError("gauss_filter failed (%d)" % rc)
    return out


def correlate_points(points, subset_size, ref, defm, params=None, update_params=False):
    """dice_correlate drop-in (DICe_api.cpp:51-153): points float32 (n, 9), modified in place."""
    points = np.ascontiguousarray(points, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    defm = np.ascontiguousarray(defm, dtype=np.float32)
    rc = lib().dice_hip_correlate_points(_p(points), points.shape[0], subset_size, _p(ref), ref.shape[1],
                                         ref.shape[0], _p(defm), defm.shape[1], defm.shape[0],
                                         ctypes.byref(params) if params is not None else None,
                                         1 if update_params else 0)
    return rc, points
