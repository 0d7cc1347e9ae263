"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  Nothing in the product imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# Field slots of the per-subset SoA store (enum oracle_field).
FIELDS = ["COORD_X", "COORD_Y", "DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY",
          "SIGMA", "GAMMA", "BETA", "MATCH", "NOISE", "CONTRAST", "ACTIVE_PIXELS",
          "STATUS", "ITERATIONS", "CONDITION", "NEIGHBOR_ID",
          "FIELD_1", "FIELD_2", "FIELD_3", "FIELD_4", "FIELD_5", "FIELD_6", "FIELD_7"]
F = {name: i for i, name in enumerate(FIELDS)}
NUM_FIELDS = len(FIELDS)


class Params(ctypes.Structure):
    """Same layout as oracle_params / dice_params."""
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


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i = ctypes.c_int
        _lib.oracle_correlate_f.argtypes = [P, i, i, i, P, P, i, i, P, P, P, P, P, i, i]
        _lib.oracle_correlate_d.argtypes = [P, i, i, i, P, P, i, i, P, P, P, P, P, i, i]
        _lib.oracle_correlate_ex_f.argtypes = [P, i, i, i, P, P, i, i, P, P, P, P, P, i, i, P, P, P, P]
        _lib.oracle_correlate_ex_d.argtypes = [P, i, i, i, P, P, i, i, P, P, P, P, P, i, i, P, P, P, P]
        _lib.oracle_subset_gamma_f.restype = ctypes.c_double
        _lib.oracle_subset_gamma_f.argtypes = [i, i, P, P, i, i, i]
        _lib.oracle_apply_transform_d.argtypes = [i, i, P, P, i, i, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _lib.oracle_apply_transform_f.argtypes = [i, i, P, P, i, i, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        _lib.oracle_fd_gradients_f.argtypes = [i, i, P, P, P]
        _lib.oracle_fd_gradients_d.argtypes = [i, i, P, P, P]
        _lib.oracle_interp_f.argtypes = [i, i, i, P, i, P, P, P]
        _lib.oracle_interp_d.argtypes = [i, i, i, P, i, P, P, P]
        _lib.oracle_polygon_pixels.argtypes = [i, P, P, i, P, P]
        _lib.oracle_polygon_pixels_f.argtypes = [i, P, P, i, P, P]
        _lib.oracle_deformed_polygon_pixels.argtypes = [i, P, P, P, i, i, ctypes.c_double, i, P, P]
        _lib.oracle_lu_inverse.argtypes = [i, P]
        _lib.oracle_smooth_gradients_f.argtypes = [i, i, P, P]
        _lib.oracle_nelder_mead.argtypes = [i, i, P, P, P, i, ctypes.c_double, ctypes.c_double, i, P, P, P, P]
        _lib.oracle_gauss_filter_f.argtypes = [i, i, i, P, i]
        _lib.oracle_gauss_filter_d.argtypes = [i, i, P, i]
        _lib.oracle_vsg_strain.argtypes = [i, P, i, P]
        _lib.oracle_uncertainty.argtypes = [i, P, P]
        _lib.oracle_objective_f.argtypes = [P, i, i, i, P, P, i, i, i, i, P, P]
        _lib.oracle_objective_d.argtypes = [P, i, i, i, P, P, i, i, i, i, P, P]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_params():
    p = Params()
    lib().oracle_default_params(ctypes.byref(p))
    return p


def tracking_default_params():
    p = Params()
    lib().oracle_tracking_default_params(ctypes.byref(p))
    return p


def new_fields(cx, cy):
    """Field store (NUM_FIELDS x n, double) with coordinates set, as a fresh Schema holds it."""
    n = len(cx)
    fields = np.zeros((NUM_FIELDS, n), dtype=np.float64)
    fields[F["COORD_X"]] = cx
    fields[F["COORD_Y"]] = cy
    return fields


def correlate(params, ref, defm, fields, subset_size=0, conformal=None, first_frame=True,
              precision="double", threads=1, force_simplex=None, masks=None):
    """One execute_correlation over the subsets in `fields` (modified in place).

    ref/defm: 2-D float32, float64 or uint16 arrays.  conformal: None or a tuple
    (offsets int32[n+1], x int32[], y int32[], active uint8[]).  force_simplex: None or
    uint8[n] (Nelder-Mead only where set).  masks: None or (offsets int32[n+1], x int32[],
    y int32[]) of the image pixels that block each subset's warped pixels.
    """
    h, w = ref.shape
    if ref.dtype == np.uint16:
        storage = 1
    elif ref.dtype == np.float64:
        storage = 2
    else:
        storage = 0
        ref = ref.astype(np.float32)
        defm = defm.astype(np.float32)
    ref = np.ascontiguousarray(ref)
    defm = np.ascontiguousarray(defm.astype(ref.dtype))
    assert fields.dtype == np.float64 and fields.flags.c_contiguous and fields.shape[0] == NUM_FIELDS
    n = fields.shape[1]
    co = cx = cy = ca = None
    if conformal is not None:
        co, cx, cy, ca = [np.ascontiguousarray(a) for a in conformal]
        co = co.astype(np.int32); cx = cx.astype(np.int32); cy = cy.astype(np.int32); ca = ca.astype(np.uint8)
    fo = None if force_simplex is None else np.ascontiguousarray(force_simplex, dtype=np.uint8)
    mo = mx = my = None
    if masks is not None:
        mo, mx, my = [np.ascontiguousarray(a, dtype=np.int32) for a in masks]
    fn = lib().oracle_correlate_ex_d if precision == "double" else lib().oracle_correlate_ex_f
    rc = fn(ctypes.byref(params), w, h, storage, _p(ref), _p(defm), n, subset_size, _p(co), _p(cx), _p(cy),
            _p(ca), _p(fields), 1 if first_frame else 0, threads, _p(fo), _p(mo), _p(mx), _p(my))
    if rc != 0:
        raise ValueError("oracle_correlate rejected its arguments")
    return fields


def fd_gradients(img):
    img = np.ascontiguousarray(img)
    h, w = img.shape
    gx = np.zeros_like(img)
    gy = np.zeros_like(img)
    fn = lib().oracle_fd_gradients_d if img.dtype == np.float64 else lib().oracle_fd_gradients_f
    fn(w, h, _p(img), _p(gx), _p(gy))
    return gx, gy


def interp(method, img, x, y):
    """method: 0 bilinear_all, 2 keys_fourth_all, 3 keys_fourth (single)."""
    img = np.ascontiguousarray(img)
    h, w = img.shape
    x = np.ascontiguousarray(x, dtype=img.dtype)
    y = np.ascontiguousarray(y, dtype=img.dtype)
    out = np.zeros(x.shape, dtype=img.dtype)
    fn = lib().oracle_interp_d if img.dtype == np.float64 else lib().oracle_interp_f
    fn(method, w, h, _p(img), x.size, _p(x), _p(y), _p(out))
    return out


def apply_transform(img, cx, cy, u, v, theta):
    img = np.ascontiguousarray(img)
    h, w = img.shape
    out = np.zeros_like(img)
    if img.dtype == np.float64:
        lib().oracle_apply_transform_d(w, h, _p(img), _p(out), cx, cy, u, v, theta)
    else:
        lib().oracle_apply_transform_f(w, h, _p(img), _p(out), cx, cy, u, v, theta)
    return out


def subset_gamma(ref, defm, cx, cy, size):
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    defm = np.ascontiguousarray(defm, dtype=np.float32)
    h, w = ref.shape
    return lib().oracle_subset_gamma_f(w, h, _p(ref), _p(defm), cx, cy, size)


def polygon_pixels(vx, vy, precision="float"):
    """Polygon::get_owned_pixels: the winding angles in float (the reference's default build) or
    double (its DICE_USE_DOUBLE build)."""
    vx = np.ascontiguousarray(vx, dtype=np.int32)
    vy = np.ascontiguousarray(vy, dtype=np.int32)
    cap = (int(vx.max() - vx.min()) + 1) * (int(vy.max() - vy.min()) + 1)
    ox = np.zeros(cap, np.int32)
    oy = np.zeros(cap, np.int32)
    fn = lib().oracle_polygon_pixels_f if precision == "float" else lib().oracle_polygon_pixels
    cnt = fn(len(vx), _p(vx), _p(vy), cap, _p(ox), _p(oy))
    return ox[:cnt], oy[:cnt]


def deformed_polygon_pixels(vx, vy, params, cx, cy, skin_factor=1.0, cap=1 << 20):
    """Polygon::get_owned_pixels(shape_function, cx, cy, skin_factor) of a blocking subset: returns (x, y)."""
    vx = np.ascontiguousarray(vx, dtype=np.int32)
    vy = np.ascontiguousarray(vy, dtype=np.int32)
    q = np.ascontiguousarray(params, dtype=np.float64)
    ox = np.zeros(cap, np.int32)
    oy = np.zeros(cap, np.int32)
    cnt = lib().oracle_deformed_polygon_pixels(len(vx), _p(vx), _p(vy), _p(q), int(cx), int(cy), skin_factor, cap,
                                               _p(ox), _p(oy))
    if cnt < 0 or cnt > cap:
        raise ValueError("deformed_polygon_pixels: bad arguments")
    return ox[:cnt], oy[:cnt]


def circle_pixels(cx, cy, radius):
    """Circle::get_owned_pixels (DICe_Shape.cpp:221-279): integer (x, y) with (x - cx)^2 + (y - cy)^2 <=
    radius^2, searched over [c - radius - 1, c + radius + 1] (extents truncated to int).  Returns (x, y) in
    (y, x) order."""
    ys, xs = [], []
    r2 = float(radius) * float(radius)
    for y in range(int(cy - radius - 1), int(cy + radius + 1) + 1):
        for x in range(int(cx - radius - 1), int(cx + radius + 1) + 1):
            if float((x - cx) * (x - cx)) + float((y - cy) * (y - cy)) <= r2:
                ys.append(y)
                xs.append(x)
    return np.array(xs, np.int32), np.array(ys, np.int32)


def rectangle_pixels(cx, cy, width, height):
    """Rectangle::get_owned_pixels without deformation (DICe_Shape.cpp:281-420): origin = centre - size // 2,
    pixels [origin, origin + size) on each axis, in (y, x) order."""
    ox, oy = cx - width // 2, cy - height // 2
    yy, xx = np.mgrid[oy:oy + height, ox:ox + width]
    return xx.ravel().astype(np.int32), yy.ravel().astype(np.int32)


def gauss_filter(img, mask_size):
    """Image_::gauss_filter on a copy of img: float32 and uint16 images in float arithmetic,
    float64 images in double arithmetic."""
    out = np.ascontiguousarray(img).copy()
    h, w = out.shape
    if out.dtype == np.float64:
        rc = lib().oracle_gauss_filter_d(w, h, _p(out), mask_size)
    elif out.dtype in (np.float32, np.uint16):
        rc = lib().oracle_gauss_filter_f(w, h, 0 if out.dtype == np.float32 else 1, _p(out), mask_size)
    else:
        raise TypeError("gauss_filter: float32, uint16 or float64 image")
    if rc != 0:
        raise ValueError("gauss_filter: bad mask size or image too small")
    return out


def vsg_strain(fields, window):
    """VSG_Strain_Post_Processor: returns (7, n) VSG fields; sets MATCH = -1 in fields where it fails."""
    assert fields.dtype == np.float64 and fields.flags.c_contiguous and fields.shape[0] == NUM_FIELDS
    n = fields.shape[1]
    out = np.zeros((7, n))
    if lib().oracle_vsg_strain(n, _p(fields), window, _p(out)) != 0:
        raise ValueError("vsg_strain: bad arguments")
    return out


def objective(params, ref, defm, subset_size, cx, cy, call, p, precision="double"):
    """One Objective_ZNSSD call (0 gamma, 1 sigma, 2 computeUpdateFast, 3 computeUpdateRobust) on the square
    subset at (cx, cy); returns (parameters after the call, result[10])."""
    h, w = ref.shape
    if precision == "double":
        ref, defm = [np.ascontiguousarray(a, dtype=np.float64) if a.dtype != np.uint16 else a for a in (ref, defm)]
        storage = 1 if ref.dtype == np.uint16 else 2
        fn = lib().oracle_objective_d
    else:
        ref, defm = [np.ascontiguousarray(a, dtype=np.float32) if a.dtype != np.uint16 else a for a in (ref, defm)]
        storage = 1 if ref.dtype == np.uint16 else 0
        fn = lib().oracle_objective_f
    q = np.ascontiguousarray(p, dtype=np.float64).copy()
    r = np.zeros(10)
    if fn(ctypes.byref(params), w, h, storage, _p(ref), _p(defm), subset_size, int(cx), int(cy), call, _p(q), _p(r)):
        raise ValueError("objective: bad arguments")
    return q, r


def uncertainty(fields):
    """Uncertainty_Post_Processor (2-D): (2, n) array of UNCERTAINTY and UNCERTAINTY_ANGLE."""
    assert fields.dtype == np.float64 and fields.flags.c_contiguous and fields.shape[0] == NUM_FIELDS
    n = fields.shape[1]
    out = np.zeros((2, n))
    if lib().oracle_uncertainty(n, _p(fields), _p(out)) != 0:
        raise ValueError("uncertainty: bad arguments")
    return out


def lu_inverse(a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    info = lib().oracle_lu_inverse(a.shape[0], _p(a))
    return a, info


def smooth_gradients(gx, gy):
    """smooth_gradients_convolution_5_point on float32 copies of gx, gy."""
    gx = np.ascontiguousarray(gx, dtype=np.float32).copy()
    gy = np.ascontiguousarray(gy, dtype=np.float32).copy()
    h, w = gx.shape
    lib().oracle_smooth_gradients_f(w, h, _p(gx), _p(gy))
    return gx, gy


def nelder_mead(kind, p0, deltas, coef=None, max_iterations=1000, tolerance=1e-6, threshold=1e-10, cap=256):
    """Simplex::minimize on a test objective (kind 0: quadratic bowl sum coef[i] (p_i - coef[nd + i])^2,
    kind 1: double well sum (p_i^2 - 1)^2).  Returns dict(status, p (last evaluated point), iterations,
    evals, stop (bits: 1 early, 2 rtol, 4 rtol difference, 8 repeats, 16 cap), trace (evals x nd + 1))."""
    p = np.ascontiguousarray(p0, dtype=np.float64).copy()
    nd = len(p)
    d = np.ascontiguousarray(deltas, dtype=np.float64)
    c = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    trace = np.zeros((cap, nd + 1))
    ne, it, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    status = lib().oracle_nelder_mead(kind, nd, _p(p), _p(d), _p(c), max_iterations, tolerance, threshold, cap,
                                      _p(trace), ctypes.byref(ne), ctypes.byref(it), ctypes.byref(st))
    if status < 0:
        raise ValueError("nelder_mead: bad arguments")
    return {"status": status, "p": p, "iterations": it.value, "evals": ne.value, "stop": st.value,
            "trace": trace[:min(ne.value, cap)]}
