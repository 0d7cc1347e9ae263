"""CPU checks of the C-ABI boundary: the library loads, exports every entry point
that include/dice_hip.h declares, and the parameter struct layout agrees between
the C header, the Python binding and the oracle.  No GPU work."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import dice_amd as dh
import oracle_py as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dice_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dice_hip_\w+)\s*\(", text)))


def test_header_declarations_match_binding_list():
    assert declared_functions() == sorted(dh.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(dh.LIB_PATH):
        pytest.skip("libdice_hip.so not built")
    lib = ctypes.CDLL(dh.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_reference_entry_points_are_exported():
    """include/dice_api.h: dice_correlate from libdice_hip.so, dice_correlate_conformal from
    libdice_schema.so, under the reference's names (DICe_api.h:129-133, 176-180)."""
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "dice_api.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(dice_correlate\w*)\s*\(", text)))
    assert names == ["dice_correlate", "dice_correlate_conformal"]
    lib_dir = os.path.join(ROOT, "dice_amd", "lib")
    if not os.path.exists(os.path.join(lib_dir, "libdice_schema.so")):
        pytest.skip("libraries not built")
    assert hasattr(ctypes.CDLL(os.path.join(lib_dir, "libdice_hip.so")), "dice_correlate")
    assert hasattr(ctypes.CDLL(os.path.join(lib_dir, "libdice_schema.so")), "dice_correlate_conformal")


def test_create_without_gpu_fails_cleanly():
    """With no visible GPU, context creation reports an error code instead of crashing."""
    if not os.path.exists(dh.LIB_PATH):
        pytest.skip("libdice_hip.so not built")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(dh.DiceHipError):
        dh.Context(0)


def test_params_struct_layout():
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "dice_hip.h"
#define P(f) printf("%s %zu\n", #f, offsetof(dice_params, f));
int main(void) {
  printf("sizeof %zu\n", sizeof(dice_params));
  P(interpolation_method) P(max_solver_iterations_robust) P(fast_solver_tolerance)
  P(momentum_factor) P(compute_def_gradients) P(gradient_method)
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = dict(line.split() for line in subprocess.check_output([exe]).decode().splitlines())
    assert int(out["sizeof"]) == ctypes.sizeof(dh.Params) == ctypes.sizeof(o.Params)
    for name in ("interpolation_method", "max_solver_iterations_robust", "fast_solver_tolerance",
                 "momentum_factor", "compute_def_gradients", "gradient_method"):
        assert int(out[name]) == getattr(dh.Params, name).offset == getattr(o.Params, name).offset


def test_default_params_match_reference_defaults():
    """dice_default_params / tracking_default_params (DICe_ParameterUtilities.cpp:297-420)."""
    if not os.path.exists(dh.LIB_PATH):
        pytest.skip("libdice_hip.so not built")
    p = dh.default_params()
    assert (p.interpolation_method, p.optimization_method, p.max_solver_iterations_fast) == (2, 4, 25)
    assert p.fast_solver_tolerance == 1e-4 and p.robust_solver_tolerance == 1e-6
    assert (p.enable_rotation, p.enable_normal_strain, p.enable_shear_strain) == (0, 0, 0)
    assert p.disp_jump_tol == 10000.0 and p.theta_jump_tol == 100.0
    t = dh.tracking_default_params()
    assert (t.optimization_method, t.enable_rotation, t.correlation_routine) == (0, 1, 1)
    assert t.disp_jump_tol == 8.0 and t.theta_jump_tol == 0.1
    # the oracle's defaults are the same numbers
    q = o.default_params()
    for name, _ in dh.Params._fields_:
        assert getattr(q, name) == getattr(p, name), name


def test_regular_grid_convention():
    """Schema(w, h, step, step, S) grid (DICe_Schema.cpp:115-133): first centre S-1, count (W-2S)/step+1."""
    cx, cy = dh.regular_grid(512, 512, 31, 10)
    assert len(cx) == 46 * 46 and cx[0] == 30 and cy[0] == 30 and cx.max() == 30 + 45 * 10
    cx, cy = dh.regular_grid(2048, 2048, 41, 5)
    assert len(cx) == 394 * 394
