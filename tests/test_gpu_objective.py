"""Objective_ZNSSD for one subset on the GPU (dice_hip_objective, include/dice_hip.h; the C++ mirror's
Objective_ZNSSD in dice_amd/host) against the oracle's restatement of the same calls
(DICe_Objective.cpp:58-76 gamma, :150-190 sigma, :430-657 computeUpdateFast, :369-388 computeUpdateRobust).

Bar: gamma within 1e-5 relative, sigma and noise within 1e-4 relative, the Gauss-Newton update within
1e-4 (parameters) with the same status and iteration count, its condition number within 1e-3 relative;
Nelder-Mead within its float/double spread (see tests/test_gpu_parity.py)."""
import os

import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from dice_amd import synth
from test_host_cpp import raw, run

pytestmark = pytest.mark.gpu

PTS = [(60, 60), (200, 250), (431, 97), (300, 440), (255, 255), (120, 400)]


@pytest.fixture(scope="module")
def case():
    ref, de = synth.config1()
    cx = np.array([p[0] for p in PTS], np.float64)
    cy = np.array([p[1] for p in PTS], np.float64)
    c = dh.Context(0)
    p = dh.default_params()
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    c.set_params(p)
    c.set_ref_image(ref)
    c.set_def_image(de)
    c.set_subsets(cx, cy, 31)
    q = o.Params()
    for name, _ in dh.Params._fields_:
        setattr(q, name, getattr(p, name))
    yield c, q, ref, de, cx, cy
    c.close()


def guess(x, y, off=0.3):
    return np.array([0.5 + 0.01 * x + off, -0.3 + 0.005 * y - off, 0.0, 0.0, 0.0, 0.0])


def test_objective_gamma_and_sigma(case):
    c, q, ref, de, cx, cy = case
    for i, (x, y) in enumerate(zip(cx, cy)):
        for g0 in (guess(x, y, 0.0), guess(x, y, 0.4)):
            _, rg = c.objective(i, dh.OBJ_GAMMA, g0)
            _, ro = o.objective(q, ref, de, 31, x, y, 0, g0)
            assert abs(rg[0] - ro[0]) <= 1e-5 * abs(ro[0]), (i, rg[0], ro[0])
            _, sg = c.objective(i, dh.OBJ_SIGMA, g0)
            _, so = o.objective(q, ref, de, 31, x, y, 1, g0)
            assert abs(sg[0] - so[0]) <= 1e-4 * abs(so[0]) and abs(sg[1] - so[1]) <= 1e-4 * abs(so[1]), i


def test_objective_update_fast(case):
    c, q, ref, de, cx, cy = case
    for i, (x, y) in enumerate(zip(cx, cy)):
        pg, rg = c.objective(i, dh.OBJ_UPDATE_FAST, guess(x, y))
        po, ro = o.objective(q, ref, de, 31, x, y, 2, guess(x, y))
        assert rg[0] == ro[0] == 0 and rg[1] == ro[1], (i, rg[:2], ro[:2])
        assert np.abs(pg - po).max() <= 1e-4, (i, pg, po)
        assert abs(rg[2] - ro[2]) <= 1e-3 * abs(ro[2])          # CONDITION_NUMBER
        assert abs(rg[3] - ro[3]) <= 1e-3 * abs(ro[3])          # FIELD_1
    # against the imposed affine field at the subset centre
    assert abs(pg[0] - (0.5 + 0.01 * cx[-1])) < 0.01 and abs(pg[3] - 0.01) < 0.002


def test_objective_update_robust(case):
    c, q, ref, de, cx, cy = case
    d = []
    for i, (x, y) in enumerate(zip(cx, cy)):
        pg, rg = c.objective(i, dh.OBJ_UPDATE_ROBUST, guess(x, y))
        po, ro = o.objective(q, ref, de, 31, x, y, 3, guess(x, y), precision="double")
        pf, rf = o.objective(q, ref, de, 31, x, y, 3, guess(x, y), precision="float")
        assert rg[0] == 0 and ro[0] == 0
        d.append((np.abs(pg[:2] - po[:2]).max(), np.abs(pf[:2] - po[:2]).max()))
    d = np.array(d)
    assert np.median(d[:, 0]) <= 2.0 * np.median(d[:, 1]) + 1e-4 and d[:, 0].max() < 2e-2


def test_cpp_objective_matches_c_abi(case, tmp_path):
    """The C++ mirror's Objective_ZNSSD on the same subset gives what the C ABI gives."""
    c, q, ref, de, cx, cy = case
    h, w = ref.shape
    g0 = guess(cx[1], cy[1])
    r = run("objective", raw(tmp_path, "r.f32", ref), raw(tmp_path, "d.f32", de), w, h, 31, int(cx[1]), int(cy[1]),
            g0[0], g0[1])
    assert r.returncode == 0, r.stderr
    out = {ln.split()[0]: [float(t) for t in ln.split()[1:]] for ln in r.stdout.splitlines()}
    _, rg = c.objective(1, dh.OBJ_GAMMA, g0)
    assert abs(out["gamma"][0] - rg[0]) <= 1e-7 * abs(rg[0])
    pf, rf = c.objective(1, dh.OBJ_UPDATE_FAST, g0)
    assert out["fast"][0] == rf[0] and out["fast"][1] == rf[1]
    assert np.abs(np.array(out["fast"][2:8]) - pf).max() <= 1e-7
