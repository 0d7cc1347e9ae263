"""USE_NEIGHBOR_VALUES on the GPU (the rounds of dice_hip_correlate, include/dice_hip.h) against the
oracle, which correlates the subsets one after another as the reference does
(Field_Value_Initializer::initial_guess, DICe_Initializer.cpp:398-434; execute_correlation's subset loop,
DICe_Schema.cpp:1720-1735).  The seeded regression decks run end to end in tests/test_host_cpp.py.

Bar: that of tests/test_gpu_parity.py (displacement / rotation 1e-4, gamma 1e-5 relative, status >= 99.9 %),
with every status identical here."""
import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from dice_amd import synth
from test_gpu_parity import MODE_NAMES, MODES, compare, run_gpu, run_oracle, simplex_spread_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dh.Context(0)
    yield c
    c.close()


def chain_case(size=512, S=31, step=12):
    ref, de = synth.config1()
    cx, cy, nb = dh.serpentine_grid(size, size, S, step)
    f = dh.new_fields(cx, cy)
    f[dh.F["NEIGHBOR_ID"]] = nb
    return ref, de, cx, cy, f


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("opt", [dh.GRADIENT_BASED_THEN_SIMPLEX, dh.SIMPLEX])
def test_neighbor_chain_parity(ctx, mode, opt):
    """The serpentine grid over the 512^2 affine speckle pair (u = 0.5 + 0.01 x, v = -0.3 + 0.005 y), every
    subset starting from zero fields except through its neighbour: one chain of 1444 subsets (Gauss-Newton,
    affine parameters on) or 225 (Nelder-Mead, u and v).  The GPU rounds against the sequential oracle;
    Nelder-Mead against the spread between the oracle's float and double builds (see
    simplex_spread_check), which the chain carries from subset to subset."""
    simplex = opt == dh.SIMPLEX
    ref, de, cx, cy, f = chain_case(step=30 if simplex else 12)
    p = dh.default_params()
    p.initialization_method = dh.USE_NEIGHBOR_VALUES
    p.optimization_method = opt
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 0 if simplex else 1
    fg = run_gpu(ctx, p, ref, de, cx, cy, 31, mode, fields=f)
    rounds = ctx.neighbor_rounds()
    fo = run_oracle(p, ref, de, cx, cy, 31, fields=f)
    print("subsets %d, rounds %d" % (len(cx), rounds))
    check = (lambda a, label: simplex_spread_check(a, run_oracle(p, ref, de, cx, cy, 31, fields=f, precision="float"), fo)) \
        if simplex else (lambda a, label: compare(a, fo, status_agree=1.0, label=label))
    check(fg, "neighbor chain")
    st = fg[dh.F["STATUS"]]
    assert st[0] == 1 and st[1] == 1 and np.mean(st == dh.INITIALIZE_USING_NEIGHBOR_VALUE_SUCCESSFUL) > 0.95
    assert 1 <= rounds <= len(cx) + 1
    # a settle tolerance of 1e-6 ends the Gauss-Newton rounds early, within the same bar (a Nelder-Mead
    # result moves by more than that when its start moves at all, so its chain still takes every round)
    p.neighbor_settle_tolerance = 1e-6
    ft = run_gpu(ctx, p, ref, de, cx, cy, 31, mode, fields=f)
    print("settle 1e-6: rounds %d" % ctx.neighbor_rounds())
    assert ctx.neighbor_rounds() <= rounds if simplex else ctx.neighbor_rounds() < rounds // 10
    check(ft, "neighbor chain, settle 1e-6")
    # from zero fields alone (USE_FIELD_VALUES) the far subsets (u up to 5 px) are not found the same way
    q = dh.default_params()
    q.optimization_method = opt
    q.enable_rotation = q.enable_normal_strain = q.enable_shear_strain = p.enable_rotation
    fz = run_gpu(ctx, q, ref, de, cx, cy, 31, mode, fields=f)
    assert ctx.neighbor_rounds() == 0
    assert not np.array_equal(fz[dh.F["DISP_X"]], fg[dh.F["DISP_X"]])


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_neighbor_ids_in_any_order(ctx, mode):
    """Random neighbour ids: lower ids give this frame's result, higher ids the neighbour's frame-start
    values, -1 and self the subset's own; frame-start guesses near the truth for some subsets, failed
    (sigma -1) for a few, so that failures travel down their chains as INITIALIZE_FAILED."""
    ref, de = synth.config1()
    cx, cy = dh.regular_grid(512, 512, 31, 16)
    n = len(cx)
    rng = np.random.default_rng(5)
    f = dh.new_fields(cx, cy)
    nb = rng.integers(-1, n, size=n).astype(np.float64)
    nb[rng.random(n) < 0.1] = -1.0
    nb[::17] = np.arange(n)[::17]  # some subsets name themselves
    f[dh.F["NEIGHBOR_ID"]] = nb
    f[dh.F["DISP_X"]] = 0.5 + 0.01 * cx + rng.normal(0, 0.3, n)
    f[dh.F["DISP_Y"]] = -0.3 + 0.005 * cy + rng.normal(0, 0.3, n)
    f[dh.F["SIGMA"], rng.random(n) < 0.03] = -1.0
    p = dh.default_params()
    p.initialization_method = dh.USE_NEIGHBOR_VALUES
    fg = run_gpu(ctx, p, ref, de, cx, cy, 31, mode, fields=f)
    fo = run_oracle(p, ref, de, cx, cy, 31, fields=f)
    compare(fg, fo, status_agree=1.0, label="neighbor ids")
    assert np.any(fg[dh.F["STATUS"]] == 5)  # INITIALIZE_FAILED reached through a neighbour


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_neighbor_first_step_only(ctx, mode):
    """USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY: neighbours on the first frame, the subset's own fields after."""
    ref, de, cx, cy, f = chain_case(step=20)
    p = dh.default_params()
    p.initialization_method = dh.USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY
    fg, fo = f.copy(), f.copy()
    for frame in range(2):
        fg = run_gpu(ctx, p, ref, de, cx, cy, 31, mode, fields=fg, first_frame=frame == 0)
        assert (ctx.neighbor_rounds() > 0) == (frame == 0)
        fo = run_oracle(p, ref, de, cx, cy, 31, fields=fo, first_frame=frame == 0)
        compare(fg, fo, status_agree=1.0, label="first step only, frame %d" % frame)
    assert np.mean(fg[dh.F["STATUS"]] == 1) > 0.95


def test_neighbor_errors(ctx):
    ref, de, cx, cy, f = chain_case(step=40)
    p = dh.default_params()
    p.initialization_method = dh.USE_NEIGHBOR_VALUES
    bad = f.copy()
    bad[dh.F["NEIGHBOR_ID"], 3] = len(cx)
    with pytest.raises(dh.DiceHipError):
        run_gpu(ctx, p, ref, de, cx, cy, 31, dh.MODE_BATCHED, fields=bad)
    bad[dh.F["NEIGHBOR_ID"], 3] = 1.5
    with pytest.raises(dh.DiceHipError):
        run_gpu(ctx, p, ref, de, cx, cy, 31, dh.MODE_BATCHED, fields=bad)
    p.initialization_method = 3
    with pytest.raises(dh.DiceHipError):
        ctx.set_params(p)
