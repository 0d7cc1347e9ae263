"""dice_hip_correlate_device_noalloc (the correlation on the library's static work area, no allocation)
against dice_hip_correlate_device on the same inputs, through the C ABI: results bit-identical, over
calls in a row whose inputs shrink, covering the overflow lists (polygon subsets larger than the LDS
slot), the fallback list (GRADIENT_BASED_THEN_SIMPLEX) and the Nelder-Mead lane results; sizes beyond
the static area are rejected."""
import numpy as np
import pytest

import dice_amd as dh
from dice_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def both_ways(ctx, cx, cy, mode, first_frame=True, fields=None):
    f0 = dh.new_fields(cx, cy) if fields is None else fields
    a = torch.from_numpy(np.ascontiguousarray(f0)).cuda()
    b = a.clone()
    torch.cuda.synchronize()
    ctx.correlate_device(a.data_ptr(), mode=mode, first_frame=first_frame)
    ctx.synchronize()
    counts_a = ctx.pass_counts()
    ctx.correlate_device_noalloc(b.data_ptr(), mode=mode, first_frame=first_frame)
    ctx.synchronize()
    counts_b = ctx.pass_counts()
    return a.cpu().numpy(), b.cpu().numpy(), counts_a, counts_b


@pytest.mark.parametrize("mode", [dh.MODE_BATCHED, dh.MODE_PLAIN], ids=["batched", "plain"])
def test_noalloc_matches_correlate_device(mode):
    ref, seq, motion, subs = synth.config3(n_subsets=600, frames=1, size=512)
    pcx, pcy, off, px, py, act = subs
    ref1, de1 = synth.config1(seed=3)
    with dh.Context(0) as ctx:
        # 1: 600 polygon subsets of 200-3000 px, GN with the simplex fallback, from a zero guess
        p = dh.default_params()
        p.optimization_method = dh.GRADIENT_BASED_THEN_SIMPLEX
        p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
        p.max_solver_iterations_fast = 2
        ctx.set_params(p)
        ctx.set_ref_image(ref)
        ctx.set_def_image(seq[0])
        ctx.set_subsets(pcx, pcy, 0, conformal=(off, px, py, act))
        a, b, ca, cb = both_ways(ctx, pcx, pcy, mode)
        assert np.array_equal(a, b)
        assert ca == cb
        if mode == dh.MODE_BATCHED:
            assert ca[1] > 0  # subsets overflowed to the PLAIN kernel through the lists
        assert ca[0] > 0      # the fallback pass ran
        # 2: fewer, smaller subsets on another pair: 15 x 15 squares, Nelder-Mead (one subset per lane)
        q = dh.default_params()
        q.optimization_method = dh.SIMPLEX
        ctx.set_params(q)
        ctx.set_ref_image(ref1)
        ctx.set_def_image(de1)
        cx, cy = dh.regular_grid(512, 512, 15, 40)
        ctx.set_subsets(cx, cy, 15)
        a2, b2, ca2, cb2 = both_ways(ctx, cx, cy, mode)
        assert len(cx) < len(pcx)
        assert np.array_equal(a2, b2)
        assert np.isin(a2[dh.F["STATUS"]], dh.CONVERGED_STATUSES).mean() > 0.9
        # 3: fewer still, Gauss-Newton on 31 x 31 squares, a later frame
        q.optimization_method = dh.GRADIENT_BASED
        ctx.set_params(q)
        cx3, cy3 = cx[::7], cy[::7]
        ctx.set_subsets(cx3, cy3, 31)
        a3, b3, _, _ = both_ways(ctx, cx3, cy3, mode, first_frame=False)
        assert np.array_equal(a3, b3)


def test_noalloc_rejects_more_subsets_than_its_area():
    ref, de = synth.config1(seed=3)
    n = (1 << 18) + 1
    cx = np.full(n, 256.0)
    cy = np.full(n, 256.0)
    with dh.Context(0) as ctx:
        ctx.set_ref_image(ref)
        ctx.set_def_image(de)
        ctx.set_subsets(cx, cy, 15)
        f = torch.zeros((dh.NUM_FIELDS, n), dtype=torch.float64, device="cuda")
        with pytest.raises(dh.DiceHipError, match="static work area"):
            ctx.correlate_device_noalloc(f.data_ptr())
