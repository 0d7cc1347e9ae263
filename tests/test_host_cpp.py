"""The C++ Schema mirror (dice_amd/host) driven by tests/cpp/test_schema, a C++
program in the shape of the reference's component tests.  CPU tests cover the
host logic (parameters, polygon raster); GPU tests run the correlation."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_schema")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args, timeout=120):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_schema is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    print(r.stderr)
    return r


def raw(tmp_path, name, img):
    p = tmp_path / name
    np.ascontiguousarray(img, dtype=np.float32).tofile(p)
    return str(p)


def test_cpp_parameter_errors_and_defaults():
    assert run("errors").returncode == 0


def test_cpp_polygon_raster_matches_oracle_and_golden():
    out = run("polygon").stdout.split()
    pts = np.array(out, dtype=int).reshape(-1, 2)  # (y, x) rows
    xs, ys = [], []
    for vx, vy in [([940, 951, 964, 980, 959], [422, 399, 410, 413, 432]),
                   ([982, 1000, 991, 978], [426, 428, 456, 445])]:
        px, py = o.polygon_pixels(vx, vy)
        xs.append(px)
        ys.append(py)
    oracle = np.stack([np.concatenate(ys), np.concatenate(xs)], axis=1)
    assert np.array_equal(pts, oracle)
    g = np.load(os.path.join(GOLD, "conformal_raster.npz"))["conformal_d"]
    union = sorted(set(map(tuple, pts.tolist())))
    min_y = min(p[0] for p in union)
    min_x = min(p[1] for p in union)
    mask = np.zeros(g.shape, bool)
    for yy, xx in union:
        mask[yy - min_y, xx - min_x] = True
    assert np.all(mask[g != 0])


def test_batch_polygon_raster_matches_oracle():
    """dice_polygon_pixels (OpenMP over polygons) against the oracle's winding-rule raster, on the
    golden's two polygons and on random convex polygons of the uneven-batching workload's generator."""
    from dice_amd import synth
    polys = [([940, 951, 964, 980, 959], [422, 399, 410, 413, 432]), ([982, 1000, 991, 978], [426, 428, 456, 445])]
    rng = np.random.default_rng(3)
    for _ in range(40):
        k = int(rng.integers(3, 10))
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = rng.uniform(3, 30)
        c = rng.uniform(40, 80, 2)
        polys.append((np.round(c[0] + r * np.cos(ang)).astype(int).tolist(),
                      np.round(c[1] + r * np.sin(ang)).astype(int).tolist()))
    vo = np.cumsum([0] + [len(p[0]) for p in polys])
    off, px, py = synth.rasterize_polygons(vo, np.concatenate([p[0] for p in polys]),
                                           np.concatenate([p[1] for p in polys]))
    for i, (vx, vy) in enumerate(polys):
        ox, oy = o.polygon_pixels(vx, vy)
        assert np.array_equal(px[off[i]:off[i + 1]], ox), i
        assert np.array_equal(py[off[i]:off[i + 1]], oy), i


def test_conformal_subset_generator():
    """Polygon subsets of the uneven-batching workload: pixel counts within [200, 3000], y-major order,
    inside the image with the interpolation margin."""
    from dice_amd import synth
    cx, cy, off, px, py, act = synth.conformal_subsets(300, 1024, 1024, seed=2)
    cnt = np.diff(off)
    assert len(cx) == 300 and cnt.min() >= 200 and cnt.max() <= 3000
    assert px.min() >= 4 and py.min() >= 4 and px.max() <= 1019 and py.max() <= 1019
    for i in range(0, 300, 37):
        key = py[off[i]:off[i + 1]].astype(np.int64) * 4096 + px[off[i]:off[i + 1]]
        assert np.all(np.diff(key) > 0)
        assert abs(px[off[i]:off[i + 1]].mean() - cx[i]) <= 0.5 and abs(py[off[i]:off[i + 1]].mean() - cy[i]) <= 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_cpp_correlate_params(tmp_path, backend):
    """DICe_TestCorrelateParams.cpp:78-140 through the C++ Schema mirror."""
    d = np.load(os.path.join(GOLD, "subset_construction.npz"))
    h, w = d["ref"].shape
    r = run("correlate_params", raw(tmp_path, "ref.f32", d["ref"]), raw(tmp_path, "def.f32", d["defm"]), w, h,
            backend)
    assert r.returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_cpp_correlate_conformal(tmp_path, backend):
    """DICe_TestCorrelateConformal.cpp:75-135 (one polygon subset, tolerance 5e-2)."""
    d = np.load(os.path.join(GOLD, "subset_construction.npz"))
    h, w = d["ref"].shape
    r = run("correlate_conformal", raw(tmp_path, "ref.f32", d["ref"]), raw(tmp_path, "def.f32", d["defm"]), w, h,
            backend)
    assert r.returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_cpp_parallel_subsets_gold(tmp_path, backend):
    """tests/regression/parallel_subsets gold through the C++ Schema mirror."""
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    h, w = d["ref"].shape
    r = run("parallel_subsets", raw(tmp_path, "ref.f32", d["ref"]), raw(tmp_path, "def.f32", d["defm"]), w, h,
            os.path.join(GOLD, "parallel_subsets_gold.txt"), backend)
    assert r.returncode == 0
