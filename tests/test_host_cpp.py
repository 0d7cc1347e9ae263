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
    # the float build's golden (the library rasterises in float, as the reference's default build)
    g = np.load(os.path.join(GOLD, "conformal_raster.npz"))["conformal_f"]
    union = sorted(set(map(tuple, pts.tolist())))
    min_y = min(p[0] for p in union)
    min_x = min(p[1] for p in union)
    mask = np.zeros(g.shape, bool)
    for yy, xx in union:
        mask[yy - min_y, xx - min_x] = True
    assert np.array_equal(mask, g != 0)


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


def seeded_grid(name):
    """test_schema grid over the deck's regions of interest: points, neighbour ids and seeds."""
    ddir = os.path.join(DECKS, name)
    r = run("grid", 512, 512, 27, 10, os.path.join(ddir, "subsets.txt"))
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines()]
    pts = np.array([[float(t) for t in ln] for ln in lines if ln[0] != "seed"])
    seeds = {int(ln[1]): (float(ln[2]), float(ln[3])) for ln in lines if ln[0] == "seed"}
    return pts, seeds


def test_seeded_grid_matches_regression_gold():
    """Regions of interest with seeds (tests/regression/parallel_subsets_seed): four 50 x 50 rectangles,
    seeds nearest (25, 25), (65, 25), (25, 65), (65, 65) with displacement (0.4, 0.4).  The grid has the
    gold's 64 centres in the gold's order: per region the seed, the serpentine to its right, then the one
    to its left; the seed's neighbour is -1, each serpentine starts from the seed."""
    pts, seeds = seeded_grid("parallel_subsets_seed")
    rows = [ln.split() for ln in open(os.path.join(DECKS, "parallel_subsets_seed", "gold", "DICe_solution_0.txt"))
            if ln[:1].isdigit()]
    assert np.array_equal(pts[:, :2], np.array([[float(t[1]), float(t[2])] for t in rows]))
    assert sorted(seeds) == [0, 9, 24, 39] and all(v == (0.4, 0.4) for v in seeds.values())
    nb = pts[:, 2].astype(int)
    assert all(nb[s] == -1 for s in seeds)
    # region 2: seed (66, 26), right serpentine from it (seed neighbour), then left serpentine from the seed
    assert list(pts[9:13, :2].ravel()) == [66, 26, 66, 36, 66, 46, 76, 46]
    assert nb[10] == 9 and nb[11] == 10
    left = [i for i in range(10, 24) if pts[i, 0] < 66]
    assert left and nb[left[0]] == 9 and list(pts[left[0], :2]) == [56, 26]


@pytest.mark.parametrize("name,frames", [("parallel_subsets_seed", 1), ("parallel_seed_first_step_only", 2)])
def test_oracle_neighbor_initialization_matches_seed_golds(name, frames):
    """USE_NEIGHBOR_VALUES (and _FIRST_STEP_ONLY) in the oracle: the subsets one after another, each from
    its neighbour's result, on the seeded grid -- the golds of both decks (double build) to the printed
    digits: u, v within 1e-5, sigma within 2e-6; neighbour-initialized subsets record status 3, the seeds
    1, and the second frame of the first-step-only deck (field values) 1 everywhere."""
    pts, seeds = seeded_grid(name)
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    p = o.default_params()
    p.initialization_method = 1 if frames == 1 else 2
    f = o.new_fields(pts[:, 0], pts[:, 1])
    f[o.F["NEIGHBOR_ID"]] = pts[:, 2]
    for sid, (u, v) in seeds.items():
        f[o.F["DISP_X"], sid], f[o.F["DISP_Y"], sid] = u, v
    for frame in range(frames):
        o.correlate(p, d["ref"], d["defm"], f, subset_size=27, first_frame=frame == 0, precision="double")
        g = np.array([[float(t) for t in ln.split()] for ln in
                      open(os.path.join(DECKS, name, "gold", "DICe_solution_%d.txt" % frame)) if ln[:1].isdigit()])
        assert np.abs(f[o.F["DISP_X"]] - g[:, 3]).max() <= 1e-5
        assert np.abs(f[o.F["DISP_Y"]] - g[:, 4]).max() <= 1e-5
        assert np.abs(f[o.F["SIGMA"]] - g[:, 5]).max() <= 2e-6
        want = np.full(len(pts), 1.0 if frame else 3.0)
        if frame == 0:
            want[list(seeds)] = 1.0
        assert np.array_equal(f[o.F["STATUS"]], want)
    # starting every subset from its own (zero) fields instead does not reproduce the chain
    q = o.default_params()
    f2 = o.new_fields(pts[:, 0], pts[:, 1])
    o.correlate(q, d["ref"], d["defm"], f2, subset_size=27, precision="double")
    assert not np.array_equal(f2[o.F["DISP_X"]], f[o.F["DISP_X"]])


VSG_DECK = os.path.join(GOLD, "regression_decks", "dic_challenge_12_sorted")


def vsg_deck_frames():
    import json
    deck = json.load(open(os.path.join(VSG_DECK, "deck.json")))
    im = np.load(os.path.join(VSG_DECK, deck["images"]))
    return deck, [im["frame_%d" % k] for k in range(3)]


def gold_csv(path):
    return np.array([[float(t) for t in ln.split(",")] for ln in open(path) if ln[:1].isdigit() or ln[:1] == "-"])


def test_oracle_vsg_deck_matches_gold():
    """tests/regression/dic_challenge_12_sorted, frames 1 and 2: the grid in a region of interest with an
    excluded circle, USE_NEIGHBOR_VALUES_FIRST_STEP_ONLY, u, v and normal strains by Gauss-Newton, then the
    VSG strains over a 360-px window; the rows sorted by x, then y.  The oracle (double build) against
    the gold to its printed digits: coordinates and MATCH exactly, u and v within 5e-5 px, the VSG strains,
    gamma within 2e-4 relative (5 significant digits are printed)."""
    deck, frames = vsg_deck_frames()
    h, w = frames[0].shape
    r = run("grid", w, h, deck["subset_size"], deck["step_size"], os.path.join(VSG_DECK, deck["subset_file"]))
    pts = np.array([[float(t) for t in ln.split()] for ln in r.stdout.splitlines() if not ln.startswith("seed")])
    p = o.default_params()
    p.initialization_method = deck["initialization_method"]
    p.optimization_method = deck["optimization_method"]
    p.enable_normal_strain = deck["enable_normal_strain"]
    f = o.new_fields(pts[:, 0], pts[:, 1])
    f[o.F["NEIGHBOR_ID"]] = pts[:, 2]
    for k, gold in zip((1, 2), deck["gold"]):
        o.correlate(p, frames[0], frames[k], f, subset_size=deck["subset_size"], first_frame=k == 1,
                    precision="double")
        vsg = o.vsg_strain(f, deck["vsg_window"])
        g = gold_csv(os.path.join(VSG_DECK, "gold", gold))
        order = np.lexsort((f[o.F["COORD_Y"]], f[o.F["COORD_X"]]))
        ours = np.stack([f[o.F["COORD_X"]], f[o.F["COORD_Y"]], f[o.F["DISP_X"]], f[o.F["DISP_Y"]], vsg[0], vsg[1],
                         f[o.F["GAMMA"]], f[o.F["MATCH"]]], 1)[order]
        assert ours.shape == g.shape
        assert np.array_equal(ours[:, [0, 1, 7]], g[:, [0, 1, 7]])
        assert np.abs(ours[:, 2:4] - g[:, 2:4]).max() <= 5e-5
        rel = np.abs(ours[:, 4:7] - g[:, 4:7]) / np.abs(g[:, 4:7])
        assert rel.max() <= 2e-4


def test_oracle_vsg_closed_forms():
    """The VSG plane fit: a linear displacement field gives its gradient exactly (to the float storage of
    u and v) and the Green-Lagrange strains of that gradient; subsets with sigma < 0, with fewer than 3
    valid neighbours, or whose valid neighbours lie on one line get zeros and MATCH = -1."""
    xs, ys = np.meshgrid(np.arange(20.0, 400.0, 10.0), np.arange(20.0, 300.0, 10.0))
    cx, cy = xs.ravel(), ys.ravel()
    f = o.new_fields(cx, cy)
    a, b, c, d = 0.01, -0.002, 0.004, 0.003
    f[o.F["DISP_X"]] = 0.3 + a * cx + b * cy
    f[o.F["DISP_Y"]] = -0.1 + c * cx + d * cy
    f[o.F["SIGMA"]] = 0.01
    f[o.F["SIGMA"], 5] = -1.0
    v = o.vsg_strain(f, 41)
    ok = np.arange(len(cx)) != 5
    for k, want in ((3, a), (4, b), (5, c), (6, d)):
        assert np.abs(v[k][ok] - want).max() < 2e-8
    assert np.allclose(v[0][ok], 0.5 * (2 * a + a * a + c * c), atol=1e-7)
    assert np.allclose(v[1][ok], 0.5 * (2 * d + b * b + d * d), atol=1e-7)
    assert np.allclose(v[2][ok], 0.5 * (b + c + a * b + c * d), atol=1e-7)
    assert np.all(v[:, 5] == 0) and f[o.F["MATCH"], 5] == -1 and np.all(f[o.F["MATCH"]][ok] == 0)
    # one row of subsets: every neighbourhood is collinear (singular normal matrix); two points: too few
    g = o.new_fields(np.arange(10.0, 200.0, 10.0), np.full(19, 50.0))
    g[o.F["SIGMA"]] = 0.01
    assert np.all(o.vsg_strain(g, 41) == 0) and np.all(g[o.F["MATCH"]] == -1)
    h = o.new_fields(np.array([10.0, 20.0]), np.array([10.0, 12.0]))
    assert np.all(o.vsg_strain(h, 41) == 0) and np.all(h[o.F["MATCH"]] == -1)


def test_oracle_vsg_k_nearest_matches_radius_on_complete_shells():
    """On a regular grid the 9 nearest subsets are the 3 x 3 block, the same set as a 30-px window (radius
    15 > 10 sqrt 2), away from the grid's edge: both forms give the same strains bit for bit."""
    xs, ys = np.meshgrid(np.arange(20.0, 300.0, 10.0), np.arange(20.0, 300.0, 10.0))
    cx, cy = xs.ravel(), ys.ravel()
    f = o.new_fields(cx, cy)
    f[o.F["DISP_X"]] = (0.01 * cx + 0.002 * cy + 1e-5 * cx * cy).astype(np.float32)
    f[o.F["DISP_Y"]] = (0.003 * cx - 2e-5 * cy * cy).astype(np.float32)
    f[o.F["SIGMA"]] = 0.01
    knn, rad = o.vsg_strain(f.copy(), -9), o.vsg_strain(f.copy(), 30)
    inner = (cx > 20) & (cx < 290) & (cy > 20) & (cy < 290)
    assert np.array_equal(knn[:, inner], rad[:, inner]) and np.any(knn[0, inner] != 0)


def test_oracle_uncertainty_post_processor():
    """Uncertainty_Post_Processor (2-D): sigma over the cosine in FIELD_1, rounded as the reference's
    1.0 / angle * sigma (double literal, stored in float); zero for failed subsets and a zero cosine."""
    f = o.new_fields(np.arange(4.0), np.zeros(4))
    f[o.F["SIGMA"]] = [0.02, 0.013, -1.0, 0.5]
    f[o.F["FIELD_1"]] = np.array([0.5, 0.7, 0.9, 0.0], np.float32)
    u = o.uncertainty(f)
    want = np.float32(1.0 / np.float64(np.float32(0.7)) * np.float64(np.float32(0.013)))
    assert u[0, 0] == np.float32(0.04) and u[0, 1] == want and u[0, 2] == 0 and u[0, 3] == 0
    assert np.array_equal(u[1], np.array([0.5, 0.7, 0.9, 0.0], np.float32))


def test_serpentine_grid_helper_matches_cpp():
    import dice_amd as dh
    for w, h, s, step in ((512, 512, 27, 75), (300, 211, 21, 17), (128, 640, 31, 9)):
        r = run("grid", w, h, s, step)
        pts = np.array(r.stdout.split(), dtype=float).reshape(-1, 3)
        cx, cy, nb = dh.serpentine_grid(w, h, s, step)
        assert np.array_equal(pts[:, 0], cx) and np.array_equal(pts[:, 1], cy)
        assert np.array_equal(pts[:, 2], nb)


def test_subset_file_seed_errors(tmp_path):
    """A seed needs a location and a displacement guess (DICe_Parser.cpp:1350-1351)."""
    bad = tmp_path / "bad.txt"
    bad.write_text("BEGIN REGION_OF_INTEREST\n BEGIN BOUNDARY\n BEGIN RECTANGLE\n UPPER_LEFT 0 0\n"
                   " LOWER_RIGHT 50 50\n END\n END\n BEGIN SEED\n LOCATION 25 25\n END\nEND\n")
    r = run("grid", 512, 512, 27, 10, bad)
    assert r.returncode == 98 and "displacement" in r.stderr


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


# ---- subset files, output files and the regression deck tests/regression/conformal_subset_def ----

DECKS = os.path.join(GOLD, "regression_decks")
DECK = os.path.join(DECKS, "conformal_subset_def")
DIFF = os.path.join(ROOT, "dice_amd", "lib", "dice_diff")


def subset_info(path):
    """Parse test_schema subset_file_info: coordinates, and per conformal subset (pixels y x active, meta)."""
    r = run("subset_file_info", path)
    assert r.returncode == 0, r.stderr
    coords, conf, cur = [], {}, None
    for line in r.stdout.splitlines():
        t = line.split()
        if t[0] == "coord":
            coords.append((float(t[1]), float(t[2])))
        elif t[0] == "conformal":
            cur = int(t[1])
            conf[cur] = {"n": int(t[2]), "active": int(t[3]), "obstructed": int(t[4]),
                         "blocking": [int(v) for v in t[5:]], "pix": []}
        else:
            conf[cur]["pix"].append(tuple(int(v) for v in t))
    return coords, conf


def union_pixels(*shapes):
    s = set()
    for px, py in shapes:
        s |= set(zip(py.tolist(), px.tolist()))
    return sorted(s)


def test_circle_and_rectangle_shapes(tmp_path):
    """Circle pixel counts are the Gauss circle numbers N(r) = #{(x, y) : x^2 + y^2 <= r^2} (5, 13, 29, 81,
    317 for r = 1, 2, 3, 5, 10); rectangles by centre/size and by corners; C++ reader and oracle agree."""
    gauss = {1: 5, 2: 13, 3: 29, 5: 81, 10: 317}
    lines = ["BEGIN SUBSET_COORDINATES"] + ["%d 50" % (20 + 30 * i) for i in range(len(gauss) + 2)] + \
            ["END SUBSET_COORDINATES"]
    for i, r in enumerate(gauss):
        lines += ["BEGIN CONFORMAL_SUBSET", "SUBSET_ID %d" % i, "BEGIN BOUNDARY", "BEGIN CIRCLE",
                  "CENTER %d 50" % (20 + 30 * i), "RADIUS %d" % r, "END CIRCLE", "END BOUNDARY", "END CONFORMAL_SUBSET"]
    k = len(gauss)
    lines += ["begin conformal_subset", "subset_id %d" % k, "begin boundary", "begin rectangle", "center 200 50",
              "width 7", "height 4", "end rectangle", "end boundary", "end conformal_subset"]
    lines += ["BEGIN CONFORMAL_SUBSET", "SUBSET_ID %d" % (k + 1), "BEGIN BOUNDARY", "BEGIN RECTANGLE",
              "UPPER_LEFT 227 45 # corners", "LOWER_RIGHT 236 52", "END RECTANGLE", "END BOUNDARY",
              "END CONFORMAL_SUBSET"]
    f = tmp_path / "shapes.txt"
    f.write_text("\n".join(lines) + "\n")
    coords, conf = subset_info(str(f))
    assert len(coords) == len(gauss) + 2
    for i, r in enumerate(gauss):
        assert conf[i]["n"] == gauss[r]
        assert [p[:2] for p in conf[i]["pix"]] == union_pixels(o.circle_pixels(20 + 30 * i, 50, r))
    assert [p[:2] for p in conf[k]["pix"]] == union_pixels(o.rectangle_pixels(200, 50, 7, 4))
    assert conf[k]["n"] == 28 and conf[k]["pix"][0][:2] == (48, 197)
    # corners 227,45 .. 236,52: width 9, height 7, centre (231, 48)
    assert [p[:2] for p in conf[k + 1]["pix"]] == union_pixels(o.rectangle_pixels(231, 48, 9, 7))


def test_regression_subset_file_reader():
    """The deck's subset file (comments, blank lines, a union of two polygons and a circle, an excluded
    polygon, an obstruction and blocking ids) read by read_subset_file, against the oracle's shapes."""
    coords, conf = subset_info(os.path.join(DECK, "subsets.txt"))
    assert coords == [(126, 157), (125, 250), (397, 139), (177, 314), (395, 405)]
    assert sorted(conf) == [2, 3]
    s2 = union_pixels(o.polygon_pixels([307, 352, 426, 421, 372], [136, 86, 109, 243, 143]),
                      o.circle_pixels(362, 201, 45), o.polygon_pixels([359, 455, 437], [228, 175, 267]))
    assert [p[:2] for p in conf[2]["pix"]] == s2 and conf[2]["active"] == len(s2)
    s3 = union_pixels(o.circle_pixels(178, 352, 64))
    ex = set(union_pixels(o.polygon_pixels([148, 160, 205], [341, 374, 340])))
    assert [p[:2] for p in conf[3]["pix"]] == s3
    assert [p[2] for p in conf[3]["pix"]] == [0 if p in ex else 1 for p in s3]
    assert 0 < len(s3) - conf[3]["active"] == len(ex & set(s3))
    assert conf[3]["obstructed"] == 1 and conf[3]["blocking"] == [0, 4]


def test_subset_file_errors(tmp_path):
    for body, what in [("BEGIN MOTION_WINDOW\nEND MOTION_WINDOW\n", "MOTION_WINDOW"),
                       ("BEGIN REGION_OF_INTEREST\nDIRICHLET_BC BOUNDARY 0 0 1 0 0.0\nEND\n", "DIRICHLET_BC"),
                       ("BEGIN CONFORMAL_SUBSET\nBEGIN BOUNDARY\nBEGIN CIRCLE\nCENTER 10 10\nRADIUS 3\nEND CIRCLE\n"
                        "END BOUNDARY\nEND CONFORMAL_SUBSET\n", "SUBSET_ID"),
                       ("BEGIN SUBSET_COORDINATES\n10\nEND SUBSET_COORDINATES\n", "not enough values")]:
        f = tmp_path / "bad.txt"
        f.write_text(body)
        r = run("subset_file_info", str(f))
        assert r.returncode == 98 and what in r.stderr, (what, r.stderr)


def dice_diff(a, b, *opts):
    if not os.path.exists(DIFF):
        pytest.fail("dice_amd/lib/dice_diff is not built (run __graft_entry__.build())")
    r = subprocess.run([DIFF, str(a), str(b)] + [str(x) for x in opts], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def test_dice_diff_semantics(tmp_path):
    """tools/DICe_Diff.cpp (serial): relative differences against the first file, the floor, tiny values,
    string tokens, the skipped masthead, and line-count mismatches."""
    a = tmp_path / "a.txt"
    b = tmp_path / "b.txt"
    a.write_text("***\n*** Digital Image Correlation Engine (DICe) v1\nX Y\n0 1.0000E+00 2.0000E-05\n")
    b.write_text("***\n*** Digital Image Correlation Engine, another build\nX Y\n0 1.0400E+00 9.0000E-05\n")
    assert dice_diff(a, b, "-t", 0.05, "-f", 5e-5)[0] == 0          # 4 % and a value below the floor
    assert dice_diff(a, b, "-t", 0.01, "-f", 5e-5)[0] == 1          # 4 % > 1 %
    assert dice_diff(a, b, "-t", 0.05)[0] == 1                      # no floor: 2e-5 vs 9e-5 fails
    b.write_text("***\n*** Digital\nX Z\n0 1.0000E+00 2.0000E-05\n")
    assert dice_diff(a, b)[0] == 1 and dice_diff(a, b, "-n")[0] == 0  # a string token differs
    b.write_text("***\n*** Digital\nX Y\n")
    code, out = dice_diff(a, b)
    assert code == 1 and "TEST FAILED" in out


def test_oracle_on_regression_deck():
    """The oracle (GENERIC routine) on the deck against its gold: the square subsets within 5e-5 (u, v) and
    2e-6 (sigma), one or two units of the last printed digit;
    every value within the regression suite's 5 % (the conformal subset 3 has an obstruction, which
    belongs to the TRACKING routine)."""
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    coords, conf = subset_info(os.path.join(DECK, "subsets.txt"))
    offs, xs, ys, acts = [0], [], [], []
    for i, (cx, cy) in enumerate(coords):
        if i in conf:
            pix = conf[i]["pix"]
        else:
            pix = [(y, x, 1) for y in range(int(cy) - 13, int(cy) + 14) for x in range(int(cx) - 13, int(cx) + 14)]
        ys += [p[0] for p in pix]
        xs += [p[1] for p in pix]
        acts += [p[2] for p in pix]
        offs.append(offs[-1] + len(pix))
    f = o.new_fields(np.array([c[0] for c in coords]), np.array([c[1] for c in coords]))
    o.correlate(o.default_params(), d["ref"], d["defm"], f,
                conformal=(np.array(offs, np.int32), np.array(xs, np.int32), np.array(ys, np.int32),
                           np.array(acts, np.uint8)), precision="double", threads=4)
    for i in range(5):
        g = [float(v) for v in open(os.path.join(DECK, "gold", "DICe_solution_%d.txt" % i)).read().split("\n")[-2].split()]
        u, v, sig = f[o.F["DISP_X"], i], f[o.F["DISP_Y"], i], f[o.F["SIGMA"], i]
        assert f[o.F["STATUS"], i] == g[10] == 1
        for val, gv, tol in ((u, g[3], 5e-5), (v, g[4], 5e-5), (sig, g[9], 2e-6)):
            assert abs(val - gv) <= 0.05 * abs(gv)
            if i in (0, 1, 2, 4):
                assert abs(val - gv) <= tol


def test_regular_grid_order_matches_regression_gold():
    """create_regular_grid_of_correlation_points: the 49 centres of tests/regression/parallel_subsets
    (512^2, S = 27, step 75) in the gold's order (the reference's serpentine: down the first column, up
    the next), with the neighbour of each point the previous one."""
    r = run("grid", 512, 512, 27, 75)
    pts = np.array(r.stdout.split(), dtype=int).reshape(-1, 3)
    rows = [ln.split() for ln in open(os.path.join(DECKS, "parallel_subsets", "gold", "DICe_solution_0.txt"))
            if ln[:1].isdigit()]
    gold = np.array([[float(t[1]), float(t[2])] for t in rows])
    assert np.array_equal(pts[:, :2], gold)
    assert list(pts[:2, 2]) == [-1, -1] and np.array_equal(pts[2:, 2], np.arange(1, len(pts) - 1))


def write_tiff_u8(path, img):
    """Little-endian, uncompressed 8-bit grey TIFF, one strip (TIFF 6.0 section 2)."""
    import struct
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    tags = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 8), (259, 3, 1, 1), (262, 3, 1, 1),
            (273, 4, 1, 8 + 2 + 12 * 8 + 4), (278, 4, 1, h), (279, 4, 1, w * h)]
    out = bytearray(b"II" + struct.pack("<HI", 42, 8) + struct.pack("<H", len(tags)))
    for tag, typ, cnt, val in tags:
        out += struct.pack("<HHI", tag, typ, cnt) + (struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val))
    out += struct.pack("<I", 0) + img.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_vsg_regression_deck(tmp_path, backend):
    """dic_challenge_12_sorted (frames 1 and 2) end to end through the Schema mirror: TIFF images read by
    name, the region-of-interest grid, neighbour initialization on the first frame, normal strains, the
    VSG post-processor on the GPU, sorted comma-separated output without row ids -- dice_diff against both
    golds with the suite's options, and the columns within the oracle test's tolerances."""
    deck, frames = vsg_deck_frames()
    h, w = frames[0].shape
    names = []
    for k, img in enumerate(frames):
        write_tiff_u8(tmp_path / ("oht_cfrp_%02d.tiff" % k), img)
        names.append(str(tmp_path / ("oht_cfrp_%02d.tiff" % k)))
    cfg = {"ref_file": names[0], "def_file": names[1], "def_files": ",".join(names[1:]), "width": w, "height": h,
           "subset_size": deck["subset_size"], "step_size": deck["step_size"],
           "subset_file": os.path.join(VSG_DECK, deck["subset_file"]), "out_dir": tmp_path,
           "prefix": deck["prefix"], "separate": 0, "backend": backend, "ref_name": deck["reference_image"],
           "def_name": deck["deformed_image"], "def_names": ",".join(deck["deformed_images"]), "routine": 0,
           "fields": ",".join(deck["fields"]), "optimization": deck["optimization_method"],
           "interpolation": deck["interpolation_method"], "initialization": deck["initialization_method"],
           "enable_normal_strain": deck["enable_normal_strain"], "vsg_window": deck["vsg_window"],
           "sort": int(deck["sort"]), "omit_row_id": int(deck["omit_row_id"]), "delimiter": deck["delimiter"],
           "frames": deck["frames"], "first_frame_index": deck["first_frame_index"],
           "num_frames": deck["num_frames"]}
    cfg_path = tmp_path / "deck.cfg"
    cfg_path.write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = run("deck", cfg_path)
    assert r.returncode == 0, r.stderr
    for gname in deck["gold"]:
        gold = os.path.join(VSG_DECK, "gold", gname)
        code, out = dice_diff(gold, tmp_path / gname, "-t", deck["diff_tolerance"], "-f", deck["diff_floor"], "-v")
        assert code == 0 and "TEST PASSED" in out, (gname, out)
        g, ours = gold_csv(gold), gold_csv(tmp_path / gname)
        assert np.array_equal(ours[:, [0, 1, 7]], g[:, [0, 1, 7]])
        assert np.abs(ours[:, 2:4] - g[:, 2:4]).max() <= 1e-4
        assert (np.abs(ours[:, 4:7] - g[:, 4:7]) / np.abs(g[:, 4:7])).max() <= 5e-3


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
@pytest.mark.parametrize("name", ["conformal_subset_def", "conformal_subset_def_override", "single_speckled_shift",
                                  "parallel_subsets", "parallel_subsets_seed", "parallel_seed_first_step_only"])
def test_regression_deck(tmp_path, name, backend):
    """A deck of tests/regression end to end through the C++ Schema mirror -- subsets from its subset file
    (read_subset_file) or its regular grid (create_regular_grid_of_correlation_points), the correlation,
    write_output in the deck's file layout and field list -- then dice_diff against every gold file with
    the regression suite's options (-t 0.05 -f 5.0E-5, header strings included).  Square subsets also
    within the tolerances of the gold tests of tests/test_gpu_parity.py: u, v 5e-5, sigma and gamma 2e-6."""
    import json
    # conformal_subset_def: subset 3 is blocked by the square subsets 0 and 4, and has an OBSTRUCTED polygon.
    # With the reference's present logic (dice_default_params: override_force_simplex false) subsets 0, 3
    # and 4 run Nelder-Mead only, and the deck passes at the suite's tolerance; its gold holds the
    # Gauss-Newton results for 0 and 4 (within 1e-5 px), so it was written before blocking subsets forced
    # the simplex: with override_force_simplex the square subsets and the polygon subset 2 match it at
    # 5e-5, and the obstructed subset 3 (whose gold sigma the obstruction reproduces to 5e-6) at the
    # suite's tolerance.
    override = name.endswith("_override")
    name = name.replace("_override", "")
    ddir = os.path.join(DECKS, name)
    deck = json.load(open(os.path.join(ddir, "deck.json")))
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    h, w = d["ref"].shape
    cfg = {"ref": raw(tmp_path, "ref.f32", d["ref"]), "def": raw(tmp_path, "def.f32", d["defm"]), "width": w,
           "height": h, "subset_size": deck["subset_size"], "out_dir": tmp_path, "prefix": deck["prefix"],
           "separate": int(deck["separate"]), "backend": backend, "ref_name": deck["reference_image"],
           "def_name": deck["deformed_image"], "routine": deck.get("correlation_routine", 0)}
    if override:
        cfg["override_force_simplex"] = 1
    if "subset_file" in deck:
        cfg["subset_file"] = os.path.join(ddir, deck["subset_file"])
    if "step_size" in deck:
        cfg["step_size"] = deck["step_size"]
    if "fields" in deck:
        cfg["fields"] = ",".join(deck["fields"])
    for k, key in (("optimization_method", "optimization"), ("interpolation_method", "interpolation"),
                   ("initialization_method", "initialization"), ("frames", "frames")):
        if k in deck:
            cfg[key] = deck[k]
    cfg["def_names"] = ",".join(deck["deformed_images"])
    cfg_path = tmp_path / "deck.cfg"
    cfg_path.write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = run("deck", cfg_path)
    assert r.returncode == 0, r.stderr
    for g in deck["gold"]:
        code, out = dice_diff(os.path.join(ddir, "gold", g), tmp_path / g, "-t", deck["diff_tolerance"],
                              "-f", deck["diff_floor"], "-v")
        assert code == 0 and "TEST PASSED" in out, (g, out)
        # the numbers of the square subsets: u, v, sigma / gamma to 5e-5 relative
        gold_rows = [ln.split() for ln in open(os.path.join(ddir, "gold", g)) if ln[:1].isdigit()]
        ours = [ln.split() for ln in open(tmp_path / g) if ln[:1].isdigit()]
        header = [ln.split() for ln in open(os.path.join(ddir, "gold", g)) if ln.startswith(("FRAME", "SUBSET_ID"))][0]
        for gr, orow in zip(gold_rows, ours):
            if name == "conformal_subset_def" and g in (("DICe_solution_3.txt",) if override else
                                                          ("DICe_solution_0.txt", "DICe_solution_3.txt",
                                                           "DICe_solution_4.txt")):
                continue  # within the suite's tolerance (above)
            for col, tol in (("DISPLACEMENT_X", 5e-5), ("DISPLACEMENT_Y", 5e-5), ("SIGMA", 2e-6), ("GAMMA", 2e-6)):
                if col in header:
                    k = header.index(col)
                    assert abs(float(orow[k]) - float(gr[k])) <= tol, (g, col, orow[k], gr[k])
            for col in ("COORDINATE_X", "COORDINATE_Y", "STATUS_FLAG", "ITERATIONS"):
                if col in header:
                    k = header.index(col)
                    assert float(orow[k]) == float(gr[k]), (g, col)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_unicode_filename_deck(tmp_path, backend):
    """tests/regression/unicode_filename end to end through the Schema mirror: the image read by its
    non-ASCII file name (written back from the fixture under the name input.xml spells, which the gold's
    header prints) and correlated against itself on the regular grid of 15-px subsets at step 50, six
    parameters, comma-separated output without row ids -- dice_diff against the gold with the suite's options;
    the displacements, which the gold holds at the 1e-14 level, within 1e-6 px of zero."""
    import json
    ddir = os.path.join(DECKS, "unicode_filename")
    deck = json.load(open(os.path.join(ddir, "deck.json"), encoding="utf-8"))
    img = np.load(os.path.join(ddir, deck["images"]))["frame_0"]
    h, w = img.shape
    path = tmp_path / deck["image_file"]
    write_tiff_u8(path, img)
    cfg = {"ref_file": str(path), "def_file": str(path), "def_files": str(path), "width": w, "height": h,
           "subset_size": deck["subset_size"], "step_size": deck["step_size"], "out_dir": tmp_path,
           "prefix": deck["prefix"], "separate": 0, "backend": backend, "ref_name": deck["reference_image"],
           "def_name": deck["deformed_image"], "def_names": ",".join(deck["deformed_images"]), "routine": 0,
           "fields": ",".join(deck["fields"]), "optimization": deck["optimization_method"],
           "interpolation": deck["interpolation_method"], "initialization": deck["initialization_method"],
           "enable_rotation": deck["enable_rotation"], "enable_normal_strain": deck["enable_normal_strain"],
           "enable_shear_strain": deck["enable_shear_strain"], "omit_row_id": int(deck["omit_row_id"]),
           "delimiter": deck["delimiter"], "frames": deck["frames"]}
    cfg_path = tmp_path / "deck.cfg"
    cfg_path.write_text("".join("%s %s\n" % kv for kv in cfg.items()), encoding="utf-8")
    r = run("deck", cfg_path)
    assert r.returncode == 0, r.stderr
    for gname in deck["gold"]:
        gold = os.path.join(ddir, "gold", gname)
        code, out = dice_diff(gold, tmp_path / gname, "-t", deck["diff_tolerance"], "-f", deck["diff_floor"], "-v")
        assert code == 0 and "TEST PASSED" in out, (gname, out)
        g, ours = gold_csv(gold), gold_csv(tmp_path / gname)
        assert g.shape == ours.shape == (100, 4)
        assert np.array_equal(ours[:, :2], g[:, :2])
        assert np.abs(ours[:, 2:4]).max() <= 1e-6


def test_read_correlation_params(tmp_path):
    """read_correlation_params (DICe_Parser.cpp:353-430): names lower-cased, enum strings to their values,
    bools to 0 / 1, output-only switches ignored; parameters this path does not read, nested lists and
    unknown enum strings are errors."""
    r = run("read_params", os.path.join(GOLD, "api", "params.xml"))
    assert r.returncode == 0, r.stderr
    got = dict(ln.split() for ln in r.stdout.splitlines())
    assert got == {"interpolation_method": "2", "optimization_method": "1", "enable_rotation": "1",
                   "max_solver_iterations_fast": "30", "use_tracking_default_params": "0"}
    f = tmp_path / "p.xml"
    f.write_text('<ParameterList>\n<Parameter name="correlation_routine" type="string" value="TRACKING_ROUTINE"/>\n'
                 '<Parameter name="use_tracking_default_params" type="bool" value="true"/>\n</ParameterList>\n')
    got = dict(ln.split() for ln in run("read_params", f).stdout.splitlines())
    assert got == {"correlation_routine": "1", "use_tracking_default_params": "1"}
    for bad in ('<Parameter name="use_incremental_formulation" type="bool" value="true"/>',
                '<Parameter name="optimization_method" type="string" value="NEWTON"/>',
                '<ParameterList name="post_process_vsg_strain"></ParameterList>',
                '<Parameter name="shape_function_type" type="string" value="QUADRATIC"/>'):
        f.write_text("<ParameterList>\n%s\n</ParameterList>\n" % bad)
        assert run("read_params", f).returncode != 0, bad


@pytest.mark.gpu
def test_c_caller_resolves_reference_entry_points(tmp_path):
    """tests/cpp/test_api.c, a plain C program, dlopens the libraries and resolves dice_correlate and
    dice_correlate_conformal by name (DICe_api.h:129-133, 176-180), then correlates the integer-shift
    sequence of DICe_TestAPI.cpp:70-150: square subsets with the tracking defaults, and two conformal
    subsets (a polygon and a rectangle) from a subset file with an XML parameter file; u, v within 1e-4
    of k, theta and gamma within 1e-4 of 0, status 1."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_api is not built (run __graft_entry__.build())")
    d = np.load(os.path.join(GOLD, "synthetic_speckled.npz"))
    imgs = [raw(tmp_path, "def%d.f32" % k, d["def%d" % k].astype(np.float32)) for k in range(6)]
    lib = os.path.join(ROOT, "dice_amd", "lib")
    r = subprocess.run([exe, os.path.join(lib, "libdice_hip.so"), os.path.join(lib, "libdice_schema.so"), imgs[0]] +
                       imgs + [os.path.join(GOLD, "api", "subsets.txt"), os.path.join(GOLD, "api", "params.xml")],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0 and "API PASSED" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_sub_image_extents_deck(tmp_path, backend):
    """Schema::update_extents (DICe_Schema.cpp:886-985) through the Schema mirror: square subsets in the
    middle of the 512^2 speckle pair read from TIFF files, so that only a sub-image of each frame is read
    and handed to the device with its offsets, against the same deck with read_full_images: identical
    output, and the subsets' u, v within 5e-5 of the whole-image oracle."""
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    write_tiff_u8(tmp_path / "ref.tif", d["ref"])
    write_tiff_u8(tmp_path / "def.tif", d["defm"])
    pts = [(230, 240), (260, 250), (250, 290), (300, 270), (280, 310)]
    (tmp_path / "subsets.txt").write_text("BEGIN SUBSET_COORDINATES\n" + "".join("%d %d\n" % p for p in pts) +
                                          "END SUBSET_COORDINATES\n")
    outs = []
    for full in (0, 1):
        od = tmp_path / ("out%d" % full)
        od.mkdir()
        cfg = {"ref_file": tmp_path / "ref.tif", "def_file": tmp_path / "def.tif", "width": 512, "height": 512,
               "subset_size": 27, "subset_file": tmp_path / "subsets.txt", "out_dir": od, "prefix": "DICe_solution",
               "separate": 0, "backend": backend, "ref_name": "ref.tif", "def_name": "def.tif",
               "def_names": "def.tif", "routine": 0, "read_full_images": full}
        cfg_path = tmp_path / ("deck%d.cfg" % full)
        cfg_path.write_text("".join("%s %s\n" % kv for kv in cfg.items()))
        r = run("deck", cfg_path)
        assert r.returncode == 0, r.stderr
        assert ("extents %d" % (1 - full)) in r.stdout
        outs.append(open(od / "DICe_solution_0.txt").read())
    assert outs[0] == outs[1]
    f = o.new_fields(np.array([p[0] for p in pts], float), np.array([p[1] for p in pts], float))
    o.correlate(o.default_params(), d["ref"], d["defm"], f, subset_size=27, precision="double", threads=4)
    rows = [ln.split() for ln in outs[0].splitlines() if ln[:1].isdigit()]
    for i, row in enumerate(rows):
        assert abs(float(row[3]) - f[o.F["DISP_X"], i]) < 5e-5 and abs(float(row[4]) - f[o.F["DISP_Y"], i]) < 5e-5


BLOCK_A = ([28, 50, 51, 29], [30, 31, 55, 54])   # blocker polygon
BLOCK_B = ([42, 64, 65, 43], [31, 30, 54, 55])   # overlaps A over x = 42..50


@pytest.mark.gpu
@pytest.mark.parametrize("blocker_first", [True, False], ids=["blocker_first", "blocked_first"])
def test_tracking_blocking_subsets(tmp_path, blocker_first):
    """TRACKING_ROUTINE with a conformal blocker (check_for_blocking_subsets, DICe_Schema.cpp:4529-4556)
    over the integer-shift frames defSyntheticSpeckled1..3 (1 px a frame): the blocked subset loses the
    warped pixels that round onto the blocker's polygon moved by the blocker's current shape function --
    this frame's result when the blocker has the lower id, its previous frame's otherwise (one pixel
    apart here).  Gauss-Newton for both (override_force_simplex), against the oracle run frame by frame
    with the blocked pixels formed by its own deformed-polygon rule."""
    d = np.load(os.path.join(GOLD, "synthetic_speckled.npz"))
    ida, idb = (0, 1) if blocker_first else (1, 0)
    polys = {ida: BLOCK_A, idb: BLOCK_B}
    cen = {k: (int(np.mean(v[0])), int(np.mean(v[1]))) for k, v in polys.items()}
    lines = ["BEGIN SUBSET_COORDINATES"] + ["%d %d" % cen[k] for k in (0, 1)] + ["END SUBSET_COORDINATES"]
    for k in (0, 1):
        vx, vy = polys[k]
        lines += ["BEGIN CONFORMAL_SUBSET", "SUBSET_ID %d" % k, "BEGIN BOUNDARY", "BEGIN POLYGON", "BEGIN VERTICES"]
        lines += ["%d %d" % xy for xy in zip(vx, vy)]
        lines += ["END VERTICES", "END POLYGON", "END BOUNDARY"]
        if k == idb:
            lines += ["BEGIN BLOCKING_SUBSETS", "%d" % ida, "END BLOCKING_SUBSETS"]
        lines += ["END CONFORMAL_SUBSET"]
    (tmp_path / "subsets.txt").write_text("\n".join(lines) + "\n")
    for k in range(4):
        write_tiff_u8(tmp_path / ("def%d.tif" % k), d["def%d" % k])
    frames = [str(tmp_path / ("def%d.tif" % k)) for k in (1, 2, 3)]
    cfg = {"ref_file": tmp_path / "def0.tif", "def_file": frames[0], "def_files": ",".join(frames), "width": 100,
           "height": 100, "subset_size": 27, "subset_file": tmp_path / "subsets.txt", "out_dir": tmp_path,
           "prefix": "DICe_solution", "separate": 0, "backend": 0, "ref_name": "r", "def_name": "d",
           "def_names": "d,d,d", "routine": 1, "override_force_simplex": 1, "optimization": 1, "frames": 3}
    cfg_path = tmp_path / "deck.cfg"
    cfg_path.write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = run("deck", cfg_path)
    assert r.returncode == 0, r.stderr
    got = {}
    for ln in r.stdout.splitlines():
        t = ln.split()
        if t and t[0] == "subset":
            got[int(t[1])] = (float(t[7]), float(t[9]), float(t[11]))
    # the oracle, frame by frame: the blocker, then the blocked subset with the blocker's moved polygon
    p = o.default_params()
    p.optimization_method = 1
    ref = d["def0"].astype(np.float32)
    names = ("DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY")
    state = {k: o.new_fields(np.array([cen[k][0]], float), np.array([cen[k][1]], float)) for k in (0, 1)}

    def solve(k, de, masks=None):
        xs, ys = o.polygon_pixels(*polys[k])
        conf = (np.array([0, len(xs)], np.int32), xs, ys, np.ones(len(xs), np.uint8))
        f = state[k].copy()
        o.correlate(p, ref, de, f, conformal=conf, precision="double", masks=masks, first_frame=False)
        return f

    moved = []
    for k in (1, 2, 3):
        de = d["def%d" % k].astype(np.float32)
        before = [state[ida][o.F[n], 0] for n in names]
        state[ida] = solve(ida, de)
        after = [state[ida][o.F[n], 0] for n in names]
        mx, my = o.deformed_polygon_pixels(*polys[ida], after if blocker_first else before, cen[ida][0], cen[ida][1],
                                           1.0)
        moved.append(int(np.min(mx)))
        state[idb] = solve(idb, de, masks=(np.array([0, len(mx)], np.int32), mx, my))
    assert moved[-1] == int(np.min(o.polygon_pixels(*BLOCK_A)[0])) + (3 if blocker_first else 2)
    for k in (0, 1):
        f = state[k]
        assert got[k][2] == f[o.F["STATUS"], 0] == 1
        assert abs(got[k][0] - f[o.F["DISP_X"], 0]) < 1e-4 and abs(got[k][1] - f[o.F["DISP_Y"], 0]) < 1e-4, (k, got[k])
