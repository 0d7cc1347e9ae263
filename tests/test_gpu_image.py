"""Image ingest on the GPU: dice_hip_gauss_filter and the filter applied on image upload
(gauss_filter_images), against the reference's filtered-image goldens and the oracle; image
files through the C++ Schema mirror's set_ref_image(file) / set_def_image(file).

Bar: the filter is bit-exact (float arithmetic, one multiply and one add per tap in the
reference's order); correlations with the filter on meet the parity bar of
tests/test_gpu_parity.py (displacement / rotation 1e-4, gamma 1e-5 relative, status >= 99.9 %)."""
import os
import struct

import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from test_gpu_parity import MODE_NAMES, MODES, compare, run_gpu, run_oracle
from test_host_cpp import DECKS, GOLD, dice_diff, run
from test_image_io import IMAGES, MASKS, read_image, read_rawi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dh.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def image_b(tmp_path_factory):
    return read_image(tmp_path_factory.mktemp("imgb"), "ImageB.tif")


def write_tiff_u8(path, img):
    """Little-endian, uncompressed 8-bit grey TIFF, one strip (TIFF 6.0 section 2)."""
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


@pytest.mark.parametrize("mask", MASKS)
def test_gpu_gauss_filter_golden_bit_exact(image_b, mask):
    g = read_rawi(os.path.join(IMAGES, "outFilter%d.rawi" % mask))
    f = dh.gauss_filter(image_b, mask)
    assert f.dtype == np.float32 and np.array_equal(f, g)


@pytest.mark.parametrize("shape", [(13, 13), (45, 77), (70, 33), (257, 300)])
def test_gpu_gauss_filter_sizes_and_uint16(shape):
    """Images that do not fill whole 32 x 32 tiles, down to the mask size itself; float and uint16
    storage (saturated pixels included) bit-exact against the oracle."""
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    f32 = (rng.random(shape) * 255).astype(np.float32)
    u16 = rng.integers(0, 65536, size=shape).astype(np.uint16)
    u16[: shape[0] // 2, : shape[1] // 2] = 65535
    for m in MASKS:
        if min(shape) < m:
            continue
        assert np.array_equal(dh.gauss_filter(f32, m), o.gauss_filter(f32, m)), (shape, m)
        assert np.array_equal(dh.gauss_filter(u16, m), o.gauss_filter(u16, m)), (shape, m)


def test_gpu_gauss_filter_errors():
    img = np.zeros((20, 20), np.float32)
    for bad in (3, 6, 15):
        with pytest.raises(dh.DiceHipError):
            dh.gauss_filter(img, bad)
    with pytest.raises(dh.DiceHipError):
        dh.gauss_filter(np.zeros((12, 40), np.float32), 13)
    c = dh.Context(0)
    try:
        p = dh.default_params()
        p.gauss_filter_images = 1
        p.gauss_filter_mask_size = 8
        with pytest.raises(dh.DiceHipError):
            c.set_params(p)
    finally:
        c.close()


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("mask", [7, 13])
def test_gpu_correlation_with_gauss_filter_parity(ctx, mode, mask):
    """gauss_filter_images on: both images filtered on upload, before their gradients; the speckle
    pair on a grid (S = 27, step 20), default method.  The filter runs in float as in the reference's
    default build (bit-exact, tested above), so the double oracle correlates the float-filtered pair."""
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    ref, de = d["ref"].astype(np.float32), d["defm"].astype(np.float32)
    cx, cy = dh.regular_grid(512, 512, 27, 20)
    p = dh.default_params()
    p.gauss_filter_images = 1
    p.gauss_filter_mask_size = mask
    fg = run_gpu(ctx, p, ref, de, cx, cy, 27, mode)
    fo = run_oracle(dh.default_params(), o.gauss_filter(ref, mask), o.gauss_filter(de, mask), cx, cy, 27)
    compare(fg, fo, label="gauss %d" % mask)
    # the same as correlating the GPU-filtered pair with the filter off
    q = dh.default_params()
    fq = run_gpu(ctx, q, dh.gauss_filter(ref, mask), dh.gauss_filter(de, mask), cx, cy, 27, mode)
    assert np.array_equal(fg, fq)


@pytest.mark.parametrize("backend", [0, 1], ids=["batched", "plain"])
def test_deck_from_tiff_files(tmp_path, backend):
    """The parallel_subsets regression deck with its images given as TIFF files (Schema::set_ref_image
    / set_def_image by name) matches its gold as the raw-array deck does; with the Gauss filter on
    (mask 7) the Schema results match the double oracle's over the float-filtered pair."""
    import json
    ddir = os.path.join(DECKS, "parallel_subsets")
    deck = json.load(open(os.path.join(ddir, "deck.json")))
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    write_tiff_u8(tmp_path / "ref.tif", d["ref"])
    write_tiff_u8(tmp_path / "def.tif", d["defm"])
    h, w = d["ref"].shape
    base = {"ref_file": tmp_path / "ref.tif", "def_file": tmp_path / "def.tif", "width": w, "height": h,
            "subset_size": deck["subset_size"], "step_size": deck["step_size"], "prefix": deck["prefix"],
            "separate": int(deck["separate"]), "backend": backend, "ref_name": deck["reference_image"],
            "def_name": deck["deformed_image"], "routine": 0}
    if "fields" in deck:
        base["fields"] = ",".join(deck["fields"])
    for k, key in (("optimization_method", "optimization"), ("interpolation_method", "interpolation")):
        if k in deck:
            base[key] = deck[k]
    out = tmp_path / "plain"
    out.mkdir()
    cfg = dict(base, out_dir=out)
    (tmp_path / "a.cfg").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = run("deck", tmp_path / "a.cfg")
    assert r.returncode == 0, r.stderr
    for g in deck["gold"]:
        code, txt = dice_diff(os.path.join(ddir, "gold", g), out / g, "-t", deck["diff_tolerance"],
                              "-f", deck["diff_floor"], "-v")
        assert code == 0 and "TEST PASSED" in txt, (g, txt)
    out = tmp_path / "filtered"
    out.mkdir()
    cfg = dict(base, out_dir=out, gauss_filter_mask_size=7)
    (tmp_path / "b.cfg").write_text("".join("%s %s\n" % kv for kv in cfg.items()))
    r = run("deck", tmp_path / "b.cfg")
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("subset ")]
    cx = np.array([float(x[3]) for x in rows])
    cy = np.array([float(x[5]) for x in rows])
    fo = run_oracle(dh.default_params(), o.gauss_filter(d["ref"].astype(np.float32), 7),
                    o.gauss_filter(d["defm"].astype(np.float32), 7), cx, cy, deck["subset_size"])
    u = np.array([float(x[7]) for x in rows])
    v = np.array([float(x[9]) for x in rows])
    st = np.array([float(x[11]) for x in rows])
    assert np.mean(st == fo[o.F["STATUS"]]) >= 0.999
    ok = np.isin(st, dh.CONVERGED_STATUSES)
    assert np.abs(u - fo[o.F["DISP_X"]])[ok].max() <= 1e-4
    assert np.abs(v - fo[o.F["DISP_Y"]])[ok].max() <= 1e-4
