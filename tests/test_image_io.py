"""Image ingest: the .rawi / TIFF readers of the C++ host mirror (dice_amd/host/dice_image_io.cpp,
driven by tests/cpp/test_schema) and the oracle's Gauss filter (Image_::gauss_filter,
DICe_Image.cpp:1192-1240) against the reference's own goldens (tests/component/images:
ImageB.tif -> outFilter<M>.rawi, DICe_TestImage.cpp:318-410).  The GPU filter is in
tests/test_gpu_image.py."""
import hashlib
import json
import os
import struct

import numpy as np
import pytest

import oracle_py as o
from test_host_cpp import GOLD, run

IMAGES = os.path.join(GOLD, "images")
MASKS = (5, 7, 9, 11, 13)
# the 1-D factors whose outer products are the reference's masks (DICe_Image.cpp:54-126)
_HALF = {5: [0.0014, 0.1574, 0.62825], 7: [0.006, 0.0606, 0.2418, 0.3831],
         9: [0.0007, 0.0108, 0.0748, 0.2384, 0.3505], 11: [0.0001, 0.0017, 0.0168, 0.087, 0.2328, 0.3231],
         13: [0.0001, 0.0012, 0.0085, 0.038, 0.1109, 0.2108, 0.2611]}
FACTORS = {m: h + h[-2::-1] for m, h in _HALF.items()}


def read_rawi(path, dtype=np.float32):
    b = open(path, "rb").read()
    w, h, nb = struct.unpack("<III", b[:12])
    assert nb == np.dtype(dtype).itemsize
    return np.frombuffer(b[12:], dtype).reshape(h, w)


def rawi_bytes(img):
    img = np.ascontiguousarray(img)
    return struct.pack("<III", img.shape[1], img.shape[0], img.dtype.itemsize) + img.tobytes()


def read_image(tmp_path, name, *region):
    out = tmp_path / "out.rawi"
    r = run("read_image", os.path.join(IMAGES, name) if not os.path.isabs(name) else name, out, *region)
    assert r.returncode == 0, r.stderr
    return read_rawi(out)


@pytest.fixture(scope="module")
def image_b(tmp_path_factory):
    return read_image(tmp_path_factory.mktemp("imgb"), "ImageB.tif")


def test_image_b_decodes_to_the_reference_dimensions(image_b):
    # DICe_TestImage.cpp checks 240 x 161 for ImageB; the filter goldens keep the
    # border band (M/2 + 1 pixels) of the decoded image untouched
    assert image_b.shape == (161, 240)
    g = read_rawi(os.path.join(IMAGES, "outFilter13.rawi"))
    band = np.ones(image_b.shape, bool)
    band[7:-7, 7:-7] = False
    assert np.array_equal(image_b[band], g[band])
    assert np.all(image_b == np.round(image_b)) and image_b.min() >= 0 and image_b.max() <= 255


@pytest.mark.parametrize("mask", MASKS)
def test_oracle_gauss_filter_float_golden_bit_exact(image_b, mask):
    """Float arithmetic (the reference's default build) reproduces outFilter<M>.rawi bit for bit; the
    reference's own test only asks for an L2 difference below 0.1."""
    g = read_rawi(os.path.join(IMAGES, "outFilter%d.rawi" % mask))
    f = o.gauss_filter(image_b, mask)
    assert f.dtype == np.float32
    assert np.array_equal(f, g)
    assert np.sqrt(np.sum((f.astype(np.float64) - g) ** 2)) < 0.1


@pytest.mark.parametrize("mask", MASKS)
def test_oracle_gauss_filter_double_golden_digest(image_b, mask):
    """Double arithmetic (the DICE_USE_DOUBLE build) reproduces outFilter<M>_d.rawi bit for bit: the
    SHA-256 of the whole .rawi file built from the oracle's output equals the golden file's."""
    digests = json.load(open(os.path.join(IMAGES, "gauss_golden.json")))
    f = o.gauss_filter(image_b.astype(np.float64), mask)
    assert hashlib.sha256(rawi_bytes(f)).hexdigest() == digests["outFilter%d_d.rawi" % mask]


def test_oracle_gauss_filter_properties():
    rng = np.random.default_rng(7)
    # a constant image maps to (sum of the mask) * c inside, c on the border band
    c = np.full((40, 37), 100.0)
    for m in MASKS:
        f = o.gauss_filter(c, m)
        hm = m // 2 + 1
        assert np.allclose(f[hm:-hm, hm:-hm], 100.0 * sum(FACTORS[m]) ** 2, rtol=1e-12)
        band = np.ones(c.shape, bool)
        band[hm:-hm, hm:-hm] = False
        assert np.all(f[band] == 100.0)
    # an impulse reproduces the mask itself (outer product of the 1-D factors, x offset first)
    imp = np.zeros((31, 31))
    imp[15, 15] = 1.0
    for m in MASKS:
        full = np.array(FACTORS[m])
        f = o.gauss_filter(imp, m)
        r = m // 2
        assert np.allclose(f[15 - r:16 + r, 15 - r:16 + r], np.outer(full, full), rtol=0, atol=1e-15)
    # linearity in double arithmetic
    a, b = rng.random((30, 33)), rng.random((30, 33))
    assert np.allclose(o.gauss_filter(a + 2 * b, 9), o.gauss_filter(a, 9) + 2 * o.gauss_filter(b, 9), atol=1e-12)
    # uint16 storage: the float result truncated to an integer, saturated at 65535
    u = rng.integers(0, 65536, size=(29, 41)).astype(np.uint16)
    u[5:20, 5:20] = 65535
    for m in MASKS:
        fu = o.gauss_filter(u, m)
        ff = o.gauss_filter(u.astype(np.float32), m)
        assert fu.dtype == np.uint16
        assert np.array_equal(fu, np.minimum(np.trunc(ff), 65535).astype(np.uint16))
    # errors: mask sizes other than 5..13 odd, images smaller than the mask
    for bad_mask in (3, 6, 15):
        with pytest.raises(ValueError):
            o.gauss_filter(c, bad_mask)
    with pytest.raises(ValueError):
        o.gauss_filter(np.zeros((12, 40)), 13)


def test_oracle_correlation_applies_the_filter_to_both_images():
    """gauss_filter_images filters both images before their gradients: correlating the unfiltered pair
    with the filter on equals correlating the pre-filtered pair with it off."""
    d = np.load(os.path.join(GOLD, "speckle_512.npz"))
    ref, de = d["ref"].astype(np.float32), d["defm"].astype(np.float32)
    cx = np.arange(60, 460, 40, dtype=np.float64)
    cy = np.full(cx.shape, 255.0)
    p = o.default_params()
    p.gauss_filter_images = 1
    p.gauss_filter_mask_size = 9
    f1 = o.new_fields(cx, cy)
    o.correlate(p, ref, de, f1, subset_size=27, precision="float", threads=4)
    q = o.default_params()
    f2 = o.new_fields(cx, cy)
    o.correlate(q, o.gauss_filter(ref, 9), o.gauss_filter(de, 9), f2, subset_size=27, precision="float", threads=4)
    assert np.array_equal(f1, f2)
    f3 = o.new_fields(cx, cy)
    o.correlate(q, ref, de, f3, subset_size=27, precision="float", threads=4)
    assert not np.array_equal(f1[o.F["GAMMA"]], f3[o.F["GAMMA"]])


CASES = ["u8_raw", "u8_lzw", "u8_lzw_pred", "u8_packbits", "u16_le", "u16_be_lzw", "u8_mm_white", "u16_mm"]


@pytest.mark.parametrize("name", CASES)
def test_tiff_reader_cases(tmp_path, name):
    """Uncompressed, LZW (with and without the horizontal predictor) and PackBits strips, both byte
    orders, 8 and 16 bits, WhiteIsZero: read_image gives the 8-bit intensities (16-bit samples scaled by
    (v * 255 + 32767) / 65535, WhiteIsZero inverted) and read_tiff_samples the stored samples."""
    exp = np.load(os.path.join(IMAGES, "tiff_cases.npz"))[name].astype(np.int64)
    got = read_image(tmp_path, name + ".tif")
    want = (exp * 255 + 32767) // 65535 if name.startswith("u16") else exp
    if "white" in name:
        want = 255 - want
    assert np.array_equal(got, want.astype(np.float32))
    out = tmp_path / "samples.bin"
    assert run("tiff_samples", os.path.join(IMAGES, name + ".tif"), out).returncode == 0
    b = open(out, "rb").read()
    w, h, bits = struct.unpack("<III", b[:12])
    assert bits == (16 if name.startswith("u16") else 8)
    assert np.array_equal(np.frombuffer(b[12:], np.uint16).reshape(h, w), exp)


def test_read_image_region_and_rawi_round_trip(tmp_path):
    exp = np.load(os.path.join(IMAGES, "tiff_cases.npz"))["u8_lzw"]
    sub = read_image(tmp_path, "u8_lzw.tif", 17, 9, 40, 31)
    assert np.array_equal(sub, exp[9:40, 17:57].astype(np.float32))
    # read_image of a .rawi file, written back by write_rawi_image: the same bytes
    src = os.path.join(IMAGES, "outFilter7.rawi")
    assert run("read_image", src, tmp_path / "copy.rawi").returncode == 0
    assert open(src, "rb").read() == open(tmp_path / "copy.rawi", "rb").read()
    # errors: a region of a .rawi file, a region beyond the image, a double .rawi into float,
    # an unsupported type, a missing file
    assert run("read_image", src, tmp_path / "x.rawi", 1, 1, 10, 10).returncode == 98
    assert run("read_image", os.path.join(IMAGES, "u8_lzw.tif"), tmp_path / "x.rawi", 100, 0, 40, 10).returncode == 98
    d = tmp_path / "d.rawi"
    d.write_bytes(rawi_bytes(np.zeros((4, 5))))
    assert run("read_image", d, tmp_path / "x.rawi").returncode == 98
    assert run("read_image", tmp_path / "img.png", tmp_path / "x.rawi").returncode == 98
    assert run("read_image", tmp_path / "none.tif", tmp_path / "x.rawi").returncode == 98
