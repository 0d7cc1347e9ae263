"""Image fixtures for the image-ingest tests.

From the reference's component tests (tests/component/images, data its own tests hold):
  ImageB.tif            240 x 161 8-bit TIFF, the input of the Gauss-filter test
                        (DICe_TestImage.cpp:318-410)
  outFilter<M>.rawi     that test's float goldens for masks M = 5, 7, 9, 11, 13
  gauss_golden.json     SHA-256 of the double goldens outFilter<M>_d.rawi (whole file), so that
                        the double arithmetic is pinned without carrying 1.5 MB of doubles
Synthetic TIFFs (written here with Pillow and, for the byte order and photometric cases Pillow
does not write, by hand from the TIFF 6.0 layout), with their pixel arrays in tiff_cases.npz:
  u8_raw, u8_lzw, u8_lzw_pred, u8_packbits, u16_le, u16_be_lzw (Pillow stores it little-endian),
  u8_mm_white (big-endian, WhiteIsZero, 6 strips), u16_mm (big-endian, 16-bit, 8 strips)

Run from the repository root: python tests/golden/make_image_fixtures.py"""
import hashlib
import json
import os
import shutil
import struct

import numpy as np
from PIL import Image

SRC = "/root/reference/tests/component/images"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "images")
MASKS = (5, 7, 9, 11, 13)


def write_tiff_mm(path, img, rows_per_strip, photometric):
    """Big-endian, uncompressed grey TIFF (8- or 16-bit) in several strips (TIFF 6.0 section 2)."""
    h, w = img.shape
    bits = img.dtype.itemsize * 8
    data = img.astype(">u%d" % img.dtype.itemsize)
    strips = [data[r:r + rows_per_strip].tobytes() for r in range(0, h, rows_per_strip)]
    n_tags = 9
    ifd_off = 8
    offsets_off = ifd_off + 2 + 12 * n_tags + 4
    counts_off = offsets_off + 4 * len(strips)
    offs, p = [], counts_off + 4 * len(strips)
    for st in strips:
        offs.append(p)
        p += len(st)
    tags = [(256, 3, 1, w), (257, 3, 1, h), (258, 3, 1, bits), (259, 3, 1, 1), (262, 3, 1, photometric),
            (273, 4, len(strips), offsets_off), (277, 3, 1, 1), (278, 3, 1, rows_per_strip),
            (279, 4, len(strips), counts_off)]
    out = bytearray(b"MM" + struct.pack(">HI", 42, ifd_off) + struct.pack(">H", n_tags))
    for tag, typ, cnt, val in tags:
        out += struct.pack(">HHI", tag, typ, cnt)
        out += struct.pack(">HH", val, 0) if typ == 3 else struct.pack(">I", val)
    out += struct.pack(">I", 0)
    out += b"".join(struct.pack(">I", o) for o in offs)
    out += b"".join(struct.pack(">I", len(st)) for st in strips)
    out += b"".join(strips)
    with open(path, "wb") as f:
        f.write(bytes(out))


def main():
    os.makedirs(DST, exist_ok=True)
    shutil.copy(os.path.join(SRC, "ImageB.tif"), DST)
    digests = {}
    for m in MASKS:
        shutil.copy(os.path.join(SRC, "outFilter%d.rawi" % m), DST)
        with open(os.path.join(SRC, "outFilter%d_d.rawi" % m), "rb") as f:
            digests["outFilter%d_d.rawi" % m] = hashlib.sha256(f.read()).hexdigest()
    with open(os.path.join(DST, "gauss_golden.json"), "w") as f:
        json.dump(digests, f, indent=1)

    rng = np.random.default_rng(20261202)
    u8 = rng.integers(0, 256, size=(97, 131), dtype=np.uint8)
    smooth = (np.add.outer(np.arange(97), np.arange(131)) % 256).astype(np.uint8)
    u16 = rng.integers(0, 65536, size=(53, 71), dtype=np.uint16)
    cases = {}

    def save(name, arr, **kw):
        Image.fromarray(arr).save(os.path.join(DST, name + ".tif"), **kw)
        cases[name] = arr

    save("u8_raw", u8)
    save("u8_lzw", u8, compression="tiff_lzw")
    save("u8_lzw_pred", smooth, compression="tiff_lzw", tiffinfo={317: 2})
    save("u8_packbits", u8, compression="packbits")
    Image.frombytes("I;16", (u16.shape[1], u16.shape[0]), u16.astype("<u2").tobytes()).save(
        os.path.join(DST, "u16_le.tif"))
    cases["u16_le"] = u16
    Image.frombytes("I;16B", (u16.shape[1], u16.shape[0]), u16.astype(">u2").tobytes()).save(
        os.path.join(DST, "u16_be_lzw.tif"), compression="tiff_lzw")
    cases["u16_be_lzw"] = u16
    write_tiff_mm(os.path.join(DST, "u8_mm_white.tif"), u8[:40, :50], rows_per_strip=7, photometric=0)
    cases["u8_mm_white"] = u8[:40, :50]
    write_tiff_mm(os.path.join(DST, "u16_mm.tif"), u16[:30, :41], rows_per_strip=4, photometric=1)
    cases["u16_mm"] = u16[:30, :41]
    np.savez_compressed(os.path.join(DST, "tiff_cases.npz"), **cases)
    for f in os.listdir(DST):
        os.chmod(os.path.join(DST, f), 0o644)


if __name__ == "__main__":
    main()
