"""pt_decode_image, no GPU: PNGs written here with Python's zlib (every color
type and bit depth the reference reads, all five row filters, stored, fixed,
dynamic and run-length deflate blocks, the data split over two IDAT chunks;
the cases of the reference's tests/image_test.py), PPMs, the reference's
own texture (models/spot/spot_texture.png against the bytes the reference's
decoder made of it, kept in the spot fixture), what the reference's decoder
refuses, and the reference's decoder itself on the same files where its
harness is built."""
import os
import random
import struct
import zlib

import numpy as np
import pytest

import oracle_bind as ob
import photon_tracer_amd as pt
import render_check as rc


def _chunk(tag, body):
    return struct.pack(">I", len(body)) + tag + body + struct.pack(">I", zlib.crc32(tag + body) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _png_of(w, h, ctype, depth, raw, palette=None, level=6, strategy=zlib.Z_DEFAULT_STRATEGY, interlace=0):
    """A PNG file around filtered image data (filter byte then row, per row)."""
    comp = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
    data = comp.compress(raw) + comp.flush()
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if palette is not None:
        png += _chunk(b"PLTE", palette)
    return png + _chunk(b"IDAT", data[:len(data) // 2]) + _chunk(b"IDAT", data[len(data) // 2:]) + _chunk(b"IEND", b"")


def _png(w, h, ctype, depth, rows, palette=None, **kw):
    """A PNG of raw sample rows, the row filter cycling 0-4 down the image,
    and the RGB bytes it holds (high byte of 16-bit samples, gray repeated,
    palette looked up, alpha dropped), both from the PNG specification."""
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = ch * depth // 8
    stride = bpp * w
    raw, prev = b"", bytes(stride)
    for y in range(h):
        row, f = rows[y], y % 5
        out = bytearray(stride)
        for x in range(stride):
            a = row[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            out[x] = (row[x] - [0, a, b, (a + b) // 2, _paeth(a, b, c)][f]) & 255
        raw += bytes([f]) + bytes(out)
        prev = row
    png = _png_of(w, h, ctype, depth, raw, palette, **kw)
    rgb = np.zeros((h, w, 3), np.uint8)
    step = depth // 8
    for y in range(h):
        for x in range(w):
            s = rows[y][x * bpp:(x + 1) * bpp][::step]
            if ctype == 3:
                rgb[y, x] = list(palette[3 * s[0]:3 * s[0] + 3])
            elif ch >= 3:
                rgb[y, x] = list(s[:3])
            else:
                rgb[y, x] = s[0]
    return png, rgb


W, H = 37, 23   # odd sizes: rows line up with nothing


def _rows(n, smooth, rng):
    return [bytes(((x * 3 + y * 5 + rng.randint(0, 20)) & 255) if smooth else rng.randint(0, 255) for x in range(n))
            for y in range(H)]


def _cases():
    rng = random.Random(1)
    pal = bytes(rng.randint(0, 255) for _ in range(3 * 16))
    return {
        "rgb8": _png(W, H, 2, 8, _rows(3 * W, True, rng)),
        "rgba8": _png(W, H, 6, 8, _rows(4 * W, True, rng)),
        "gray8": _png(W, H, 0, 8, _rows(W, True, rng)),
        "graya8": _png(W, H, 4, 8, _rows(2 * W, False, rng)),
        "rgb16": _png(W, H, 2, 16, _rows(6 * W, False, rng)),
        "rgba16": _png(W, H, 6, 16, _rows(8 * W, False, rng)),
        "gray16": _png(W, H, 0, 16, _rows(2 * W, True, rng)),
        "pal8": _png(W, H, 3, 8, [bytes(rng.randint(0, 15) for _ in range(W)) for _ in range(H)], palette=pal),
        "stored": _png(W, H, 2, 8, _rows(3 * W, False, rng), level=0),
        "fixed": _png(W, H, 2, 8, _rows(3 * W, True, rng), strategy=zlib.Z_FIXED),
        "rle": _png(W, H, 2, 8, _rows(3 * W, True, rng), strategy=zlib.Z_RLE),
        "huffman only": _png(W, H, 2, 8, _rows(3 * W, False, rng), strategy=zlib.Z_HUFFMAN_ONLY),
        "one pixel": _png(1, 1, 2, 8, [bytes([9, 8, 7])]),
        "wide": _png(3000, 2, 2, 8, [bytes((7 * x + y) & 255 for x in range(9000)) for y in range(2)], level=9),
    }


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_png_decodes_to_its_pixels(name):
    png, rgb = CASES[name]
    assert np.array_equal(pt.decode_image(png), rgb)


@pytest.mark.parametrize("name", sorted(CASES))
def test_png_decodes_as_the_reference_decodes(name, ref_lib):
    png, _ = CASES[name]
    assert np.array_equal(pt.decode_image(png), ob.ref_decode_image(png))


def test_reference_texture_decodes_to_the_reference_bytes():
    # models/spot/spot_texture.png, 1024 x 1024, against what the reference's
    # decoder made of it (the texture record of the spot fixture).
    png = open(os.path.join(rc.GOLDEN, "spot_texture.png"), "rb").read()
    (want,) = rc.ptsd_textures(rc.load_spot("spot"))
    assert np.array_equal(pt.decode_image(png), want)


def test_ppm():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    p6 = b"P6\n# a comment\n7 5\n255\n" + img.tobytes()
    assert np.array_equal(pt.decode_image(p6), img)
    p3 = ("P3 7 # width\n 5 255\n" + " ".join(str(v) for v in img.reshape(-1)) + "\n").encode()
    assert np.array_equal(pt.decode_image(p3), img)
    # a maximum value below 255 scales each value by 255 / max, in integers
    low = (img // 16).astype(np.uint8)
    p6 = b"P6 7 5 15 " + low.tobytes()
    assert np.array_equal(pt.decode_image(p6), (low.astype(np.int64) * 255 // 15).astype(np.uint8))
    p3 = b"P3\n2 1\n3\n0 1 2 3 2 1\n"
    assert pt.decode_image(p3).tolist() == [[[0, 85, 170], [255, 170, 85]]]


def _corrupt(png, offset, value):
    b = bytearray(png)
    b[offset] = value
    return bytes(b)


def _refusals():
    png, _ = CASES["rgb8"]
    rng = random.Random(3)
    ihdr = 8 + 8   # IHDR body
    return {
        "not an image": b"GIF89a....",
        "empty": b"",
        "interlaced": _png(4, 4, 2, 8, _rows(12, True, rng)[:4], interlace=1)[0],
        "4-bit gray": _corrupt(png, ihdr + 8, 4),
        "16-bit palette": _png_of(4, 2, 3, 16, bytes(18), palette=bytes(48)),
        "color type 5": _corrupt(png, ihdr + 9, 5),
        "bad filter": _png_of(4, 2, 0, 8, bytes([0, 1, 2, 3, 4, 7, 1, 2, 3, 4])),
        "cut short": png[:len(png) // 2],
        "chunk past the end": png[:-20],
        "image data short": _png_of(4, 2, 0, 8, bytes(9)),
        "palette index out of range": _png_of(3, 1, 3, 8, bytes([0, 0, 1, 5]), palette=bytes(9)),
        "zero width": _png_of(0, 1, 0, 8, bytes(1)),
        "bad zlib header": _png_of(4, 1, 0, 8, bytes(5))[:41] + b"\x00" + _png_of(4, 1, 0, 8, bytes(5))[42:],
        "P5": b"P5 1 1 255 \x00",
        "16-bit PPM": b"P6 1 1 65535 \x00\x00\x00\x00\x00\x00",
        "PPM cut short": b"P6 2 2 255 \x00\x00\x00",
        "PPM bad header": b"P3 x 1 255 0 0 0",
    }


REFUSALS = _refusals()


@pytest.mark.parametrize("name", sorted(REFUSALS))
def test_what_cannot_be_read_is_refused(name):
    with pytest.raises(pt.PTError):
        pt.decode_image(REFUSALS[name])


def test_the_reference_refuses_the_same(ref_lib):
    # The cases the reference's decoder refuses too (it reads a zero-width
    # image as empty, where a texture needs a pixel).
    for name, data in REFUSALS.items():
        if name == "zero width":
            continue
        with pytest.raises(RuntimeError):
            ob.ref_decode_image(data)
            pytest.fail(name)


def test_add_obj_takes_encoded_images():
    # An image given as a file's bytes is decoded as decode_image decodes it.
    obj = b"mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nusemtl t\nf 1/1 2/1 3/1\n"
    mtl = b"newmtl t\nKd 1 1 1\nmap_Kd tex.png\n"
    png, rgb = CASES["rgba8"]
    base = pt.generate(0, 0, 0.05)
    out = pt.add_obj(base, obj, (0, 0, 0), 1.0, 0, mtl_text=mtl, images={"tex.png": png})
    assert int(out[17]) == 1 and np.array_equal(rc.ptsd_textures(out)[0], rgb)
