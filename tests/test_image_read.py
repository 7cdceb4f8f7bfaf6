"""Frame ingest (qlp_image_read) on the CPU: the reference's own test frames
(resources/fftBenchmark/rankine_vortex01_{0,1}.tif, 8-bit PackBits) against
their decoded pixels, and frames written by PIL in every layout the reader
accepts, read back exactly; unsupported layouts are refused."""
import os

import numpy as np
import pytest
from PIL import Image

import qlpiv

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_tiffs_decode_to_their_pixels(rankine_pair):
    a, b = rankine_pair
    assert np.array_equal(qlpiv.read_image(os.path.join(GOLDEN, "rankine_vortex01_0.tif")), a)
    assert np.array_equal(qlpiv.read_image(os.path.join(GOLDEN, "rankine_vortex01_1.tif")), b)


def frame(seed, H=97, W=131):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, 256, (H, W)).astype(np.uint8)
    f[10:30, 40:90] = 7  # runs for the run-length and dictionary coders
    f[50] = np.arange(W) % 256
    return f


@pytest.mark.parametrize("compression", [None, "packbits", "tiff_lzw", "tiff_adobe_deflate"])
@pytest.mark.parametrize("rows_per_strip", [None, 8])
def test_tiff_layouts_round_trip(tmp_path, compression, rows_per_strip):
    f = frame(1)
    path = tmp_path / "f.tif"
    kw = {}
    if rows_per_strip:
        kw["tiffinfo"] = {278: rows_per_strip}
    Image.fromarray(f).save(path, compression=compression, **kw)
    assert np.array_equal(qlpiv.read_image(path), f)


def test_tiff_horizontal_predictor(tmp_path):
    f = frame(2)
    path = tmp_path / "p.tif"
    Image.fromarray(f).save(path, compression="tiff_lzw", tiffinfo={317: 2})
    if Image.open(path).tag_v2.get(317, 1) != 2:
        pytest.skip("PIL wrote no predictor")
    assert np.array_equal(qlpiv.read_image(path), f)


@pytest.mark.parametrize("optimize", [False, True])
def test_png_round_trip(tmp_path, optimize):
    f = frame(3, 64, 200)
    path = tmp_path / "f.png"
    Image.fromarray(f).save(path, optimize=optimize)
    assert np.array_equal(qlpiv.read_image(path), f)


def particles(seed, H, W):
    """A PIV-like frame: Gaussian particles on a dark background."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:H, 0:W]
    f = np.zeros((H, W))
    for cy, cx, a in zip(rng.uniform(0, H, 300), rng.uniform(0, W, 300), rng.uniform(80, 255, 300)):
        f += a * np.exp(-((y - cy) ** 2 + (x - cx) ** 2) / 2.0)
    return np.clip(f + rng.normal(0, 3, (H, W)), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("quality", [5, 50, 75, 95, 100])
@pytest.mark.parametrize("shape", [(64, 64), (97, 131), (8, 8), (1, 13)])
def test_jpeg_decodes_as_libjpeg(tmp_path, quality, shape):
    """Baseline grayscale JPEG: the samples equal those of PIL's decoder
    (libjpeg-turbo, integer inverse DCT), bit for bit, at every quality and
    for sizes that are not multiples of 8."""
    f = particles(quality + shape[1], *shape)
    path = tmp_path / "f.jpg"
    Image.fromarray(f).save(path, quality=quality)
    ref = np.asarray(Image.open(path))
    assert np.array_equal(qlpiv.read_image(path), ref)


@pytest.mark.parametrize("kw", [{"optimize": True}, {"restart_marker_blocks": 3},
                                {"restart_marker_rows": 1}, {"qtables": [[1] * 64]},
                                {"qtables": [[255] * 64]}])
def test_jpeg_tables_and_restart_intervals(tmp_path, kw):
    """Optimised Huffman tables, restart intervals, and extreme quantisation
    (all 1: the widest coefficients; all 255: the largest overshoot)."""
    f = particles(7, 75, 90)
    f[20:40, 20:60] = 255
    f[50:60] = 0
    path = tmp_path / "f.jpg"
    Image.fromarray(f).save(path, **kw)
    ref = np.asarray(Image.open(path))
    assert np.array_equal(qlpiv.read_image(path), ref)


def frame16(seed, H=61, W=77):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, 65536, (H, W)).astype(np.uint16)
    f[H // 10:H // 3, W // 8:W // 2] = 4095  # runs for the run-length and dictionary coders
    f[H // 2] = np.arange(W) * 811 % 65536
    return f


@pytest.mark.parametrize("compression", [None, "packbits", "tiff_lzw", "tiff_adobe_deflate"])
@pytest.mark.parametrize("order", ["<", ">"])
def test_tiff16_round_trip(tmp_path, compression, order):
    """16-bit grayscale TIFF strips in both byte orders and every accepted
    compression read back as the written samples."""
    f = frame16(3)
    path = tmp_path / "f16.tif"
    mode = "I;16" if order == "<" else "I;16B"
    Image.frombytes(mode, (f.shape[1], f.shape[0]), f.astype(order + "u2").tobytes()).save(
        path, compression=compression, tiffinfo={278: 16})
    assert np.array_equal(qlpiv.read_image16(path), f)


def test_tiff16_horizontal_predictor(tmp_path):
    """Predictor 2 on 16-bit samples: strips written by hand (uncompressed,
    differences of whole samples modulo 2^16, little endian)."""
    import struct

    f = frame16(4, 9, 11)
    d = f.astype(np.int64)
    diff = np.concatenate([d[:, :1], np.diff(d, axis=1)], axis=1) % 65536
    data = diff.astype("<u2").tobytes()
    H, W = f.shape
    tags = [(256, 3, 1, W), (257, 3, 1, H), (258, 3, 1, 16), (259, 3, 1, 1), (262, 3, 1, 1),
            (273, 4, 1, 0), (277, 3, 1, 1), (278, 3, 1, H), (279, 4, 1, len(data)), (317, 3, 1, 2)]
    ifd_off = 8
    data_off = ifd_off + 2 + 12 * len(tags) + 4
    out = bytearray(b"II*\x00" + struct.pack("<I", ifd_off) + struct.pack("<H", len(tags)))
    for tag, typ, cnt, val in tags:
        val = data_off if tag == 273 else val
        out += struct.pack("<HHI", tag, typ, cnt) + (struct.pack("<HH", val, 0) if typ == 3
                                                      else struct.pack("<I", val))
    out += struct.pack("<I", 0) + data
    path = tmp_path / "p16.tif"
    path.write_bytes(bytes(out))
    assert np.array_equal(qlpiv.read_image16(path), f)


def test_png16_round_trip(tmp_path):
    """16-bit grayscale PNG (big-endian samples, every row filter chosen by
    the encoder) reads back exactly."""
    f = frame16(5)
    path = tmp_path / "f16.png"
    Image.frombytes("I;16", (f.shape[1], f.shape[0]), f.astype("<u2").tobytes()).save(path)
    with Image.open(path) as im:
        assert im.mode.startswith("I")
    assert np.array_equal(qlpiv.read_image16(path), f)


def test_16bit_reads_refuse_8bit_files(tmp_path):
    """8-bit files are widened through Java2D by the reference's 16-bit
    reader: refused; 16-bit files are refused by the 8-bit reader."""
    f8 = frame(9)
    for name in ("g8.png", "g8.tif", "g8.jpg"):
        Image.fromarray(f8).save(tmp_path / name)
        with pytest.raises(qlpiv.QlpError) as e:
            qlpiv.read_image16(tmp_path / name)
        assert e.value.code == qlpiv._lib.QLP_ERR_UNSUPPORTED
    f = frame16(6)
    Image.frombytes("I;16", (f.shape[1], f.shape[0]), f.astype("<u2").tobytes()).save(tmp_path / "h.tif")
    with pytest.raises(qlpiv.QlpError):
        qlpiv.read_image(tmp_path / "h.tif")


def test_unsupported_layouts_are_refused(tmp_path):
    rgb = np.zeros((8, 8, 3), np.uint8)
    for name in ("c.png", "c.tif"):
        Image.fromarray(rgb).save(tmp_path / name)
        with pytest.raises(qlpiv.QlpError) as e:
            qlpiv.read_image(tmp_path / name)
        assert e.value.code == qlpiv._lib.QLP_ERR_UNSUPPORTED
    Image.fromarray(np.zeros((8, 8), np.uint16)).save(tmp_path / "d16.png")
    with pytest.raises(qlpiv.QlpError):
        qlpiv.read_image(tmp_path / "d16.png")
    for name, kw in (("p.jpg", {"progressive": True}), ("rgb.jpg", {})):
        img = Image.fromarray(rgb if name == "rgb.jpg" else np.zeros((16, 16), np.uint8))
        img.save(tmp_path / name, **kw)
        with pytest.raises(qlpiv.QlpError) as e:
            qlpiv.read_image(tmp_path / name)
        assert e.value.code == qlpiv._lib.QLP_ERR_UNSUPPORTED
    with pytest.raises(qlpiv.QlpError):
        qlpiv.read_image(tmp_path / "missing.tif")
