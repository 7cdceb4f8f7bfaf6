"""Hostile input and sanitizer builds of the host-side code (CPU only).

* Crafted frame files whose headers declare impossible sizes, and a JPEG whose
  DC prediction overflows, are refused with QLP_ERR_ARG instead of throwing
  through the C ABI or overflowing an int.
* tests/sanitize/: the library's frame decoders and MATLAB writer built with
  AddressSanitizer + UndefinedBehaviorSanitizer and driven by a bounded,
  seeded mutation run over TIFF / PNG / JPEG seeds; the C oracle built the same
  way and run through its own CPU tests.
"""
import os
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest
from PIL import Image

import qlpiv
from qlpiv import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def tiff_header(W, H, strip=b"\0" * 16):
    """A little-endian 8-bit grayscale TIFF: one uncompressed strip."""
    entries = [(256, 4, 1, W), (257, 4, 1, H), (258, 3, 1, 8), (259, 3, 1, 1), (262, 3, 1, 1),
               (273, 4, 1, 0), (277, 3, 1, 1), (278, 4, 1, H), (279, 4, 1, len(strip))]
    ifd = 8
    data_at = ifd + 2 + 12 * len(entries) + 4
    out = bytearray(b"II*\0" + struct.pack("<I", ifd) + struct.pack("<H", len(entries)))
    for tag, typ, cnt, val in entries:
        if tag == 273:
            val = data_at
        out += struct.pack("<HHI", tag, typ, cnt)
        out += struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val)
    out += b"\0\0\0\0" + strip
    return bytes(out)


def png_header(W, H):
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 0, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(b"\0" * 64)) + chunk(b"IEND", b""))


def jpeg_dc_overflow(blocks=40):
    """Baseline one-component JPEG: 16-bit quantisers of 65535, a DC table whose
    only code is category 11 and an AC table whose only code is end-of-block;
    every block adds the maximum DC difference 2047 to the prediction."""
    seg = lambda m, d: b"\xff" + bytes([m]) + struct.pack(">H", len(d) + 2) + d
    dqt = seg(0xDB, bytes([0x10]) + b"\xff\xff" * 64)
    sof = seg(0xC0, bytes([8]) + struct.pack(">HH", 8, 8 * blocks) + bytes([1, 1, 0x11, 0]))
    one = bytes([1] + [0] * 15)
    dht = seg(0xC4, bytes([0x00]) + one + bytes([11])) + seg(0xC4, bytes([0x10]) + one + bytes([0]))
    sos = seg(0xDA, bytes([1, 1, 0x00, 0, 63, 0]))
    bits = ("0" + "1" * 11 + "0") * blocks
    bits += "1" * (-len(bits) % 8)
    data = bytearray()
    for k in range(0, len(bits), 8):
        byte = int(bits[k:k + 8], 2)
        data.append(byte)
        if byte == 0xFF:
            data.append(0)
    return b"\xff\xd8" + dqt + sof + dht + sos + bytes(data) + b"\xff\xd9"


@pytest.mark.parametrize("name,blob", [
    ("tiff_4g_square", tiff_header(0xFFFFFFFF, 0xFFFFFFFF)),
    ("tiff_2g_row", tiff_header(0x80000000, 1)),
    ("tiff_strips_short", tiff_header(65536, 65536)),
    ("png_4g_square", png_header(0xFFFFFFFF, 0xFFFFFFFF)),
    ("png_2g_row", png_header(0x7FFFFFFF, 2)),
    ("jpeg_dc_overflow", jpeg_dc_overflow()),
])
def test_hostile_frames_are_refused(tmp_path, name, blob):
    path = tmp_path / name
    path.write_bytes(blob)
    for reader in (qlpiv.read_image, qlpiv.read_image16):
        with pytest.raises(qlpiv.QlpError) as e:
            reader(path)
        assert e.value.code in (_lib.QLP_ERR_ARG, _lib.QLP_ERR_UNSUPPORTED), str(e.value)


def test_jpeg_dc_overflow_message(tmp_path):
    path = tmp_path / "dc.jpg"
    path.write_bytes(jpeg_dc_overflow())
    with pytest.raises(qlpiv.QlpError, match="DC coefficient out of range"):
        qlpiv.read_image(path)
    # below the bound the same stream decodes (16 blocks: |pred| <= 32752)
    path.write_bytes(jpeg_dc_overflow(16))
    assert qlpiv.read_image(path).shape == (8, 128)


@pytest.fixture(scope="module")
def sanitizer_builds():
    subprocess.check_call(["make", "-s", "-C", SAN])
    return os.path.join(SAN, "_build")


def seeds(d):
    rng = np.random.default_rng(0)
    f = rng.integers(0, 256, (40, 56)).astype(np.uint8)
    f[5:20, 10:30] = 9
    out = []
    for c in (None, "packbits", "tiff_lzw", "tiff_adobe_deflate"):
        p = os.path.join(d, f"s_{c}.tif")
        Image.fromarray(f).save(p, compression=c)
        out.append(p)
    for name, img, kw in (("s.png", f, {}), ("s.jpg", f, {"quality": 80}),
                          ("s16.png", f.astype(np.uint16) * 200, {})):
        p = os.path.join(d, name)
        Image.fromarray(img).save(p, **kw)
        out.append(p)
    out.append(os.path.join(ROOT, "tests", "golden", "rankine_vortex01_0.tif"))
    return out


def test_host_decoders_under_sanitizers(tmp_path, sanitizer_builds):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(sanitizer_builds, "host_fuzz"), "1000", str(tmp_path)] +
                       seeds(str(tmp_path)), env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host_fuzz ok" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]


def test_oracle_under_sanitizers(sanitizer_builds):
    asan = subprocess.check_output(["gcc", "-print-file-name=libasan.so"], text=True).strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1",
               QLPIV_ORACLE_LIB=os.path.join(sanitizer_builds, "liboracle_san.so"))
    tests = ["test_oracle.py", "test_oracle_validation.py", "test_oracle_inherit.py",
             "test_oracle_splinewarp.py", "test_oracle_warp_orders.py", "test_filter.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu"] + [os.path.join(ROOT, "tests", t) for t in tests],
                       env=env, capture_output=True, text=True, timeout=900, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
