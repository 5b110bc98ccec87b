"""Display transform (trt/tonemap.h) and image writers (image_io.cpp).

Tone curves are checked at the points their published constructions fix (Lottes: mid grey
0.18 -> 0.267 and 8 -> 1; Uncharted 2: the white point; Uchimura: continuity of its toe,
linear and shoulder sections; Reinhard variants in closed form); the writers by reading
the files back with independent decoders (numpy for PFM, zlib for PNG)."""
import ctypes
import struct
import zlib

import numpy as np
import pytest

from oracle.oracle_py import tone_curve, tonemap
from trayracer_amd import _lib


def test_reinhard_and_identity():
    x = np.array([0.0, 1.0, 3.0, 4.0], np.float32)
    np.testing.assert_allclose(tone_curve(x, 3), x / (1 + x), rtol=1e-7)
    np.testing.assert_allclose(tone_curve(x, 4), x * (1 + x / 16) / (1 + x), rtol=1e-7)
    assert tone_curve(np.float32([4.0]), 4)[0] == pytest.approx(1.0, abs=1e-7)  # white point 4
    np.testing.assert_array_equal(tone_curve(x, 7), x)


def test_aces_and_filmic_shape():
    x = np.linspace(0, 20, 2001).astype(np.float32)
    for m in (0, 1):
        y = tone_curve(x, m)
        assert y[0] == 0.0 and (np.diff(y) >= -1e-7).all() and y.max() <= 1.0 + 1e-6
    assert tone_curve(np.float32([0.004]), 1)[0] == 0.0  # filmic black offset
    assert tone_curve(np.float32([1e6]), 1)[0] == pytest.approx(1.0, abs=1e-4)
    assert tone_curve(np.float32([1e6]), 0)[0] == 1.0  # ACES clamps


def test_lottes_anchor_points():
    y = tone_curve(np.float32([0.18, 8.0]), 2)
    np.testing.assert_allclose(y, [0.267, 1.0], rtol=2e-5)


def test_uncharted2_white_point():
    assert tone_curve(np.float32([5.6]), 6)[0] == pytest.approx(1.0, abs=1e-6)  # 2 x 5.6 = W = 11.2
    y = tone_curve(np.linspace(0, 5.6, 200).astype(np.float32), 6)
    assert (np.diff(y) > 0).all()


def test_uchimura_sections_join():
    m, l0 = 0.22, (1.0 - 0.22) * 0.4
    y = tone_curve(np.float32([m, m + l0, 0.3, 1e3]), 5)
    np.testing.assert_allclose(y[:3], [m, m + l0, 0.3], rtol=2e-6)  # toe meets the line at m; line is identity
    assert y[3] == pytest.approx(1.0, abs=1e-6)  # shoulder approaches P = 1
    x = np.linspace(0, 4, 4001).astype(np.float32)
    assert (np.diff(tone_curve(x, 5)) >= -1e-6).all()


def test_quantisation_and_gamma():
    img = np.array([[[1.0, 0.0, 3.0]]], np.float32)
    out = tonemap(img, method=3, exposure=1.0, gamma=1.0)
    assert out[0, 0].tolist() == [128, 0, 191, 255]  # 0.5, 0, 0.75 -> round(255 v)
    out = tonemap(img, method=7, exposure=0.25, gamma=2.0)
    assert out[0, 0].tolist() == [128, 0, 221, 255]  # sqrt(0.25) = .5, sqrt(.75) = .866


def _lib_or_skip():
    return _lib.load()


def test_pfm_round_trip(tmp_path):
    lib = _lib_or_skip()
    rng = np.random.default_rng(0)
    img = rng.uniform(-1, 10, (7, 5, 4)).astype(np.float32)
    p = tmp_path / "a.pfm"
    assert lib.trt_write_pfm(str(p).encode(), 5, 7, img.ctypes.data_as(_lib.c_float_p)) == 0
    data = p.read_bytes()
    head = data.split(b"\n", 3)
    assert head[0] == b"PF" and head[1] == b"5 7" and float(head[2]) < 0  # little endian
    px = np.frombuffer(head[3], "<f4").reshape(7, 5, 3)
    np.testing.assert_array_equal(px, img[..., :3])


def test_png_round_trip(tmp_path):
    lib = _lib_or_skip()
    rng = np.random.default_rng(1)
    for (w, h) in ((3, 2), (300, 70)):  # the second needs several stored deflate blocks
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        p = tmp_path / f"a{w}.png"
        assert lib.trt_write_png(str(p).encode(), w, h, img.ctypes.data_as(_lib.c_u8_p)) == 0
        data = p.read_bytes()
        assert data[:8] == b"\x89PNG\r\n\x1a\n"
        pos, chunks = 8, {}
        while pos < len(data):
            n = struct.unpack(">I", data[pos:pos + 4])[0]
            typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
            crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
            assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
            chunks[typ] = body
            pos += 12 + n
        assert struct.unpack(">IIBBBBB", chunks[b"IHDR"]) == (w, h, 8, 6, 0, 0, 0)
        raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(h, 1 + 4 * w)
        assert (raw[:, 0] == 0).all()
        np.testing.assert_array_equal(raw[:, 1:].reshape(h, w, 4), img[::-1])  # file rows top-down
