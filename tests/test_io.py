"""Host I/O of the engine library (no device): `.bt` files (input.rs:86-185),
the heatmap PNG (output/png.rs:5-37)."""
import ctypes
import struct
import zlib

import numpy as np
import pytest

import tvs_amd
from tvs_amd import _lib, bt


def test_bt_round_trip_int16(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.integers(-500, 4000, (37, 37)).astype(np.int16)
    p = tmp_path / "x.bt"
    bt.write(str(p), a, left=-2.0, right=-1.0, bottom=51.0, top=52.0)
    h = bt.read_header(str(p))
    assert (h.width, h.height, h.data_size, h.is_float) == (37, 37, 2, 0)
    np.testing.assert_array_equal(bt.read(str(p)), a.reshape(-1))
    np.testing.assert_array_equal(bt.read(str(p), as_float=True), a.reshape(-1).astype(np.float32))


def test_bt_round_trip_float(tmp_path):
    a = np.linspace(-10, 10, 25 * 25, dtype=np.float32).reshape(25, 25)
    p = tmp_path / "f.bt"
    bt.write(str(p), a)
    h = bt.read_header(str(p))
    assert (h.data_size, h.is_float) == (4, 1)
    np.testing.assert_array_equal(bt.read(str(p)), a.reshape(-1))


def test_bt_layout_is_the_reference_one(tmp_path):
    """Header fields at input.rs:90-113 offsets, samples from byte 256."""
    a = np.arange(9, dtype=np.int16).reshape(3, 3)
    p = tmp_path / "l.bt"
    bt.write(str(p), a, left=1.5, right=2.5, bottom=3.5, top=4.5)
    raw = p.read_bytes()
    assert raw[:10] == b"binterr1.3"
    assert struct.unpack_from("<IIHH", raw, 10) == (3, 3, 2, 0)
    assert struct.unpack_from("<dddd", raw, 28) == (1.5, 2.5, 3.5, 4.5)
    assert len(raw) == 256 + 18
    assert np.frombuffer(raw[256:], "<i2").tolist() == list(range(9))


def test_bt_errors(tmp_path):
    p = tmp_path / "bad.bt"
    p.write_bytes(b"binterr1.2" + bytes(300))
    with pytest.raises(tvs_amd.TvsError) as e:
        bt.read_header(str(p))
    assert e.value.status == _lib.ERR_IO and "Binary Terrain" in str(e.value)
    raw = bytearray(b"binterr1.3" + bytes(300))
    struct.pack_into("<IIHH", raw, 10, 4, 4, 8, 0)          # int data of size 8
    p.write_bytes(bytes(raw))
    with pytest.raises(tvs_amd.TvsError) as e:
        bt.read_header(str(p))
    assert "data size" in str(e.value)
    raw = bytearray(b"binterr1.3" + bytes(256))
    struct.pack_into("<IIHH", raw, 10, 40, 40, 2, 0)        # truncated samples
    p.write_bytes(bytes(raw))
    with pytest.raises(tvs_amd.TvsError) as e:
        bt.read(str(p))
    assert e.value.status == _lib.ERR_IO
    with pytest.raises(tvs_amd.TvsError):
        bt.read_header(str(tmp_path / "missing.bt"))


def test_bt_scale_is_the_meridian_haversine():
    """input.rs:178-185 reads (top, right)/(top, left) as (lon, lat): the
    distance is R * |right - left| in radians whatever the latitude."""
    h = _lib.BtHeader()
    h.width, h.left, h.right, h.top, h.bottom = 1201, -2.0, -1.0, 52.0, 51.0
    want = 6371008.8 * np.radians(1.0) / 1201
    assert abs(bt.scale(h) - want) / want < 1e-12


def read_png_grey(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(raw):
        n, = struct.unpack_from(">I", raw, pos)
        kind = raw[pos + 4:pos + 8]
        body = raw[pos + 8:pos + 8 + n]
        crc, = struct.unpack_from(">I", raw, pos + 8 + n)
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack_from(">IIBB", body)
            assert (depth, ctype) == (8, 0)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    px = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, w + 1)
    assert np.all(px[:, 0] == 0)
    return px[:, 1:]


@pytest.mark.parametrize("shape", [(7, 7), (300, 300), (1, 1)])
def test_heatmap_png(tmp_path, shape):
    rng = np.random.default_rng(1)
    data = (rng.random(shape) * 1000).astype(np.float32)
    p = tmp_path / "h.png"
    bt.heatmap_png(str(p), data)
    mn, mx = data.min(), data.max()
    rng_ = np.float32(1.0) if mx - mn == 0 else np.float32(mx - mn)
    want = np.clip((data - mn) / rng_ * np.float32(255.0), 0, 255).astype(np.uint8)
    np.testing.assert_array_equal(read_png_grey(str(p)), want)


def _png_of(tmp_path, data):
    p = tmp_path / "k.png"
    bt.heatmap_png(str(p), np.ascontiguousarray(data, np.float32))
    return read_png_grey(str(p))


def test_heatmap_known_answers(tmp_path):
    """Cases the arithmetic fixes (output/png.rs:5-37), not a restated
    formula: with a power-of-two range every step is exact, so a pixel is
    floor(255 (v - min) / range); the minimum maps to 0 and the maximum to
    255 (range / range = 1); constant data is all 0 (range 1, v - min = 0);
    the map is monotone."""
    v = np.arange(0, 65, dtype=np.float64).reshape(5, 13) - 20.0   # range 64
    np.testing.assert_array_equal(_png_of(tmp_path, v),
                                  np.floor(255.0 * (v + 20.0) / 64.0).astype(np.uint8))
    rng = np.random.default_rng(3)
    data = (rng.random((40, 40)) * 1e4 - 3e3).astype(np.float32)
    px = _png_of(tmp_path, data)
    assert px.flat[np.argmin(data)] == 0 and px.flat[np.argmax(data)] == 255
    order = np.argsort(data, axis=None, kind="stable")
    assert np.all(np.diff(px.reshape(-1)[order].astype(int)) >= 0)
    assert np.all(_png_of(tmp_path, np.full((3, 4), 7.5)) == 0)


def test_heatmap_nan_and_infinities(tmp_path):
    """IEEE arithmetic and Rust's float-to-int cast (NaN to 0, saturating)
    fix these: a NaN pixel is 0 (f32::min/max skip it for the range); an
    infinite value makes the range infinite, so finite pixels divide to 0
    and the infinite one to inf - inf = NaN, hence 0."""
    data = np.array([[0.0, 1.0], [np.nan, 2.0]], np.float32)
    np.testing.assert_array_equal(_png_of(tmp_path, data), [[0, 127], [0, 255]])
    for inf in (np.inf, -np.inf):
        data = np.array([[0.0, 1.0], [inf, 2.0]], np.float32)
        assert np.all(_png_of(tmp_path, data) == 0)
