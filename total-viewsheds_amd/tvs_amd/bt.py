"""`.bt` files, heatmap PNGs and the band-data cache through the C ABI
(input.rs:86-185, output/png.rs:5-37, cache.rs:6-76)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BtHeader, _check_io, _ptr, load


def read_header(path: str) -> BtHeader:
    h = BtHeader()
    _check_io(load().tvs_bt_read_header(path.encode(), ctypes.byref(h)))
    return h


def read(path: str, as_float: bool | None = None) -> np.ndarray:
    """Samples as int16 (int16 files) or float32, flat, in file order."""
    h = read_header(path)
    n = h.width * h.height
    if h.is_float or as_float:
        out = np.empty(n, np.float32)
        _check_io(load().tvs_bt_read_f32(path.encode(), _ptr(out), n))
    else:
        out = np.empty(n, np.int16)
        _check_io(load().tvs_bt_read_i16(path.encode(), _ptr(out), n))
    return out


def write(path: str, elevations: np.ndarray, left=0.0, right=0.0, bottom=0.0,
          top=0.0, horizontal_units: int = 1) -> None:
    a = np.ascontiguousarray(elevations)
    h = BtHeader()
    h.height, h.width = a.shape if a.ndim == 2 else (int(np.sqrt(a.size)),) * 2
    h.is_float = 0 if a.dtype == np.int16 else 1
    if h.is_float:
        a = np.ascontiguousarray(a, np.float32)
    h.data_size = 2 if a.dtype == np.int16 else 4
    h.horizontal_units = horizontal_units
    h.left, h.right, h.bottom, h.top = left, right, bottom, top
    h.vertical_scale = 1.0
    _check_io(load().tvs_bt_write(path.encode(), ctypes.byref(h), _ptr(a)))


def scale(h: BtHeader) -> float:
    return float(load().tvs_bt_scale(ctypes.byref(h)))


def heatmap_png(path: str, data: np.ndarray) -> None:
    a = np.ascontiguousarray(data, np.float32)
    hgt, wid = a.shape
    _check_io(load().tvs_write_heatmap_png(path.encode(), _ptr(a), wid, hgt))
