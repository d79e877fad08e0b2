"""The walk's fast divide (csrc/fastdiv.hpp: split reciprocal + one
Markstein correction) equals IEEE f32 division on its whole proven domain.
The header is compiled for the host here, so the code checked is the code
the kernel inlines.  CPU only."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fd(tmp_path_factory):
    out = tmp_path_factory.mktemp("fastdiv") / "libfastdiv.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-fPIC", "-shared", "-o", str(out),
                    os.path.join(HERE, "native", "fastdiv_check.cpp")], check=True)
    lib = ctypes.CDLL(str(out))
    lib.fastdiv_random.restype = ctypes.c_uint64
    lib.fastdiv_random.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.fastdiv_edges.restype = ctypes.c_uint64
    lib.fastdiv_edges.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.fastdiv_hard.restype = ctypes.c_uint64
    lib.fastdiv_hard.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                 ctypes.POINTER(ctypes.c_uint64)]
    lib.fastdiv_midpoints.restype = ctypes.c_uint64
    lib.fastdiv_midpoints.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint64)]
    lib.fastdiv_pair.restype = None
    lib.fastdiv_pair.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    lib.fastdiv_heights.restype = ctypes.c_uint64
    lib.fastdiv_heights.argtypes = [ctypes.c_int32] * 5 + [ctypes.c_float, ctypes.c_void_p,
                                                          ctypes.c_uint64]
    return lib


def test_random_operands(fd):
    assert fd.fastdiv_random(20_000_000, 1) == 0


def test_hard_divisors(fd):
    assert fd.fastdiv_edges(5_000_000, 2) == 0


def test_near_midpoint_quotients(fd):
    found = ctypes.c_uint64(0)
    assert fd.fastdiv_hard(20_000, 3, ctypes.byref(found)) == 0
    assert found.value == 20_000


def test_quotients_next_to_midpoints(fd):
    """Where the first quotient can round the wrong way and only the
    residual step decides."""
    found = ctypes.c_uint64(0)
    assert fd.fastdiv_midpoints(2_000_000, 4, ctypes.byref(found)) == 0
    assert found.value == 2_000_000


def test_planted_error_is_caught(fd, tmp_path):
    """A wrong sign on the residual correction must make the check fail:
    the test can see a wrong divide."""
    src = open(os.path.join(HERE, "native", "fastdiv_check.cpp")).read()
    hdr = open(os.path.join(HERE, "..", "total-viewsheds_amd", "csrc", "fastdiv.hpp")).read()
    broken = hdr.replace("std::fma(r, y_hi, q0)", "std::fma(-r, y_hi, q0)")
    assert broken != hdr
    (tmp_path / "fastdiv.hpp").write_text(broken)
    (tmp_path / "check.cpp").write_text(
        src.replace('"../../total-viewsheds_amd/csrc/fastdiv.hpp"', '"fastdiv.hpp"'))
    so = tmp_path / "broken.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-o", str(so), str(tmp_path / "check.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    lib.fastdiv_midpoints.restype = ctypes.c_uint64
    lib.fastdiv_midpoints.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint64)]
    found = ctypes.c_uint64(0)
    assert lib.fastdiv_midpoints(20_000, 4, ctypes.byref(found)) > 0


@pytest.mark.parametrize("w,scale,los,obs,hmax", [
    (201, 100.0, 6700, 1.65, 1500),
    (1201, 90.0, 36000, 1.65, 3000),
    (257, 30.0, 2550, 1.8, 4000),
])
def test_workload_domain(fd, oracle, w, scale, los, obs, hmax):
    """Every band distance of 12 sectors x every integer height difference
    the workload's height range can produce (observer heights sampled)."""
    dem = oracle.dem_new(w, scale, los)
    _, dists = oracle.band_data_all(dem, range(0, 180, 15))
    d = np.unique(dists[dists > 0]).astype(np.float32)
    bad = fd.fastdiv_heights(0, hmax, 0, hmax, 97, obs, d.ctypes.data, d.size)
    assert bad == 0


@pytest.mark.parametrize("a_bits,d_bits", [(0xc4b6f4d0, 0x42fcf556), (0xc77240a6, 0x437bc406)])
def test_short_divide_can_be_wrong(fd, a_bits, d_bits):
    """The two-operation short divide misses the correctly rounded quotient
    for some pairs (found among the height differences of a full int16 DEM
    and the band distances of the 1201 x 1201 workload), which is why the
    walk uses it only for sectors the device has verified; the split-
    reciprocal divide is right on them."""
    out = np.zeros(3, np.float32)
    fd.fastdiv_pair(a_bits, d_bits, out.ctypes.data)
    assert out[1].view(np.uint32) == out[0].view(np.uint32)
    assert out[2].view(np.uint32) != out[0].view(np.uint32)
