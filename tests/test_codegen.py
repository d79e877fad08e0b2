"""Shape of the batched walk's inner loop in the gfx950 assembly (CPU only:
hipcc cross-compiles).  Guards the properties the kernel's speed rests on:
8 height loads with scalar offsets per chunk (typed loads that convert the
int16 heights, so no conversion instruction), the step tables read with
scalar dwordx8 loads, the exec-masked visibility update, two FMAs fewer per
step in the short-divide loop, and the only branches in the loop those
that skip a band's masked update when no lane of the wave sees the point."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "total-viewsheds_amd", "csrc", "walk.hip")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    out = tmp_path_factory.mktemp("asm") / "walk.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", str(out),
                    SRC], check=True, capture_output=True)
    return out.read_text()


def kernel_body(asm: str, mangled_part: str) -> str:
    m = re.search(r"^(_Z\S*walk_tiles_kernel" + mangled_part + r"\S*):", asm, re.M)
    assert m, f"kernel {mangled_part} not found"
    start = m.end()
    return asm[start:asm.index("s_endpgm", start)]


def unrolled_loops(body: str) -> list[str]:
    """The loops whose bodies hold the unrolled height loads: one per divide
    variant (split reciprocal, short, IEEE), in the kernel's order."""
    lines = body.splitlines()
    out = []
    for h, line in enumerate(lines):
        if "Inner Loop Header" not in line:
            continue
        end = next((i for i in range(h, len(lines)) if re.search(r"s_cbranch_scc\d", lines[i])),
                   len(lines) - 1)
        chunk = "\n".join(lines[h:end + 1])
        if len(re.findall(r"buffer_load_(format_x|dword)\s", chunk)) >= 8:
            out.append(chunk)
    assert out, "no unrolled loop found"
    return out


def valu(loop: str) -> int:
    return len(re.findall(r"^\s+v_", loop, re.M))


def check_updates(loop: str, band_steps: int) -> None:
    """One compare per band-step, each followed by the branch that skips the
    band's masked update (exec set from vcc, add, move, exec restored) when
    no lane sees the point; no other divergent branch."""
    assert "s_cbranch_execz" not in loop and "v_cmpx" not in loop
    assert len(re.findall(r"v_cmp_gt_f32\s+vcc", loop)) == band_steps
    assert len(re.findall(r"v_cmp_gt_f32\s+vcc, v\d+, v\d+\n\s+s_cbranch_vccz", loop)) == band_steps
    assert len(re.findall(r"s_mov_b64\s+exec, vcc", loop)) == band_steps


def check_loop(loop: str, load: str, band_steps: int) -> None:
    check_updates(loop, band_steps)
    assert len(re.findall(load + r"\s+v\d+, v\d+, s\[\d+:\d+\], s\d+ offen", loop)) == band_steps
    assert "s_load_dwordx8" in loop
    assert "v_cvt_f32_i32" not in loop
    assert "scratch" not in loop and "buffer_store" not in loop


@pytest.mark.parametrize("heights,load", [("Is", "buffer_load_format_x"),
                                          ("If", "buffer_load_dword")])
@pytest.mark.parametrize("pipe", ["Lb0E", "Lb1E"])
def test_walk_loop_shape(asm, heights, load, pipe):
    """One band per lane, 8 steps per chunk: the split-reciprocal loop spends
    9 VALU per step, the short-divide loop 7, the IEEE loop more."""
    loops = unrolled_loops(kernel_body(asm, heights + "Lb0ELi256ELi8E" + pipe + "Lb0ELb0ELi0ELb0E"))
    fast = [l for l in loops if "v_div_scale" not in l]
    assert len(fast) == 2
    for loop in fast:
        check_loop(loop, load, 8)
    counts = sorted(valu(l) for l in fast)
    assert counts[0] <= 7 * 8 + 8 and counts[1] <= 9 * 8 + 8, counts
    assert counts[1] - counts[0] >= 16                   # two FMAs fewer per step


def test_paired_walk_loop_shape(asm):
    """Both bands of a point per lane: 16 height loads and 16 masked updates
    per 8-step chunk, one set of table reads (pos_f, pos_b + the shared
    constants), 9 (split reciprocal) or 7 (short divide) VALU per band-step."""
    loops = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi8ELb0ELb0ELb1ELi0ELb0E"))
    fast = [l for l in loops if "v_div_scale" not in l]
    assert len(fast) == 2
    for loop in fast:
        check_loop(loop, "buffer_load_format_x", 16)
    short, split = sorted(fast, key=valu)
    assert valu(short) <= 7 * 16 + 8 and valu(split) <= 9 * 16 + 8, (valu(short), valu(split))
    assert len(re.findall(r"s_load_dwordx8", split)) == 7
    assert len(re.findall(r"s_load_dwordx8", short)) == 6      # no distance needed


def test_two_point_walk_loop_shape(asm):
    """Two adjacent points per lane: one dword load per step fetches both
    int16 heights (8 loads for 16 band-steps), each converted by one SDWA
    instruction; 8 (short divide) or 10 VALU per band-step."""
    loops = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi8ELb0ELb0ELb0ELi1ELb0E"))
    twin = [l for l in loops if "v_div_scale" not in l and "buffer_load_dword" in l]
    assert len(twin) == 2
    for loop in twin:
        check_updates(loop, 16)
        assert len(re.findall(r"buffer_load_dword\s+v\d+, v\d+, s\[\d+:\d+\], s\d+ offen",
                              loop)) == 8
        assert len(re.findall(r"_sdwa", loop)) == 16
        assert "scratch" not in loop and "buffer_store" not in loop
    short, split = sorted(twin, key=valu)
    assert valu(short) <= 8 * 16 + 8 and valu(split) <= 10 * 16 + 8, (valu(short), valu(split))


def scalar_operand_valu(loop: str) -> int:
    """Vector instructions of a loop that read a scalar register."""
    return sum(1 for args in re.findall(r"^\s+v_\w+\s+([^\n]*)", loop, re.M)
               if re.search(r"\bs\d+\b|\bs\[\d", args))


def test_eight_band_walk_loop_shape(asm):
    """Two pairs of adjacent points per lane (rows 2r and 2r + 1), both
    directions, 4 steps per chunk: 16 dword loads (one per pair, direction
    and step) and 32 masked updates per chunk; the step's constants are
    copied into vector registers once per step, so only those copies (4 per
    step with the short divide, 5 with the split reciprocal) read a scalar
    register, against half the instructions of the four-band loop; 8.5
    vector instructions per band-step with the short divide."""
    loops = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi4ELb0ELb0ELb1ELi2ELb0E"))
    # (the kernel also holds the one-point loops of odd widths' left-over
    # column, which load with buffer_load_format_x)
    fast = [l for l in loops if "v_div_scale" not in l and "buffer_load_format_x" not in l]
    assert len(fast) == 2
    for loop in fast:
        check_updates(loop, 32)
        assert len(re.findall(r"buffer_load_dword\s+v\d+, v\d+, s\[\d+:\d+\], s\d+ offen",
                              loop)) == 16
        assert "scratch" not in loop and "buffer_store" not in loop
    short, split = sorted(fast, key=valu)
    assert valu(short) <= 272 + 8 and valu(split) <= 336 + 8, (valu(short), valu(split))
    assert scalar_operand_valu(short) <= 16 and scalar_operand_valu(split) <= 20
    four = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi8ELb0ELb0ELb1ELi1ELb0E"))
    four_short = min((l for l in four if "v_div_scale" not in l and
                      "buffer_load_format_x" not in l and "v_fma_mix_f32" not in l), key=valu)
    assert valu(four_short) <= 8 * 32 + 8 + 8
    assert scalar_operand_valu(four_short) >= valu(four_short) // 3
    # The four-band loop copies the curvature term into a vector register
    # once per step (8 per chunk), nothing else.
    assert len(re.findall(r"v_mov_b32(?:_e32)?\s+v\d+, s\d+", four_short)) == 8


def test_half_height_loop_shape(asm):
    """int16 heights with their f16 form, two points and both directions per
    lane: the height difference of every band-step is one v_fma_mix_f32 from
    the loaded dword (no conversion, no separate subtraction), 16 dword loads
    per 8-step chunk; 7.5 vector instructions per band-step with the short
    divide (the masked update's two included), three quarters of an
    instruction fewer than the int16 loop of the same kernel."""
    loops = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi8ELb0ELb0ELb1ELi1ELb0E"))
    fast = [l for l in loops if "v_div_scale" not in l and "buffer_load_format_x" not in l]
    half = [l for l in fast if "v_fma_mix_f32" in l]
    cvt = [l for l in fast if "v_fma_mix_f32" not in l]
    assert len(half) == 2 and len(cvt) == 2
    for loop in half:
        check_updates(loop, 32)
        assert len(re.findall(r"v_fma_mix_f32\s", loop)) == 32
        assert "v_cvt_f32_i32" not in loop
        assert len(re.findall(r"buffer_load_dword\s+v\d+, v\d+, s\[\d+:\d+\], s\d+ offen",
                              loop)) == 16
    # 32 conversions and 32 subtractions become 32 mixed FMAs; the step's
    # curvature term and reciprocal high part are copied into vector
    # registers (16 copies per chunk, 8 in the int16 loop).
    assert min(valu(l) for l in half) + 32 - 8 == min(valu(l) for l in cvt)
    assert min(valu(l) for l in half) <= 7 * 32 + 16 + 8
    for loop in half:
        assert len(re.findall(r"v_mov_b32(?:_e32)?\s+v\d+, s\d+", loop)) == 16


def test_eight_band_kernel_holds_five_waves(asm):
    info = kernel_info(asm, "IsLb0ELi64ELi4ELb0ELb0ELb1ELi2ELb0E")
    assert info["Occupancy"] >= 5 and info["ScratchSize"] == 0, info


def kernel_info(asm: str, mangled_part: str) -> dict:
    """The compiler's register and occupancy summary of one kernel."""
    m = re.search(r"^(_Z\S*walk_tiles_kernel" + mangled_part + r"\S*):", asm, re.M)
    assert m, f"kernel {mangled_part} not found"
    start = asm.index("; Kernel info:", m.end())
    block = asm[start:asm.index("; Occupancy:", start) + 40]
    return {k: int(v) for k, v in re.findall(r"; (\w+): (\d+)", block)}


@pytest.mark.parametrize("shape", ["Li64ELi4ELb0ELb0ELb1ELi1E", "Li64ELi8ELb0ELb0ELb1ELi1E",
                                   "Li128ELi4ELb0ELb0ELb1ELi1E", "Li64ELi8ELb0ELb0ELb0ELi1E",
                                   "Li256ELi8ELb0ELb0ELb0ELi0E"])
def test_fast_divide_kernels_keep_seven_waves(asm, shape):
    """The fast-divide kernels walk a tile without the fast divides one IEEE
    step at a time, so their registers are those of the fast loops: at most
    72 VGPRs, seven or more waves per SIMD (the unrolled IEEE loop alone
    needs ~94 VGPRs, five waves)."""
    info = kernel_info(asm, "IsLb0E" + shape + "Lb0E")
    assert info["TotalNumVgprs"] <= 72 and info["Occupancy"] >= 7, info
    assert info["ScratchSize"] == 0


def test_ieee_only_kernel_unrolls_the_ieee_loop(asm):
    """The IEEE-only instantiation holds the unrolled IEEE loop (16 loads of
    two heights, IEEE divide sequences) and no fast-divide loop; the
    fast-divide instantiation holds both fast loops and no unrolled IEEE loop."""
    ieee = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi8ELb0ELb0ELb0ELi1ELb1E"))
    assert ieee and all("v_div_scale" in l for l in ieee)
    fast = unrolled_loops(kernel_body(asm, "IsLb0ELi64ELi8ELb0ELb0ELb0ELi1ELb0E"))
    assert not any("v_div_scale" in l for l in fast)
    assert len([l for l in fast if "buffer_load_dword" in l]) == 2


def test_no_packed_fp32_and_no_spills(asm):
    assert not re.search(r"\bv_pk_(add|mul|fma)_f32\b", asm)
    sizes = re.findall(r"\.private_segment_fixed_size:\s+(\d+)", asm)
    assert sizes and all(int(x) == 0 for x in sizes)


DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                "-fhip-fp32-correctly-rounded-divide-sqrt", "--cuda-device-only", "-S"]
NO_PACKED = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def test_no_packed_fp32_in_band_kernels(tmp_path):
    """bands.hip (sort, selection, step tables) compiled as the library is:
    no packed-FP32 instruction and no scratch."""
    out = tmp_path / "bands.s"
    subprocess.run([HIPCC, *DEVICE_FLAGS, *NO_PACKED, "-o", str(out),
                    os.path.join(ROOT, "total-viewsheds_amd", "csrc", "bands.hip")],
                   check=True, capture_output=True)
    text = out.read_text()
    assert not re.search(r"\bv_pk_(add|mul|fma)_f32\b", text)
    sizes = re.findall(r"\.private_segment_fixed_size:\s+(\d+)", text)
    assert sizes and all(int(x) == 0 for x in sizes)


def test_packed_fp32_flag_is_effective(tmp_path):
    """The flag the library is built with takes effect on the device pass:
    a float2 add/multiply kernel uses v_pk_*_f32 without it and none with
    it (so the absence checks above are not vacuous)."""
    src = tmp_path / "pk.hip"
    src.write_text(
        "#include <hip/hip_runtime.h>\n"
        "__global__ void k(float2 *a, const float2 *b, const float2 *c, int n) {\n"
        "  int i = blockIdx.x * blockDim.x + threadIdx.x;\n"
        "  if (i < n) { float2 x = b[i], y = c[i];\n"
        "    a[i] = make_float2(x.x + y.x, x.y + y.y);\n"
        "    a[i + n] = make_float2(x.x * y.x, x.y * y.y); }\n"
        "}\n")
    counts = []
    for extra in ([], NO_PACKED):
        out = tmp_path / f"pk{len(extra)}.s"
        subprocess.run([HIPCC, *DEVICE_FLAGS, *extra, "-o", str(out), str(src)], check=True,
                       capture_output=True)
        counts.append(len(re.findall(r"\bv_pk_(?:add|mul|fma)_f32\b", out.read_text())))
    assert counts[0] > 0 and counts[1] == 0, counts
