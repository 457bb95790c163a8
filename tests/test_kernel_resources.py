"""Register budgets the Newton-set schedule and the K + jwM kernel rely on, read
from the built library's gfx950 code objects (AMDGPU metadata, .vgpr_count).

A SIMD holds 512 VGPRs per lane, allocated in granules of 8.
- FP32-local Newton set: the listed-tile pass on the side stream (one fused
  launch by default, two launches as the alternative) runs beside the
  one-plane pass only where a SIMD has room for one of its waves next to two
  one-plane waves. A build in which the one-plane pass took 168 VGPRs
  instead of 158 (same occupancy) starved the side passes until the one-plane
  grid drained: 0.655 -> 0.72 ms on the 4M-tet set.
- FP64 Newton set: the three-plane side pass fits beside one one-plane wave.
- FP32-local K + jwM: 3 waves per SIMD need at most 168 VGPRs (165 today;
  172 ran 1.238 -> 1.293 ms on the 128^3 box).
No GPU is needed: the code objects are unbundled from the .so with the ROCm
LLVM tools."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "magpackage_amd" / "libmpk_assembly.so"
LLVM = Path("/opt/rocm/llvm/bin")
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def _vgpr_counts(tmp: Path) -> dict:
    subprocess.run([str(LLVM / "llvm-objcopy"), f"--dump-section=.hip_fatbin={tmp / 'fb.bin'}",
                    str(SO), str(tmp / "copy.so")], check=True, capture_output=True)
    blob = (tmp / "fb.bin").read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(b"__CLANG_OFFLOAD_BUNDLE__"), blob)]
    counts = {}
    for i, s in enumerate(starts):
        e = starts[i + 1] if i + 1 < len(starts) else len(blob)
        (tmp / f"b{i}.bin").write_bytes(blob[s:e])
        subprocess.run([str(LLVM / "clang-offload-bundler"), "--unbundle", "--type=o",
                        f"--targets={TARGET}", f"--input={tmp / f'b{i}.bin'}",
                        f"--output={tmp / f'co{i}.o'}"], check=True, capture_output=True)
        notes = subprocess.run([str(LLVM / "llvm-readelf"), "--notes", str(tmp / f"co{i}.o")],
                               check=True, capture_output=True, text=True).stdout
        name = None
        for line in notes.splitlines():
            m = re.match(r"\s*\.name:\s+(\S+)", line)  # the kernel's .name follows its .args
            if m:
                name = m.group(1)
            m = re.match(r"\s*\.vgpr_count:\s+(\d+)", line)
            if m and name:
                counts[name] = int(m.group(1))
    return counts


@pytest.fixture(scope="module")
def vgprs(tmp_path_factory):
    if not SO.exists():
        pytest.fail(f"{SO} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    if not (LLVM / "clang-offload-bundler").exists():
        pytest.skip("ROCm LLVM tools not installed")
    counts = _vgpr_counts(tmp_path_factory.mktemp("co"))
    assert len(counts) > 100
    return counts


def _one(counts, *parts):
    hits = [v for k, v in counts.items() if all(p in k for p in parts)]
    assert len(hits) == 1, (parts, len(hits))
    return hits[0]


def _alloc(v):
    return (v + 7) // 8 * 8


def test_fp32_newton_side_passes_fit_beside_two_one_plane_waves(vgprs):
    lin = _one(vgprs, "k_csc_ring_batch", "FrameNewtonLinILi2E", "Li6ELb1E")
    kd = _one(vgprs, "k_csc_ring_list", "FrameNewtonILi2ELb0E", "Li6ELb1E")
    mass = _one(vgprs, "k_csc_ring_list", "FrameMassILi2EEELi3ELb1E")
    assert 2 * _alloc(lin) + _alloc(kd) <= 512, (lin, kd)
    assert 2 * _alloc(lin) + _alloc(mass) <= 512, (lin, mass)


def test_fp32_newton_fused_side_pass_fits_beside_two_one_plane_waves(vgprs):
    # the default listed-tile launch: the two-plane op (4 incidences per
    # gather round) and the mass op in one kernel
    lin = _one(vgprs, "k_csc_ring_batch", "FrameNewtonLinILi2E", "Li6ELb1E")
    fused = _one(vgprs, "k_csc_ring_list2", "FrameNewtonILi2ELb0E", "Li4ELi3ELb1E")
    assert 2 * _alloc(lin) + _alloc(fused) <= 512, (lin, fused)


def test_fp64_newton_side_pass_fits_beside_one_one_plane_wave(vgprs):
    lin = _one(vgprs, "k_csc_ring_batch", "FrameNewtonLinILi1E", "Li6ELb1E")
    kmd = _one(vgprs, "k_csc_ring_list", "FrameNewtonILi1ELb1E", "Li6ELb1E")
    assert _alloc(lin) + _alloc(kmd) <= 512, (lin, kmd)


def test_fp32_local_k_jwm_keeps_three_waves_per_simd(vgprs):
    km = _one(vgprs, "k_csc_ring_batch", "FrameKMILi2EEELi6ELb1E")
    assert _alloc(km) <= 168, km
