"""Register and scratch use of the built kernels (CPU only: reads the code
object metadata of libss.so's gfx950 bundles with llvm-readelf).

A kernel that keeps values in scratch memory (per-lane private memory in
HBM) pays a load and a store through the memory system for each such value;
in the particle updates that is tens of bytes per particle on top of the 64
the update is measured against.  The persistent ensemble update once copied
its whole job structure (about 540 bytes) into scratch this way, which made
it several times slower without changing a result, so no parity test could
see it.
"""
import os
import re
import struct
import subprocess

import pytest

from stochasticsampling_amd import build as b

READELF = "/opt/rocm/llvm/bin/llvm-readelf"
OBJCOPY = "/opt/rocm/llvm/bin/llvm-objcopy"
BUNDLE_MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"

# Kernels allowed to use scratch, with the reason: neither is on a default
# path.  k_fft_axis<T, true> is the axis pass with the generic radix stage
# (axis lengths with prime factors above the built-in radices; its
# run-time-length butterfly array lives in scratch); k_spec_pencils2 is the
# opt-in paired spectral stage.
ALLOWED = ("k_fft_axis", "k_spec_pencils2")


def _code_objects(lib, tmp_path):
    fat = tmp_path / "fatbin.bin"
    subprocess.run([OBJCOPY, f"--dump-section=.hip_fatbin={fat}", lib, str(tmp_path / "lib.copy")], check=True,
                   capture_output=True)
    data = fat.read_bytes()
    out = []
    pos = 0
    while True:
        p = data.find(BUNDLE_MAGIC, pos)
        if p < 0:
            break
        (count,) = struct.unpack_from("<Q", data, p + 24)
        off = p + 32
        for _ in range(count):
            o, size, tlen = struct.unpack_from("<QQQ", data, off)
            off += 24
            triple = data[off:off + tlen].decode()
            off += tlen
            if "gfx950" in triple and size > 0:
                path = tmp_path / f"co{len(out)}.elf"
                path.write_bytes(data[p + o:p + o + size])
                out.append(path)
        pos = p + 1
    return out


def _kernels(lib, tmp_path):
    """name -> metadata fields of every kernel in the library."""
    res = {}
    for co in _code_objects(lib, tmp_path):
        notes = subprocess.run([READELF, "--notes", str(co)], check=True, capture_output=True, text=True).stdout
        for block in notes.split("- .agpr_count")[1:]:
            name = re.search(r"\.name:\s+(\S+)", block)
            if not name:
                continue
            fields = {}
            for key in ("private_segment_fixed_size", "vgpr_count", "sgpr_spill_count", "vgpr_spill_count", "group_segment_fixed_size", "agpr_count", "sgpr_count"):
                m = re.search(r"\." + key + r":\s+(\d+)", block)
                fields[key] = int(m.group(1)) if m else None
            res[name.group(1)] = fields
    return res


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    if not (os.path.exists(READELF) and os.path.exists(OBJCOPY)):
        pytest.skip("llvm-readelf / llvm-objcopy not installed")
    if not os.path.exists(b.LIB):
        pytest.skip("libss.so not built")
    return _kernels(b.LIB, tmp_path_factory.mktemp("co"))


def test_metadata_read(kernels):
    assert any("k_tile_update_batched" in k for k in kernels)
    assert any("k_tile_update" in k and "batched" not in k for k in kernels)
    assert all(v["private_segment_fixed_size"] is not None for v in kernels.values())


def test_no_scratch_outside_allowed_kernels(kernels):
    bad = sorted(k for k, v in kernels.items()
                 if v["private_segment_fixed_size"] and not any(a in k for a in ALLOWED))
    assert not bad, f"kernels using scratch memory: {bad[:8]}"


def test_fp32_updates_fit_four_waves(kernels):
    """The default-physics FP32 updates (single system and ensemble) stay
    within 128 VGPRs, i.e. four waves per SIMD, with no vector spills."""
    for name, v in kernels.items():
        if re.search(r"k_tile_updateIfLi6ELb1ELi0ELb0ELb0ELb0EE|k_tile_update_batchedIfLi6ELb1ELi0ELb0ELb0EE", name):
            assert v["vgpr_count"] <= 128, (name, v)
            assert v["vgpr_spill_count"] == 0, (name, v)


def test_fp64_tiled_update_fits_two_waves(kernels):
    """The default-physics FP64 tiled update stays within 256 VGPRs (two
    waves per SIMD, which its LDS window also allows).  Inlining a
    straight-line binary64 atan2 for the two orientation angles once took it
    to 280 and one wave, and the update from 0.77 to 1.24 ms at 2^24
    particles."""
    found = False
    for name, v in kernels.items():
        if re.search(r"k_tile_updateIdLi6ELb1ELi0ELb0ELb[01]ELb0EE", name):
            found = True
            assert v["vgpr_count"] <= 256, (name, v)
            assert v["private_segment_fixed_size"] == 0, (name, v)
    assert found
