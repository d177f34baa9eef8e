"""Checks on the gfx950 code objects inside the built library (no GPU needed).

Every kernel must run from registers and LDS alone: no private (scratch)
segment and no VGPR spills (SGPR spills land in VGPR lanes, not memory).  A
value placed in scratch inside a pixel loop costs a memory round trip at each
use; a conditional that picks either a local struct or a table reference is
enough to make the compiler put the local there.
"""
import os
import re
import subprocess

import pytest

from fractura_amd import build as hip_build

LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def _kernel_metadata(tmp_path):
    lib = hip_build.LIB
    if not os.path.exists(lib):
        pytest.fail("libapeiron_hip.so is not built")
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf"):
        if not os.path.exists(os.path.join(LLVM, tool)):
            pytest.skip(f"{tool} not available")
    fat = tmp_path / "fat.bin"
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fat}", lib,
                    os.devnull], check=True, capture_output=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)] + [len(data)]
    kernels = {}
    for i in range(len(starts) - 1):
        part = tmp_path / f"bundle{i}.bin"
        part.write_bytes(data[starts[i]:starts[i + 1]])
        co = tmp_path / f"co{i}.elf"
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                        f"--targets={TARGET}", f"--input={part}", f"--output={co}"],
                       check=True, capture_output=True)
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(co)],
                               check=True, capture_output=True, text=True).stdout
        for blk in notes.split(".agpr_count:")[1:]:
            name = re.search(r"\.name:\s+(\S+)", blk)
            if not name:
                continue
            fields = {}
            for key in ("private_segment_fixed_size", "vgpr_spill_count", "sgpr_spill_count",
                        "vgpr_count"):
                m = re.search(rf"\.{key}:\s+(\d+)", blk)
                fields[key] = int(m.group(1)) if m else 0
            kernels[name.group(1)] = fields
    return kernels


def test_no_kernel_uses_scratch_or_spills_vgprs(tmp_path):
    kernels = _kernel_metadata(tmp_path)
    hot = [k for k in kernels if "tile_pass" in k or "queue_pass" in k]
    assert len(hot) >= 40, f"expected the pass kernels of every engine, found {len(hot)}"
    assert any("bla_low_kernel" in k for k in kernels)
    bad = {k: f for k, f in kernels.items() if f["private_segment_fixed_size"] or f["vgpr_spill_count"]}
    assert not bad, f"kernels with scratch or spills: {bad}"


def test_fp32_centre_pass_keeps_six_waves(tmp_path):
    # At most 80 VGPRs leaves room for 6 waves per SIMD, enough to hide the
    # latency of the FP32 pass's table loads.
    kernels = _kernel_metadata(tmp_path)
    fexp32 = {k: f["vgpr_count"] for k, f in kernels.items()
              if "tile_pass" in k and "EngFexp32ILi0E" in k and "Lb0E" in k}
    assert fexp32, "FP32 centre pass kernel not found"
    assert max(fexp32.values()) <= 80, fexp32
