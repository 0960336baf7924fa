"""The device code's matrix instructions, checked in the gfx950 assembly that
hipcc makes from each HIP source with the library's own flags (CPU only: hipcc
cross-compiles).

An MFMA's destination registers must either be exactly its C (accumulator)
operand or not overlap it: the matrix core reads C and writes the
destination over several passes.  With VGPR-form accumulators
(-amdgpu-mfma-vgpr-form) the register coalescer has produced a destination
that overlaps C by two registers (v[18:21] from C = v[20:23]) for
accumulators copied at a loop exit, and the kernel's results were wrong.
This test rejects any such instruction in every MFMA kernel of the library.
"""
import os
import re
import shutil
import subprocess

import pytest

from kalnajs_log_spirals_amd import build_lib

HIPCC = build_lib.HIPCC
MFMA = re.compile(r"^\s*(v_mfma\S*)\s+([av])\[(\d+):(\d+)\],\s*[^,]+,\s*[^,]+,\s*([av])\[(\d+):(\d+)\]",
                  re.M)
KERNEL = re.compile(r"^(_Z\S*|k_\S*):\s*(?:;.*)?$", re.M)


def device_assembly(src: str, extra, tmp_path) -> str:
    out = tmp_path / (os.path.basename(src) + ".s")
    cmd = ([HIPCC] + build_lib.COMMON + build_lib.DEVICE + ["-x", "hip"] + list(extra)
           + ["--cuda-device-only", "-S", "-o", str(out), src])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return out.read_text()


def partial_overlaps(asm: str):
    """(kernel, instruction) for every MFMA whose destination overlaps its C
    operand without being equal to it."""
    bad = []
    starts = [(m.start(), m.group(1)) for m in KERNEL.finditer(asm)]
    for k, (pos, name) in enumerate(starts):
        end = starts[k + 1][0] if k + 1 < len(starts) else len(asm)
        for m in MFMA.finditer(asm, pos, end):
            dfile, d0, d1 = m.group(2), int(m.group(3)), int(m.group(4))
            cfile, c0, c1 = m.group(5), int(m.group(6)), int(m.group(7))
            if dfile != cfile or (d0, d1) == (c0, c1):
                continue
            if not (d1 < c0 or d0 > c1):
                bad.append((name, m.group(0).strip()))
    return bad


def test_overlap_detector_on_known_forms():
    asm = ("_Z1kv:                                   ; @k\n"
           "\tv_mfma_f32_16x16x32_bf16 v[18:21], v[82:85], v[24:27], v[20:23]\n"
           "\tv_mfma_f32_16x16x32_bf16 v[20:23], v[92:95], v[28:31], v[20:23]\n"
           "\tv_mfma_f64_16x16x4_f64 a[0:7], v[0:1], v[2:3], a[0:7]\n"
           "\tv_mfma_f32_16x16x32_bf16 v[0:3], v[4:7], v[8:11], v[12:15]\n")
    bad = partial_overlaps(asm)
    assert len(bad) == 1 and "v[18:21]" in bad[0][1]


MFMA_SOURCES = [(s, extra) for s, is_hip, extra in build_lib.SOURCES
                if is_hip and s.startswith("klsp_contract")] + [("klsp_device.hip", [])]


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None,
                    reason="hipcc not installed")
@pytest.mark.parametrize("src,extra", MFMA_SOURCES, ids=[s for s, _ in MFMA_SOURCES])
def test_no_mfma_destination_partly_overlaps_its_accumulator(src, extra, tmp_path):
    asm = device_assembly(os.path.join(build_lib.CSRC, src), extra, tmp_path)
    n = len(MFMA.findall(asm))
    assert n > 0, f"{src}: no MFMA found (the pattern no longer matches the assembly)"
    bad = partial_overlaps(asm)
    assert not bad, f"{src}: {len(bad)} MFMAs with a destination partly overlapping C: {bad[:3]}"
