"""Build libklsp.so in-tree with hipcc for gfx950.

    python -m kalnajs_log_spirals_amd.build_lib

The shared library lands in kalnajs_log_spirals_amd/lib/ (git-ignored, but it
travels with the source tree).  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libklsp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no implicit FMA contraction, so scalar FP64 code rounds
# like the reference; MFMA builtins are unaffected.  Packed FP32 ops are
# disabled for all device code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
    "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# (source, is HIP device code, extra flags).  The contractions with at most
# ~8 accumulator blocks per wave keep them in VGPRs (-amdgpu-mfma-vgpr-form):
# with the default AGPR form the compiler shuffles some of them between the
# two register files every loop trip, and those moves take issue slots from
# the MFMAs.  The one-omega FP64 contraction (16 complex blocks) and the
# symmetric-basis FP64 contraction (32 real blocks: P and Q of a 4 x 4 tile)
# fill the AGPRs and keep the default form.
SOURCES = [
    ("klsp_host.cpp", False, []),
    ("klsp_device.hip", True, []),
    ("klsp_contract_f64.hip", True, []),
    ("klsp_contract_sym64.hip", True, []),
    ("klsp_contract_f64_batch.hip", True, ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("klsp_contract_f32.hip", True, ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("klsp_contract_bf16.hip", True, ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("klsp_contract_symbf16.hip", True, ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("klsp_contract_sym32.hip", True, ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("klsp_lu.hip", True, []),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    """Compile every source (in parallel, MAX_JOBS or up to 8 at once) and link."""
    os.makedirs(LIBDIR, exist_ok=True)
    objs, cmds = [], []
    for src, is_hip, extra in SOURCES:
        obj = os.path.join(LIBDIR, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if is_hip else ["-x", "c++"]) + extra
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
