"""Build ``libzosimos_hip.so`` in-tree with hipcc for gfx950.

Usage: ``python -m zosimos_amd.build`` (or ``build()`` from Python).  Object
files go to ``zosimos_amd/_build/``; the shared library lands next to this
file so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libzosimos_hip.so")

SOURCES = [
    "zh_kernels.hip",
    "zh_spec.hip",
    "zh_warp.hip",
    "zh_mlp.hip",
    "zh_runtime.cpp",
    "zh_planner.cpp",
    "zh_command.cpp",
    "zh_surrogate.cpp",
    "zh_pool.cpp",
    "zh_dynamic.cpp",
    "zh_jit.cpp",
]
HEADERS = ["zh_stage.h", "zh_codec.h", "zh_device.h", "zh_spec.h", "zh_prog.h", "zh_spec_kernels.h",
           "zh_warp_kernels.h", "zh_internal.h",
           "zh_mlp.h"]
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is disabled for every device compile.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
          f"--offload-arch={ARCH}"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# Per-source device flags.  The MLP kernels keep their MFMA accumulators in
# ordinary VGPRs (no AGPR round trip: v_accvgpr_read per result element).
SOURCE_FLAGS = {"zh_mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, "-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd += DEVICE_FLAGS + SOURCE_FLAGS.get(src, [])
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


SHADER_DIR = os.path.join(HERE, "shaders")


def build_shaders() -> list:
    """Compile each user stage kernel under ``shaders/`` into a gfx950 code
    object next to its source (``hipcc --genco``), the form
    ``zh_shader_create`` takes."""
    out = []
    inc_dir = os.path.join(os.path.dirname(HERE), "include")
    for f in sorted(os.listdir(SHADER_DIR)):
        if not f.endswith(".hip"):
            continue
        co = os.path.join(SHADER_DIR, os.path.splitext(f)[0] + ".co")
        # Contraction as fast as usual, but a shader's own
        # `#pragma clang fp contract(off)` is honoured.
        cmd = [hipcc(), "--genco", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=fast-honor-pragmas",
               *DEVICE_FLAGS, "-I", inc_dir,
               os.path.join(SHADER_DIR, f), "-o", co]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc --genco failed for {f}:\n{res.stderr}")
        out.append(co)
    return out


def build(force: bool = True, jobs: int = 6) -> str:
    """Compile every source (in parallel) and link the shared library."""
    build_shaders()
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    hdr_mtime = max(os.path.getmtime(os.path.join(CSRC, h)) for h in HEADERS if os.path.exists(os.path.join(CSRC, h)))
    inc_dir = os.path.join(os.path.dirname(HERE), "include")
    hdr_mtime = max([hdr_mtime] + [os.path.getmtime(os.path.join(inc_dir, f)) for f in os.listdir(inc_dir)])

    def stale(s: str) -> bool:
        obj = os.path.join(OUT_DIR, os.path.splitext(s)[0] + ".o")
        if force or not os.path.exists(obj):
            return True
        return os.path.getmtime(obj) < max(os.path.getmtime(os.path.join(CSRC, s)), hdr_mtime)

    todo = [s for s in srcs if stale(s)]
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OUT_DIR, os.path.splitext(s)[0] + ".o") for s in srcs]
    tmp = LIB + ".tmp"
    # hiprtc compiles the run-time kernels of unregistered signatures (zh_jit.cpp).
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    force = "--incremental" not in sys.argv
    print(build(force=force))
