"""Builds the native libraries in-tree with hipcc for gfx950.

    python -m tilestream_amd.build

produces tilestream_amd/_native/libtilestream_amd.so (host API, C-ABI and HIP
kernels) and tilestream_amd/_native/libtsgen.so (seeded workload generators).
Every HIP source is compiled with -ffp-contract=off (calibration rounds twice),
without fast-math or denormal flushing, and with packed FP32 instructions
disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tilestream_amd")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
OUT = os.path.join(PKG, "_native")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host")]
HIP_FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-fno-gpu-rdc",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# Host-only sources: plain C++ against the HIP runtime API, no device pass.
HOST_FLAGS = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-Wall", "-Wextra"]

# The tile decoder's kernels: one source per kind, compiled once per lane-group
# width (the object name carries the width).
DECODE_KINDS = ["mixed", "u16", "f32", "u16_cal", "f32_cal"]
GROUPS = [1, 2, 4, 8, 16]
HIP_SOURCES = ([f"hip/decode_{k}.hip:{g}" for k in DECODE_KINDS for g in GROUPS] +
               ["hip/decode_block.hip", "hip/index.hip", "hip/decode.hip", "hip/encode.hip", "hip/calibrate.hip"])
HOST_SOURCES = ["host/sizing.cpp", "host/parser.cpp", "host/cpu_codec.cpp", "host/tiff.cpp", "host/plan.cpp",
                "host/queue.cpp", "host/serve.cpp", "host/capi.cpp"]


def _compile(entry: str) -> str:
    src, _, group = entry.partition(":")
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src.replace("/", "_") + (f"_g{group}" if group else "") + ".o")
    flags = COMMON + (HIP_FLAGS if src.endswith(".hip") else HOST_FLAGS)
    if group:
        flags = flags + [f"-DTSK_DECODE_GROUP={group}"]
    if src.startswith("hip/decode_"):
        # The work queue's look-ahead draw is one lane's atomic whose result is
        # awaited later: keep it a plain returning atomic, not a wave-combined
        # one that waits for its result (and every earlier store) at once.
        flags = flags + ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]
    cmd = [HIPCC] + flags + ["-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (res.stderr or "").splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> None:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(OUT, exist_ok=True)
    sources = HIP_SOURCES + HOST_SOURCES
    workers = min(len(sources), max(1, min(os.cpu_count() or 4, 16)))
    with cf.ThreadPoolExecutor(workers) as pool:
        objs = list(pool.map(_compile, sources))
    lib = os.path.join(OUT, "libtilestream_amd.so")
    tmp = lib + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lz"],
                   check=True)
    os.replace(tmp, lib)
    gen = os.path.join(OUT, "libtsgen.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", gen + ".tmp",
                    os.path.join(CSRC, "gen", "generate.cpp")], check=True)
    os.replace(gen + ".tmp", gen)
    if verbose:
        print(f"built {lib} and {gen}")


if __name__ == "__main__":
    build()
