"""Builds libraym0nade_hip.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with
  -ffp-contract=off                    bit-exact fp32 traversal (no FMA contraction)
  -fno-gpu-flush-denormals-to-zero     IEEE fp32 like the host
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed fp32 instructions
and HIP's default correctly rounded fp32 division and square root.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libraym0nade_hip.so")
CLI = os.path.join(HERE, "raym0nade_cli")
OBJDIR = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

SOURCES = ["api.cpp", "scene.cpp", "scenegen.cpp", "assets.cpp", "fbx.cpp", "frame.cpp", "trace.hip", "render.hip", "post.hip", "dof.hip", "denoise.hip", "adjoint.hip"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
DEVICE = [
    "--offload-arch=gfx950", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    cmds = []
    for s in srcs:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJDIR, s + ".o")
        objs.append(obj)
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip" if s.endswith(".hip") else "c++"]
        if s.endswith(".cpp"):
            # host-only translation units still go through hipcc for the HIP headers
            cmd = [HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-x", "c++"]
        cmd += ["-c", src, "-o", obj]
        cmds.append(cmd)
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lpthread", "-lz"])
    os.replace(tmp, LIB)
    # the console host (reference CLI grammar) linked against the library
    tmp = CLI + ".tmp"
    _run([HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-x", "c++",
                              os.path.join(CSRC, "cli.cpp"), "-x", "none", "-o", tmp, "-L" + HERE,
                              "-lraym0nade_hip", "-Wl,-rpath,$ORIGIN", "-lz"])
    os.replace(tmp, CLI)
    if verbose:
        print("built", LIB, CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
