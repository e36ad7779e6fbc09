"""Build the gfx950 shared library ``liblumol_gpu.so`` in-tree with hipcc.

The library holds the HIP kernels and the C ABI of ``include/lumol_gpu.h``.
It is built here (hipcc cross-compiles gfx950 without a GPU) and shipped
as a file next to this module.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liblumol_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", "lg_api.hip")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f) for f in ("lg_device.hpp", "lg_kernels.hpp")] + [
    os.path.join(ROOT, "include", "lumol_gpu.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Werror=misleading-indentation",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose=False):
    """Compile liblumol_gpu.so (always recompiles)."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    tmp = LIB + ".tmp"
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc"] + FLAGS + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and res.stderr.strip():
        # warnings (the host pass reports the unknown packed-fp32 feature)
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines[:40]))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
