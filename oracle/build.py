"""Build recipe of the oracle (test infrastructure only).

``liblumol_oracle.so`` is the CPU FP64 restatement of the reference's
algorithm in ``lumol_oracle.cpp``.  It is compiled with -ffp-contract=off so
that, like the reference's Rust, no multiply-add is fused.

The reference itself (Rust, a Cargo workspace with crates.io dependencies
and no lockfile) cannot be built here: there is no cargo/rustc and no
network, so no ``oracle/_ref`` build exists.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblumol_oracle.so")
SRC = os.path.join(HERE, "lumol_oracle.cpp")

FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared", "-pthread",
         "-Wall", "-Wextra"]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [os.environ.get("CXX", "g++")] + FLAGS + [SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
