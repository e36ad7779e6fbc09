"""Host bounds check of the coordinate-derived index arithmetic of the
device kernels (cell lists, cell tags, Monte Carlo cells, erfcx pieces) for
non-finite and huge coordinates, compiled from the kernels' own headers and
run on the CPU (tests/native/index_bounds.cpp).  A simulation that blows up
must keep every kernel in bounds; the MD run then fails with LG_E_NONFINITE."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_index_math_in_bounds_for_any_coordinate(tmp_path):
    exe = str(tmp_path / "index_bounds")
    src = os.path.join(HERE, "native", "index_bounds.cpp")
    # the functions under test are __host__ __device__; main() runs on the CPU
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", src, "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
