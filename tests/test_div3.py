"""The message kernel's division by 3 (hbp_common.h div3_nonneg) equals x / 3 for every
non-negative float: exhaustive, on the host (the function is __host__ __device__)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc (HIP headers) not available")
def test_div3_nonneg_exhaustive():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "div3_check")
        subprocess.run([hipcc, "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "hbp-isp_amd", "csrc"),
                        os.path.join(ROOT, "tests", "div3_check.cpp"), "-o", exe], check=True,
                       capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr

