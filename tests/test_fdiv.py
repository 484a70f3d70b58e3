"""div_r, the shared-reciprocal division of the physics kernels, against IEEE
division bit for bit on the host (the device evaluates the same IEEE
multiplications and fused multiply-adds)."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_div_r_is_correctly_rounded():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "fdiv_check")
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-ffp-contract=off", "-o", exe,
                        os.path.join(HERE, "fdiv_check.hip")], check=True, capture_output=True, cwd=d)
        out = subprocess.run([exe, "20000000"], check=False, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "checked 20000000 mismatches 0" in out.stdout
