"""The device kernels' fast division by run-time constants, checked on the
host against integer division (exhaustive for small divisors, sampled up to
2^31 and around multiples for the brick/stride divisors that occur)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fastdiv_matches_integer_division(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.check_call(["hipcc", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "macaw_amd", "csrc"),
                           os.path.join(ROOT, "tests", "native", "fastdiv_check.cpp"), "-o", exe],
                          stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "fastdiv ok" in out.stdout
