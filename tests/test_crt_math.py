"""The integer arithmetic of the int8 CRT form (libxm_amd/csrc/crt_math.h),
compiled for the host and checked against exact 128-bit integers: the
modular reduction over the whole int32 range, the limb split and residues
of integers up to 2^50, and Garner reconstructions of int32 residue-product
sums (what the int8 matrix cores compute per modulus)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_crt_arithmetic(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "crtcheck")
    subprocess.check_call([gxx, "-O2", "-std=c++17",
                           "-I" + os.path.join(ROOT, "libxm_amd", "csrc"),
                           "-o", exe,
                           os.path.join(ROOT, "tests", "crt_math_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK")
