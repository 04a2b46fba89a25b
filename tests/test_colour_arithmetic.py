"""The product's colour arithmetic (pixel.hpp: integer luma quotient and the
three-lane similarity subtract) against the oracle on the host: every colour,
every single-field value pair, 50 M colour pairs and the class bits.  A C++
program built with hipcc from the package's header, linked with the oracle."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "pixel-art-remaster-gpu_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.timeout(600)
def test_colour_arithmetic_equals_oracle(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    subprocess.run(["make", "-C", ORACLE], check=True, capture_output=True)
    exe = tmp_path / "colour_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, "-I", ORACLE,
                    os.path.join(HERE, "native", "colour_check.cpp"), "-L", ORACLE, "-loracle",
                    f"-Wl,-rpath,{ORACLE}", "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=500)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout.splitlines()[-1]
