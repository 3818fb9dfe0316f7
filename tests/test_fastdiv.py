"""Correct rounding of the FP64 division by a run-constant divisor used in the
step kernels (phinum_amd/csrc/pn_fastdiv.h), checked bit for bit against IEEE
division on the host: random dividends over the whole fast range, dividends
built to put the quotient next to a rounding midpoint, range ends, zero."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fast_division_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "fastdiv_check.cpp")], check=True)
    r = subprocess.run([exe, "200000", "5"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().startswith("mismatches 0 checked")
    assert int(r.stdout.split()[-1]) > 50_000_000
