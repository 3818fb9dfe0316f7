"""Correct rounding of the FP64 division by a run-constant divisor used in the
step kernels (phinum_amd/csrc/pn_fastdiv.h), four- and two-operation forms,
checked bit for bit against IEEE division on the host: random dividends over
the whole fast range, dividends built to put the quotient next to a rounding
midpoint, range ends, zero; and the divisor test that admits the two-operation
form, on divisors where it must fail."""
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
    out = r.stdout.split()
    vals = dict(zip(out[0::2], (int(v) for v in out[1::2])))
    assert vals["mismatches"] == 0, r.stdout
    assert vals["checked"] > 50_000_000
    # 6, 120 and most random divisors take the two-operation form; the two
    # planted counterexamples are rejected
    assert vals["bmr_divisors"] >= 30 and vals["bmr_rejected"] >= 2
