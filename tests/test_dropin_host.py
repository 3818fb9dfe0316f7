"""The drop-in layer's host side without a device: its parameter parsing and
completion (GetParameters, CompleteParameters) and the reference's parameter
summary (Print_Info_Console, Print_Info_File), byte for byte against the
reference's own output for the same command line (fixture from
oracle/_ref/ref_harness --info).  Also runs where no GPU is present."""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "phinum_amd")
INFO = os.path.join(HERE, "golden", "info", "driver_info.json")
CASES = os.path.join(HERE, "golden", "info", "parse_cases.json")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(os.path.join(PKG, "libphinum_dropin.so")):
        pytest.skip("libphinum_dropin.so is not built")
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_info_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", ROOT, "-I", os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(HERE, "dropin_info_check.cpp"), "-L", PKG, "-lphinum_dropin", "-lphinum_gpu",
                    "-Wl,-rpath," + PKG], check=True)
    return exe


def test_parameter_summary_matches_reference(checker, tmp_path):
    g = json.load(open(INFO))
    r = subprocess.run([checker] + g["args"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == g["console"]
    assert (tmp_path / "output" / "0" / "info.txt").read_text() == g["info_txt"]


def test_parameter_summary_sees_each_parameter(checker, tmp_path):
    # a changed time step, lattice size and seed field: the summary follows
    # (dt = a_t * dtBaseTrans on the console, dtBT and dtBL in the file)
    g = json.load(open(INFO))
    args = list(g["args"])
    args[args.index("-dt") + 1] = "0.2"
    args[args.index("-Nt") + 1] = "24"
    args[args.index("-phi0_0") + 1] = "1.5"
    r = subprocess.run([checker] + args, cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "N_s=24\n" in r.stdout and "dt=0.1\n" in r.stdout and "phi0_0=1.5\n" in r.stdout
    info = (tmp_path / "output" / "0" / "info.txt").read_text()
    assert "dtBT=0.2\ndtBL=0.2\n" in info and r.stdout != g["console"]


@pytest.mark.parametrize("name", sorted(json.load(open(CASES))))
def test_parameter_parsing_edge_cases_match_reference(checker, tmp_path, name):
    # the reference's own summary for command lines with a duplicate key, an
    # unknown key, a trailing key without a value, exponent notation, a
    # negative value first, an over-long key, a fraction for an integer, and
    # no arguments (fixture from oracle/_ref/ref_harness --info)
    c = json.load(open(CASES))[name]
    r = subprocess.run([checker] + c["args"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == c["console"]
    assert (tmp_path / "output" / "0" / "info.txt").read_text() == c["info_txt"]
