"""The reference-named host layer and driver (phinum_main, main.cpp:21-117)
against the oracle, plus SetInitialConditions on the device."""
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu
MAIN = os.path.join(os.path.dirname(pn.__file__), "phinum_main")
INFO = os.path.join(os.path.dirname(__file__), "golden", "info", "driver_info.json")


@pytest.mark.parametrize("N,seed,prec", [(16, 1234, pn.PN_FP64), (32, 7, pn.PN_FP64),
                                         (16, 3, pn.PN_FP32)])
def test_initial_conditions_match_oracle(N, seed, prec):
    m = O.model(N)
    phi0 = float(np.sqrt(6.0))
    ophi, opi = O.initial_conditions(m, seed, phi0)
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"],
                    precision=prec) as L:
        L.initial_conditions(pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=phi0))
        gphi, gpi = L.get_fields()
    # rocFFT and numpy's c2r differ in rounding only
    tol = 1e-12 if prec == pn.PN_FP64 else 2e-7
    assert np.max(np.abs(gphi - ophi)) <= tol * np.max(np.abs(ophi))
    assert np.max(np.abs(gpi - opi)) <= tol * np.max(np.abs(opi))
    assert np.std(gphi) > 0.01  # fluctuations are there


def _cadence(T0, TMax, dtbase, a, TSO):
    """Step indices and LastSOoutput values at which Print_AveragesExtrema writes
    (fieldanalysis.cpp:753-780) inside for(reset(); tau<TMax; evolve())."""
    dtau = dtbase * a
    tau = T0
    last = None
    out = []
    step = 0
    while tau < TMax:
        if step == 0:
            last = tau
            out.append((step, last))
        elif not tau < last + TSO:
            if tau >= last + TSO:
                last = last + TSO
            out.append((step, last))
        tau = tau + 0.5 * dtau
        tau = tau + 0.5 * dtau
        step += 1
    return out


def test_driver_matches_oracle(tmp_path):
    N, seed = 16, 1234
    m = O.model(N)
    phi0 = float(np.sqrt(6.0))
    args = [MAIN, "-runs", "1", "-Nt", str(N), "-at", "0.5", "-M", "1", "-Mmode", "1", "-L", "6",
            "-Lmode", "-1", "-Kappa", "%.17g" % m["kappa"], "-T0", "0", "-TMax", "2.0", "-dt", "0.1",
            "-TSO", "0.5", "-TMO", "1.0", "-Noise", "0.5", "-Ampl", "0", "-QS", "%.17g" % (np.sqrt(3.0) / 0.5),
            "-Q2", "%.17g" % (np.sqrt(3.0) / 0.5), "-phi0_0", "%.17g" % phi0, "-seed", str(seed), "-islands", "1"]
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    so = np.loadtxt(tmp_path / "output" / "1" / "SO-1.txt")
    cad = _cadence(0.0, 2.0, 0.1, 0.5, 0.5)
    assert so.shape[0] == len(cad)
    m2 = dict(m)
    ophi, opi = O.initial_conditions(m2, seed, phi0)
    prev = 0
    for (step, last), row in zip(cad, so):
        if step > prev:
            ophi, opi = _cont(m, ophi, opi, prev, step)
            prev = step
        avg = O.volume_averages(m, ophi, opi)
        vals, _ = O.extrema(N, ophi)
        ref = np.concatenate([[last], avg, vals])
        np.testing.assert_allclose(row, ref, rtol=2e-5, atol=1e-9)
    dist = sorted(os.listdir(tmp_path / "output" / "1"))
    assert any(d.startswith("Distrib-1-") for d in dist)
    # the spectrum file at tau = 0 (Print_fp, momentumspace.cpp:307-381) from
    # the initial fields: the oracle's rows, to the 6 digits the file keeps
    phi_ic, pi_ic = O.initial_conditions(m2, seed, phi0)
    F, K, _ = O.spectrum(N, 0.5, phi_ic, pi_ic)
    rows, valid = O.print_fp_rows(N, 0.5, O.bin_table(N, 0.5), F, K)
    got = np.loadtxt(tmp_path / "output" / "1" / "Distrib-1-0.txt")
    want = rows[valid]
    assert got.shape == want.shape
    np.testing.assert_allclose(got[:, :2], want[:, :2], rtol=2e-6)  # p and the shell counts
    # shells above the IC cutoff hold only the transforms' round-off (1e-32
    # against 1e3, rocFFT and numpy's differently): compare the others
    phys = want[:, 3] > 1e-20 * want[:, 3].max()
    assert phys.sum() >= 4
    np.testing.assert_allclose(got[phys], want[phys], rtol=2e-5)


def test_driver_stops_when_the_false_vacuum_decays(tmp_path):
    # false_vacuum_decayed (fieldanalysis.cpp:887-893) ends the run at the
    # output where the average interaction turns negative (main.cpp:85-89).
    # With g = 0.1 (kappa^2 = 3) the potential phi^2/2 - phi^4/24 +
    # phi^6/(720 kappa^2) is negative around phi = 7.3: a run started there
    # stops at its first output, after one SO row
    N, seed = 16, 5
    m = O.model(N, g=0.1)
    phi0 = 7.3
    args = [MAIN, "-runs", "1", "-Nt", str(N), "-at", "0.5", "-M", "1", "-Mmode", "1", "-L", "6",
            "-Lmode", "-1", "-Kappa", "%.17g" % m["kappa"], "-T0", "0", "-TMax", "2.0", "-dt", "0.1",
            "-TSO", "0.5", "-TMO", "1.0", "-Noise", "0.5", "-Ampl", "0", "-QS", "%.17g" % (np.sqrt(3.0) / 0.5),
            "-Q2", "%.17g" % (np.sqrt(3.0) / 0.5), "-phi0_0", "%.17g" % phi0, "-seed", str(seed)]
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "false vacuum decayed at tau=0" in r.stdout
    so = np.loadtxt(tmp_path / "output" / "1" / "SO-1.txt", ndmin=2)
    assert so.shape[0] == 1
    ophi, opi = O.initial_conditions(dict(m), seed, phi0)
    avg = O.volume_averages(m, ophi, opi)
    assert avg[8] < 0  # the interaction average the driver tested
    np.testing.assert_allclose(so[0, 1:11], avg, rtol=2e-5, atol=1e-9)


def _cont(m, phi, pi, start, stop):
    """oracle steps start..stop-1 with the factors those steps get in one run"""
    dtk, dtd, _ = O.time_factors(stop)
    import ctypes as C
    N = m["N"]
    rphi, rpi = O.to_ref(phi, N), O.to_ref(pi, N)
    p = O.params(m)
    k = np.ascontiguousarray(dtk[start:stop])
    d = np.ascontiguousarray(dtd[start:stop])
    O.lib().or_evolve(C.byref(p), O.dptr(rphi), O.dptr(rpi), stop - start, O.dptr(k), O.dptr(d))
    return O.from_ref(rphi, N), O.from_ref(rpi, N)


def test_driver_checkpoint_round_trip(tmp_path):
    # -saveconfig writes the reference's field-<n>-<ID>.out / moment-<n>.out
    # (fieldanalysis.cpp:252-282) for the state the oracle reaches; a second
    # run with -extractconfig resumes from it and saves the same state
    N, seed, TMax = 16, 77, 1.0
    m = O.model(N)
    phi0 = float(np.sqrt(6.0))
    base = [MAIN, "-runs", "1", "-Nt", str(N), "-at", "0.5", "-M", "1", "-Mmode", "1", "-L", "6",
            "-Lmode", "-1", "-Kappa", "%.17g" % m["kappa"], "-T0", "0", "-TMax", "%g" % TMax,
            "-dt", "0.1", "-TSO", "0.5", "-TMO", "1.0", "-Noise", "0.5", "-Ampl", "0",
            "-QS", "%.17g" % (np.sqrt(3.0) / 0.5), "-Q2", "%.17g" % (np.sqrt(3.0) / 0.5),
            "-phi0_0", "%.17g" % phi0, "-seed", str(seed), "-output", "0"]
    r = subprocess.run(base + ["-saveconfig", "7"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, taus = O.time_factors(100)
    steps = int(np.argmax(taus >= TMax)) + 1  # for (reset(); tau < TMax; evolve())
    ophi, opi = O.initial_conditions(dict(m), seed, phi0)
    ophi, opi = O.evolve(m, ophi, opi, steps)
    field = (tmp_path / "field-7-0.out").read_text()
    vals = np.array([float(t) for t in field.split()]).reshape(-1, 2)
    # the format exactly (%.17e round-trips doubles); the state to 1e-10
    # (the driver's IC goes through rocFFT, the oracle's through numpy)
    assert field == O.config_text(vals[:, 0], vals[:, 1])
    np.testing.assert_allclose(vals[:, 0], ophi, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(vals[:, 1], opi, rtol=1e-10, atol=1e-12)
    assert (tmp_path / "moment-7.out").read_text() == O.moment_text(taus[steps - 1])
    r = subprocess.run(base + ["-extractconfig", "7", "-saveconfig", "8"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "field-8-0.out").read_text() == field
    assert (tmp_path / "moment-8.out").read_text() == (tmp_path / "moment-7.out").read_text()


def test_driver_queued_steps_bitwise(tmp_path):
    # evolve() queues its device work; 300 steps without an output in between
    # pass the 256-step flush inside the run.  Resumed from exact fields
    # (-extractconfig) and saved at the end (-saveconfig), the fields are the
    # oracle's bit for bit.
    N, TMax = 16, 15.0
    m = O.model(N)
    rng = np.random.default_rng(5)
    phi = np.sqrt(6.0) + 0.4 * rng.standard_normal(N ** 3)
    pi = 0.2 * rng.standard_normal(N ** 3)
    (tmp_path / "field-5-0.out").write_text(O.config_text(phi, pi))
    (tmp_path / "moment-5.out").write_text(O.moment_text(0.0))
    args = [MAIN, "-runs", "1", "-Nt", str(N), "-at", "0.5", "-M", "1", "-Mmode", "1", "-L", "6",
            "-Lmode", "-1", "-Kappa", "%.17g" % m["kappa"], "-T0", "0", "-TMax", "%g" % TMax,
            "-dt", "0.1", "-TSO", "1000", "-TMO", "1000", "-Noise", "0.5", "-Ampl", "0",
            "-phi0_0", "%.17g" % np.sqrt(6.0), "-seed", "1", "-output", "0",
            "-extractconfig", "5", "-saveconfig", "9"]
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, taus = O.time_factors(400)
    steps = int(np.argmax(taus >= TMax)) + 1
    assert steps > 256
    ophi, opi = O.evolve(m, phi, pi, steps)
    assert (tmp_path / "field-9-0.out").read_text() == O.config_text(ophi, opi)


def test_driver_info_matches_reference(tmp_path):
    # Print_Info_Console and Print_Info_File (fieldanalysis.cpp:81-247): the
    # reference's own text for the same parameters (the fixture comes from
    # oracle/_ref/ref_harness --info), byte for byte
    g = json.load(open(INFO))
    r = subprocess.run([MAIN] + g["args"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert g["console"] in r.stdout
    assert (tmp_path / "output" / "1" / "info.txt").read_text() == g["info_txt"]
