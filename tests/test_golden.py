"""The oracle against the committed golden fixtures (tests/golden/*.npz),
which hold seeded inputs and the outputs of the reference's own sources
compiled with strict IEEE flags (tests/golden/make_golden.py).  Runs
everywhere, including where the reference tree is absent."""
import glob
import os

import numpy as np
import pytest

import _oracle as O

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_oracle_reproduces_reference_fixture(path):
    z = np.load(path)
    N = int(z["N"])
    m = dict(N=N, a=float(z["a"]), mass_mode=int(z["mass_mode"]), m2=float(z["m2"]),
             lambda_mode=int(z["lambda_mode"]), kappa=float(z["kappa"]), Lambda=float(z["Lambda"]))
    steps = int(z["steps"])
    phi, pi = O.evolve(m, z["phi_in"], z["pi_in"], steps, T0=float(z["T0"]), dtbase=float(z["dtbase"]))
    assert np.array_equal(phi, z["phi_out"]) and np.array_equal(pi, z["pi_out"])
    assert np.array_equal(O.volume_averages(m, phi, pi), z["avg"])
    vals, idx = O.extrema(N, phi)
    assert np.array_equal(vals, z["ext"]) and np.array_equal(idx, z["extpos"])
    if int(z["islands"]):
        cat, labels = O.islands(m, 0, float(z["thr"]), phi, pi)
        assert np.array_equal(cat, z["catalog"])
        assert np.array_equal(labels, z["labels"])


def test_fixtures_present():
    assert len(FIX) >= 4
