"""Pins the oracle bitwise against the reference's own sources, compiled
unmodified with strict IEEE flags (oracle/_ref/ref_harness: dynamics.cpp,
comm.cpp, fieldanalysis.cpp, islands.cpp and their includes).  Skipped where
the harness has not been built (it needs the reference tree)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref/ref_harness not built")


def _fields(N, seed, phi0=2.449489742783178, amp=0.4):
    rng = np.random.default_rng(seed)
    return phi0 + amp * rng.standard_normal(N ** 3), 0.2 * rng.standard_normal(N ** 3)


@pytest.mark.parametrize("N,steps,seed", [(8, 40, 1), (12, 25, 2), (16, 10, 3), (6, 100, 4)])
def test_evolve_bitwise(N, steps, seed):
    m = O.model(N)
    phi, pi = _fields(N, seed)
    ref = O.run_ref(m, phi, pi, steps)
    ophi, opi = O.evolve(m, phi, pi, steps)
    assert np.array_equal(ophi.view(np.uint64), ref["phi"].view(np.uint64))
    assert np.array_equal(opi.view(np.uint64), ref["pi"].view(np.uint64))


def test_evolve_bitwise_other_params():
    # non power-of-two a^2, negative mass mode, nonzero T0, other dt base
    m = dict(N=10, a=0.37, mass_mode=-1, m2=0.7, lambda_mode=1, kappa=1.9, Lambda=3.0)
    phi, pi = _fields(10, 7, phi0=0.3)
    ref = O.run_ref(m, phi, pi, 30, T0=1.25, dtbase=0.07)
    ophi, opi = O.evolve(m, phi, pi, 30, T0=1.25, dtbase=0.07)
    assert np.array_equal(ophi, ref["phi"]) and np.array_equal(opi, ref["pi"])


@pytest.mark.parametrize("N,seed", [(8, 11), (16, 12)])
def test_averages_and_extrema_bitwise(N, seed):
    m = O.model(N)
    phi, pi = _fields(N, seed)
    ref = O.run_ref(m, phi, pi, 0)
    avg = O.volume_averages(m, phi, pi)
    assert np.array_equal(avg, ref["avg"])
    vals, idx = O.extrema(N, phi)
    assert np.array_equal(vals, ref["ext"])
    assert np.array_equal(idx, ref["extpos"])


@pytest.mark.parametrize("N,seed,thr", [(8, 21, -1.0), (12, 22, 2.2), (16, 23, -1.0)])
def test_islands_bitwise(N, seed, thr):
    m = O.model(N)
    phi, pi = _fields(N, seed, phi0=0.0, amp=1.0)
    ref = O.run_ref(m, phi, pi, 0, islands=True, thr=thr)
    cat, labels = O.islands(m, 0, ref["thr"], phi, pi)
    assert cat.shape[0] == ref["cat"].shape[0] > 0
    assert np.array_equal(cat, ref["cat"])
    assert np.array_equal(labels, ref["labels"])


def test_time_replay_matches_reference_steps():
    # the oracle's replayed factors reproduce the reference evolve() exactly:
    # one step through the harness equals one oracle kick+drift with dtk[0], dtd[0]
    m = O.model(8)
    phi, pi = _fields(8, 31)
    for T0 in (0.0, 3.7):
        ref = O.run_ref(m, phi, pi, 3, T0=T0)
        ophi, opi = O.evolve(m, phi, pi, 3, T0=T0)
        assert np.array_equal(ophi, ref["phi"]) and np.array_equal(opi, ref["pi"])


def test_fast_math_build_differs_only_at_roundoff():
    # the reference's own flags (run_PC.pbs:1) change the bits but not the physics
    if not O.have_ref(fast=True):
        pytest.skip("fast harness not built")
    m = O.model(8)
    phi, pi = _fields(8, 41)
    strict = O.run_ref(m, phi, pi, 200)
    fast = O.run_ref(m, phi, pi, 200, fast=True)
    d = np.max(np.abs(strict["phi"] - fast["phi"])) / np.max(np.abs(strict["phi"]))
    assert d < 1e-11
