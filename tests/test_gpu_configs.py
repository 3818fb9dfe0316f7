"""The batching and differentiation forms at their configured sizes.

Batch: 64 realizations, 32 x 64^3 + 24 x 128^3 + 8 x 192^3 (115,343,360
sites), seeds 0..63, lambda = 1 and g drawn uniformly from [0.1, 0.4] with
numpy's default_rng(0) -- the `runs` loop of main.cpp:52-96 stepped together.
103 steps with graphs on (one graph replay of 96 or 100 steps and the rest
launched singly) in the three-step, two-step and one-step modes: every member
bitwise equal to the same realization stepped alone, and one 128^3 and one
192^3 member bitwise equal to the oracle (dynamics.cpp:52-237).

Adjoint: 128^3, 200 steps in two checkpoint segments (K = 100), gradients of
the oscillon fraction against the CPU reverse pass and a tangent-linear
directional derivative (the linearisation of dynamics.cpp:52-194)."""
import hashlib

import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu

SIZES = [64] * 32 + [128] * 24 + [192] * 8
STEPS = 103


def _reals():
    g = np.random.default_rng(0).uniform(0.1, 0.4, len(SIZES))
    return [dict(N=N, a=0.5, mass_mode=1, m2=1.0, lambda_mode=-1, kappa=float(np.sqrt(36.0 / (120.0 * gg))))
            for N, gg in zip(SIZES, g)]


def _ic(r):
    return pn.ic_params(r, a=0.5, Lambda=6.0, phi0=float(np.sqrt(6.0)))


def _digest(phi, pi):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(phi).tobytes())
    h.update(np.ascontiguousarray(pi).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def alone():
    """Each realization stepped on its own lattice: digests of its fields, and
    the fields of the first 128^3 and 192^3 members with their initial state."""
    reals = _reals()
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, STEPS)
    dig, keep = [], {}
    probe = {SIZES.index(128), SIZES.index(192)}
    for r, re in enumerate(reals):
        with pn.Lattice(re["N"], 0.5, 1, 1.0, -1, re["kappa"]) as L:
            L.initial_conditions(_ic(r))
            if r in probe:
                keep[r] = L.get_fields()
            L.evolve(dtk, dtd)
            f, p = L.get_fields()
        dig.append(_digest(f, p))
        if r in probe:
            keep[r] = (keep[r], (f, p))
    return reals, dig, keep


@pytest.mark.parametrize("mode", ["three-step", "two-step", "one-step"])
def test_batch_of_64_at_configured_sizes(alone, mode):
    reals, dig, keep = alone
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, STEPS)
    with pn.Batch(reals) as B:
        if mode == "one-step":
            B.two_step(False)
        elif mode == "two-step":
            B.three_step(False)
        assert B.three_step() == (mode == "three-step")
        assert B.two_step() == (mode != "one-step")
        for r in range(len(reals)):
            B.view(r).initial_conditions(_ic(r))
        B.autotune(trial_steps=5)  # restores the state; the choice changes no bit
        B.evolve(dtk, dtd, use_graph=True)
        got = []
        for r in range(len(reals)):
            f, p = B.view(r).get_fields()
            got.append(_digest(f, p))
            if r in keep and mode == "three-step":
                (f0, p0), (fa, pa) = keep[r]
                m = dict(reals[r], N=reals[r]["N"])
                of, op = O.evolve(m, f0, p0, STEPS)
                assert np.array_equal(of, f) and np.array_equal(op, p), r
                assert np.array_equal(fa, f) and np.array_equal(pa, p), r
    bad = [r for r in range(len(reals)) if got[r] != dig[r]]
    assert not bad, bad


def test_adjoint_at_128_cube_two_segments():
    N, S, K = 128, 200, 100
    m = O.model(N)
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"]) as L:
        L.initial_conditions(pn.ic_params(7, a=0.5, Lambda=6.0, phi0=float(np.sqrt(6.0))))
        phi0, pi0 = L.get_fields()
    s = float(np.sqrt(6.0))
    phi0, pi0 = phi0 / s, pi0 / s  # physical units, V = phi^2/2 - phi^4/4 + 0.2 phi^6/6
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, S)
    F, gphi, gpi, gc, res = pn.adjoint(pn.adj_params(N, checkpoint_every=K), phi0, pi0, dtk, dtd)
    op = O.adj_params(N)
    oF, ogphi, ogpi, ogc = O.adj_reverse(op, phi0, pi0, dtk, dtd)
    # the checker sums the observable's 2M terms left to right (relative error
    # up to ~1e-12 here), the GPU in a compensated tree
    assert abs(F - oF) <= 1e-11 * abs(oF)
    assert np.max(np.abs(gphi - ogphi)) <= 1e-10 * np.max(np.abs(ogphi))
    assert np.max(np.abs(gpi - ogpi)) <= 1e-10 * np.max(np.abs(ogpi))
    np.testing.assert_allclose(gc, ogc, rtol=1e-10)
    rng = np.random.default_rng(128)
    v, w, dc = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3), rng.standard_normal(3)
    tl = O.adj_tlm(op, phi0, pi0, dtk, dtd, v, w, dc)
    ad = float(np.dot(gphi, v) + np.dot(gpi, w) + np.dot(gc, dc))
    assert abs(tl - ad) <= 1e-10 * abs(tl)
    assert res.peak_bytes > 0 and res.backward_ms > 0
