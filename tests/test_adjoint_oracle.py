"""Pins the CPU checker of the differentiable leapfrog: closed-form dF/drho
against finite differences, the dot-product identity between the
tangent-linear and adjoint passes, and central finite differences of the
whole forward run.  CPU only."""
import numpy as np
import pytest

import _oracle as O


def _state(N, seed, amp=0.6):
    rng = np.random.default_rng(seed)
    return 1.0 + amp * rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)


def test_observable_derivative_closed_form():
    N = 6
    p = O.adj_params(N)
    phi, pi = _state(N, 1, amp=1.2)
    F, rho, g = O.adj_observable(p, phi, pi)
    assert 0.0 < F < 1.0
    # perturb rho through pi at a few sites: dF = g_j * pi_j * dpi_j
    for j in (0, 17, 101, 200):
        eps = 1e-6
        pp, pm = pi.copy(), pi.copy()
        pp[j] += eps
        pm[j] -= eps
        fd = (O.adj_observable(p, phi, pp)[0] - O.adj_observable(p, phi, pm)[0]) / (2 * eps)
        assert abs(fd - g[j] * pi[j]) <= 1e-7 * max(1e-3, abs(fd))


def test_dot_product_tlm_vs_adjoint():
    N, S = 6, 60
    p = O.adj_params(N)
    phi0, pi0 = _state(N, 2)
    dtk, dtd, _ = O.time_factors(S)
    F, gphi, gpi, gc = O.adj_reverse(p, phi0, pi0, dtk, dtd)
    rng = np.random.default_rng(3)
    for _ in range(3):
        v = rng.standard_normal(N ** 3)
        w = rng.standard_normal(N ** 3)
        dc = rng.standard_normal(3)
        tl = O.adj_tlm(p, phi0, pi0, dtk, dtd, v, w, dc)
        ad = float(np.dot(gphi, v) + np.dot(gpi, w) + np.dot(gc, dc))
        assert abs(tl - ad) <= 1e-12 * max(abs(tl), 1e-300) + 1e-15


def test_dot_product_tlm_vs_adjoint_128_cube():
    # the same identity at the adjoint configuration's lattice, 128^3, over
    # 100 steps (the linearisation of dynamics.cpp:52-194 against the
    # checker's reverse pass): 7.6e-14 relative here.  A sign error in the
    # coupling gradient, or a gradient for phi0 that misses pi0's share,
    # breaks it by orders of magnitude more than the tolerance.
    N, S = 128, 100
    p = O.adj_params(N)
    rng = np.random.default_rng(128)
    phi0 = 1.0 + 0.6 * rng.standard_normal(N ** 3)
    pi0 = 0.3 * rng.standard_normal(N ** 3)
    dtk, dtd, _ = O.time_factors(S)
    F, gphi, gpi, gc = O.adj_reverse(p, phi0, pi0, dtk, dtd)
    v, w, dc = rng.standard_normal(N ** 3), rng.standard_normal(N ** 3), rng.standard_normal(3)
    tl = O.adj_tlm(p, phi0, pi0, dtk, dtd, v, w, dc)
    ad = float(np.dot(gphi, v) + np.dot(gpi, w) + np.dot(gc, dc))
    assert abs(tl - ad) <= 1e-11 * abs(tl)
    wrong = (float(np.dot(gphi, v) + np.dot(gpi, w) - np.dot(gc, dc)),
             float(np.dot(gphi, v) + np.dot(gc, dc)))
    for bad in wrong:
        assert abs(tl - bad) > 1e3 * 1e-11 * abs(tl)


@pytest.mark.parametrize("which", ["phi", "c4", "c6"])
def test_adjoint_matches_central_differences(which):
    N, S = 6, 40
    p = O.adj_params(N)
    phi0, pi0 = _state(N, 4)
    dtk, dtd, _ = O.time_factors(S)
    F, gphi, gpi, gc = O.adj_reverse(p, phi0, pi0, dtk, dtd)
    eps = 1e-6
    if which == "phi":
        v = np.random.default_rng(5).standard_normal(N ** 3)
        fp = O.adj_forward_F(p, phi0 + eps * v, pi0, dtk, dtd)
        fm = O.adj_forward_F(p, phi0 - eps * v, pi0, dtk, dtd)
        exact = float(np.dot(gphi, v))
    else:
        k = {"c4": 1, "c6": 2}[which]
        pp = O.adj_params(N, c4=p.c4 + (eps if k == 1 else 0), c6=p.c6 + (eps if k == 2 else 0))
        pm = O.adj_params(N, c4=p.c4 - (eps if k == 1 else 0), c6=p.c6 - (eps if k == 2 else 0))
        fp = O.adj_forward_F(pp, phi0, pi0, dtk, dtd)
        fm = O.adj_forward_F(pm, phi0, pi0, dtk, dtd)
        exact = gc[k]
    fd = (fp - fm) / (2 * eps)
    assert abs(fd - exact) <= 1e-6 * abs(exact) + 1e-9
