"""The oracle's restatement against what the mathematics fixes, on the CPU:
the free-field leapfrog's exact dispersion, second-order convergence of the
homogeneous field to the ODE solution, kick/drift time reversal, a single
plane wave's shell and Parseval's identity for the binned spectrum, and the
binary32 restatement (or_evolve_f32) against the binary64 one.  The binary64
functions are also bitwise equal to the reference build (test_golden,
test_oracle_vs_reference); these checks pin them without it, chosen so that a
dropped term, a wrong sign or a missing half kick fails."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import _oracle as O


def _free(N, a=0.5, m2=1.0):
    # lambda_mode = 0 and kappa so large that (q/kappa)^2/120 is lost against
    # m2 in double: a free field of mass^2 m2
    return dict(N=N, a=a, mass_mode=1, m2=m2, lambda_mode=0, kappa=1e150)


def _plane(N, a, n):
    ix = np.arange(N)
    z, y, x = np.meshgrid(ix, ix, ix, indexing="ij")
    return np.cos(2 * np.pi * (n[0] * x + n[1] * y + n[2] * z) / N).ravel()


@pytest.mark.parametrize("n", [(1, 2, 3), (0, 0, 4), (3, 0, 1)])
def test_free_field_leapfrog_dispersion(n):
    # phi'' = lap phi - m2 phi on a plane wave: the leapfrog (half kick first,
    # dynamics.cpp:196-237) gives phi_s = A cos(s w dt) exactly, with
    # sin(w dt / 2) = omega dt / 2 and omega^2 = m2 + sum_i (2/a sin(pi n_i / N))^2
    N, a, A = 8, 0.5, 0.7
    m = _free(N, a)
    steps = 300
    phi0 = A * _plane(N, a, n)
    phi, pi = O.evolve(m, phi0, np.zeros_like(phi0), steps)
    dtk, dtd, _ = O.time_factors(steps, 0.0, 0.1, a)
    dt = 0.1 * a
    assert abs(dtk[0] - dt / 2) < 1e-15 and np.all(np.abs(dtd - dt) < 1e-15)
    omega2 = m["m2"] + sum((2.0 / a * math.sin(math.pi * k / N)) ** 2 for k in n)
    w = 2.0 / dt * math.asin(math.sqrt(omega2) * dt / 2.0)
    want_phi = math.cos(steps * w * dt) * phi0
    # Pi lags half a step: pi_{s-1/2} = (phi_s - phi_{s-1}) / dt
    want_pi = (math.cos(steps * w * dt) - math.cos((steps - 1) * w * dt)) / dt * phi0
    assert np.max(np.abs(phi - want_phi)) <= 1e-11 * A
    assert np.max(np.abs(pi - want_pi)) <= 1e-10 * A
    # the continuum frequency (no lattice dispersion) is off by far more
    w_wrong = math.sqrt(omega2)
    assert np.max(np.abs(math.cos(steps * w_wrong * dt) * phi0 - want_phi)) > 1e-4 * A


def _homogeneous_error(dtbase, T, phi0=2.0):
    m = O.model(4)
    N = 4
    steps = int(round(T / (dtbase * m["a"])))
    phi, pi = O.evolve(m, np.full(N ** 3, phi0), np.zeros(N ** 3), steps, dtbase=dtbase)
    assert np.ptp(phi) == 0.0  # the Laplacian of a constant field is zero

    def rhs(t, u):
        f = u[0]
        q = f * f
        force = (m["mass_mode"] * m["m2"] + m["lambda_mode"] * (q / 6.0) + (q / m["kappa"]) ** 2 / 120.0) * f
        return [u[1], -force]

    sol = solve_ivp(rhs, (0.0, T), [phi0, 0.0], method="DOP853", rtol=1e-13, atol=1e-14)
    return abs(phi[0] - sol.y[0, -1])


def test_homogeneous_field_converges_at_second_order():
    # a uniform field follows phi'' = -V'(phi) (dynamics.cpp:84 with lap = 0);
    # the leapfrog's error at fixed T falls 4x per halving of dt (a missing
    # half kick or a sign error in the force breaks the ratio or the size)
    T = 4.0
    e1 = _homogeneous_error(0.1, T)
    e2 = _homogeneous_error(0.05, T)
    e3 = _homogeneous_error(0.025, T)
    assert e2 < 1e-3
    assert 3.6 < e1 / e2 < 4.4 and 3.6 < e2 / e3 < 4.4, (e1, e2, e3)


def test_kick_and_drift_reverse_to_the_start():
    # kick depends on phi only and drift on pi only: undoing the steps in
    # reverse order with negated factors recovers the start to roundoff
    N = 6
    m = O.model(N)
    rng = np.random.default_rng(3)
    phi0 = 2.449 + 0.3 * rng.standard_normal(N ** 3)
    pi0 = 0.1 * rng.standard_normal(N ** 3)
    steps = 50
    dtk, dtd, _ = O.time_factors(steps, 0.0, 0.1, m["a"])
    phi = O.to_ref(phi0, N)
    pi = O.to_ref(pi0, N)
    p = O.params(m)
    lib = O.lib()
    lib.or_evolve(C.byref(p), O.dptr(phi), O.dptr(pi), steps, O.dptr(dtk), O.dptr(dtd))
    moved = np.max(np.abs(O.from_ref(phi, N) - phi0))
    for s in range(steps - 1, -1, -1):
        lib.or_update_fields(N, O.dptr(phi), O.dptr(pi), -dtd[s])
        lib.or_exchange(N, O.dptr(phi))
        lib.or_update_momenta(C.byref(p), O.dptr(phi), O.dptr(pi), -dtk[s])
    assert moved > 0.1
    assert np.max(np.abs(O.from_ref(phi, N) - phi0)) <= 1e-12 * np.max(np.abs(phi0))
    assert np.max(np.abs(O.from_ref(pi, N) - pi0)) <= 1e-11


@pytest.mark.parametrize("n", [(2, 1, 0), (0, 3, 1), (1, 0, 0), (0, 8, 0)])
def test_single_plane_wave_lands_in_its_shell(n):
    # cos(k.x) puts all of sum_shell |phi_k|^2 = a^6 N^3 sum phi^2 (Parseval,
    # with the half-spectrum's py double count) into the shell of
    # |p_lat| = sqrt(sum (2/a sin(pi n_i / N))^2) in units of (2/a) sin(pi/N)
    N, a = 16, 0.5
    phi = _plane(N, a, n)
    F, K, Q = O.spectrum(N, a, phi, np.zeros_like(phi))
    total = a ** 6 * N ** 3 * np.sum(phi * phi)
    p = math.sqrt(sum((2.0 / a * math.sin(math.pi * k / N)) ** 2 for k in n))
    b = int(math.floor(p / ((2.0 / a) * math.sin(math.pi / N)) + 0.5))
    assert abs(F[b] - total) <= 1e-12 * total
    assert np.sum(np.abs(np.delete(F, b))) <= 1e-20 * total
    assert np.all(K == 0) and np.all(Q == 0)


def test_print_fp_rows_on_a_single_mode():
    # Print_fp (momentumspace.cpp:318-369) on phi = A cos(k.x), pi = B cos(k.x):
    # one shell written, with omega = |B/A| and the norm 1/((N a)^3 count):
    # FF norm = a^3 N^3 A^2 / (2 count), and the other columns from it
    N, a, A, B = 16, 0.5, 0.7, -1.3
    n = (2, 1, 0)
    wave = _plane(N, a, n)
    F, K, Q = O.spectrum(N, a, A * wave, B * wave)
    counts = O.bin_table(N, a)
    rows, valid = O.print_fp_rows(N, a, counts, F, K)
    b = int(np.argmax(F))
    # the other shells hold only the transform's round-off
    others = valid.copy()
    others[b] = False
    assert valid[b] and np.all(rows[others, 3] <= 1e-20 * rows[b, 3])
    binsize = (2.0 / a) * math.sin(math.pi / N)
    ff = a ** 3 * N ** 3 * A * A / (2 * counts[b])
    kk = a ** 3 * N ** 3 * B * B / (2 * counts[b])
    want = [binsize * b, counts[b], abs(B / A), ff, kk, math.sqrt(ff * kk),
            ff * (binsize * b) ** 3 / (2 * math.pi ** 2)]
    np.testing.assert_allclose(rows[b], want, rtol=1e-12)


def test_parseval_for_random_fields():
    N, a = 12, 0.5
    rng = np.random.default_rng(8)
    phi = 1.0 + rng.standard_normal(N ** 3)
    pi = rng.standard_normal(N ** 3)
    F, K, Q = O.spectrum(N, a, phi, pi)
    s = a ** 6 * N ** 3
    assert abs(F.sum() - s * np.sum(phi * phi)) <= 1e-12 * s * np.sum(phi * phi)
    assert abs(K.sum() - s * np.sum(pi * pi)) <= 1e-12 * s * np.sum(pi * pi)
    assert abs(Q.sum() - s * np.sum(phi * pi)) <= 1e-11 * s * np.sqrt(np.sum(phi * phi) * np.sum(pi * pi))
    assert O.bin_table(N, a).sum() == N ** 3


def _f32(m, phi0, pi0, steps, c4=None, c6=None):
    N, a = m["N"], m["a"]
    coef = [np.float32(1.0 / (a * a)), np.float32(m["mass_mode"] * m["m2"]),
            np.float32(m["lambda_mode"] / 6.0) if c4 is None else np.float32(c4),
            np.float32(1.0 / (120.0 * m["kappa"] * m["kappa"])) if c6 is None else np.float32(c6)]
    phi = O.to_ref(phi0.astype(np.float32).astype(np.float64), N).astype(np.float32)
    pi = O.to_ref(pi0.astype(np.float32).astype(np.float64), N).astype(np.float32)
    dtk, dtd, _ = O.time_factors(steps, 0.0, 0.1, a)
    dtk = dtk.astype(np.float32)
    dtd = dtd.astype(np.float32)
    O.lib().or_evolve_f32(N, *[float(v) for v in coef], O.fptr(phi), O.fptr(pi), steps, O.fptr(dtk), O.fptr(dtd))
    P = N * N
    return phi[P:P * (N + 1)].astype(np.float64), pi[P:P * (N + 1)].astype(np.float64)


def test_fp32_restatement_tracks_the_fp64_oracle():
    # or_evolve_f32 writes the force as ((c2 + c4 q) + c6 q^2) phi with
    # host-rounded coefficients: over 20 steps it stays within binary32
    # rounding of the binary64 oracle; a force with its phi^6 term dropped,
    # or with the phi^4 sign flipped, leaves that bound by orders of magnitude
    N = 12
    m = O.model(N)
    rng = np.random.default_rng(21)
    phi0 = 2.449 + 0.3 * rng.standard_normal(N ** 3)
    pi0 = 0.1 * rng.standard_normal(N ** 3)
    steps = 20
    ref_phi, ref_pi = O.evolve(m, phi0.astype(np.float32).astype(np.float64),
                               pi0.astype(np.float32).astype(np.float64), steps)
    scale = np.max(np.abs(ref_phi))
    phi, pi = _f32(m, phi0, pi0, steps)
    assert np.max(np.abs(phi - ref_phi)) <= 2e-6 * scale
    assert np.max(np.abs(pi - ref_pi)) <= 5e-6 * scale
    for wrong in (dict(c6=0.0), dict(c4=-m["lambda_mode"] / 6.0)):
        wphi, _ = _f32(m, phi0, pi0, steps, **wrong)
        assert np.max(np.abs(wphi - ref_phi)) > 1e-3 * scale, wrong
