"""More of the oracle pinned by what the mathematics fixes, on the CPU: the
initial condition's normalisation (initialconditions.cpp:636-703 with the c2r
of clusterfft.cpp:148-182), the island energy density Delta
(fieldanalysis.cpp:447,514-515) and the shell assignment of the spectrum
(momentumspace.cpp:126).  Each check is chosen so that the plausible mistake
named beside it fails: the same check is run on the planted mistake and must
reject it."""
import math

import numpy as np

import _oracle as O


def _full_weights(N):
    # the half spectrum [pz][px][py <= N/2] stands for the full one: modes with
    # py not in {0, N/2} have a conjugate partner outside it
    w = np.full(N // 2 + 1, 2.0)
    w[0] = 1.0
    w[N // 2] = 1.0
    return w


def _lattice_p(N, a):
    # |p_lat| of every half-spectrum mode, in the reference's per-axis
    # expressions (momentumspace.cpp:102-125; negative frequencies above N/2)
    H = N // 2 + 1
    pi_ = math.acos(-1.)

    def lat(k):
        return (2.0 / a) * math.sin(k * pi_ / N) if k < N // 2 + 1 else -(2.0 / a) * math.sin(k * pi_ / N)

    tx = np.array([lat(k) for k in range(N)])
    ty = np.array([(2.0 / a) * math.sin(k * pi_ / N) for k in range(H)])
    tz = np.array([0.0 if k == 0 else lat(k) for k in range(N)])
    return np.sqrt((tx[None, :, None] ** 2 + ty[None, None, :] ** 2) + tz[:, None, None] ** 2)


def test_ic_fill_normalisation_by_parseval():
    # phi - phi0 = c2r(A) / Norm * NormFactor with Norm = (aN)^3 = V and
    # NormFactor = sqrt(V): Parseval for the unnormalised c2r gives
    # sum_x (phi - phi0)^2 = N^3 sum_p |A_p|^2 / V over the full spectrum, and
    # likewise for pi and E.  A NormFactor of V or 1, or a missing /Norm, is
    # off by a power of V = 64 and fails.
    N, a, seed, phi0, pi0 = 16, 0.5, 4321, 2.449489742783178, 0.25
    m = O.model(N)
    A, E = O.ic_modes(N, a, seed, m2_long=m["m2"], Lambda=m["Lambda"])
    phi, pi = O.initial_conditions(m, seed, phi0, pi0)
    V = (a * N) ** 3
    w = _full_weights(N)[None, None, :]
    for field, modes, base in ((phi, A, phi0), (pi, E, pi0)):
        want = N ** 3 * np.sum(w * np.abs(modes) ** 2) / V
        got = math.fsum((field - base) ** 2)
        assert abs(got - want) <= 1e-11 * want, (got, want)
        X = O.c2r_unnormalised(modes, N)
        for wrong in (X / V * V, X / V, X * math.sqrt(V)):
            bad = math.fsum(wrong ** 2)
            assert abs(bad - want) > 0.5 * want


def test_ic_variance_matches_occupation_closed_form():
    # <|A_p|^2> = Occ / Omega <g^2> and <|E_p|^2> = Occ Omega <g^2> with
    # <g^2> = 1 for |Gauss| (initialconditions.cpp:346-349,563-580): the
    # per-site variances of phi and pi are sum_p Occ/(Omega V) and
    # sum_p Occ Omega / V.  Occ = Noise*Lambda below QS, Omega =
    # sqrt(p^2 + M2_long).  Eight seeds at 32^3 (3845 occupied modes each,
    # about 2.5 % scatter per seed): the mean within 4 %.  A missing Lambda in
    # Occ (6x), Omega^(-1/2) in place of Omega^(-1) for the field, or the two
    # spectra swapped, each miss by more.
    N, a = 32, 0.5
    m = O.model(N)
    Lam, noise = m["Lambda"], 0.5
    p = _lattice_p(N, a)
    QS = math.sqrt(3.0 * (4.0 / (a * a))) / 2.0
    occ = np.where(p * p < QS * QS, noise * Lam, 0.0)
    omega = np.sqrt(p * p + m["m2"])
    w = _full_weights(N)[None, None, :]
    V = (a * N) ** 3
    var_phi_cf = np.sum(w * occ / omega) / V
    var_pi_cf = np.sum(w * occ * omega) / V
    vp, vq = [], []
    for seed in range(11, 19):
        phi, pi = O.initial_conditions(m, seed, 0.0, 0.0)
        vp.append(np.mean(phi * phi))
        vq.append(np.mean(pi * pi))
    got_phi, got_pi = np.mean(vp), np.mean(vq)
    assert abs(got_phi / var_phi_cf - 1.0) < 0.04, (got_phi, var_phi_cf)
    assert abs(got_pi / var_pi_cf - 1.0) < 0.04, (got_pi, var_pi_cf)
    wrong = {"no Lambda": np.sum(w * (occ / Lam) / omega) / V,
             "Omega^-1/2": np.sum(w * occ / np.sqrt(omega)) / V,
             "swapped": var_pi_cf}
    for name, cf in wrong.items():
        assert abs(got_phi / cf - 1.0) > 0.2, name


def test_delta_sum_identity():
    # Delta = pi^2/2 + sum_i (grad+_i phi)^2 + V(phi) per site (the gradient
    # without the 1/2 that Energy has, fieldanalysis.cpp:447 against :625), so
    # sum Delta = vol (E + G/(2 vol)) and G/vol = 3 (E + Charge - <pi^2>)
    # (Charge, :624): sum Delta = vol (E + 1.5 (E + Charge - <pi^2>)).  E,
    # Charge and <pi^2> are pinned to the reference build; Delta with the
    # gradient halved (as in Energy) fails by the size of the gradient term.
    N = 10
    m = O.model(N)
    rng = np.random.default_rng(31)
    phi = 2.449 + 0.5 * rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    avg, delta = O.volume_averages(m, phi, pi, want_delta=True)
    E, charge, pi2 = avg[7], avg[9], avg[3]
    vol = N ** 3
    want = vol * (E + 1.5 * (E + charge - pi2))
    got = math.fsum(delta)
    assert abs(got - want) <= 1e-12 * math.fsum(np.abs(delta)), (got, want)
    f = phi.reshape(N, N, N)
    a2 = m["a"] * m["a"]
    g = sum((np.roll(f, -1, ax) - f) ** 2 for ax in (0, 1, 2)).ravel() / a2
    half = math.fsum(delta - 0.5 * g)
    assert abs(half - want) > 1e3 * 1e-12 * math.fsum(np.abs(delta))


def _shell_of_each_mode(N, a):
    # the oracle's binning (or_spectrum_bins) run on one mode at a time: the
    # shell that receives its power
    H = N // 2 + 1
    nb = O.lib().or_bin_count(N, a, None)
    zero = np.zeros(2 * N * N * H)
    shell = np.empty((N, N, H), dtype=np.int64)
    weight = np.empty((N, N, H))
    for pz in range(N):
        for px in range(N):
            for py in range(H):
                pk = zero.copy()
                pk[2 * (py + px * H + pz * N * H)] = 1.0
                F = np.zeros(nb)
                K = np.zeros(nb)
                Q = np.zeros(nb)
                O.lib().or_spectrum_bins(N, a, O.dptr(pk), O.dptr(zero), O.dptr(F), O.dptr(K), O.dptr(Q))
                (hit,) = np.nonzero(F)
                assert hit.size == 1
                shell[pz, px, py] = hit[0]
                weight[pz, px, py] = F[hit[0]]
    return shell, weight


def test_shell_is_the_nearest_multiple_of_the_bin_width():
    # momentumspace.cpp:126 puts a mode in shell int(|p|/pTmin + 0.5): the
    # nearest multiple of the bin width, so |p|/pTmin lies within half a bin
    # of the shell index for every mode (and the printed p = k pTmin,
    # :318-323, is the shell's centre).  Truncation, int(|p|/pTmin), puts
    # modes up to a whole bin above their shell and fails.  The py double
    # count (weight 2 for py not in {0, N/2}) is fixed by Parseval.
    N, a = 12, 0.5
    shell, weight = _shell_of_each_mode(N, a)
    b = (2.0 / a) * math.sin(math.acos(-1.) / N)
    r = _lattice_p(N, a) / b
    assert np.all(np.abs(r - shell) <= 0.5 + 1e-12)
    assert np.array_equal(weight, np.broadcast_to(_full_weights(N)[None, None, :], weight.shape))
    assert np.array_equal(np.bincount(shell.ravel(), weights=weight.ravel()).astype(np.int64),
                          O.bin_table(N, a)[:shell.max() + 1])
    truncated = np.floor(r).astype(np.int64)
    assert np.max(np.abs(r - truncated)) > 0.6
