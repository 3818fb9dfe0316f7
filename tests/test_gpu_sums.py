"""The FP64 sums of the volume averages (fieldanalysis.cpp:399-699) and of the
shell spectrum (momentumspace.cpp:196-247) against exact sums (math.fsum) of
the same per-site and per-mode terms.  The device accumulates (hi, lo) pairs
(two-sum) and merges them in a fixed order, so the results are within a few
ulp of the exact sum and identical from call to call; the reference sums
sequentially."""
import math

import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu


def _lat(N, prec, a=0.5):
    return pn.Lattice(N, a, 1, 1.0, -1, 2.0, precision=prec)


@pytest.mark.parametrize("prec,scale", [(pn.PN_FP64, 1e12), (pn.PN_FP32, 1e5)])
def test_volume_sums_match_exact_sums(prec, scale):
    # pairs of +-scale cancel; a unit-size remainder is all that is left, so a
    # sequential sum over the lattice loses most of its digits
    N = 32
    n = N ** 3
    rng = np.random.default_rng(9)
    big = np.where(np.arange(n) % 2 == 0, scale, -scale)[rng.permutation(n)]
    phi = big + rng.standard_normal(n)
    pi = -1e-3 * big + rng.standard_normal(n)
    with _lat(N, prec) as L:
        L.set_fields(phi, pi)
        f, p = L.get_fields()  # the values the lattice holds, in double
        avg = L.volume_averages()
    for got, terms in ((avg.phi, f), (avg.pi, p), (avg.phipi, f * p)):
        exact = math.fsum(terms) / n  # n is a power of two: the division is exact
        assert exact != 0.0
        assert abs(got - exact) <= 4 * np.spacing(abs(exact)), (got, exact)
    if prec == pn.PN_FP64:
        # left to right, as the reference's loop: far from exact on these
        # fields (FP32-held values of this size add exactly in double)
        seq = np.cumsum(f)[-1] / n
        assert abs(seq - math.fsum(f) / n) > 1e3 * np.spacing(abs(math.fsum(f) / n))


def _bins(N, a):
    """Bin of every mode [pz][px][py] from the reference's expressions
    (momentumspace.cpp:40,102-125,222-228) in the same double operations."""
    H = N // 2 + 1
    pi_ = math.acos(-1.)

    def lat(k):
        return (2.0 / a) * math.sin(k * pi_ / N) if k < N // 2 + 1 else -(2.0 / a) * math.sin(k * pi_ / N)

    tx = np.array([lat(k) ** 2 for k in range(N)])
    ty = np.array([((2.0 / a) * math.sin(k * pi_ / N)) ** 2 for k in range(H)])
    tz = np.array([0.0 if k == 0 else lat(k) ** 2 for k in range(N)])
    binsize = (2.0 / a) * math.sin(pi_ / N)
    psqr = (tx[None, :, None] + ty[None, None, :]) + tz[:, None, None]
    return np.floor(np.sqrt(psqr) / binsize + 0.5).astype(np.int64)


@pytest.mark.parametrize("N,prec", [(32, pn.PN_FP64), (33, pn.PN_FP64), (48, pn.PN_FP64), (32, pn.PN_FP32)])
def test_spectrum_sums_match_exact_sums(N, prec):
    a = 0.5
    H = N // 2 + 1
    rng = np.random.default_rng(N)
    phi = 3.0 + rng.standard_normal(N ** 3) * np.exp(rng.uniform(-6, 0, N ** 3))
    pi = rng.standard_normal(N ** 3)
    with _lat(N, prec, a) as L:
        L.set_fields(phi, pi)
        s1 = L.spectrum()
        s2 = L.spectrum()
        A = L.fft_modes(0)  # a^3-scaled modes, as the binning kernel forms them
        B = L.fft_modes(1)
    # the same call gives the same bits
    for key in ("F", "K", "Q", "counts"):
        assert np.array_equal(s1[key], s2[key])
    assert np.array_equal(s1["counts"], O.bin_table(N, a))
    b = _bins(N, a)
    twice = np.ones(H, dtype=bool)  # the reference's py != 0 && py != N/2
    twice[0] = False
    twice[N // 2] = False
    ff = A.real * A.real + A.imag * A.imag
    kk = B.real * B.real + B.imag * B.imag
    qq = A.real * B.real + A.imag * B.imag
    for key, v in (("F", ff), ("K", kk), ("Q", qq)):
        v = np.where(twice[None, None, :], v + v, v)
        for k in range(s1["nbins"]):
            terms = v[b == k]
            exact = math.fsum(terms)
            tol = 4 * np.spacing(abs(exact)) if key != "Q" else 4 * np.spacing(math.fsum(np.abs(terms)))
            assert abs(s1[key][k] - exact) <= tol, (key, k, s1[key][k], exact)


@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_island_sums_match_exact_sums(prec):
    # per-island phi, potential, kinetic and gradient sums (islands.cpp:131-142)
    # against math.fsum over the island's sites; the same bits on a second call
    N, a = 24, 0.5
    m = O.model(N)
    rng = np.random.default_rng(17)
    phi = 2.4 + 0.8 * rng.standard_normal(N ** 3)
    pi = 0.5 * rng.standard_normal(N ** 3)
    thr = 3.0
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"], precision=prec) as L:
        L.set_fields(phi, pi)
        f, p = L.get_fields()
        rows1, lab, _ = L.islands(0, thr, labels=True)
        rows2, _, _ = L.islands(0, thr)
    assert len(rows1) > 20
    fields = ("avg_phi", "pot_en", "kin_en", "grad_en", "peak_value", "cx", "cy", "cz")
    assert [[getattr(r, k) for k in fields] for r in rows1] == [[getattr(r, k) for k in fields] for r in rows2]
    F = f.reshape(N, N, N)
    fx = np.roll(F, -1, axis=2).ravel()
    fy = np.roll(F, -1, axis=1).ravel()
    fz = np.roll(F, -1, axis=0).ravel()
    a2 = a * a
    grad = 0.5 * ((fx - f) * (fx - f) + (fy - f) * (fy - f) + (fz - f) * (fz - f)) / a2
    kin = 0.5 * (p * p)
    q2 = (f * f) * (f * f)
    cm = m["mass_mode"] * 0.5 * m["m2"]
    d6 = 720.0 * m["kappa"] * m["kappa"]
    pot = cm * (f * f) + m["lambda_mode"] * (q2 / 24.0) + (q2 * (f * f) / d6)
    a3 = (a * a) * a
    for r in rows1:
        sel = lab == r.index
        nn = int(sel.sum())
        assert nn == r.nn
        for got, terms, post in ((r.avg_phi, f[sel], lambda s: s / nn), (r.pot_en, pot[sel], lambda s: s * a3),
                                 (r.kin_en, kin[sel], lambda s: s * a3), (r.grad_en, grad[sel], lambda s: s * a3)):
            want = post(math.fsum(terms))
            assert abs(got - want) <= 4 * np.spacing(math.fsum(np.abs(post(terms)))), (r.index, got, want)
