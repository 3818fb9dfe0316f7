"""Checks at the configs' full sizes: bitwise steps at 256^3 against the
oracle, Parseval's identity for the binned spectrum, time reversal of the
leapfrog, bounded energy, and the island catalog at 128^3."""
import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu


def _lat(m, precision=pn.PN_FP64):
    return pn.Lattice(m["N"], m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"],
                      precision=precision)


def _ic(L, m, seed=1234):
    L.initial_conditions(pn.ic_params(seed, a=m["a"], Lambda=m["Lambda"], phi0=float(np.sqrt(m["Lambda"]))))


def test_256_cube_steps_bitwise():
    m = O.model(256)
    with _lat(m) as L:
        _ic(L, m)
        phi, pi = L.get_fields()
        dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 2)
        L.evolve(dtk, dtd)
        gphi, gpi = L.get_fields()
    ophi, opi = O.evolve(m, phi, pi, 2)
    assert np.array_equal(gphi, ophi) and np.array_equal(gpi, opi)


@pytest.mark.parametrize("N,prec", [(256, pn.PN_FP64), (256, pn.PN_FP32), (192, pn.PN_FP64)])
def test_parseval_binned_spectrum(N, prec):
    # sum over shells of |phi_k|^2 (with the half-spectrum double count) equals
    # a^6 N^3 sum_x phi^2 for the a^3-scaled unnormalised DFT
    m = O.model(N)
    with _lat(m, prec) as L:
        _ic(L, m, seed=5)
        phi, pi = L.get_fields()
        s = L.spectrum()
    a6 = m["a"] ** 6
    tol = 1e-12 if prec == pn.PN_FP64 else 1e-6
    assert abs(s["F"].sum() - a6 * N ** 3 * np.sum(phi * phi)) <= tol * a6 * N ** 3 * np.sum(phi * phi)
    assert abs(s["K"].sum() - a6 * N ** 3 * np.sum(pi * pi)) <= tol * a6 * N ** 3 * np.sum(pi * pi)
    H = N // 2 + 1
    assert s["counts"].sum() == N ** 3 and s["counts"].size == pn.lib().pn_spectrum_nbins(N, m["a"])
    del H


def test_time_reversal_256():
    # inverse of (kick k, drift d) x n is drift(-d) then (kick -k, drift -d)
    # pairs shifted by one, then kick(-k0): recovers the start to roundoff
    m = O.model(256)
    n = 60
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], n)
    with _lat(m) as L:
        _ic(L, m, seed=9)
        phi0, pi0 = L.get_fields()
        L.evolve(dtk, dtd)
        L.drift(-dtd[n - 1])
        bk = -dtk[1:][::-1]
        bd = -dtd[:-1][::-1]
        L.evolve(bk, bd)
        L.kick(-dtk[0])
        phi1, pi1 = L.get_fields()
    assert np.max(np.abs(phi1 - phi0)) <= 1e-10 * np.max(np.abs(phi0))
    assert np.max(np.abs(pi1 - pi0)) <= 1e-9 * np.max(np.abs(pi0))


def test_energy_bounded_1000_steps_256():
    m = O.model(256)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 1000)
    with _lat(m) as L:
        _ic(L, m, seed=3)
        e = [L.volume_averages().energy]
        for c in range(5):
            L.evolve(dtk[200 * c:200 * (c + 1)], dtd[200 * c:200 * (c + 1)])
            e.append(L.volume_averages().energy)
    e = np.array(e)
    assert np.all(np.isfinite(e)) and np.max(np.abs(e - e[0])) <= 1e-2 * abs(e[0])


def test_islands_128_cube_match_bfs():
    m = O.model(128)
    with _lat(m) as L:
        _ic(L, m, seed=11)
        dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 100)
        L.evolve(dtk, dtd)
        phi, pi = L.get_fields()
        thr = 2.57583 * np.sqrt(L.volume_averages().phi2)
        rows, lab, _ = L.islands(0, thr, capacity=1 << 20, labels=True)
    cat, labels = O.islands(m, 0, thr, phi, pi, cap=1 << 20)
    assert len(rows) == cat.shape[0] and np.array_equal(lab, labels)
    nn = np.array([r.nn for r in rows])
    assert np.array_equal(nn, cat[:, 1].astype(np.int64))
    assert np.array_equal(np.array([r.peak_pos for r in rows]), cat[:, 3].astype(np.int64))


def test_island_catalog_capacity_error_and_growth():
    # a catalog larger than the caller's array: PN_ERR_CAPACITY with the
    # count reported, and the wrapper's retry returns the whole catalog
    m = O.model(32)
    rng = np.random.default_rng(4)
    with _lat(m) as L:
        L.set_fields(rng.standard_normal(32 ** 3), rng.standard_normal(32 ** 3))
        full, _, _ = L.islands(0, 1.5, capacity=1 << 16)
        assert len(full) > 8
        with pytest.raises(pn.PhiNumError) as e:
            L.islands(0, 1.5, capacity=8, grow=False)
        assert e.value.code == pn.PN_ERR_CAPACITY
        grown, _, _ = L.islands(0, 1.5, capacity=8)
        arr, _, _ = L.islands(0, 1.5, capacity=8, as_array=True)
    assert [(r.index, r.nn, r.peak_pos) for r in grown] == [(r.index, r.nn, r.peak_pos) for r in full]
    assert arr["nn"].tolist() == [r.nn for r in full] and arr["peak_pos"].tolist() == [r.peak_pos for r in full]
