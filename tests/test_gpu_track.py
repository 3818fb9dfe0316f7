"""Island tracking over time on the GPU (pn_tracker_*, through the C ABI)
against the reference fixtures (tests/golden/track/, the reference's own
Full_Analysis run repeatedly) and the CPU checker: island numbers, statuses,
history lengths, peak sites, times and volumes exactly; energy and mean-field
sums to 1e-12 (their summation order differs from the BFS order); the text
files row by row to the 6 significant digits they are printed with."""
import glob
import os
import tempfile

import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "track", "*.npz")))


def _run_gpu(m, phi, pi, analyses, steps, thr, mode=0):
    dtk, dtd, taus = pn.time_factors(0.0, 0.1, m["a"], steps * (analyses - 1))
    with pn.Lattice(m["N"], m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"]) as L, \
            pn.Tracker() as T, tempfile.TemporaryDirectory() as td:
        L.set_fields(phi, pi)
        for k in range(analyses):
            if k > 0:
                sl = slice(steps * (k - 1), steps * k)
                L.evolve(dtk[sl], dtd[sl])
            T.update(L, mode, thr, 0.0 if k == 0 else taus[steps * k - 1], td)
        files = {fn: open(os.path.join(td, fn)).read() for fn in sorted(os.listdir(td))}
        return T.islands(), files


def _same_track(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[:3] == w[:3], (g[:3], w[:3])
        assert len(g[3]) == len(w[3])
        for eg, ew in zip(g[3], w[3]):
            # at_time, peak pos, peak value, volume exact
            assert (eg[0], eg[1], eg[2], eg[4]) == (ew[0], ew[1], ew[2], ew[4])
            np.testing.assert_allclose(np.array(eg)[[3, 5, 6, 7]], np.array(ew)[[3, 5, 6, 7]],
                                       rtol=1e-12, atol=1e-13)


def _same_files(got, want):
    assert sorted(got) == sorted(want)
    for name in want:
        lg, lw = got[name].splitlines(), want[name].splitlines()
        assert len(lg) == len(lw), name
        for a, b in zip(lg, lw):
            ta, tb = a.split(), b.split()
            assert len(ta) == len(tb), (name, a, b)
            np.testing.assert_allclose([float(t) for t in ta], [float(t) for t in tb], rtol=2e-5,
                                       atol=1e-12, err_msg=name)


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_tracking_matches_reference_fixture(path):
    m, z, track, files = O.load_track_fixture(path)
    got, gfiles = _run_gpu(m, z["phi_in"], z["pi_in"], int(z["analyses"]), int(z["steps"]),
                           float(z["thr"]))
    _same_track(got, track)
    _same_files(gfiles, files)


@pytest.mark.parametrize("mode,thr", [(0, 2.0), (1, 60.0)])
def test_tracking_matches_oracle(mode, thr):
    # mode 1 (Delta islands) has no reference build to compare with
    N = 16
    m = O.model(N)
    rng = np.random.default_rng(40 + mode)
    phi = 1.0 * rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    want, wfiles = O.track_replay(m, phi, pi, 8, 5, thr, mode=mode)
    got, gfiles = _run_gpu(m, phi, pi, 8, 5, thr, mode=mode)
    assert len(want) > 5
    _same_track(got, want)
    _same_files(gfiles, wfiles)


@pytest.mark.parametrize("mode", [0, 1])
def test_tracking_with_tied_peaks(mode):
    # quantized fields (quarter units) give many islands equal maxima; a
    # tracked island re-grows from its last peak (islands.cpp:798-817), so its
    # new peak is the first of the equal maxima in breadth-first order from
    # there, and that peak is the seed of the next analysis.  Three analyses
    # on related fields, against the checker's Full_Analysis
    N = 20
    m = O.model(N)
    rng = np.random.default_rng(300 + mode)
    base = 0.9 * rng.standard_normal(N ** 3)
    pi = np.round(4.0 * 0.3 * rng.standard_normal(N ** 3)) / 4.0
    thr = 1.6 if mode == 0 else 2.0 * O.volume_averages(m, np.round(4 * base) / 4, pi)[7]
    T = O.Tracker(m)
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"]) as L, \
            pn.Tracker() as G, tempfile.TemporaryDirectory() as gd, tempfile.TemporaryDirectory() as od:
        for k in range(3):
            phi = np.round(4.0 * (base + 0.1 * k * rng.standard_normal(N ** 3))) / 4.0
            T.update(mode, thr, phi, pi, 1.0 * k, od)
            L.set_fields(phi, pi)
            G.update(L, mode, thr, 1.0 * k, gd)
        got, want = G.islands(), T.islands()
        gfiles = {fn: open(os.path.join(gd, fn)).read() for fn in sorted(os.listdir(gd))}
        wfiles = {fn: open(os.path.join(od, fn)).read() for fn in sorted(os.listdir(od))}
    T.close()
    assert len(want) > 20
    _same_track(got, want)
    _same_files(gfiles, wfiles)


def test_tracker_catalog_is_the_update_components():
    # pn_tracker_catalog returns the catalog the update computed, as
    # pn_islands gives it for the same fields and threshold
    N = 24
    m = O.model(N)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"]) as L:
        L.set_fields(phi, pi)
        T = pn.Tracker()
        T.update(L, 0, 1.2, 0.0)
        got = T.catalog()
        want, _, _ = L.islands(0, 1.2, as_array=True)
    assert len(got) == len(want) > 100
    assert got.tobytes() == want.tobytes()


def test_lookup_matches_catalog_labels():
    N = 12
    m = O.model(N)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    with pn.Lattice(N, m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"]) as L:
        L.set_fields(phi, pi)
        rows, lab, _ = L.islands(0, 1.5, labels=True)
        pos = np.arange(N ** 3, dtype=np.int64) + N * N  # every interior site, pos() order
        out = np.empty(N ** 3, dtype=np.int32)
        import ctypes as C
        rc = pn.lib().pn_islands_lookup(L.h, N ** 3, pos.ctypes.data_as(C.POINTER(C.c_longlong)),
                                        out.ctypes.data_as(C.POINTER(C.c_int)))
        assert rc == 0
        assert np.array_equal(out, lab)
        bad = np.array([N * N - 1], dtype=np.int64)  # a halo site
        rc = pn.lib().pn_islands_lookup(L.h, 1, bad.ctypes.data_as(C.POINTER(C.c_longlong)),
                                        out.ctypes.data_as(C.POINTER(C.c_int)))
        assert rc == pn.PN_ERR_ARG
