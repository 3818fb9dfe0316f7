"""GPU parity of the fused leapfrog step, reductions, spectrum and islands
against the oracle (bitwise where the reference's arithmetic is reproduced,
stated tolerances where the summation order necessarily differs)."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu


def _fields(N, seed, phi0=2.449489742783178, amp=0.4):
    rng = np.random.default_rng(seed)
    return phi0 + amp * rng.standard_normal(N ** 3), 0.2 * rng.standard_normal(N ** 3)


def _lat(m, precision=pn.PN_FP64):
    return pn.Lattice(m["N"], m["a"], m["mass_mode"], m["m2"], m["lambda_mode"], m["kappa"],
                      precision=precision)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("N,steps,seed", [(8, 30, 1), (12, 20, 2), (16, 20, 3), (33, 5, 4),
                                          (64, 10, 5), (100, 3, 6)])
def test_fused_step_bitwise_fp64(N, steps, seed):
    m = O.model(N)
    phi, pi = _fields(N, seed)
    ophi, opi = O.evolve(m, phi, pi, steps)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        L.evolve(dtk, dtd)
        gphi, gpi = L.get_fields()
    assert np.array_equal(_bits(gphi), _bits(ophi))
    assert np.array_equal(_bits(gpi), _bits(opi))


@pytest.mark.parametrize("N,steps,prec", [(4, 7, pn.PN_FP64), (6, 9, pn.PN_FP64), (10, 5, pn.PN_FP64),
                                          (16, 11, pn.PN_FP64), (34, 6, pn.PN_FP64), (130, 3, pn.PN_FP64),
                                          (256, 4, pn.PN_FP64), (260, 2, pn.PN_FP64),
                                          (4, 3, pn.PN_FP32), (8, 9, pn.PN_FP32), (36, 5, pn.PN_FP32),
                                          (256, 4, pn.PN_FP32), (260, 2, pn.PN_FP32), (520, 2, pn.PN_FP32)])
def test_two_step_kernel_bitwise_equals_single_steps(N, steps, prec):
    # temporal blocking changes only the evaluation order across sites
    m = O.model(N)
    phi, pi = _fields(N, 70 + N)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    out = []
    for two in (True, False):
        with _lat(m, prec) as L:
            if two:
                assert L.two_step(True)
            else:
                L.two_step(False)
            L.set_fields(phi, pi)
            L.evolve(dtk, dtd)
            out.append(L.get_fields())
    assert np.array_equal(_bits(out[0][0]), _bits(out[1][0]))
    assert np.array_equal(_bits(out[0][1]), _bits(out[1][1]))
    if prec == pn.PN_FP64 and N <= 34:
        ophi, opi = O.evolve(m, phi, pi, steps)
        assert np.array_equal(_bits(out[0][0]), _bits(ophi)) and np.array_equal(_bits(out[0][1]), _bits(opi))


@pytest.mark.parametrize("N,steps,prec", [(516, 3, pn.PN_FP64), (1024, 2, pn.PN_FP64), (1024, 2, pn.PN_FP32)])
def test_two_step_kernel_bitwise_equals_single_steps_large(N, steps, prec):
    # the N > 512 shapes, compared on the device: both runs start from the
    # same initial condition and max |d| over both fields must be exactly 0
    m = O.model(N)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    ic = pn.ic_params(5, a=m["a"], Lambda=m["Lambda"], phi0=float(np.sqrt(m["Lambda"])))
    print("two-step vs single steps N=%d prec=%d" % (N, prec), flush=True)
    with _lat(m, prec) as A, _lat(m, prec) as B:
        assert A.two_step(True)
        B.two_step(False)
        for L in (A, B):
            L.initial_conditions(ic)
            L.evolve(dtk, dtd)
        d = A.field_diff(B)
    assert d["max_dphi"] == 0.0 and d["max_dpi"] == 0.0
    assert d["sum_phi2"] > 0 and d["max_pi"] > 0


@pytest.mark.parametrize("N,prec", [(40, pn.PN_FP64), (40, pn.PN_FP32), (130, pn.PN_FP64)])
def test_two_step_zblocks_and_tuning_bitwise(N, prec):
    # every z block gives the bits of single steps, and so does the first
    # long evolve call, which times its own pairs of steps with each z block
    m = O.model(N)
    phi, pi = _fields(N, 90 + N)
    steps = 140  # the timed choice completes within these steps
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    with _lat(m, prec) as L:
        L.two_step(False)
        L.set_fields(phi, pi)
        L.evolve(dtk, dtd)
        ref = L.get_fields()
    runs = [8, 12, 24, N, 0]  # 0: two- or three-step kernel and z block chosen by timing
    for zb in runs:
        with _lat(m, prec) as L:
            assert L.two_step(True)
            if zb:
                L.three_step(False)
            L.step2_zblock(zb)
            L.set_fields(phi, pi)
            L.evolve(dtk, dtd)
            got = L.get_fields()
            used = L.step2_zblock()
            used3 = L.three_step()
        assert np.array_equal(_bits(got[0]), _bits(ref[0])) and np.array_equal(_bits(got[1]), _bits(ref[1])), zb
        if zb:
            assert used == zb and not used3
        else:
            assert used in (8, 12, 16, 24, 32, 48, 64, N)


@pytest.mark.parametrize("N,steps,prec,zb", [(8, 7, pn.PN_FP64, 0), (10, 8, pn.PN_FP64, 0), (16, 11, pn.PN_FP64, 0),
                                             (34, 6, pn.PN_FP64, 8), (130, 5, pn.PN_FP64, 0),
                                             (256, 6, pn.PN_FP64, 0), (256, 3, pn.PN_FP64, 24),
                                             (8, 9, pn.PN_FP32, 0), (36, 7, pn.PN_FP32, 16),
                                             (256, 6, pn.PN_FP32, 0), (260, 3, pn.PN_FP32, 0),
                                             (512, 3, pn.PN_FP32, 0)])
def test_three_step_kernel_bitwise_equals_single_steps(N, steps, prec, zb):
    # three steps per pass change only the evaluation order across sites
    m = O.model(N)
    phi, pi = _fields(N, 170 + N)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    out = []
    for three in (True, False):
        with _lat(m, prec) as L:
            if three:
                assert L.three_step(True, zblock=zb)
            else:
                L.two_step(False)
            L.set_fields(phi, pi)
            L.evolve(dtk, dtd)
            out.append(L.get_fields())
    assert np.array_equal(_bits(out[0][0]), _bits(out[1][0]))
    assert np.array_equal(_bits(out[0][1]), _bits(out[1][1]))
    if prec == pn.PN_FP64 and N <= 34:
        ophi, opi = O.evolve(m, phi, pi, steps)
        assert np.array_equal(_bits(out[0][0]), _bits(ophi)) and np.array_equal(_bits(out[0][1]), _bits(opi))


@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_multistep_choice_uses_only_the_callers_steps(prec):
    # the timed choice runs inside the caller's steps: with the three-step
    # kernel forced, and chosen by timing, a call of 130 steps from the
    # initial condition equals 130 single steps (the choice once ran past the
    # arrays)
    N = 64
    m = O.model(N)
    ic = pn.ic_params(1234, a=m["a"], Lambda=m["Lambda"], phi0=float(np.sqrt(m["Lambda"])))
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 130)
    out = []
    for mode in ("single", True, "auto"):
        with _lat(m, prec) as L:
            L.initial_conditions(ic)
            if mode == "single":
                L.two_step(False)
            else:
                L.three_step(mode)
            L.evolve(dtk, dtd)
            out.append(L.get_fields())
    for got in out[1:]:
        assert np.array_equal(_bits(got[0]), _bits(out[0][0])) and np.array_equal(_bits(got[1]), _bits(out[0][1]))
        assert np.all(np.isfinite(got[0]))


@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_multistep_choice_across_short_calls(prec):
    # calls of 7 to 50 steps (the drop-in flushes at every output): the timing
    # goes on from call to call, finishes, and every call's steps are the
    # single steps' bits; a call shorter than 12 steps does not time
    N = 64
    m = O.model(N)
    ic = pn.ic_params(1234, a=m["a"], Lambda=m["Lambda"], phi0=float(np.sqrt(m["Lambda"])))
    calls = [7, 13, 20, 30, 12, 50, 25, 40, 33, 28, 45]
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], sum(calls))
    with _lat(m, prec) as L:
        L.initial_conditions(ic)
        L.two_step(False)
        L.evolve(dtk, dtd)
        ref = L.get_fields()
    with _lat(m, prec) as L:
        L.initial_conditions(ic)
        s = 0
        timed = []
        for n in calls:
            L.evolve(dtk[s:s + n], dtd[s:s + n])
            s += n
            timed.append(L.kernel_timed())
        got = L.get_fields()
    assert not timed[0] and timed[-1]
    assert np.array_equal(_bits(got[0]), _bits(ref[0])) and np.array_equal(_bits(got[1]), _bits(ref[1]))


def test_two_step_unavailable_is_reported():
    m = O.model(33)  # odd N: no vector width fits
    with _lat(m) as L:
        assert not L.two_step()
        with pytest.raises(pn.PhiNumError):
            L.two_step(True)


def test_gpu_matches_reference_fixtures():
    # the HIP path against the reference's own outputs (tests/golden), no oracle
    import glob
    import os
    for path in sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))):
        z = np.load(path)
        N = int(z["N"])
        steps = int(z["steps"])
        dtk, dtd, _ = pn.time_factors(float(z["T0"]), float(z["dtbase"]), float(z["a"]), steps)
        with pn.Lattice(N, float(z["a"]), int(z["mass_mode"]), float(z["m2"]), int(z["lambda_mode"]),
                        float(z["kappa"])) as L:
            L.set_fields(z["phi_in"], z["pi_in"])
            L.evolve(dtk, dtd)
            gphi, gpi = L.get_fields()
            assert np.array_equal(_bits(gphi), _bits(z["phi_out"])), path
            assert np.array_equal(_bits(gpi), _bits(z["pi_out"])), path
            ext = L.extrema()
            assert (ext.min, ext.max, ext.maxabs) == tuple(z["ext"])
            assert (ext.minpos, ext.maxpos, ext.maxabspos) == tuple(int(i) for i in z["extpos"])
            np.testing.assert_allclose(L.volume_averages().as_array(), z["avg"], rtol=1e-13, atol=1e-15)
            if int(z["islands"]):
                rows, lab, _ = L.islands(0, float(z["thr"]), labels=True)
                assert np.array_equal(lab, z["labels"])
                cat = z["catalog"]
                assert len(rows) == cat.shape[0]
                for r, c in zip(rows, cat):
                    assert (r.index, r.nn, r.peak_pos) == (int(c[0]), int(c[1]), int(c[3]))
                    assert r.volume == c[2] and r.peak_value == c[4]
                    np.testing.assert_allclose([r.avg_phi, r.pot_en, r.kin_en, r.grad_en], c[5:9],
                                               rtol=1e-12, atol=1e-14)


def test_non_pow2_spacing_and_other_couplings_bitwise():
    m = dict(N=20, a=0.37, mass_mode=-1, m2=0.7, lambda_mode=1, kappa=1.9, Lambda=3.0)
    phi, pi = _fields(20, 9, phi0=0.3)
    ophi, opi = O.evolve(m, phi, pi, 15, T0=1.25, dtbase=0.07)
    dtk, dtd, _ = pn.time_factors(1.25, 0.07, m["a"], 15)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        L.evolve(dtk, dtd)
        gphi, gpi = L.get_fields()
    assert np.array_equal(_bits(gphi), _bits(ophi)) and np.array_equal(_bits(gpi), _bits(opi))


@pytest.mark.parametrize("N,huge", [(16, False), (16, True), (15, False), (15, True), (64, False),
                                    (64, True)])
def test_division_outside_fast_range_bitwise(N, huge):
    # sites whose |phi| is 0, -0, below 2^-100 (tiny, subnormal) or above
    # 2^100 make their wave take the IEEE divisions; the rest of the lattice
    # keeps the four-FMA form: both must give the oracle's bits (a tiny time
    # step keeps the huge sites finite over both steps)
    m = O.model(N)
    phi, pi = _fields(N, 90 + N)
    rng = np.random.default_rng(N)
    idx = rng.choice(N ** 3, 12, replace=False)
    phi[idx[:8]] = [0.0, -0.0, 1e-40, -3e-35, 5e-324, 1e-310, 2.0 ** -100, -2.0 ** -101]
    dtbase = 0.1
    if huge:
        phi[idx[8:10]] = [1.5e30, -2.0 ** 100 * 1.5]
        dtbase = 1e-160
    ophi, opi = O.evolve(m, phi, pi, 2, dtbase=dtbase)
    assert np.all(np.isfinite(ophi)) and np.all(np.isfinite(opi))
    dtk, dtd, _ = pn.time_factors(0.0, dtbase, m["a"], 2)
    for two in (True, False):
        with _lat(m) as L:
            if two and not L.two_step():
                continue  # odd N: single steps only
            L.two_step(two)
            L.set_fields(phi, pi)
            L.evolve(dtk, dtd)
            gphi, gpi = L.get_fields()
            assert np.array_equal(_bits(gphi), _bits(ophi)) and np.array_equal(_bits(gpi), _bits(opi))


@pytest.mark.parametrize("N,steps,seed", [(16, 20, 61), (18, 10, 62), (64, 10, 63), (100, 4, 64)])
def test_fused_step_bitwise_fp32(N, steps, seed):
    # the FP32 form against its binary32 oracle (vector path when N % 4 == 0)
    m = O.model(N)
    phi, pi = _fields(N, seed)
    ophi, opi = O.evolve_f32(m, phi, pi, steps)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], steps)
    with _lat(m, pn.PN_FP32) as L:
        L.set_fields(phi, pi)
        L.evolve(dtk, dtd)
        gphi, gpi = L.get_fields()
    assert np.array_equal(gphi.astype(np.float32), ophi)
    assert np.array_equal(gpi.astype(np.float32), opi)


def test_fp32_energy_tracks_fp64():
    # FP32 fields with FP64 compensated sums: relative energy difference to
    # the FP64 run stays below 1e-6 over 2000 steps at 32^3
    N = 32
    m = O.model(N)
    phi, pi = _fields(N, 65, amp=0.2)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 2000)
    with _lat(m) as L64, _lat(m, pn.PN_FP32) as L32:
        L64.set_fields(phi, pi)
        L32.set_fields(phi, pi)
        for c in range(10):
            sl = slice(200 * c, 200 * (c + 1))
            L64.evolve(dtk[sl], dtd[sl])
            L32.evolve(dtk[sl], dtd[sl])
            e64 = L64.volume_averages().energy
            e32 = L32.volume_averages().energy
            assert abs(e32 - e64) <= 1e-6 * abs(e64)


def test_kick_then_drift_equals_reference_functions():
    # updateMomenta / updateFields / exchange called one by one
    N = 16
    m = O.model(N)
    phi, pi = _fields(N, 12)
    dtk, dtd, _ = O.time_factors(4)
    rphi, rpi = O.to_ref(phi, N), O.to_ref(pi, N)
    p = O.params(m)
    import ctypes as C
    with _lat(m) as L:
        L.set_fields(rphi, rpi, layout=pn.PN_LAYOUT_REF)
        for n in range(4):
            O.lib().or_update_momenta(C.byref(p), O.dptr(rphi), O.dptr(rpi), dtk[n])
            O.lib().or_update_fields(N, O.dptr(rphi), O.dptr(rpi), dtd[n])
            O.lib().or_exchange(N, O.dptr(rphi))
            L.kick(dtk[n])
            L.drift(dtd[n])
        gphi, gpi = L.get_fields(layout=pn.PN_LAYOUT_REF)
    assert np.array_equal(_bits(gphi), _bits(rphi))
    # the reference never exchanges Pi: compare its interior only
    P = N * N
    assert np.array_equal(_bits(gpi[P:P * (N + 1)]), _bits(rpi[P:P * (N + 1)]))


def test_layout_roundtrip():
    N = 12
    m = O.model(N)
    phi, pi = _fields(N, 13)
    with _lat(m) as L:
        L.set_fields(O.to_ref(phi, N), O.to_ref(pi, N), layout=pn.PN_LAYOUT_REF)
        dphi, dpi = L.get_fields(layout=pn.PN_LAYOUT_DENSE)
        rphi, _ = L.get_fields(layout=pn.PN_LAYOUT_REF)
    assert np.array_equal(dphi, phi) and np.array_equal(dpi, pi)
    assert np.array_equal(rphi, O.to_ref(phi, N))


def test_wrap_unique_values():
    # a unique value per site: one kick with only the Laplacian must see the
    # periodic neighbours of every site (comm.cpp:63-78, util.cpp:12-18)
    N = 4
    m = dict(N=N, a=1.0, mass_mode=0, m2=0.0, lambda_mode=0, kappa=1e300, Lambda=1.0)
    phi = np.arange(N ** 3, dtype=np.float64) * 1.0
    pi = np.zeros(N ** 3)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        L.kick(1.0)
        _, gpi = L.get_fields()
    f = phi.reshape(N, N, N)
    lap = (np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1)
           + np.roll(f, 1, 2) + np.roll(f, -1, 2) - 6 * f)
    assert np.array_equal(gpi, lap.ravel())


@pytest.mark.parametrize("N,seed", [(8, 21), (16, 22), (48, 23)])
def test_volume_averages_and_extrema(N, seed):
    m = O.model(N)
    phi, pi = _fields(N, seed)
    ref = O.volume_averages(m, phi, pi)
    vals, idx = O.extrema(N, phi)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        avg = L.volume_averages().as_array()
        ext = L.extrema()
    # sums are compensated on the GPU and sequential in the reference: 1e-13
    np.testing.assert_allclose(avg, ref, rtol=1e-13, atol=1e-15)
    assert (ext.min, ext.max, ext.maxabs) == tuple(vals)
    assert (ext.minpos, ext.maxpos, ext.maxabspos) == tuple(int(i) for i in idx)


@pytest.mark.parametrize("N,prec", [(16, pn.PN_FP64), (48, pn.PN_FP64), (32, pn.PN_FP32)])
def test_volume_analysis_is_the_two_calls(N, prec):
    # pn_volume_analysis: averages and extrema from one pass, the same bits as
    # pn_volume_averages and pn_extrema_get called one after the other; with
    # no averages argument it still fills the extrema
    import ctypes as C
    m = O.model(N)
    phi, pi = _fields(N, 40 + N)
    with _lat(m, prec) as L:
        L.set_fields(phi, pi)
        avg, ext = L.volume_analysis()
        want_avg, want_ext = L.volume_averages(), L.extrema()
        only = pn.PnExtrema()
        L._chk(L.L.pn_volume_analysis(L.h, None, C.byref(only)))
    assert np.array_equal(_bits(avg.as_array()), _bits(want_avg.as_array()))
    for e in (ext, only):
        assert (e.min, e.max, e.maxabs) == (want_ext.min, want_ext.max, want_ext.maxabs)
        assert (e.minpos, e.maxpos, e.maxabspos) == (want_ext.minpos, want_ext.maxpos, want_ext.maxabspos)
    if prec == pn.PN_FP64:
        np.testing.assert_allclose(avg.as_array(), O.volume_averages(m, phi, pi), rtol=1e-13, atol=1e-15)
        vals, idx = O.extrema(N, phi)
        assert (ext.min, ext.max, ext.maxabs) == tuple(vals)
        assert (ext.minpos, ext.maxpos, ext.maxabspos) == tuple(int(i) for i in idx)


@pytest.mark.parametrize("N,prec,kappa", [(33, pn.PN_FP64, None), (64, pn.PN_FP64, None),
                                           (64, pn.PN_FP32, None), (40, pn.PN_FP64, 1e300)])
def test_volume_sums_outside_fast_range(N, prec, kappa):
    # waves holding zeros, -0, subnormals, tiny (< 2^-100) or huge (> 2^100)
    # sites take the IEEE divisions and square root in a second pass; kappa =
    # 1e300 puts the phi^6 divisor outside the fast form's range, so every
    # site does; values against the oracle, extrema bitwise
    m = O.model(N)
    if kappa is not None:
        m = dict(m, kappa=kappa)
    phi, pi = _fields(N, 70 + N)
    rng = np.random.default_rng(N)
    sites = rng.choice(N ** 3, 40, replace=False)
    special = [0.0, -0.0, 5e-320, -1e-310, 1e-120, 2.0 ** 120, -3.0 ** 70, 1e-35]
    for k, i in enumerate(sites):
        phi[i] = special[k % len(special)]
    if prec == pn.PN_FP32:
        phi = phi.astype(np.float32).astype(np.float64)
        pi = pi.astype(np.float32).astype(np.float64)
    ref = O.volume_averages(m, phi, pi)
    vals, idx = O.extrema(N, phi)
    with _lat(m, prec) as L:
        L.set_fields(phi, pi)
        avg, ext = L.volume_analysis()
        again, _ = L.volume_analysis()
    assert np.array_equal(_bits(avg.as_array()), _bits(again.as_array()))
    np.testing.assert_allclose(avg.as_array(), ref, rtol=1e-13, atol=1e-15)
    assert _bits(np.array([ext.min, ext.max, ext.maxabs])).tolist() == _bits(np.array(vals)).tolist()
    assert (ext.minpos, ext.maxpos, ext.maxabspos) == tuple(int(i) for i in idx)


def test_extrema_of_signed_zeros():
    # all zeros with -0 at the first site: the reference's strict comparisons
    # keep the first site's -0 as both extremes (fieldanalysis.cpp:331-355)
    N = 16
    m = O.model(N)
    phi = np.zeros(N ** 3)
    phi[0] = -0.0
    phi[N ** 3 // 3] = -0.0
    pi = np.zeros(N ** 3)
    vals, idx = O.extrema(N, phi)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        ext = L.extrema()
    assert _bits(np.array([ext.min, ext.max, ext.maxabs])).tolist() == _bits(np.array(vals)).tolist()
    assert (ext.minpos, ext.maxpos, ext.maxabspos) == tuple(int(i) for i in idx)


@pytest.mark.parametrize("N,steps", [(16, 7), (33, 4)])
def test_synced_energy_is_the_half_kicked_averages(N, steps):
    # pn_volume_averages_synced: the averages of (phi, pi + half kick), the
    # kick of updateMomenta with factor dtau/2 (dynamics.cpp:52-175); the
    # oracle's kick on a copy of pi, then its averages; the fields unchanged
    # and the next steps the same bits as without the call
    import ctypes as C
    m = O.model(N)
    phi0, pi0 = _fields(N, 90 + N)
    dtk, dtd = pn.time_factors(0.0, 0.1, m["a"], 2 * steps)[:2]
    half = 0.5 * 0.1 * m["a"]
    with _lat(m) as L:
        L.set_fields(phi0, pi0)
        L.evolve(dtk[:steps], dtd[:steps])
        f, p = L.get_fields()
        syn = L.volume_averages_synced(half).as_array()
        f2, p2 = L.get_fields()
        L.evolve(dtk[steps:], dtd[steps:])
        fa, pa = L.get_fields()
    assert np.array_equal(_bits(f), _bits(f2)) and np.array_equal(_bits(p), _bits(p2))
    phi = O.to_ref(f, N)
    pik = O.to_ref(p, N)
    prm = O.params(m)
    O.lib().or_update_momenta(C.byref(prm), O.dptr(phi), O.dptr(pik), half)
    want = O.volume_averages(m, f, O.from_ref(pik, N))
    np.testing.assert_allclose(syn, want, rtol=1e-13, atol=1e-15)
    ef, ep = O.evolve_tables(m, f, p, dtk[steps:], dtd[steps:])
    assert np.array_equal(_bits(fa), _bits(ef)) and np.array_equal(_bits(pa), _bits(ep))


def test_extrema_ties_pick_first_site():
    N = 8
    m = O.model(N)
    phi = np.zeros(N ** 3)
    phi[[100, 7, 300]] = 3.0
    phi[[50, 400]] = -3.0
    pi = np.zeros(N ** 3)
    vals, idx = O.extrema(N, phi)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        ext = L.extrema()
    assert (ext.minpos, ext.maxpos, ext.maxabspos) == tuple(int(i) for i in idx)
    assert (ext.min, ext.max, ext.maxabs) == tuple(vals)


@pytest.mark.parametrize("N", [16, 64, 128])
def test_spectrum_counts_and_sums(N):
    m = O.model(N)
    phi, pi = _fields(N, 31 + N)
    counts = O.bin_table(N, m["a"])
    F, K, Q = O.spectrum(N, m["a"], phi, pi)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        s = L.spectrum()
    assert s["nbins"] == counts.size
    assert np.array_equal(s["counts"], counts)
    np.testing.assert_allclose(s["F"], F, rtol=1e-12, atol=1e-12 * F.max())
    np.testing.assert_allclose(s["K"], K, rtol=1e-12, atol=1e-12 * K.max())
    np.testing.assert_allclose(s["Q"], Q, rtol=1e-11, atol=1e-12 * np.abs(Q).max())


def test_fft_modes_match_reference_layout():
    N = 16
    m = O.model(N)
    phi, pi = _fields(N, 41)
    ref = O.ref_modes(phi, N, m["a"])
    with _lat(m) as L:
        L.set_fields(phi, pi)
        got = L.fft_modes(0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("N,seed,mode", [(8, 51, 0), (16, 52, 0), (32, 53, 0), (16, 54, 1),
                                         (32, 55, 1)])
def test_islands_match_bfs(N, seed, mode):
    m = O.model(N)
    phi, pi = _fields(N, seed, phi0=0.0, amp=1.0)
    if mode == 0:
        thr = 2.57583 * np.sqrt(O.volume_averages(m, phi, pi)[2])
    else:
        thr = 4.0 * O.volume_averages(m, phi, pi)[7]
    cat, labels = O.islands(m, mode, thr, phi, pi)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        rows, glab, _ = L.islands(mode, thr, labels=True)
    assert len(rows) == cat.shape[0] > 0
    assert np.array_equal(glab, labels)
    P = N * N
    for r, c in zip(rows, cat):
        assert r.index == int(c[0]) and r.nn == int(c[1])
        assert r.volume == c[2]
        assert r.peak_pos == int(c[3])
        assert r.peak_value == c[4]
        np.testing.assert_allclose([r.avg_phi, r.pot_en, r.kin_en, r.grad_en], c[5:9],
                                   rtol=1e-12, atol=1e-14)
        members = np.nonzero(labels == r.index)[0]
        assert r.first_pos == members.min() + P


def _check_centroids(rows, lab, N):
    P = N * N
    for r in rows:
        sites = np.nonzero(lab == r.index)[0]
        pk = r.peak_pos - P
        px, py, pz = pk % N, (pk // N) % N, pk // P
        x, y, z = sites % N, (sites // N) % N, sites // P
        dx = (x - px + N // 2) % N - N // 2
        dy = (y - py + N // 2) % N - N // 2
        assert (r.nn, r.sum_dx, r.sum_dy, r.sum_dz) == (sites.size, dx.sum(), dy.sum(), (z - pz).sum())
        assert r.cz == pz + (z - pz).sum() / sites.size
        assert abs(r.cx - (px + dx.sum() / sites.size) % N) < 1e-12
        assert abs(r.cy - (py + dy.sum() / sites.size) % N) < 1e-12


def test_island_centroids_brute_force():
    # the centroid is a new output (the reference keeps only the peak site):
    # offsets from the peak, minimum image in x and y (periodic), plain in z
    # (open), checked against a numpy restatement on a random lattice and on
    # a planted island that straddles the periodic x and y edges
    N = 16
    m = O.model(N)
    P = N * N
    rng = np.random.default_rng(81)
    phi = rng.standard_normal(N ** 3)
    pi = 0.1 * rng.standard_normal(N ** 3)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        rows, lab, _ = L.islands(0, 1.5, labels=True)
        assert len(rows) > 20
        _check_centroids(rows, lab, N)
        bg = np.clip(phi, -1.0, 1.0).reshape(N, N, N)  # background stays below thr = 3
        for x in (14, 15, 0, 1):
            for y in (15, 0):
                bg[9, y, x] = 5.0
        bg[9, 0, 15] = 6.0  # the peak
        L.set_fields(bg.ravel(), pi)
        rows, lab, _ = L.islands(0, 3.0, labels=True)
        _check_centroids(rows, lab, N)
    assert len(rows) == 1 and rows[0].nn == 8 and rows[0].peak_pos == P + 9 * P + 15
    # x sites 14, 15, 0, 1 around 15: mean 15.5; y sites 15, 0 around 0: mean -0.5 -> 15.5
    assert (rows[0].cx, rows[0].cy, rows[0].cz) == (15.5, 15.5, 9.0)


@pytest.mark.parametrize("N,seed", [(32, 91), (48, 92), (64, 93)])
def test_island_threshold_from_each_sides_averages(N, seed):
    # the island threshold 2.57583 sqrt(<phi^2>) (main.cpp:70) taken from each
    # side's own sums: the device's compensated tree and the reference's
    # sequential loop may differ in the last bits, so the masks agree except
    # at sites whose phi^2/2 + pi^2/2 (islands.cpp:76) lies between the two
    # thresholds' squares over 2
    m = O.model(N)
    phi, pi = _fields(N, seed, phi0=0.0, amp=1.0)
    thr_ref = 2.57583 * np.sqrt(O.volume_averages(m, phi, pi)[2])
    with _lat(m) as L:
        L.set_fields(phi, pi)
        thr_gpu = 2.57583 * np.sqrt(L.volume_averages().phi2)
        rows, glab, _ = L.islands(0, thr_gpu, labels=True)
    _, olab = O.islands(m, 0, thr_ref, phi, pi)
    assert abs(thr_gpu - thr_ref) <= 1e-13 * thr_ref
    assert len(rows) > 0
    e = phi * phi * 0.5 + pi * pi * 0.5
    diff = (glab >= 0) != (olab >= 0)
    lo, hi = sorted((thr_gpu, thr_ref))
    band = (e > lo * lo * 0.5 * (1 - 1e-14)) & (e <= hi * hi * 0.5 * (1 + 1e-14))
    assert np.all(band[diff])
    if thr_gpu == thr_ref:
        assert np.array_equal(glab, olab)


@pytest.mark.parametrize("mode", [0, 1])
def test_large_islands_match_bfs(mode):
    # a threshold low enough that one island spans most of a 40^3 lattice: its
    # statistics come from chunks of 4096 sites (one atomic per chunk for the
    # integers, an ordered merge of the chunks' sums); the small islands from
    # one wave each.  Catalog and labels against the oracle's BFS, centroids by
    # brute force, and two planted equal maxima in the large island: one tie,
    # the peak the one the BFS from the island's first site reaches first
    N = 40
    m = O.model(N)
    P = N * N
    phi, pi = _fields(N, 61 + mode, phi0=0.0, amp=1.0)
    if mode == 0:
        thr = 0.6 * np.sqrt(O.volume_averages(m, phi, pi)[2])
        # a site of the largest island and three x neighbours after it: two
        # equal maxima (9.0) joined through two sites of 8.0
        cat0, lab0 = O.islands(m, 0, thr, phi, pi)
        big0 = int(np.argmax(cat0[:, 1]))
        i1 = int(np.nonzero(lab0 == big0)[0][0])
        row = i1 - i1 % N
        i2 = row + (i1 % N + 3) % N
        phi = phi.copy()
        phi[i1] = phi[i2] = 9.0
        phi[row + (i1 % N + 1) % N] = phi[row + (i1 % N + 2) % N] = 8.0
    else:
        thr = 0.5 * O.volume_averages(m, phi, pi)[7]
    cat, labels = O.islands(m, mode, thr, phi, pi)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        rows, glab, _ = L.islands(mode, thr, labels=True)
    sizes = np.array([r.nn for r in rows])
    assert sizes.max() > 4096 and sizes.min() <= 4096
    assert len(rows) == cat.shape[0]
    assert np.array_equal(glab, labels)
    for r, c in zip(rows, cat):
        assert r.index == int(c[0]) and r.nn == int(c[1])
        np.testing.assert_allclose([r.avg_phi, r.pot_en, r.kin_en, r.grad_en], c[5:9],
                                   rtol=1e-12, atol=1e-14)
        members = np.nonzero(labels == r.index)[0]
        assert r.first_pos == members.min() + P
        assert r.peak_pos == int(c[3]) and r.peak_value == c[4]
    _check_centroids(rows, glab, N)
    if mode == 0:
        big = rows[int(np.argmax(sizes))]
        assert glab[i1] == big.index and glab[i2] == big.index
        assert big.peak_ties == 1 and big.peak_pos in (i1 + P, i2 + P) and big.peak_value == 9.0


def test_catalog_into_pageable_and_page_locked_arrays():
    # the same rows, byte for byte, into a caller's pageable array and into
    # page-locked memory from pn_host_alloc, through the C ABI; then a
    # catalog larger than the array: the count, PN_ERR_CAPACITY and the
    # first rows only
    N = 32
    m = O.model(N)
    phi, pi = _fields(N, 57, phi0=0.0, amp=1.0)
    sz = C.sizeof(pn.PnIsland)
    n = N ** 3
    with _lat(m) as L:
        L.set_fields(phi, pi)
        buf = np.zeros(n * sz, np.uint8)
        cat = pn.PnCatalog(n, 0, 0.0, C.cast(buf.ctypes.data, C.POINTER(pn.PnIsland)), None)
        assert L.L.pn_islands(L.h, 0, 1.5, C.byref(cat)) == pn.PN_OK
        assert cat.count > 100
        p = C.c_void_p()
        assert L.L.pn_host_alloc(n * sz, C.byref(p)) == pn.PN_OK
        try:
            catp = pn.PnCatalog(n, 0, 0.0, C.cast(p, C.POINTER(pn.PnIsland)), None)
            assert L.L.pn_islands(L.h, 0, 1.5, C.byref(catp)) == pn.PN_OK
            assert catp.count == cat.count
            assert C.string_at(p, cat.count * sz) == buf[:cat.count * sz].tobytes()
            few = pn.PnCatalog(7, 0, 0.0, C.cast(p, C.POINTER(pn.PnIsland)), None)
            C.memset(p, 0, n * sz)
            assert L.L.pn_islands(L.h, 0, 1.5, C.byref(few)) == pn.PN_ERR_CAPACITY
            assert few.count == cat.count
            assert C.string_at(p, 7 * sz) == buf[:7 * sz].tobytes()
            assert C.string_at(p.value + 7 * sz, sz) == bytes(sz)
        finally:
            assert L.L.pn_host_free(p) == pn.PN_OK
        rows, _, _ = L.islands(0, 1.5, as_array=True)
        assert rows.tobytes() == buf[:cat.count * sz].tobytes()


@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_tied_peak_is_the_first_in_bfs_order(prec):
    # an L-shaped island in one plane: a row y = 3, x = 2..8, and a column
    # x = 2, y = 3..9.  Its first site (2, 3) seeds the reference's BFS; equal
    # phi^2 at (8, 3) (phi = +9, the smaller index) and (2, 5) (phi = -9, two
    # steps from the seed against six): the BFS keeps (2, 5) and Phi = -9
    # (islands.cpp:104-116), where the smallest index would give (8, 3)
    N = 16
    m = O.model(N)
    P = N * N
    phi = np.zeros((N, N, N))
    z = 5
    for x in range(2, 9):
        phi[z, 3, x] = 4.0
    for y in range(3, 10):
        phi[z, y, 2] = 4.0
    phi[z, 3, 8] = 9.0
    phi[z, 5, 2] = -9.0
    phi = phi.ravel()
    pi = np.zeros(N ** 3)
    cat, labels = O.islands(m, 0, 3.0, phi, pi)
    with _lat(m, prec) as L:
        L.set_fields(phi, pi)
        rows, glab, _ = L.islands(0, 3.0, labels=True)
    assert len(rows) == 1 and cat.shape[0] == 1
    r = rows[0]
    assert r.peak_ties == 1
    assert r.peak_pos == int(cat[0][3]) == P + z * P + 5 * N + 2
    assert r.peak_value == cat[0][4] == -9.0
    _check_centroids(rows, glab, N)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_quantized_fields_peaks_match_bfs(mode, prec):
    # phi and pi on a grid of quarter units: many islands hold equal maxima;
    # every peak site and value against the oracle's BFS (binary32 holds these
    # values exactly, so both lattices see the same fields)
    N = 24
    m = O.model(N)
    rng = np.random.default_rng(77 + mode)
    phi = np.round(4.0 * (0.8 * rng.standard_normal(N ** 3))) / 4.0
    pi = np.round(4.0 * (0.3 * rng.standard_normal(N ** 3))) / 4.0
    if mode == 0:
        thr = 1.6
    else:
        thr = 2.0 * O.volume_averages(m, phi, pi)[7]
    cat, labels = O.islands(m, mode, thr, phi, pi)
    with _lat(m, prec) as L:
        L.set_fields(phi, pi)
        rows, glab, _ = L.islands(mode, thr, labels=True)
    assert np.array_equal(glab, labels)
    assert len(rows) == cat.shape[0] > 20
    if mode == 0:
        assert sum(r.peak_ties > 0 for r in rows) > 5
    for r, c in zip(rows, cat):
        assert (r.peak_pos, r.peak_value) == (int(c[3]), c[4]), r.index
    _check_centroids(rows, glab, N)


@pytest.mark.parametrize("kappa", [1.2247448713915889, 0.46775209018414249])
@pytest.mark.parametrize("N", [16, 15, 64])
def test_kappa_division_paths_bitwise(N, kappa):
    # q/kappa takes the two-operation form when kappa passes pn_bmr_exact
    # (sqrt(1.5) does) and the corrected four-operation form otherwise (the
    # second value fails it); both must give the oracle's bits
    m = dict(O.model(N), kappa=kappa)
    phi, pi = _fields(N, 95 + N, phi0=0.8, amp=0.5)
    ophi, opi = O.evolve(m, phi, pi, 6)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, m["a"], 6)
    with _lat(m) as L:
        L.set_fields(phi, pi)
        L.evolve(dtk, dtd)
        gphi, gpi = L.get_fields()
    assert np.array_equal(_bits(gphi), _bits(ophi)) and np.array_equal(_bits(gpi), _bits(opi))
