"""Batched realizations (mixed sizes, per-realization couplings, one launch
per step, HIP graph replay, autotuned tiles) against single-lattice runs and
the oracle: every realization must be bitwise what it would be alone."""
import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu


def _reals(sizes, gs):
    out = []
    for N, g in zip(sizes, gs):
        m = O.model(N, lam=1.0, g=g)
        out.append(dict(N=N, a=0.5, mass_mode=1, m2=1.0, lambda_mode=-1, kappa=m["kappa"]))
    return out


def _init(N, seed):
    rng = np.random.default_rng(seed)
    return 2.449 + 0.3 * rng.standard_normal(N ** 3), 0.1 * rng.standard_normal(N ** 3)


@pytest.mark.parametrize("use_graph,two_step", [(True, True), (False, True), (True, False),
                                                (False, False)])
def test_batch_bitwise_equals_oracle(use_graph, two_step):
    sizes = [8, 16, 12, 24, 8, 20]
    gs = [0.1, 0.2, 0.3, 0.4, 0.25, 0.15]
    reals = _reals(sizes, gs)
    steps = 231  # two graph replays plus an odd remainder
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, steps)
    with pn.Batch(reals) as B:
        assert B.two_step(two_step) == two_step
        init = [_init(N, 100 + i) for i, N in enumerate(sizes)]
        for i in range(len(sizes)):
            B.view(i).set_fields(*init[i])
        B.evolve(dtk, dtd, use_graph=use_graph)
        for i, N in enumerate(sizes):
            gphi, gpi = B.view(i).get_fields()
            m = dict(reals[i])
            ophi, opi = O.evolve(m, init[i][0], init[i][1], steps)
            assert np.array_equal(gphi, ophi), i
            assert np.array_equal(gpi, opi), i


def test_batch_fp32_equals_single_fp32():
    sizes = [16, 32, 8]
    reals = _reals(sizes, [0.2, 0.3, 0.1])
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, 120)
    with pn.Batch(reals, precision=pn.PN_FP32) as B:
        init = [_init(N, 7 + i) for i, N in enumerate(sizes)]
        for i in range(len(sizes)):
            B.view(i).set_fields(*init[i])
        B.evolve(dtk, dtd)
        for i, N in enumerate(sizes):
            r = reals[i]
            with pn.Lattice(N, 0.5, 1, 1.0, -1, r["kappa"], precision=pn.PN_FP32) as L:
                L.set_fields(*init[i])
                L.evolve(dtk, dtd)
                sphi, spi = L.get_fields()
            bphi, bpi = B.view(i).get_fields()
            assert np.array_equal(bphi, sphi) and np.array_equal(bpi, spi)


def test_per_realization_time_tables():
    sizes = [8, 8]
    reals = _reals(sizes, [0.2, 0.2])
    k0, d0, _ = pn.time_factors(0.0, 0.1, 0.5, 40)
    k1, d1, _ = pn.time_factors(2.5, 0.07, 0.5, 40)
    init = _init(8, 3)
    with pn.Batch(reals) as B:
        for i in range(2):
            B.view(i).set_fields(*init)
        B.evolve(np.stack([k0, k1]), np.stack([d0, d1]), use_graph=False)
        a = B.view(0).get_fields()
        b = B.view(1).get_fields()
    oa = O.evolve(dict(reals[0]), init[0], init[1], 40)
    ob = O.evolve(dict(reals[1]), init[0], init[1], 40, T0=2.5, dtbase=0.07)
    assert np.array_equal(a[0], oa[0]) and np.array_equal(b[0], ob[0])


def test_batch_three_step_is_the_default_and_tunes():
    # three steps per launch where every realization fits; the autotuner then
    # picks planes per three-step tile and the bits stay the oracle's
    sizes = [16, 32, 24]
    reals = _reals(sizes, [0.1, 0.2, 0.3])
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, 200)
    init = [_init(N, 80 + i) for i, N in enumerate(sizes)]
    with pn.Batch(reals) as B:
        assert B.two_step() and B.three_step()
        for i in range(len(sizes)):
            B.view(i).set_fields(*init[i])
        chosen = B.autotune(trial_steps=4)
        assert sorted(c[0] for c in chosen) == [16, 24, 32]
        B.evolve(dtk, dtd)
        got = [B.view(i).get_fields() for i in range(len(sizes))]
    for i, N in enumerate(sizes):
        ophi, opi = O.evolve(dict(reals[i]), init[i][0], init[i][1], 200)
        assert np.array_equal(got[i][0], ophi) and np.array_equal(got[i][1], opi)


def test_autotune_keeps_bits_and_state():
    sizes = [16, 32, 16, 24]
    reals = _reals(sizes, [0.1, 0.2, 0.3, 0.4])
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, 50)
    init = [_init(N, 40 + i) for i, N in enumerate(sizes)]
    with pn.Batch(reals) as B:
        for i in range(len(sizes)):
            B.view(i).set_fields(*init[i])
        chosen = B.autotune(trial_steps=3)
        assert sorted(c[0] for c in chosen) == [16, 24, 32]
        for i in range(len(sizes)):  # state restored
            f, p = B.view(i).get_fields()
            assert np.array_equal(f, init[i][0]) and np.array_equal(p, init[i][1])
        B.evolve(dtk, dtd)
        got = [B.view(i).get_fields() for i in range(len(sizes))]
    for i, N in enumerate(sizes):
        ophi, opi = O.evolve(dict(reals[i]), init[i][0], init[i][1], 50)
        assert np.array_equal(got[i][0], ophi) and np.array_equal(got[i][1], opi)


def test_decayed_realization_is_frozen():
    sizes = [8, 12]
    reals = _reals(sizes, [0.2, 0.2])
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, 101)
    init = [_init(N, 60 + i) for i, N in enumerate(sizes)]
    with pn.Batch(reals) as B:
        for i in range(2):
            B.view(i).set_fields(*init[i])
        B.set_decayed([1, 0])
        B.evolve(dtk, dtd)
        f0, p0 = B.view(0).get_fields()
        f1, _ = B.view(1).get_fields()
    assert np.array_equal(f0, init[0][0]) and np.array_equal(p0, init[0][1])
    ophi, _ = O.evolve(dict(reals[1]), init[1][0], init[1][1], 101)
    assert np.array_equal(f1, ophi)


@pytest.mark.parametrize("steps", [20, 22])
def test_views_analysis_matches_single_context(steps):
    # 22 steps = 11 two-step launches: Pi ends in the second buffer
    sizes = [16, 32]
    reals = _reals(sizes, [0.2, 0.35])
    with pn.Batch(reals) as B:
        for i, N in enumerate(sizes):
            v = B.view(i)
            v.initial_conditions(pn.ic_params(10 + i, a=0.5, Lambda=6.0, phi0=np.sqrt(6.0)))
        dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, steps)
        B.evolve(dtk, dtd, use_graph=False)
        for i, N in enumerate(sizes):
            r = reals[i]
            v = B.view(i)
            with pn.Lattice(N, 0.5, 1, 1.0, -1, r["kappa"]) as L:
                L.initial_conditions(pn.ic_params(10 + i, a=0.5, Lambda=6.0, phi0=np.sqrt(6.0)))
                L.evolve(dtk, dtd)
                vf, lf = v.get_fields(), L.get_fields()
                assert np.array_equal(vf[0], lf[0]) and np.array_equal(vf[1], lf[1])
                a1, a2 = v.volume_averages().as_array(), L.volume_averages().as_array()
                assert np.array_equal(a1, a2)
                s1, s2 = v.spectrum(), L.spectrum()
                assert np.array_equal(s1["counts"], s2["counts"])
                np.testing.assert_allclose(s1["F"], s2["F"], rtol=1e-13)
                thr = 2.57583 * np.sqrt(a2[2])
                i1, l1, _ = v.islands(0, thr, labels=True)
                i2, l2, _ = L.islands(0, thr, labels=True)
                assert np.array_equal(l1, l2) and len(i1) == len(i2)


def test_view_cannot_step_alone():
    with pn.Batch(_reals([8], [0.2])) as B:
        v = B.view(0)
        with pytest.raises(pn.PhiNumError) as e:
            v.evolve([0.01], [0.01])
        assert e.value.code == pn.PN_ERR_STATE


@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_batch_two_step_equals_single_steps_with_decay(prec):
    # mixed sizes, a realization frozen part-way, odd step counts, graph
    # replays and eager launches: three steps, two steps and single steps per
    # launch give the same bits
    sizes = [8, 16, 12, 24, 64, 20, 36]
    reals = _reals(sizes, [0.1, 0.2, 0.3, 0.4, 0.25, 0.15, 0.35])
    init = [_init(N, 300 + i) for i, N in enumerate(sizes)]
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, 333)
    out = []
    for two, three in ((True, True), (True, False), (False, False)):
        with pn.Batch(reals, precision=prec) as B:
            assert B.two_step(two) == two
            assert B.three_step(three) == three
            for i in range(len(sizes)):
                B.view(i).set_fields(*init[i])
            B.evolve(dtk[:151], dtd[:151])
            B.set_decayed([0, 1, 0, 0, 1, 0, 0])
            B.evolve(dtk[151:], dtd[151:], use_graph=False)
            out.append([B.view(i).get_fields() for i in range(len(sizes))])
    for mode in (0, 1):
        for i in range(len(sizes)):
            assert np.array_equal(out[mode][i][0], out[2][i][0]), (mode, i)
            assert np.array_equal(out[mode][i][1], out[2][i][1]), (mode, i)
