"""pn_field_diff_get, the precision form's field error against an FP64 run
(no reference counterpart), against numpy on fields set from the host."""
import math

import numpy as np
import pytest

import phinum_amd as pn

pytestmark = pytest.mark.gpu


def _lattice(N, prec):
    return pn.Lattice(N, 0.5, 1, 1.0, -1, 2.0, precision=prec)


@pytest.mark.parametrize("N,pa,pb", [(16, pn.PN_FP32, pn.PN_FP64), (12, pn.PN_FP64, pn.PN_FP64),
                                     (20, pn.PN_FP64, pn.PN_FP32), (8, pn.PN_FP32, pn.PN_FP32)])
def test_field_diff_matches_numpy(N, pa, pb):
    rng = np.random.default_rng(N)
    phi = 1.0 + 0.3 * rng.standard_normal(N ** 3)
    pi = 0.2 * rng.standard_normal(N ** 3)
    dphi = 1e-6 * rng.standard_normal(N ** 3)
    with _lattice(N, pa) as A, _lattice(N, pb) as B:
        A.set_fields(phi + dphi, pi - dphi)
        B.set_fields(phi, pi)
        d = A.field_diff(B)
        # the values each lattice holds, in double
        fa, ga = A.get_fields()
        fb, gb = B.get_fields()
    df, dg = fa - fb, ga - gb
    assert d["max_dphi"] == np.max(np.abs(df))
    assert d["max_dpi"] == np.max(np.abs(dg))
    assert d["max_phi"] == np.max(np.abs(fb))
    assert d["max_pi"] == np.max(np.abs(gb))
    for key, v in (("sum_dphi2", df * df), ("sum_phi2", fb * fb), ("sum_dpi2", dg * dg), ("sum_pi2", gb * gb)):
        assert abs(d[key] - math.fsum(v)) <= 1e-14 * math.fsum(v), key
    if pa == pb == pn.PN_FP64:
        assert d["l2_dphi"] == pytest.approx(np.sqrt(np.sum(dphi ** 2) / np.sum(phi ** 2)), rel=1e-12)


def test_field_diff_identical_and_mismatched():
    N = 16
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(N ** 3)
    with _lattice(N, pn.PN_FP64) as A, _lattice(N, pn.PN_FP64) as B, _lattice(8, pn.PN_FP64) as C8:
        A.set_fields(phi, -phi)
        B.set_fields(phi, -phi)
        d = A.field_diff(B)
        assert d["max_dphi"] == 0.0 and d["max_dpi"] == 0.0 and d["l2_dphi"] == 0.0
        with pytest.raises(pn.PhiNumError):
            A.field_diff(C8)
        phi[5] = np.nan
        A.set_fields(phi, -phi)
        assert np.isnan(A.field_diff(B)["max_dphi"])
