"""initialize==3 modes drawn on the device (pn_ic_modes_device, the default
of pn_initial_conditions) against the host generator pn_ic_modes, which is
bitwise equal to the oracle's restatement of the reference's draws
(test_ic_host.py; initialconditions.cpp:12-42,83-594, clusterfft.cpp:55-145).

A mode's values come from its own draws only, so a mode that started at a
wrong stream position, or took a different polar attempt, would differ at
O(1).  Agreement of every mode to a few ulp therefore shows the device found
every mode's draws exactly; the remaining ulps are log, sincos and hypot of
the device library against the host C library."""
import numpy as np
import pytest

import phinum_amd as pn

pytestmark = pytest.mark.gpu

REL = 8e-15  # a few ulp of log / sincos / hypot


def _check(dev, host):
    d = dev.view(np.float64)
    h = host.view(np.float64)
    assert d.shape == h.shape
    # modes outside the occupied shell, and the zeroed imaginary parts of the
    # self-conjugate modes, are exact zeros on both sides
    assert np.array_equal(d == 0, h == 0)
    bad = np.abs(d - h) > REL * np.abs(h)
    assert not bad.any(), (np.argwhere(bad)[:5], d[bad][:5], h[bad][:5])


# odd N exercise the Hermitian planes' edge cases; 128 and 256 span several
# groups of chunks in the chaining passes
@pytest.mark.parametrize("N,seed", [(5, 1), (4, 1234), (7, 3), (8, 1234), (12, 0), (16, 7),
                                    (33, 5), (64, 1234), (128, 99), (256, 1234)])
def test_device_modes_match_host_draws(N, seed):
    a = 0.5
    ic = pn.ic_params(seed, a=a, Lambda=6.0, phi0=2.0)
    print("ic modes N=%d seed=%d" % (N, seed), flush=True)
    A0, E0 = pn.ic_modes(N, a, ic)
    with pn.Lattice(N, a, 1, 1.0, -1, 2.0) as L:
        A1, E1 = L.ic_modes_device(ic)
    _check(A1, A0)
    _check(E1, E0)


def test_device_modes_other_parameters():
    # amplitude, justnoise and scale2 branches of the occupation, another a
    N, a = 16, 0.7
    for kw in (dict(amplitude=2.0), dict(justnoise=1, scale2=1.5), dict(scale2=0.0, justnoise=1)):
        ic = pn.ic_params(21, a=a, Lambda=3.0, m2=0.5, **kw)
        A0, E0 = pn.ic_modes(N, a, ic)
        with pn.Lattice(N, a, 1, 0.5, -1, 2.0) as L:
            A1, E1 = L.ic_modes_device(ic)
        _check(A1, A0)
        _check(E1, E0)


@pytest.mark.parametrize("prec", [pn.PN_FP64, pn.PN_FP32])
def test_initial_conditions_device_and_host_modes_agree(prec):
    N, a = 32, 0.5
    ic = pn.ic_params(1234, a=a, Lambda=6.0, phi0=float(np.sqrt(6.0)))
    out = []
    for dev in (True, False):
        with pn.Lattice(N, a, 1, 1.0, -1, 2.0, precision=prec) as L:
            assert L.ic_on_device() is True  # the default
            L.ic_on_device(dev)
            assert L.ic_on_device() is dev
            L.initial_conditions(ic)
            out.append(L.get_fields())
    (p1, q1), (p0, q0) = out
    tol = 1e-13 if prec == pn.PN_FP64 else 1.2e-7
    assert np.max(np.abs(p1 - p0)) <= tol * np.max(np.abs(p0))
    assert np.max(np.abs(q1 - q0)) <= tol * np.max(np.abs(q0))
