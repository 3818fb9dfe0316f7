"""Initial conditions: the host half of SetInitialConditions (initialize==3)
in the product library against the oracle, and the oracle's RNG against
glibc's drand48 (POSIX).  No GPU needed."""
import ctypes
import ctypes.util

import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn


def test_oracle_drand48_matches_libc():
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 1234, 2 ** 31 + 77, -5):
        libc.srand48(seed)
        O.lib().or_srand48(seed)
        a = [libc.drand48() for _ in range(2000)]
        b = [O.lib().or_drand48() for _ in range(2000)]
        assert a == b


def test_oracle_gauss_pairs():
    # polar method: both values of a pair come from one accepted (x, y) and
    # g2/g1 = s/c = y/x; the cached second value is returned next
    O.lib().or_srand48(99)
    g = [O.lib().or_gauss() for _ in range(20000)]
    g = np.array(g)
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1.0) < 0.02
    O.lib().or_srand48(99)
    u = [O.lib().or_drand48() for _ in range(8)]
    x, y = 2.0 * u[0] - 1.0, 2.0 * u[1] - 1.0
    k = 0
    while np.sqrt(x * x + y * y) >= 1.0:
        k += 2
        x, y = 2.0 * u[k] - 1.0, 2.0 * u[k + 1] - 1.0
    r = np.sqrt(x * x + y * y)
    g1 = (x / r) * np.sqrt(-2.0 * np.log(r * r))
    assert g[0] == g1
    assert g[1] == ((y / r) / (x / r)) * g1


@pytest.mark.parametrize("N,seed", [(8, 1234), (16, 7), (12, 0)])
def test_product_modes_bitwise_equal_oracle(N, seed):
    a = 0.5
    A0, E0 = O.ic_modes(N, a, seed)
    A1, E1 = pn.ic_modes(N, a, pn.ic_params(seed, a=a))
    assert np.array_equal(A0.view(np.float64), A1.view(np.float64))
    assert np.array_equal(E0.view(np.float64), E1.view(np.float64))


def test_modes_are_hermitian_on_the_halved_planes():
    N, a = 16, 0.5
    A, E = pn.ic_modes(N, a, pn.ic_params(5, a=a))
    for F in (A, E):
        for py in (0, N // 2):
            for pz in range(N):
                for px in range(N):
                    assert F[(N - pz) % N, (N - px) % N, py] == np.conj(F[pz, px, py])
        # so the c2r output is real up to roundoff: the full-cube transform of
        # the Hermitian extension has no imaginary part
        full = np.zeros((N, N, N), complex)
        full[:, :, :N // 2 + 1] = F
        for py in range(N // 2 + 1, N):
            full[:, :, py] = np.conj(F[(-np.arange(N)) % N][:, (-np.arange(N)) % N, N - py])
        x = np.fft.ifftn(full)
        assert np.max(np.abs(x.imag)) < 1e-13 * np.max(np.abs(x.real))


def test_ic_spectrum_matches_occupation():
    # <|A_k|^2> = Occ/Omega * <g^2> with <g^2> = 1 for |Gauss|: average over
    # many modes of one shell against the closed form
    N, a = 32, 0.5
    A, E = pn.ic_modes(N, a, pn.ic_params(11, a=a))
    H = N // 2 + 1
    p = np.array([2.0 / a * np.sin(k * np.pi / N) for k in range(N)])
    pz, px, py = np.meshgrid(p, p, p[:H], indexing="ij")
    psqr = px ** 2 + py ** 2 + pz ** 2
    omega = np.sqrt(psqr + 1.0)
    occ = np.where(np.sqrt(psqr) < np.sqrt(3.0) / a, 0.5 * 6.0, 0.0)
    sel = (occ > 0) & (py > 0) & (py < p[N // 2])
    r = (np.abs(A[sel]) ** 2 * omega[sel] / occ[sel]).mean()
    s = (np.abs(E[sel]) ** 2 / omega[sel] / occ[sel]).mean()
    assert abs(r - 1.0) < 0.05 and abs(s - 1.0) < 0.05
    assert np.all(A[~(occ > 0)] == 0)
