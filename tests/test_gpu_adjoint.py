"""Differentiable leapfrog on the GPU: forward and gradients against the CPU
checker (tangent-linear, full reverse) and central finite differences, plus a
torch.autograd.Function harness over the C-ABI reverse entry point."""
import numpy as np
import pytest

import _oracle as O
import phinum_amd as pn

pytestmark = pytest.mark.gpu


def _state(N, seed, amp=0.6):
    rng = np.random.default_rng(seed)
    return 1.0 + amp * rng.standard_normal(N ** 3), 0.3 * rng.standard_normal(N ** 3)


@pytest.mark.parametrize("N,S,K", [(8, 60, 7), (12, 50, 10), (16, 40, 40)])
def test_gradients_match_cpu_reverse(N, S, K):
    phi0, pi0 = _state(N, N)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, S)
    F, gphi, gpi, gc, res = pn.adjoint(pn.adj_params(N, checkpoint_every=K), phi0, pi0, dtk, dtd)
    oF, ogphi, ogpi, ogc = O.adj_reverse(O.adj_params(N), phi0, pi0, dtk, dtd)
    assert abs(F - oF) <= 1e-13 * abs(oF)
    scale = np.max(np.abs(ogphi))
    assert np.max(np.abs(gphi - ogphi)) <= 1e-11 * scale
    assert np.max(np.abs(gpi - ogpi)) <= 1e-11 * np.max(np.abs(ogpi))
    np.testing.assert_allclose(gc, ogc, rtol=1e-10)
    assert F == pytest.approx(pn.forward_observable(pn.adj_params(N), phi0, pi0, dtk, dtd), rel=1e-14)


def test_gradients_match_finite_differences_and_tlm():
    N, S = 16, 200
    phi0, pi0 = _state(N, 5)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, S)
    p = pn.adj_params(N, checkpoint_every=50)
    F, gphi, gpi, gc, _ = pn.adjoint(p, phi0, pi0, dtk, dtd)
    v = np.random.default_rng(9).standard_normal(N ** 3)
    eps = 1e-6
    fd = (pn.forward_observable(p, phi0 + eps * v, pi0, dtk, dtd)
          - pn.forward_observable(p, phi0 - eps * v, pi0, dtk, dtd)) / (2 * eps)
    ad = float(np.dot(gphi, v))
    assert abs(fd - ad) <= 1e-6 * abs(ad)
    tl = O.adj_tlm(O.adj_params(N), phi0, pi0, dtk, dtd, v, np.zeros(N ** 3), np.zeros(3))
    assert abs(tl - ad) <= 1e-10 * abs(tl)
    for j, name in ((1, "c4"), (2, "c6")):
        kw = dict(c4=p.c4, c6=p.c6)
        kw[name] += eps
        fp = pn.forward_observable(pn.adj_params(N, **kw), phi0, pi0, dtk, dtd)
        kw[name] -= 2 * eps
        fm = pn.forward_observable(pn.adj_params(N, **kw), phi0, pi0, dtk, dtd)
        assert abs((fp - fm) / (2 * eps) - gc[j]) <= 1e-6 * abs(gc[j])


def test_resident_store_equals_checkpoints():
    # checkpoint_every >= nsteps keeps every phi_n from the forward pass (no
    # recompute); the same kernels in the same order otherwise, so F and
    # every gradient are bitwise those of checkpointed runs (K = 7 and 1)
    N, S = 12, 30
    phi0, pi0 = _state(N, 31)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, S)
    runs = [pn.adjoint(pn.adj_params(N, checkpoint_every=K), phi0, pi0, dtk, dtd) for K in (S, 100, 7, 1)]
    F0, gphi0, gpi0, gc0, r0 = runs[0]
    for F, gphi, gpi, gc, r in runs[1:]:
        assert F == F0
        assert np.array_equal(gphi, gphi0) and np.array_equal(gpi, gpi0) and np.array_equal(gc, gc0)
    # the resident run holds S planes of phi_n and no checkpoints
    assert r0.peak_bytes > runs[2][4].peak_bytes


def test_torch_autograd_harness():
    pn.lib()  # the library first: it refuses a process whose HIP runtime came from torch's wheel
    torch = pytest.importorskip("torch")

    class OscillonFraction(torch.autograd.Function):
        """Test harness only: forward and backward call the C-ABI entry points."""

        @staticmethod
        def forward(ctx, phi0, pi0, couplings, dtk, dtd, N):
            c = couplings.detach().cpu().numpy()
            p = pn.adj_params(N, c2=c[0], c4=c[1], c6=c[2], checkpoint_every=25)
            F, gphi, gpi, gc, _ = pn.adjoint(p, phi0.detach().cpu().numpy(), pi0.detach().cpu().numpy(),
                                             dtk, dtd)
            ctx.save_for_backward(torch.from_numpy(gphi), torch.from_numpy(gpi), torch.from_numpy(gc))
            return torch.tensor(F, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gphi, gpi, gc = ctx.saved_tensors
            return gout * gphi, gout * gpi, gout * gc, None, None, None

    N, S = 8, 50
    phi0, pi0 = _state(N, 21)
    dtk, dtd, _ = pn.time_factors(0.0, 0.1, 0.5, S)
    tphi = torch.tensor(phi0, requires_grad=True)
    tpi = torch.tensor(pi0, requires_grad=True)
    tc = torch.tensor([1.0, -1.0, 0.2], dtype=torch.float64, requires_grad=True)
    F = OscillonFraction.apply(tphi, tpi, tc, dtk, dtd, N)
    F.backward()
    oF, ogphi, ogpi, ogc = O.adj_reverse(O.adj_params(N), phi0, pi0, dtk, dtd)
    assert abs(F.item() - oF) <= 1e-13 * abs(oF)
    np.testing.assert_allclose(tphi.grad.numpy(), ogphi, rtol=0, atol=1e-11 * np.max(np.abs(ogphi)))
    np.testing.assert_allclose(tc.grad.numpy(), ogc, rtol=1e-10)
