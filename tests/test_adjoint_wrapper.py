"""Host logic of the torch.autograd wrapper (libcugpe_amd/adjoint.py), with a
stand-in engine so it runs without a GPU: one live forward per engine, the
gradient scaled by the incoming gradient, and the Gaussian-bump potential
against a direct numpy sum."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from libcugpe_amd import adjoint as A  # noqa: E402


class StandInEngine:
    """Records calls; its 'gradient' is the potential it was given."""

    def __init__(self):
        self.v = None
        self.adjoint_generation = 0

    def set_external_potential(self, v):
        self.v = np.array(v)

    def set_psi(self, psi):
        pass

    def adjoint_forward(self, nt, psi_target, ckpt=0):
        self.adjoint_generation += 1
        return float(np.sum(self.v))

    def adjoint_backward(self):
        return self.v.copy()


def test_backward_scales_the_engine_gradient():
    eng = StandInEngine()
    V = torch.arange(8, dtype=torch.float64).reshape(2, 2, 2).requires_grad_(True)
    L = A.overlap_loss(V, eng, None, None, 3)
    (3.0 * L).backward()
    assert torch.equal(V.grad, 3.0 * V.detach())


def test_second_forward_on_one_engine_voids_the_first_backward():
    eng = StandInEngine()
    V1 = torch.ones((2, 2, 2), dtype=torch.float64, requires_grad=True)
    V2 = torch.full((2, 2, 2), 2.0, dtype=torch.float64, requires_grad=True)
    L1 = A.overlap_loss(V1, eng, None, None, 3)
    L2 = A.overlap_loss(V2, eng, None, None, 3)
    with pytest.raises(RuntimeError, match="another adjoint forward"):
        (L1 + L2).backward()
    # one engine per live loss works
    e1, e2 = StandInEngine(), StandInEngine()
    V1.grad = None
    V2.grad = None
    L1 = A.overlap_loss(V1, e1, None, None, 3)
    L2 = A.overlap_loss(V2, e2, None, None, 3)
    (L1 + L2).backward()
    assert torch.equal(V1.grad, V1.detach()) and torch.equal(V2.grad, V2.detach())


def test_gaussian_bumps_match_a_direct_sum():
    shape = (8, 6, 10)
    centres = A.bump_centres(shape, 5, 3)
    amp = torch.tensor(np.random.default_rng(1).standard_normal(5))
    got = A.gaussian_bumps(amp, centres, 1.5, shape).numpy()
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    want = np.zeros(shape)
    for j in range(5):
        d2 = 0.0
        for x, n, c in zip((ix, iy, iz), shape, centres[j]):
            d = x - c
            d = d - n * np.round(d / n)
            d2 = d2 + d * d
        want += amp[j].item() * np.exp(-d2 / (2 * 1.5 ** 2))
    assert np.max(np.abs(got - want)) < 1e-13
