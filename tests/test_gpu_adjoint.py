"""GPU adjoint of the real-time split step (gpe_engine_adjoint_forward /
_backward) against the FP64 oracle's adjoint (itself pinned by finite
differences in test_oracle_adjoint.py), against central finite differences of
the complex128 engine, and through the torch.autograd wrapper."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def inputs(shape, seed):
    rng = np.random.default_rng(seed)
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    v = 0.2 * np.cos(2 * np.pi * ix / shape[0]) + 0.1 * np.sin(2 * np.pi * iy / shape[1]) + 0.02 * rng.random(shape)
    psi0 = (1 + 0.2 * rng.random(shape)) * np.exp(1j * rng.uniform(-0.5, 0.5, shape))
    psit = (1 + 0.2 * rng.random(shape)) * np.exp(1j * rng.uniform(-0.5, 0.5, shape))
    return v, psi0, psit


def gpu_engine(shape, inter, v, psi0, prec, dt=0.05, a=0.05):
    g = E.Engine(shape, 1.0, 0.0, dt, 50.0, inter, E.VARRAY, 0, prec)
    g.set_user_params([0.0, 0.0, 0.0, a])
    g.set_external_potential(v)
    g.set_psi(psi0)
    return g


@pytest.mark.parametrize("inter", [E.CONTACT, E.UFG])
@pytest.mark.parametrize("ckpt", [2, 0])
def test_adjoint_matches_oracle(oracle, inter, ckpt):
    say("adjoint vs oracle", inter, ckpt)
    shape = (16, 12, 8)
    v, psi0, psit = inputs(shape, 4)
    nt = 6
    g = gpu_engine(shape, inter, v, psi0, E.C128)
    Lg = g.adjoint_forward(nt, psit, ckpt)
    gg = g.adjoint_backward()
    pg, _ = g.get_psi()
    o = oracle.OracleEngine(shape, 1.0, 0.0, 0.05, 50.0, inter, 3, 0)
    o.set_user_params([0.0, 0.0, 0.0, 0.05])
    o.set_vext_array(v)
    o.set_psi(psi0)
    Lo, go = o.adjoint_overlap(nt, psit, ckpt)
    po, _ = o.get_psi()
    o.close()
    assert abs(Lg - Lo) < 1e-12
    assert np.max(np.abs(gg - go)) < 1e-10 * np.max(np.abs(go))
    # psi is left at psi(T)
    assert np.max(np.abs(pg - po)) < 1e-12 * np.max(np.abs(po))
    g.close()


def test_adjoint_c64_against_c128():
    say("adjoint c64 vs c128")
    shape = (32, 32, 32)
    v, psi0, psit = inputs(shape, 6)
    out = {}
    for prec in (E.C64, E.C128):
        g = gpu_engine(shape, E.UFG, v, psi0, prec, dt=0.01)
        L = g.adjoint_forward(60, psit)
        out[prec] = (L, g.adjoint_backward())
        g.close()
    (L64, g64), (L128, g128) = out[E.C64], out[E.C128]
    rel = np.linalg.norm(g64 - g128) / np.linalg.norm(g128)
    say("c64 adjoint: |dL| %.2e, rel L2 grad error %.2e" % (abs(L64 - L128), rel))
    assert abs(L64 - L128) < 1e-6
    assert rel < 1e-4


def test_torch_wrapper_matches_finite_differences():
    """dL/dA for Gaussian-bump amplitudes through autograd (torch ops for the
    bumps, the engine's adjoint for dL/dV) against central differences of the
    complex128 engine with h = 1e-4.  The complex128 adjoint must meet them to
    1e-6 of the largest component.  The complex64 adjoint differentiates the
    complex64 trajectory, whose overlap (L ~ 1e-3 here) carries absolute
    errors ~1e-6 from FP32 rounding over the run; its bound is 5e-4 of the
    largest component (about 2e-8 absolute here)."""
    import torch

    from libcugpe_amd import adjoint as A

    say("adjoint torch + finite differences")
    shape = (24, 24, 24)
    rng = np.random.default_rng(5)
    nb = 16
    centres = A.bump_centres(shape, nb, 5)
    amp = torch.tensor(0.05 * rng.standard_normal(nb), dtype=torch.float64, requires_grad=True)
    psi0 = (1 + 0.1 * rng.random(shape)) * np.exp(1j * 0.2 * rng.standard_normal(shape))
    psit = (1 + 0.1 * rng.random(shape)) * np.exp(1j * 0.2 * rng.standard_normal(shape))
    nt, dt = 100, 0.01
    grads = {}
    for prec in (E.C128, E.C64):
        eng = E.Engine(shape, 1.0, 0.0, dt, float(np.prod(shape)), E.UFG, E.VARRAY, 0, prec)
        a1 = amp.detach().clone().requires_grad_(True)
        L = A.overlap_loss(A.gaussian_bumps(a1, centres, 1.5, shape), eng, psi0, psit, nt)
        L.backward()
        grads[prec] = a1.grad.numpy().copy()
        eng.close()

    eng = E.Engine(shape, 1.0, 0.0, dt, float(np.prod(shape)), E.UFG, E.VARRAY, 0, E.C128)

    def loss(a):
        V = A.gaussian_bumps(torch.tensor(a), centres, 1.5, shape).numpy()
        eng.set_external_potential(V)
        eng.set_psi(psi0)
        eng.evolve_qf(nt)
        p, _ = eng.get_psi()
        ov = np.vdot(psit, p)
        return 1.0 - abs(ov) ** 2 / (np.vdot(psit, psit).real * np.vdot(psi0, psi0).real)

    h = 1e-4
    a0 = amp.detach().numpy().copy()
    scale = np.max(np.abs(grads[E.C128]))
    for j in (0, 5, 11):
        ap, am = a0.copy(), a0.copy()
        ap[j] += h
        am[j] -= h
        fd = (loss(ap) - loss(am)) / (2 * h)
        say("bump %d: fd %.9e adj128 %.9e adj64 %.9e" % (j, fd, grads[E.C128][j], grads[E.C64][j]))
        assert abs(fd - grads[E.C128][j]) < 1e-6 * scale
        assert abs(fd - grads[E.C64][j]) < 5e-4 * scale
    eng.close()


@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_torch_wrapper_on_device_tensors(prec):
    """A potential tensor on the GPU crosses the C ABI as device memory
    (gpe_engine_set_external_potential_device, gpe_engine_adjoint_backward_device):
    the loss and dL/dV equal those of the host-array path bit for bit."""
    import torch

    from libcugpe_amd import adjoint as A

    say("adjoint torch device tensors", prec)
    shape = (16, 16, 24)
    rng = np.random.default_rng(9)
    V0 = torch.tensor(0.05 * rng.random(shape), dtype=torch.float64)
    psi0 = (1 + 0.1 * rng.random(shape)) * np.exp(1j * 0.2 * rng.standard_normal(shape))
    psit = (1 + 0.1 * rng.random(shape)) * np.exp(1j * 0.2 * rng.standard_normal(shape))
    out = {}
    for dev in ("cpu", "cuda"):
        eng = E.Engine(shape, 1.0, 0.0, 0.01, float(np.prod(shape)), E.UFG, E.VARRAY, 0, prec)
        V = V0.detach().clone().to(dev).requires_grad_(True)
        L = A.overlap_loss(V, eng, psi0, psit, 30)
        L.backward()
        assert V.grad.device.type == dev
        out[dev] = (float(L.detach().cpu()), V.grad.detach().cpu().numpy().copy())
        eng.close()
    assert out["cpu"][0] == out["cuda"][0]
    assert np.array_equal(out["cpu"][1], out["cuda"][1])
    # the device entry points refuse a missing pointer or a lattice without a VEXT 3 array
    g = E.Engine(shape, 1.0, 0.0, 0.01, 10.0, E.UFG, E.HARMONIC, 0, prec)
    with pytest.raises(E.GpeError):
        g.set_external_potential_device(V0.to("cuda").data_ptr())
    g.close()


def test_adjoint_rejects_unsupported_runs():
    g = E.Engine((16, 16, 16), 0.0, 1.0, 0.01, 10.0, E.CONTACT, E.HARMONIC, 0, E.C64)
    with pytest.raises(E.GpeError) as ex:
        g.adjoint_forward(2, np.ones((16, 16, 16)))
    assert ex.value.code == -12
    with pytest.raises(E.GpeError):
        g.adjoint_backward()
    g.close()
