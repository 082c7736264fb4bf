"""Pipelined strided passes (launcher variant bit 2, complex64): persistent
workgroups that copy the next tile into LDS while transforming the current
one.  Their per-element arithmetic is that of the one-tile-per-workgroup
passes up to the compiler's choice of fused multiply-adds, so results agree
with the other launcher choices at complex64 rounding level; each choice is
held to the complex128 engine (itself at 1e-12 of the FP64 oracle) with the
same accuracy, and runs are reproducible bit for bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

PARAMS = [0.1, 2.0, 0.0, 0.01]  # cosine lattice A=0.1, P=2; contact a=0.01


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def state(shape, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.1 * rng.standard_normal(shape)) * np.exp(0.3j * rng.standard_normal(shape))


def run(shape, mode, variant, prec=E.C64, steps=5):
    alpha, beta = (0.0, 1.0) if mode == "ite" else (1.0, 0.0)
    g = E.Engine(shape, alpha, beta, 0.02, float(np.prod(shape)), E.CONTACT, E.COSINE, 0, prec)
    g.set_user_params(PARAMS)
    if mode == "qf":
        g.set_quantum_friction_coeff(0.2)
    g.set_variant(variant)
    g.set_psi(state(shape, 3))
    g.normalize_psi()
    mu = g.evolve_qf(steps, want_mu=(mode == "ite"))
    psi = g.get_psi()[0]
    e = g.energy()
    g.close()
    return psi, sum(e[1:]), mu


@pytest.mark.parametrize("shape", [(256, 64, 32), (128, 256, 32), (64, 128, 16), (256, 256, 16)])
@pytest.mark.parametrize("mode", ["rte", "ite", "qf"])
def test_pipelined_passes_track_complex128(shape, mode):
    say("pipelined", shape, mode)
    ref_psi, ref_e, ref_mu = run(shape, mode, 0, E.C128)
    errs = {}
    for variant in (0, 4):
        psi, e, mu = run(shape, mode, variant)
        mu_err = abs(mu / ref_mu - 1) if mode == "ite" else 0.0
        errs[variant] = (float(np.max(np.abs(psi - ref_psi))), abs(e / ref_e - 1), mu_err)
    say("max |psi - c128|, rel energy, rel mu error: variant 0 %.2e %.2e %.2e, variant 4 %.2e %.2e %.2e"
        % (errs[0] + errs[4]))
    # complex64 rounding: |psi| ~ 1, 5 steps
    assert errs[4][0] < 1e-5 and errs[4][0] < 2.0 * errs[0][0] + 1e-7
    assert errs[4][1] < 1e-6 and errs[4][1] < 2.0 * errs[0][1] + 1e-8
    assert errs[4][2] < 5e-5 and errs[4][2] < 2.0 * errs[0][2] + 1e-7


@pytest.mark.parametrize("shape", [(256, 256, 16), (128, 256, 32)])
def test_runs_are_reproducible(shape):
    """The same launcher choice twice gives the same bits."""
    for variant in (0, 4):
        a, b = run(shape, "rte", variant)[0], run(shape, "rte", variant)[0]
        assert np.array_equal(a, b), variant


def test_pipelined_adjoint_and_fft():
    """The adjoint's conjugate kinetic passes and the plain transform."""
    say("pipelined adjoint / fft")
    shape = (256, 32, 32)
    rng = np.random.default_rng(4)
    v = 0.05 * rng.random(shape)
    res = []
    for variant in (0, 4):
        g = E.Engine(shape, 1.0, 0.0, 0.01, float(np.prod(shape)), E.UFG, E.VARRAY, 0, E.C64)
        g.set_variant(variant)
        g.set_external_potential(v)
        g.set_psi(state(shape, 5))
        L = g.adjoint_forward(6, state(shape, 6))
        grad = g.adjoint_backward()
        x = state(shape, 7)
        g.set_psi(x)
        g.fft(-1)
        f = g.get_psi()[0]
        res.append((L, grad, f))
        g.close()
    want = np.fft.fftn(state(shape, 7))
    for L, grad, f in res:
        assert np.max(np.abs(f - want)) < 2e-6 * np.max(np.abs(want))
    assert abs(res[0][0] - res[1][0]) < 1e-6
    assert np.max(np.abs(res[0][1] - res[1][1])) < 1e-3 * np.max(np.abs(res[0][1]))
