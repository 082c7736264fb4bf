"""Dipolar path of the oracle (INTERACTIONS -2, dev/gpe_engine.cu:1490-1605,
1716-1810, dev/gpe_user_defined.h:422-438) against closed forms: the
convolution potential of single density waves along z and x (which isolate
the (3 kz^2/k^2 - 1) factor and the k = 0 branch), the uniform-state step
phase and chemical potential, and the two dipolar energies against each other
(equal up to the factor 1/2 by Parseval) and against the uniform closed form."""
import numpy as np
import pytest

SHAPE = (12, 10, 16)
AC, AD = 0.03, 0.02


def make(oracle, alpha=1.0, beta=0.0, dt=0.01, npart=100.0):
    o = oracle.OracleEngine(SHAPE, alpha, beta, dt, npart, -2, 0, 0)
    o.set_user_params([0.0, 0.0, 0.0, AC, AD])
    return o


@pytest.mark.parametrize("axis,factor", [(2, AC + 2 * AD), (0, AC - AD), (1, AC - AD)])
def test_convolution_of_a_density_wave(oracle, axis, factor):
    rho0, eps, m = 0.8, 0.1, 2
    idx = np.meshgrid(*[np.arange(n) for n in SHAPE], indexing="ij")
    wave = np.cos(2 * np.pi * m * idx[axis] / SHAPE[axis])
    o = make(oracle)
    o.set_psi(np.sqrt(rho0 + eps * wave))  # gamma = 1 for bosons
    v = o.dipolar_potential()
    want = 4 * np.pi * AC * rho0 + 4 * np.pi * factor * eps * wave
    assert np.max(np.abs(v - want)) < 1e-13
    o.close()


def test_energies_uniform_and_parseval(oracle):
    rho0 = 0.7
    n = int(np.prod(SHAPE))
    o = make(oracle)
    o.set_psi(np.full(SHAPE, np.sqrt(rho0), dtype=np.complex128))
    _, e1 = o.energy_dipolar(1)
    _, e2 = o.energy_dipolar(2)
    assert abs(e1 - 4 * np.pi * AC * rho0 ** 2 * n) < 1e-10 * e1
    assert abs(e2 - 0.5 * e1) < 1e-10 * e1
    rng = np.random.default_rng(3)
    o.set_psi(rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE))
    _, e1 = o.energy_dipolar(1)
    _, e2 = o.energy_dipolar(2)
    assert abs(e2 - 0.5 * e1) < 1e-12 * abs(e1)
    o.close()


def test_uniform_state_step_phase_and_mu(oracle):
    rho0, dt, nt = 0.5, 0.01, 7
    o = make(oracle, dt=dt)
    psi0 = np.full(SHAPE, np.sqrt(rho0), dtype=np.complex128)
    o.set_psi(psi0)
    o.evolve_dipolar(nt)
    p, t = o.get_psi()
    # two potential half steps of exp(-i dt/2 * 4 pi a_con rho0) per step, kinetic trivial at k = 0
    want = psi0 * np.exp(-1j * dt * 4 * np.pi * AC * rho0 * nt)
    assert np.max(np.abs(p - want)) < 1e-13
    assert abs(t - nt * dt) < 1e-15
    o.close()
    npart = rho0 * np.prod(SHAPE)
    o = make(oracle, alpha=0.0, beta=1.0, dt=dt, npart=npart)
    o.set_psi(psi0)
    mu = o.evolve_dipolar(3, want_mu=True)
    # imaginary time: the second half step sees the density already damped by
    # exp(-dt V1), so the norm decays by exp(-dt (V1 + V2)) with V2 = V1 exp(-dt V1)
    v1 = 4 * np.pi * AC * rho0
    assert abs(mu - 0.5 * (v1 + v1 * np.exp(-dt * v1))) < 1e-12
    assert np.max(np.abs(o.get_psi()[0] - psi0)) < 1e-13
    o.close()


def test_real_time_dipolar_step_is_unitary(oracle):
    rng = np.random.default_rng(9)
    o = make(oracle)
    psi = (1 + 0.2 * rng.random(SHAPE)) * np.exp(1j * rng.uniform(-1, 1, SHAPE))
    o.set_psi(psi)
    n0 = o.get_norm()
    o.evolve_dipolar(5)
    assert abs(o.get_norm() / n0 - 1) < 1e-13
    o.close()
