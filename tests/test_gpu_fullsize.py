"""Full-size (256^3) GPU checks through size-independent properties: the
oracle is too slow at this size, so these pin the engine by what the
mathematics fixes at any size (Parseval, transform round trips, single
Fourier coefficients summed directly, the exact phase of a plane wave, norm
conservation of the unitary real-time step) and by the integer plaquette
windings, which the oracle computes at full size in a fraction of a second.
The workload is the headline one: 256^3, contact bosons, cosine potential,
20 vortex lines."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

SHAPE = (256, 256, 256)


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def direct_coefficient(x, k):
    """X[k] = sum_n x[n] exp(-2 pi i k.n / N) as three FP64 contractions."""
    e = [np.exp(-2j * np.pi * kk * np.arange(n) / n) for kk, n in zip(k, x.shape)]
    return complex(e[0] @ ((x @ e[2]) @ e[1]))


@pytest.mark.parametrize("prec", [E.C64, E.C128])
def test_fft_256cube_parseval_coefficients_round_trip(prec):
    """Forward transform: Parseval (sum |X|^2 = N sum |x|^2) and three single
    coefficients against direct FP64 sums (a wrong sign, axis order or
    twiddle index moves them by O(|x|)); then the inverse transform returns
    N x.  Parseval alone would pass a unitary transform with misplaced
    twiddles; the coefficients would not.  complex64 bounds: relative
    rounding of a few 1e-7 per transform of length 256^3 (FP32 butterflies
    with compensated twiddles) for the coefficients and the round trip, 1e-8
    for Parseval (the transform is unitary up to that rounding); complex128
    bounds 1e-13."""
    say("fft 256^3", prec)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(SHAPE) + 1j * rng.standard_normal(SHAPE)
    g = E.Engine(SHAPE, 1.0, 0.0, 0.01, 1.0, precision=prec)
    g.set_psi(x)
    x = g.get_psi()[0]  # the values on the device (rounded to complex64 there)
    g.fft(-1)
    X, _ = g.get_psi()
    n = x.size
    e_in = float(np.vdot(x, x).real)
    e_out = float(np.vdot(X, X).real)
    par = abs(e_out / (n * e_in) - 1)
    norm = np.sqrt(e_in)
    coef = []
    for k in ((0, 0, 0), (1, 255, 17), (128, 37, 200)):
        coef.append(abs(X[k] - direct_coefficient(x, k)) / norm)
    g.fft(+1)
    b, _ = g.get_psi()
    rt = float(np.max(np.abs(b / n - x)) / np.max(np.abs(x)))
    say("parseval %.2e, coefficient errors / |x| %s, round trip %.2e" % (par, ["%.2e" % c for c in coef], rt))
    tol_p, tol_c, tol_r = (1e-8, 3e-6, 3e-6) if prec == E.C64 else (1e-13, 1e-13, 1e-13)
    assert par < tol_p
    assert max(coef) < tol_c
    assert rt < tol_r
    g.close()


def test_plane_wave_256cube_exact_phase():
    """A plane wave in a uniform potential only turns its phase:
    psi(t) = psi(0) exp(-i t (k^2/2 + 4 pi a rho)) (contact bosons, gamma 1).
    complex64: 10 steps of 4 transforms each, bound 1e-5 relative."""
    say("plane wave 256^3")
    nx, ny, nz = SHAPE
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij", sparse=True)
    k = (2 * np.pi * 3 / nx, -2 * np.pi * 2 / ny, 2 * np.pi * 5 / nz)
    psi = 0.7 * np.exp(1j * (k[0] * ix + k[1] * iy + k[2] * iz))
    dt, a, steps = 0.05, 0.013, 10
    g = E.Engine(SHAPE, 1.0, 0.0, dt, 1.0, E.CONTACT, E.HARMONIC, 0, E.C64)
    g.set_user_params([0.0, 0.0, 0.0, a])
    g.set_psi(psi)
    g.evolve_qf(steps)
    out, t = g.get_psi()
    ph = -dt * (sum(x * x for x in k) / 2 + 4 * np.pi * a * 0.49)
    want = psi * np.exp(1j * steps * ph)
    err = float(np.max(np.abs(out - want)) / 0.7)
    say("plane wave: max rel error %.2e after %d steps" % (err, steps))
    assert abs(t - steps * dt) < 1e-12
    assert err < 1e-5
    g.close()


def test_headline_workload_norm_energy_and_windings(oracle):
    """The headline workload in complex64 and complex128 for 50 steps: the
    real-time step is unitary, so the norm stays put (complex64 within 1e-6,
    complex128 within 1e-12); the complex64 energy stays within 1e-6 of the
    complex128 one; and the plaquette windings of the evolved complex64 state
    equal the oracle's bit for bit (20 straight lines pierce each of the 256
    xy planes at the start, so at least 20 xy plaquettes per plane wind)."""
    import bench

    say("headline workload 256^3")
    psi0 = bench.initial_state(SHAPE)
    res = {}
    for prec in (E.C64, E.C128):
        g = bench.make_engine(E, SHAPE, prec, psi0)
        n0 = g.get_norm()
        if prec == E.C64:
            w0, c0 = g.detect_vortices(0.05)
            # component 0: xy plaquettes (normal to z)
            assert np.count_nonzero(w0[0]) >= 20 * SHAPE[2]
        g.evolve_qf(50)
        n1 = g.get_norm()
        e = g.energy()
        res[prec] = (n1 / n0 - 1, sum(e[1:]))
        if prec == E.C64:
            wg, cg = g.detect_vortices(0.05)
            on_device = g.get_psi()[0].astype(np.complex64)
            wo, co = oracle.plaquette_windings(on_device, 0.05)
            say("windings: %d plaquettes (oracle %d)" % (cg, co))
            assert cg == co and cg > 0
            assert np.array_equal(wg, wo)
        g.close()
    (d64, E64), (d128, E128) = res[E.C64], res[E.C128]
    rel = abs(E64 - E128) / abs(E128)
    say("norm drift c64 %.2e c128 %.2e, energy c64 vs c128 %.2e" % (d64, d128, rel))
    assert abs(d64) < 1e-6
    assert abs(d128) < 1e-12
    assert rel < 1e-6
