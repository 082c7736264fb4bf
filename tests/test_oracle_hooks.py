"""Pins the oracle's user-hook path (INTERACTIONS / VEXT 100, the hook file
include/gpe_user_hooks.h compiled into both the oracle and the GPU library):
where the reference calls gpe_modify_psi, gpe_external_potential and
gpe_dEDFdn with the time index it or it + 1 (orginal/gpe_engine.cu:737-739,
792, 826-830, 798, 838, 865), checked against closed forms."""
import numpy as np
import pytest

USER = 100
CONTACT, HARMONIC = -1, 0
# slots of the example hook file
OMEGA, A_SCAT, TRAP_EPS, TRAP_W, G3, KICK, A_RAMP, V_RAMP = 0, 3, 4, 5, 6, 7, 8, 9


def params(**kw):
    p = np.zeros(10)
    p[:3] = kw.get("omega", (0.0, 0.0, 0.0))
    for name, slot in (("a", A_SCAT), ("eps", TRAP_EPS), ("w", TRAP_W), ("g3", G3), ("kick", KICK),
                       ("a_ramp", A_RAMP), ("v_ramp", V_RAMP)):
        p[slot] = kw.get(name, 0.0)
    return p


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.0, 1.0)])
def test_hooks_with_zero_extras_are_the_builtin_contact_gas(oracle, alpha, beta):
    """With slots 4..9 zero the example hooks are the contact gas in the
    harmonic trap: the same trajectory bit for bit."""
    shape = (16, 12, 8)
    rng = np.random.default_rng(2)
    psi = (1 + 0.2 * rng.random(shape)) * np.exp(0.3j * rng.standard_normal(shape))
    p = params(omega=(0.2, 0.25, 0.3), a=0.03)
    out = []
    for inter, vext in ((USER, USER), (CONTACT, HARMONIC)):
        e = oracle.OracleEngine(shape, alpha, beta, 0.05, 100.0, inter, vext, 1)
        e.set_user_params(p)
        e.set_psi(psi)
        e.normalize_psi()
        mu = e.evolve_qf(5, want_mu=beta != 0)
        out.append((e.get_psi()[0], e.energy(), mu))
        e.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("gam", [1, 2])
def test_time_dependent_hooks_uniform_state_exact_phase(oracle, gam):
    """A uniform state with a time-dependent uniform potential V = b t and a
    ramped contact coupling a(t) = a0 (1 + r t) plus a three-body term: each
    step multiplies psi by exp(-i gamma dt (V(t_n) + V(t_n+1))/2) (predictor
    potential at it, corrector average with it + 1), the trapezoidal rule, exact
    for a potential linear in t.  So after N steps from t0 the phase is
    -gamma times the integral of b t + 4 pi a(t) rho + g3 rho^2 from t0 to t0 + T.
    Evaluating the corrector at it instead of it + 1 would miss it by
    gamma (b + 4 pi a0 r rho) T dt / 2."""
    shape = (8, 6, 4)
    amp, dt, nsteps, t0 = 0.8, 0.04, 25, 0.3
    a0, r, b, g3 = 0.02, 0.7, 0.15, 0.05
    e = oracle.OracleEngine(shape, 1.0, 0.0, dt, 1.0, USER, USER, gam)
    e.set_user_params(params(a=a0, a_ramp=r, v_ramp=b, g3=g3))
    e.set_psi(np.full(shape, amp, dtype=complex), t=t0)
    e.evolve_qf(nsteps)
    out, t = e.get_psi()
    e.close()
    T = nsteps * dt
    t1 = t0 + T
    rho = gam * amp * amp
    phase = -gam * (b * (t1 * t1 - t0 * t0) / 2 + 4 * np.pi * a0 * rho * (T + r * (t1 * t1 - t0 * t0) / 2)
                    + g3 * rho * rho * T)
    assert abs(t - t1) < 1e-15
    assert np.max(np.abs(out - amp * np.exp(1j * phase))) < 1e-13


def _kick_step_expected(nx, amp, kappa, w, t, dt, gam):
    """One step from a uniform state with only the phase kick: the kicked state
    amp exp(i kappa sin(2 pi ix/nx - w t)) = amp sum_m J_m(kappa) e^{i m (2 pi ix/nx - w t)}
    (Jacobi-Anger) sampled on the lattice, harmonic m landing on the lattice
    mode q = m mod nx, then the free step exp(-i dt k_q^2 / (2 gamma))."""
    from scipy.special import jv

    coef = np.zeros(nx, dtype=complex)
    for m in range(-60, 61):
        coef[m % nx] += jv(m, kappa) * np.exp(-1j * m * w * t)
    k = 2 * np.pi * np.fft.fftfreq(nx)
    coef *= np.exp(-1j * dt * k * k / (2 * gam))
    ix = np.arange(nx)
    return amp * np.exp(2j * np.pi * np.outer(ix, np.arange(nx)) / nx) @ coef


@pytest.mark.parametrize("gam", [1, 2])
def test_modify_psi_applied_at_the_step_time(oracle, gam):
    """The reference applies gpe_modify_psi to psi0 with the step's own time
    index it before the kinetic step of the corrector (orginal/gpe_engine.cu:
    798).  A phase kick exp(i kappa sin(2 pi ix/nx - w t)) on a uniform state
    with no potential and no interaction then gives, after one step, the
    Jacobi-Anger series of the kick at time t0 + it dt propagated freely for dt.
    Checked for the first step and for a step taken at index 2."""
    shape = (32, 4, 4)
    amp, dt, t0, kappa, w = 0.6, 0.3, 0.45, 0.35, 1.3
    e = oracle.OracleEngine(shape, 1.0, 0.0, dt, 1.0, USER, USER, gam)
    e.set_user_params(params(kick=kappa, w=w))
    e.set_psi(np.full(shape, amp, dtype=complex), t=t0)
    e.evolve_qf(1)
    out, _ = e.get_psi()
    want = _kick_step_expected(32, amp, kappa, w, t0, dt, gam)
    assert np.max(np.abs(out - want[:, None, None])) < 1e-12
    # two steps with no kick (the uniform state is unchanged), then one kicked step at it = 2
    e.set_user_params(params(w=w))
    e.set_psi(np.full(shape, amp, dtype=complex), t=t0)
    e.evolve_qf(2)
    e.set_user_params(params(kick=kappa, w=w))
    e.evolve_qf(1)
    out, _ = e.get_psi()
    e.close()
    want = _kick_step_expected(32, amp, kappa, w, t0 + 2 * dt, dt, gam)
    assert np.max(np.abs(out - want[:, None, None])) < 1e-12


def test_user_energy_uses_the_current_time(oracle):
    """gpe_energy evaluates gpe_EDF and gpe_external_potential at the current
    time index (orginal/gpe_engine.cu:1095): on a uniform state E_int =
    N (2 pi a(t) rho^2 + g3 rho^3 / 3) and E_ext = N rho b t."""
    shape = (8, 6, 4)
    amp, dt, t0 = 0.7, 0.05, 0.2
    a0, r, b, g3 = 0.03, 0.5, 0.2, 0.04
    e = oracle.OracleEngine(shape, 0.0, 1.0, dt, float(np.prod(shape)) * amp * amp, USER, USER, 1)
    e.set_user_params(params(a=a0, a_ramp=r, v_ramp=b, g3=g3))
    e.set_psi(np.full(shape, amp, dtype=complex), t=t0)
    e.evolve_qf(7)  # imaginary time keeps the uniform state and its norm
    t, ek, ei, ee = e.energy()
    e.close()
    N = np.prod(shape)
    rho = amp * amp
    assert abs(t - (t0 + 7 * dt)) < 1e-15
    assert abs(ek) < 1e-12
    assert abs(ei - N * (2 * np.pi * a0 * (1 + r * t) * rho * rho + g3 * rho ** 3 / 3)) < 1e-12 * abs(ei)
    assert abs(ee - N * rho * b * t) < 1e-12 * abs(ee)
