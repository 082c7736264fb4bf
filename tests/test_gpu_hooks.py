"""The user hook boundary on the GPU (INTERACTIONS / VEXT 100: the hook file
include/gpe_user_hooks.h compiled into the library, as the reference compiles
its gpe_user_defined.h): a time-dependent potential, a ramped and three-body
functional and a non-identity gpe_modify_psi against the FP64 oracle built
with the same file, with graphs on and off, in every step variant; and
parameter sweeps, one d_user_param block per lattice of an ensemble."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

SHAPE = (32, 24, 16)


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def hook_params(scale=1.0):
    # omega x/y/z, a, trap modulation eps and W, g3, kick, a ramp, V ramp
    return [0.05, 0.06, 0.08, 0.01 * scale, 0.3, 1.7, 0.002 * scale, 0.2, 0.4, 0.05]


def state(shape, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.2 * rng.random(shape)) * np.exp(0.4j * rng.standard_normal(shape))


def pair(oracle, alpha, beta, prec=E.C128, dt=0.02, t0=0.25, qf=0.0, graph=False, seed=1):
    n = float(np.prod(SHAPE))
    g = E.Engine(SHAPE, alpha, beta, dt, n, E.USER, E.USER, 0, prec)
    o = oracle.OracleEngine(SHAPE, alpha, beta, dt, n, E.USER, E.USER, 0)
    g.set_graph(graph)
    for x in (g, o):
        x.set_user_params(hook_params())
        if qf:
            x.set_quantum_friction_coeff(qf)
        x.set_psi(state(SHAPE, seed), t=t0)
        x.normalize_psi()
    return g, o


def rel_err(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("alpha,beta,qf", [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (1.0, 0.0, 0.3), (0.7, 0.2, 0.0)])
def test_user_hooks_match_oracle(oracle, alpha, beta, qf, graph):
    say("user hooks", alpha, beta, qf, graph)
    g, o = pair(oracle, alpha, beta, qf=qf, graph=graph)
    want_mu = beta != 0.0
    for nt in (5, 4):  # two evolve calls: the time index carries over
        mg = g.evolve_qf(nt, want_mu=want_mu)
        mo = o.evolve_qf(nt, want_mu=want_mu)
        pg, tg = g.get_psi()
        po, to = o.get_psi()
        assert tg == to
        assert rel_err(pg, po) < 1e-12, rel_err(pg, po)
        if want_mu:
            assert abs(mg - mo) < 1e-12 * abs(mo)
    eg, eo = g.energy(), o.energy()
    assert abs(sum(eg[1:]) - sum(eo[1:])) < 1e-12 * abs(sum(eo[1:]))
    g.close()
    o.close()


@pytest.mark.parametrize("variant", ["vortex", "vortex2", "vortex3", "phase"])
def test_user_hooks_in_pinned_variants(oracle, variant):
    say("user hooks pinned", variant)
    g, o = pair(oracle, 0.0, 1.0, dt=0.01)
    for x in (g, o):
        if variant == "phase":
            x.set_phase(np.angle(state(SHAPE, 9)))
        else:
            x.set_vortex(0.5, -0.5, 1)
    name = {"vortex": "evolve_vortex", "vortex2": "evolve_vortex2", "vortex3": "evolve_vortex3",
            "phase": "evolve_enforced_phase"}[variant]
    mg = getattr(g, name)(6, want_mu=True)
    mo = getattr(o, name)(6, want_mu=True)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < 1e-12
    assert abs(mg - mo) < 1e-12 * abs(mo)
    g.close()
    o.close()


def test_user_hooks_graph_replays_are_direct_launches():
    """Captured step graphs read the time index from device memory: replays
    over several evolve calls equal direct launches bit for bit."""
    say("user hooks graph vs direct")
    out = []
    for graph in (False, True):
        g = E.Engine(SHAPE, 1.0, 0.0, 0.02, float(np.prod(SHAPE)), E.USER, E.USER, 0, E.C64)
        g.set_graph(graph)
        g.set_user_params(hook_params())
        g.set_psi(state(SHAPE, 3), t=0.1)
        g.evolve_qf(4)
        g.evolve_qf(7)
        out.append(g.get_psi()[0])
        g.close()
    assert np.array_equal(out[0], out[1])


def test_user_hooks_complex64(oracle):
    say("user hooks c64")
    g, o = pair(oracle, 1.0, 0.0, prec=E.C64)
    g.evolve_qf(20)
    o.evolve_qf(20)
    err = rel_err(g.get_psi()[0], o.get_psi()[0])
    say("complex64 vs oracle after 20 steps: %.2e" % err)
    assert err < 1e-5
    g.close()
    o.close()


@pytest.mark.parametrize("prec", [E.C64, E.C128])
def test_parameter_sweep_ensemble(oracle, prec):
    """Each lattice of an ensemble carries its own d_user_param block
    (gpe_ensemble_set_instance_params): every lattice equals a single engine
    with its parameters bit for bit, and complex128 matches the oracle."""
    say("parameter sweep", prec)
    shapes = [SHAPE, SHAPE, (16, 16, 24), SHAPE]
    scales = [1.0, 2.0, 0.5, 3.0]
    ens = E.Ensemble(shapes, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0, prec)
    ens.set_user_params(hook_params())
    for i, (s, k) in enumerate(zip(shapes, scales)):
        ens.set_instance_params(i, hook_params(k))
        ens.set_psi(i, state(s, 20 + i))
    ens.evolve(3)
    ens.evolve(4)
    for i, (s, k) in enumerate(zip(shapes, scales)):
        g = E.Engine(s, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0, prec)
        g.set_user_params(hook_params(k))
        g.set_variant(ens.groups()[ens.instance(i)[0]][2])
        g.set_psi(state(s, 20 + i))
        g.evolve_qf(7)
        got, want = ens.get_psi(i)[0], g.get_psi()[0]
        assert np.array_equal(got, want), i
        g.close()
        if prec == E.C128:
            o = oracle.OracleEngine(s, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0)
            o.set_user_params(hook_params(k))
            o.set_psi(state(s, 20 + i))
            o.evolve_qf(7)
            assert rel_err(got, o.get_psi()[0]) < 1e-12
            o.close()
    # built-in physics reads one parameter set per ensemble
    b = E.Ensemble([SHAPE], 1.0, 0.0, 0.02, 100.0, E.CONTACT, E.HARMONIC, 0, prec)
    with pytest.raises(E.GpeError):
        b.set_instance_params(0, [0.1])
    b.close()
    ens.close()
