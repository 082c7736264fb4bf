"""Whole real-time steps in one kernel (launcher variant bit 16): cubes whose
state fits in one workgroup's LDS run every step of an evolve call inside
fused_steps_kernel.  Its transforms are a different decomposition from the
axis passes (a lane pair per pencil), so it is held to the FP64 oracle
directly: complex128 to the 1e-12 parity bound in every physics kind the
kernel takes (contact, UFG, ETF, user hooks; harmonic, cosine, Woods-Saxon,
array, user potential), complex64 to a stated rounding bound; and ensemble
members on it equal single engines on it bit for bit.  Runs it does not cover
(imaginary time, quantum friction, pinning) fall back to the passes of the
lower variant bits, bit for bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

FUSED = 16


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def state(shape, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.2 * rng.standard_normal(shape)) * np.exp(0.4j * rng.standard_normal(shape))


def hook_params(scale=1.0):
    # omega x/y/z, a, trap modulation eps and W, g3, kick, a ramp, V ramp
    return [0.05, 0.06, 0.08, 0.01 * scale, 0.3, 1.7, 0.002 * scale, 0.2, 0.4, 0.05]


def varray(shape, seed):
    rng = np.random.default_rng(seed)
    return 0.05 * rng.random(shape)


# name, interactions, vext, params
PHYSICS = [
    ("contact_cosine", E.CONTACT, E.COSINE, [0.1, 2.0, 0.0, 0.02]),
    ("ufg_harmonic", E.UFG, E.HARMONIC, [0.1, 0.12, 0.15]),
    ("etf_woods_saxon", E.ETF, E.WOODS_SAXON, [1.0, 0.5, 3.0, 1.5]),
    ("contact_array", E.CONTACT, E.VARRAY, [0.0, 0.0, 0.0, 0.02]),
    ("user_hooks", E.USER, E.USER, hook_params()),
]


def make(cls, L, prec, phys, seed, t0=0.0, **kw):
    name, inter, vext, params = phys
    shape = (L, L, L)
    x = cls(shape, 1.0, 0.0, 0.02, float(L ** 3), inter, vext, 0, *([prec] if cls is E.Engine else []))
    x.set_user_params(params)
    if vext == E.VARRAY:
        if cls is E.Engine:
            x.set_external_potential(varray(shape, seed))
        else:
            x.set_vext_array(varray(shape, seed))
    x.set_psi(state(shape, seed), t=t0)
    return x


@pytest.mark.parametrize("phys", PHYSICS, ids=[p[0] for p in PHYSICS])
@pytest.mark.parametrize("L", [8, 12, 16])
def test_fused_c128_matches_oracle(oracle, L, phys):
    say("fused c128", L, phys[0])
    g = make(E.Engine, L, E.C128, phys, 3, t0=0.25)
    o = make(oracle.OracleEngine, L, None, phys, 3, t0=0.25)
    g.set_variant(FUSED)
    for nt in (1, 9):  # two calls: the time index carries over
        g.evolve_qf(nt)
        o.evolve_qf(nt)
        pg, tg = g.get_psi()
        po, to = o.get_psi()
        assert tg == to
        assert rel_err(pg, po) < 1e-12, (L, phys[0], nt, rel_err(pg, po))
    eg, eo = g.energy(), o.energy()
    assert abs(sum(eg[1:]) - sum(eo[1:])) < 1e-12 * abs(sum(eo[1:]))
    g.close()
    o.close()


@pytest.mark.parametrize("phys", [PHYSICS[0], PHYSICS[1], PHYSICS[4]], ids=["contact_cosine", "ufg_harmonic",
                                                                          "user_hooks"])
@pytest.mark.parametrize("L", [8, 12, 16, 24])
def test_fused_c64_matches_oracle(oracle, L, phys):
    """complex64 after 20 steps: within 2e-5 of max |psi| of the FP64 oracle,
    and no further from it than 3x the axis passes' complex64 error (plus
    1e-6)."""
    say("fused c64", L, phys[0])
    o = make(oracle.OracleEngine, L, None, phys, 5)
    o.evolve_qf(20)
    po = o.get_psi()[0]
    o.close()
    err = {}
    for v in (0, FUSED):
        g = make(E.Engine, L, E.C64, phys, 5)
        g.set_variant(v)
        g.evolve_qf(20)
        err[v] = rel_err(g.get_psi()[0], po)
        g.close()
    say("c64 vs oracle: passes %.2e fused %.2e" % (err[0], err[FUSED]))
    assert err[FUSED] < 2e-5
    assert err[FUSED] < 3 * err[0] + 1e-6


@pytest.mark.parametrize("prec,L", [(E.C64, 24), (E.C128, 16)])
def test_fused_long_runs_split_into_launches(prec, L):
    """An evolve call of more than 256 steps runs in several launches: 300
    steps in one call equal 120 + 180 steps in two, bit for bit, also with the
    time-dependent user hooks."""
    say("fused long run", prec, L)
    for phys in (PHYSICS[0], PHYSICS[4]):
        out = []
        for calls in ((300,), (120, 180)):
            g = make(E.Engine, L, prec, phys, 8)
            g.set_variant(FUSED)
            for nt in calls:
                g.evolve_qf(nt)
            out.append(g.get_psi())
            g.close()
        assert out[0][1] == out[1][1]
        assert np.array_equal(out[0][0], out[1][0]), phys[0]


@pytest.mark.parametrize("what", ["ite", "qf", "vortex", "complex_time"])
def test_fused_variant_falls_back_to_passes(what):
    """Runs the fused kernel does not cover take the passes of the lower bits:
    variant 16 + b gives the results of variant b bit for bit."""
    say("fused fallback", what)
    L = 16
    out = []
    for v in (2, FUSED | 2):
        alpha, beta = {"ite": (0.0, 1.0), "complex_time": (0.8, 0.2)}.get(what, (1.0, 0.0))
        g = E.Engine((L, L, L), alpha, beta, 0.02, float(L ** 3), E.CONTACT, E.COSINE, 0, E.C64)
        g.set_user_params(PHYSICS[0][3])
        g.set_variant(v)
        if what == "qf":
            g.set_quantum_friction_coeff(0.1)
        g.set_psi(state((L, L, L), 9))
        if what == "vortex":
            g.set_vortex(0.3, -0.2, 1)
            g.evolve_vortex(6)
        else:
            g.evolve_qf(6)
        out.append(g.get_psi()[0])
        g.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("prec", [E.C64, E.C128])
def test_fused_ensemble_matches_single_engines(prec, graph):
    """An ensemble whose small groups run the fused kernel on their own
    streams beside the passes of a 32^3 group: every lattice equals a single
    engine of the same variant bit for bit."""
    say("fused ensemble", prec, graph)
    small = 24 if prec == E.C64 else 16
    shapes = [(small,) * 3, (32, 32, 32), (small,) * 3, (8, 8, 8), (small,) * 3, (8, 8, 8)]
    params = PHYSICS[0][3]
    ens = E.Ensemble(shapes, 1.0, 0.0, 0.02, 500.0, E.CONTACT, E.COSINE, 0, prec)
    ens.set_user_params(params)
    ens.set_graph(graph)
    groups = ens.groups()
    for gi, (shape, _, v) in enumerate(groups):
        if shape[0] != 32:
            ens.set_variant(gi, FUSED | v)
    variants = {shape: v for (shape, _, v) in ens.groups()}
    for i, s in enumerate(shapes):
        ens.set_psi(i, state(s, 40 + i))
    ens.evolve(3)
    ens.evolve(5)
    for i, s in enumerate(shapes):
        g = E.Engine(s, 1.0, 0.0, 0.02, 500.0, E.CONTACT, E.COSINE, 0, prec)
        g.set_user_params(params)
        g.set_variant(variants[s])
        g.set_psi(state(s, 40 + i))
        g.evolve_qf(8)
        want, tw = g.get_psi()
        got, t = ens.get_psi(i)
        assert t == tw
        assert np.array_equal(got, want), (i, s)
        g.close()
    ens.close()


def test_fused_parameter_sweep_ensemble(oracle):
    """User hooks with one d_user_param block per lattice on the fused kernel
    (complex128): every lattice matches the oracle with its own parameters."""
    say("fused parameter sweep")
    L = 12
    scales = [1.0, 2.0, 0.5]
    shape = (L, L, L)
    ens = E.Ensemble([shape] * 3, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0, E.C128)
    ens.set_user_params(hook_params())
    ens.set_variant(0, FUSED)
    for i, k in enumerate(scales):
        ens.set_instance_params(i, hook_params(k))
        ens.set_psi(i, state(shape, 60 + i))
    ens.evolve(6)
    for i, k in enumerate(scales):
        o = oracle.OracleEngine(shape, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0)
        o.set_user_params(hook_params(k))
        o.set_psi(state(shape, 60 + i))
        o.evolve_qf(6)
        assert rel_err(ens.get_psi(i)[0], o.get_psi()[0]) < 1e-12, i
        o.close()
    ens.close()


def test_fused_autotune_keeps_a_valid_variant():
    say("fused autotune")
    ens = E.Ensemble([(24, 24, 24)] * 64, 1.0, 0.0, 0.02, 500.0, E.CONTACT, E.COSINE, 0, E.C64)
    ens.set_user_params(PHYSICS[0][3])
    for i in range(64):
        ens.set_psi(i, state((24, 24, 24), i))
    ens.autotune(3)
    v = ens.groups()[0][2]
    say("autotuned 24^3 x 64 variant", v)
    assert 0 <= v <= 127
    ens.evolve(2)
    ens.close()
