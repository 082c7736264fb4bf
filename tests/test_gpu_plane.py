"""The y-z plane plan (launcher variant bit 32): where ny = nz = 32, 48 or 64,
a real-time step is two plane passes (z and y transforms of a whole plane in
LDS with the physics between them) and two x turn-arounds.  Held to the FP64
oracle (complex128 to the 1e-12 parity bound, every physics kind the passes
take, time-dependent user hooks included), complex64 to a stated rounding
bound, ensemble members bit-equal to single engines of the same variant, and
bit-equal fallback to the lower variant bits for imaginary time, quantum
friction and pinning."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

PLANE = 32


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def state(shape, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.2 * rng.standard_normal(shape)) * np.exp(0.4j * rng.standard_normal(shape))


def hook_params(scale=1.0):
    return [0.05, 0.06, 0.08, 0.01 * scale, 0.3, 1.7, 0.002 * scale, 0.2, 0.4, 0.05]


PHYSICS = [
    ("contact_cosine", E.CONTACT, E.COSINE, [0.1, 2.0, 0.0, 0.02]),
    ("ufg_harmonic", E.UFG, E.HARMONIC, [0.1, 0.12, 0.15]),
    ("etf_woods_saxon", E.ETF, E.WOODS_SAXON, [1.0, 0.5, 6.0, 1.5]),
    ("contact_array", E.CONTACT, E.VARRAY, [0.0, 0.0, 0.0, 0.02]),
    ("user_hooks", E.USER, E.USER, hook_params()),
]


def make(cls, shape, prec, phys, seed, t0=0.0):
    name, inter, vext, params = phys
    x = cls(shape, 1.0, 0.0, 0.02, float(np.prod(shape)), inter, vext, 0, *([prec] if cls is E.Engine else []))
    x.set_user_params(params)
    if vext == E.VARRAY:
        v = 0.05 * np.random.default_rng(seed).random(shape)
        if cls is E.Engine:
            x.set_external_potential(v)
        else:
            x.set_vext_array(v)
    x.set_psi(state(shape, seed), t=t0)
    return x


@pytest.mark.parametrize("phys", PHYSICS, ids=[p[0] for p in PHYSICS])
@pytest.mark.parametrize("shape", [(16, 32, 32), (48, 48, 48), (24, 64, 64)])
def test_plane_c128_matches_oracle(oracle, shape, phys):
    say("plane c128", shape, phys[0])
    g = make(E.Engine, shape, E.C128, phys, 3, t0=0.25)
    o = make(oracle.OracleEngine, shape, None, phys, 3, t0=0.25)
    g.set_variant(PLANE)
    for nt in (1, 5):
        g.evolve_qf(nt)
        o.evolve_qf(nt)
        pg, tg = g.get_psi()
        po, to = o.get_psi()
        assert tg == to
        assert rel_err(pg, po) < 1e-12, (shape, phys[0], nt, rel_err(pg, po))
    g.close()
    o.close()


@pytest.mark.parametrize("phys", [PHYSICS[0], PHYSICS[1], PHYSICS[4]], ids=["contact_cosine", "ufg_harmonic",
                                                                          "user_hooks"])
@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 48, 48), (64, 64, 64)])
def test_plane_c64_matches_oracle(oracle, shape, phys):
    """complex64 after 12 steps: within 2e-5 of max |psi| of the FP64 oracle,
    and no further from it than 3x the six-pass plan's complex64 error (plus
    1e-6)."""
    say("plane c64", shape, phys[0])
    o = make(oracle.OracleEngine, shape, None, phys, 5)
    o.evolve_qf(12)
    po = o.get_psi()[0]
    o.close()
    err = {}
    for v in (0, PLANE):
        g = make(E.Engine, shape, E.C64, phys, 5)
        g.set_variant(v)
        g.evolve_qf(12)
        err[v] = rel_err(g.get_psi()[0], po)
        g.close()
    say("c64 vs oracle: six passes %.2e plane %.2e" % (err[0], err[PLANE]))
    assert err[PLANE] < 2e-5
    assert err[PLANE] < 3 * err[0] + 1e-6


@pytest.mark.parametrize("what", ["ite", "qf", "vortex", "complex_time"])
def test_plane_variant_falls_back_to_passes(what):
    say("plane fallback", what)
    shape = (16, 32, 32)
    out = []
    for v in (2, PLANE | 2):
        alpha, beta = {"ite": (0.0, 1.0), "complex_time": (0.8, 0.2)}.get(what, (1.0, 0.0))
        g = E.Engine(shape, alpha, beta, 0.02, float(np.prod(shape)), E.CONTACT, E.COSINE, 0, E.C64)
        g.set_user_params(PHYSICS[0][3])
        g.set_variant(v)
        if what == "qf":
            g.set_quantum_friction_coeff(0.1)
        g.set_psi(state(shape, 9))
        if what == "vortex":
            g.set_vortex(0.3, -0.2, 1)
            g.evolve_vortex(5)
        else:
            g.evolve_qf(5)
        out.append(g.get_psi()[0])
        g.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("prec", [E.C64, E.C128])
def test_plane_ensemble_matches_single_engines(prec, graph):
    say("plane ensemble", prec, graph)
    shapes = [(32, 32, 32), (24, 24, 24), (48, 48, 48), (32, 32, 32), (16, 64, 64)]
    params = PHYSICS[0][3]
    ens = E.Ensemble(shapes, 1.0, 0.0, 0.02, 500.0, E.CONTACT, E.COSINE, 0, prec)
    ens.set_user_params(params)
    ens.set_graph(graph)
    for gi, (shape, _, v) in enumerate(ens.groups()):
        if shape[1] in (32, 48, 64):
            ens.set_variant(gi, PLANE | v)
    variants = {shape: v for (shape, _, v) in ens.groups()}
    for i, s in enumerate(shapes):
        ens.set_psi(i, state(s, 70 + i))
    ens.evolve(2)
    ens.evolve(4)
    for i, s in enumerate(shapes):
        g = E.Engine(s, 1.0, 0.0, 0.02, 500.0, E.CONTACT, E.COSINE, 0, prec)
        g.set_user_params(params)
        g.set_variant(variants[s])
        g.set_psi(state(s, 70 + i))
        g.evolve_qf(6)
        want, tw = g.get_psi()
        got, t = ens.get_psi(i)
        assert t == tw
        assert np.array_equal(got, want), (i, s)
        g.close()
    ens.close()


@pytest.mark.parametrize("graph", [False, True])
def test_plane_user_hooks_parameter_sweep(oracle, graph):
    """Time-dependent user hooks with one d_user_param block per lattice, on
    the plane plan inside the ensemble graphs: every complex128 lattice
    matches the oracle with its own parameters."""
    say("plane parameter sweep", graph)
    shape = (8, 32, 32)
    scales = [1.0, 2.5]
    ens = E.Ensemble([shape] * 2, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0, E.C128)
    ens.set_graph(graph)
    ens.set_user_params(hook_params())
    ens.set_variant(0, PLANE)
    for i, k in enumerate(scales):
        ens.set_instance_params(i, hook_params(k))
        ens.set_psi(i, state(shape, 80 + i))
    ens.evolve(3)
    ens.evolve(4)
    for i, k in enumerate(scales):
        o = oracle.OracleEngine(shape, 1.0, 0.0, 0.02, 100.0, E.USER, E.USER, 0)
        o.set_user_params(hook_params(k))
        o.set_psi(state(shape, 80 + i))
        o.evolve_qf(7)
        assert rel_err(ens.get_psi(i)[0], o.get_psi()[0]) < 1e-12, i
        o.close()
    ens.close()


def test_plane_graph_replays_match_direct_launches():
    say("plane graph")
    shape = (64, 64, 64)
    out = []
    for graph in (False, True):
        g = make(E.Engine, shape, E.C64, PHYSICS[0], 11)
        g.set_variant(PLANE)
        g.set_graph(graph)
        g.evolve_qf(5)
        g.evolve_qf(7)
        out.append(g.get_psi()[0])
        g.close()
    assert np.array_equal(out[0], out[1])
