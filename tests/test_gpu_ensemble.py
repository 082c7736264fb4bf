"""Ensembles of uneven lattices (gpe_ensemble_*): every lattice must follow
exactly the split step of a single engine of its shape.  The batched passes
run the same arithmetic per point as the single-lattice passes, so results
are compared bit for bit; complex128 lattices are also checked against the
FP64 oracle to the 1e-12 parity bound."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

SHAPES = [(24, 24, 24), (32, 32, 32), (24, 24, 24), (16, 8, 24), (32, 32, 32), (24, 24, 24), (16, 8, 24)]
PARAMS = [0.1, 4.0, 0.0, 0.02]  # cosine lattice A=0.1, P=4; contact a=0.02


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def state(shape, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.1 * rng.standard_normal(shape)) * np.exp(1j * 0.3 * rng.standard_normal(shape))


def run_single(shape, psi, prec, steps, qf=0.0):
    g = E.Engine(shape, 1.0, 0.0, 0.02, float(np.prod(shape)), E.CONTACT, E.COSINE, 0, prec)
    g.set_user_params(PARAMS)
    if qf:
        g.set_quantum_friction_coeff(qf)
    g.set_psi(psi)
    g.evolve_qf(steps)
    out, t = g.get_psi()
    g.close()
    return out, t


def make_ensemble(prec, qf=0.0, graph=True):
    ens = E.Ensemble(SHAPES, 1.0, 0.0, 0.02, 1000.0, E.CONTACT, E.COSINE, 0, prec)
    ens.set_user_params(PARAMS)
    if qf:
        ens.set_quantum_friction_coeff(qf)
    ens.set_graph(graph)
    for i, s in enumerate(SHAPES):
        ens.set_psi(i, state(s, 100 + i))
    return ens


@pytest.mark.parametrize("prec", [E.C64, E.C128])
@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_matches_single_engines(prec, graph):
    say("ensemble vs single", prec, graph)
    steps = 5
    ens = make_ensemble(prec, graph=graph)
    groups = ens.groups()
    assert [g[0] for g in groups] == [(24, 24, 24), (32, 32, 32), (16, 8, 24)]
    assert [g[1] for g in groups] == [3, 2, 2]
    ens.evolve(steps)
    for i, s in enumerate(SHAPES):
        want, tw = run_single(s, state(s, 100 + i), prec, steps)
        got, t = ens.get_psi(i)
        assert t == tw
        assert np.array_equal(got, want), (i, s)
    ens.close()


def test_ensemble_quantum_friction_and_replays():
    """QF runs inside the graphs; two evolve calls (two rounds of graph
    replays with host work between them) equal one single-engine run."""
    say("ensemble qf")
    ens = make_ensemble(E.C64, qf=0.2)
    ens.evolve(3)
    ens.evolve(4)
    for i in (0, 1, 3):
        s = SHAPES[i]
        want, _ = run_single(s, state(s, 100 + i), E.C64, 7, qf=0.2)
        assert np.array_equal(ens.get_psi(i)[0], want), i
    ens.close()


def test_ensemble_c128_matches_oracle(oracle):
    say("ensemble vs oracle")
    ens = make_ensemble(E.C128)
    ens.evolve(4)
    for i in (0, 3):
        s = SHAPES[i]
        o = oracle.OracleEngine(s, 1.0, 0.0, 0.02, float(np.prod(s)), E.CONTACT, E.COSINE, 0)
        o.set_user_params(PARAMS)
        o.set_psi(state(s, 100 + i))
        o.evolve_qf(4)
        po, _ = o.get_psi()
        got, _ = ens.get_psi(i)
        assert np.max(np.abs(got - po)) / np.max(np.abs(po)) < 1e-12
    ens.close()


def test_ensemble_mean_density_and_autotune():
    say("ensemble mean density")
    ens = make_ensemble(E.C64)
    ens.autotune(2)
    variants = [g[2] for g in ens.groups()]
    # bits 0-2 of the passes, bit 3 as created, bit 4 the fused whole-step
    # kernel, bit 5 the y-z plane plan, bit 6 the alternative radix schedule
    assert all(0 <= v <= 127 for v in variants)
    ens.evolve(3)
    for gi, (shape, count, _) in enumerate(ens.groups()):
        members = [i for i in range(len(SHAPES)) if ens.instance(i)[0] == gi]
        assert len(members) == count
        rho = np.mean([np.abs(ens.get_psi(i)[0]) ** 2 for i in members], axis=0)
        got = ens.mean_density(gi)
        # gamma = 1 for contact bosons
        assert np.max(np.abs(got - rho)) <= 1e-12 * np.max(rho)
    ens.close()


def test_ensemble_rejects_imaginary_time():
    ens = E.Ensemble([(16, 16, 16), (16, 16, 16)], 0.0, 1.0, 0.01, 10.0, E.CONTACT, E.HARMONIC, 0, E.C64)
    with pytest.raises(E.GpeError) as ex:
        ens.evolve(1)
    assert ex.value.code == -12
    ens.close()
