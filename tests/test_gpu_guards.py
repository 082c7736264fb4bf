"""Entry points that launch step or energy kernels refuse, on the host and
before any launch, an engine whose VEXT 3 potential array was never set
(GPE_ERR_ARG = -12); the adjoint backward run refuses a parameter change
made after its forward run."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

ERR_ARG = -12


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def expect_arg_error(fn, *args):
    with pytest.raises(E.GpeError) as ex:
        fn(*args)
    assert ex.value.code == ERR_ARG


def test_missing_potential_array_is_refused_everywhere():
    say("vext=3 guards")
    shape = (16, 16, 16)
    g = E.Engine(shape, 1.0, 0.0, 0.01, 10.0, E.CONTACT, E.VARRAY, 0, E.C64)
    g.set_psi(np.ones(shape))
    expect_arg_error(g.evolve_qf, 2)
    expect_arg_error(g.energy)
    expect_arg_error(g.autotune, 1)
    expect_arg_error(g.adjoint_forward, 2, np.ones(shape))
    g.close()
    d = E.Engine(shape, 1.0, 0.0, 0.01, 10.0, E.DIPOLAR, E.VARRAY, 0, E.C64)
    d.set_psi(np.ones(shape))
    expect_arg_error(d.evolve_dipolar, 1)
    expect_arg_error(d.energy_dipolar, 1)
    d.close()
    ens = E.Ensemble([shape, shape, (8, 8, 16)], 1.0, 0.0, 0.01, 10.0, E.CONTACT, E.VARRAY, 0, E.C64)
    for i, s in enumerate(ens.shapes):
        ens.set_psi(i, np.ones(s))
    expect_arg_error(ens.evolve, 2)
    expect_arg_error(ens.autotune, 1)
    # one lattice of a group still without its array
    ens.set_external_potential(0, np.zeros(shape))
    ens.set_external_potential(2, np.zeros((8, 8, 16)))
    expect_arg_error(ens.evolve, 2)
    ens.set_external_potential(1, np.zeros(shape))
    ens.evolve(2)
    ens.close()


def test_adjoint_backward_after_a_parameter_change_is_refused():
    say("adjoint generation")
    shape = (16, 16, 16)
    g = E.Engine(shape, 1.0, 0.0, 0.01, 10.0, E.CONTACT, E.VARRAY, 0, E.C64)
    g.set_user_params([0.0, 0.0, 0.0, 0.01])
    g.set_external_potential(np.zeros(shape))
    g.set_psi(np.ones(shape))
    g.adjoint_forward(3, np.ones(shape))
    g.set_user_params([0.0, 0.0, 0.0, 0.02])
    expect_arg_error(g.adjoint_backward)
    g.adjoint_forward(3, np.ones(shape))
    assert g.adjoint_backward().shape == shape
    g.close()


def test_per_lattice_potential_arrays_of_an_ensemble():
    """Each lattice of a VEXT 3 ensemble steps with its own potential: bit for
    bit the single engine given that potential."""
    say("ensemble per-lattice potential")
    shape = (16, 8, 24)
    rng = np.random.default_rng(3)
    vs = [0.1 * rng.random(shape) for _ in range(3)]
    psis = [np.exp(0.3j * rng.standard_normal(shape)) for _ in range(3)]
    ens = E.Ensemble([shape] * 3, 1.0, 0.0, 0.02, 10.0, E.CONTACT, E.VARRAY, 0, E.C64)
    ens.set_user_params([0.0, 0.0, 0.0, 0.02])
    for i in range(3):
        ens.set_external_potential(i, vs[i])
        ens.set_psi(i, psis[i])
    ens.evolve(4)
    for i in range(3):
        g = E.Engine(shape, 1.0, 0.0, 0.02, 10.0, E.CONTACT, E.VARRAY, 0, E.C64)
        g.set_user_params([0.0, 0.0, 0.0, 0.02])
        g.set_external_potential(vs[i])
        g.set_psi(psis[i])
        g.evolve_qf(4)
        assert np.array_equal(ens.get_psi(i)[0], g.get_psi()[0]), i
        g.close()
    ens.close()
