"""The oracle's discrete adjoint (gpeo_adjoint_overlap) against central finite
differences of its own forward step: single lattice points and a random
direction, for contact bosons and the unitary Fermi gas, with and without
checkpoint recomputation.  These pin the adjoint's derivation; the GPU
adjoint is then held to the oracle."""
import numpy as np
import pytest

SHAPE = (8, 6, 4)


def setup(oracle, inter, v, psi0, dt=0.05):
    o = oracle.OracleEngine(SHAPE, 1.0, 0.0, dt, 50.0, inter, 3, 0)
    o.set_user_params([0.0, 0.0, 0.0, 0.05])
    o.set_vext_array(v)
    o.set_psi(psi0)
    return o


def loss_of(oracle, inter, v, psi0, psit, nt):
    o = setup(oracle, inter, v, psi0)
    o.evolve_qf(nt)
    p, _ = o.get_psi()
    o.close()
    ov = np.vdot(psit, p)
    return 1.0 - abs(ov) ** 2 / (np.vdot(psit, psit).real * np.vdot(psi0, psi0).real)


def inputs(seed):
    rng = np.random.default_rng(seed)
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in SHAPE], indexing="ij")
    v = 0.3 * np.cos(2 * np.pi * ix / SHAPE[0]) + 0.2 * np.sin(2 * np.pi * (iy + iz) / 6) + 0.05 * rng.random(SHAPE)
    psi0 = (1 + 0.3 * rng.random(SHAPE)) * np.exp(1j * rng.uniform(-1, 1, SHAPE))
    psit = (1 + 0.3 * rng.random(SHAPE)) * np.exp(1j * rng.uniform(-1, 1, SHAPE))
    return v, psi0, psit


@pytest.mark.parametrize("inter", [-1, 0])
def test_adjoint_matches_finite_differences(oracle, inter):
    v, psi0, psit = inputs(3)
    nt = 6
    o = setup(oracle, inter, v, psi0)
    L, g = o.adjoint_overlap(nt, psit, ckpt=2)
    o.close()
    assert abs(L - loss_of(oracle, inter, v, psi0, psit, nt)) < 1e-14
    h = 1e-5
    rng = np.random.default_rng(1)
    for _ in range(5):
        idx = tuple(int(rng.integers(0, n)) for n in SHAPE)
        vp, vm = v.copy(), v.copy()
        vp[idx] += h
        vm[idx] -= h
        fd = (loss_of(oracle, inter, vp, psi0, psit, nt) - loss_of(oracle, inter, vm, psi0, psit, nt)) / (2 * h)
        assert abs(fd - g[idx]) < 1e-7 * max(abs(fd), 1e-3), (idx, fd, g[idx])
    d = rng.standard_normal(SHAPE)
    fd = (loss_of(oracle, inter, v + h * d, psi0, psit, nt) - loss_of(oracle, inter, v - h * d, psi0, psit, nt)) / (2 * h)
    assert abs(fd - np.sum(g * d)) < 1e-7 * abs(fd)


def test_adjoint_independent_of_checkpointing(oracle):
    v, psi0, psit = inputs(5)
    out = []
    for ck in (1, 3, 0, 7):
        o = setup(oracle, -1, v, psi0)
        out.append(o.adjoint_overlap(7, psit, ckpt=ck))
        final = o.get_psi()[0]
        o.close()
    for L, g in out[1:]:
        assert L == out[0][0]
        assert np.array_equal(g, out[0][1])
    # psi is left at psi(T)
    o = setup(oracle, -1, v, psi0)
    o.evolve_qf(7)
    assert np.array_equal(o.get_psi()[0], final)
    o.close()


def test_adjoint_rejects_unsupported(oracle):
    v, psi0, psit = inputs(2)
    o = setup(oracle, 1, v, psi0)  # ETF: no derivative of the functional here
    with pytest.raises(ValueError):
        o.adjoint_overlap(2, psit)
    o.close()
