"""Pins the oracle's plaquette phase-winding detector (a new component: the
reference's host tracker, dev/gpe_vortex_detect.h:61-174, finds one vortex per
plane by an 8-neighbour ordering test) with analytic vortex states and an
independent angle-sum computation, including the reference imprint's
atan2(x, y) orientation (dev/gpe_engine.cu:503)."""
import numpy as np
import pytest


def angle_sum_windings(psi, rho_min):
    """Independent restatement: sum of principal phase differences / 2 pi."""
    nx, ny, nz = psi.shape
    out = np.zeros((3,) + psi.shape, dtype=np.int8)
    roll = lambda a, s, ax: np.roll(a, -s, axis=ax)
    loops = [
        [psi, roll(psi, 1, 0), roll(roll(psi, 1, 0), 1, 1), roll(psi, 1, 1)],
        [psi, roll(psi, 1, 1), roll(roll(psi, 1, 1), 1, 2), roll(psi, 1, 2)],
        [psi, roll(psi, 1, 2), roll(roll(psi, 1, 2), 1, 0), roll(psi, 1, 0)],
    ]
    for o, c in enumerate(loops):
        tot = np.zeros(psi.shape)
        mask = np.ones(psi.shape, dtype=bool)
        for q in range(4):
            a, b = c[q], c[(q + 1) % 4]
            d = np.angle(b * np.conj(a))  # principal value in (-pi, pi]
            tot += d
            mask &= np.abs(a) ** 2 >= rho_min
        out[o] = np.where(mask, np.rint(tot / (2 * np.pi)), 0).astype(np.int8)
    return out


def vortex_lines_z(shape, cores, background=1.0):
    nx, ny, nz = shape
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ph = np.zeros((nx, ny))
    for (x0, y0, q) in cores:
        ph += q * np.arctan2(iy - y0, ix - x0)  # standard orientation: winding +q in x->y
    psi = background * np.exp(1j * ph)
    return np.repeat(psi[:, :, None], nz, axis=2)


def test_single_line_standard_orientation(oracle):
    shape = (16, 14, 6)
    psi = vortex_lines_z(shape, [(7.5, 6.5, 1)])
    w, cnt = oracle.plaquette_windings(psi, 1e-6)
    # the phase field is not periodic: plaquettes that wrap around the box close
    # the seam, so only the interior ones are checked
    assert np.count_nonzero(w[0, :-1, :-1, :]) == 6
    assert np.all(w[0, 7, 6, :] == 1)
    assert np.count_nonzero(w[1][:, :-1, :]) == 0 and np.count_nonzero(w[2][:-1, :, :]) == 0


def test_reference_imprint_orientation_is_negative_winding(oracle):
    # phase Q*atan2(x, y) winds -Q counter-clockwise in the xy plane
    nx, ny, nz = 12, 12, 4
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x0, y0 = 0.5, 0.5
    ph = 1 * np.arctan2(ix - nx // 2 - x0, iy - ny // 2 - y0)
    psi = np.repeat(np.exp(1j * ph)[:, :, None], nz, axis=2)
    w, cnt = oracle.plaquette_windings(psi, 1e-6)
    assert np.count_nonzero(w[0, :-1, :-1, :]) == nz
    assert np.all(w[0, nx // 2, ny // 2, :] == -1)


def test_many_vortices_positions_and_signs(oracle):
    rng = np.random.default_rng(3)
    shape = (64, 64, 4)
    cores = []
    for _ in range(12):
        cores.append((rng.integers(4, 60) + 0.5, rng.integers(4, 60) + 0.5, int(rng.choice([-1, 1]))))
    # keep cores apart so that every plaquette contains at most one
    cores = [c for i, c in enumerate(cores) if all(abs(c[0] - d[0]) + abs(c[1] - d[1]) > 3 for d in cores[:i])]
    psi = vortex_lines_z(shape, cores)
    w, cnt = oracle.plaquette_windings(psi, 1e-6)
    for (x0, y0, q) in cores:
        assert np.all(w[0, int(x0), int(y0), :] == q)
    # the field is not periodic, so the wrap-around seams may carry windings too;
    # every interior winding must be one of the imprinted cores
    interior = np.argwhere(w[0, :-1, :-1, 0] != 0)
    assert len(interior) == len(cores)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_random_field_matches_angle_sum(oracle, dtype):
    rng = np.random.default_rng(21)
    shape = (10, 12, 8)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    w, cnt = oracle.plaquette_windings(psi, 0.05)
    ref = angle_sum_windings(psi.astype(np.complex128), 0.05)
    assert np.array_equal(w, ref)
    assert cnt == np.count_nonzero(ref)
    # total topological charge through every closed surface is zero:
    # the windings of the six faces of each unit cell sum to zero
    unmasked = np.all(np.abs(psi) ** 2 >= 0.05)
    if unmasked:
        div = (np.roll(w[1], -1, 0) - w[1]) + (np.roll(w[2], -1, 1) - w[2]) + (np.roll(w[0], -1, 2) - w[0])
        assert np.count_nonzero(div) == 0


def test_closed_surface_charge_is_zero(oracle):
    rng = np.random.default_rng(4)
    shape = (8, 10, 6)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    w, _ = oracle.plaquette_windings(psi, 0.0)
    div = (np.roll(w[1], -1, 0) - w[1]) + (np.roll(w[2], -1, 1) - w[2]) + (np.roll(w[0], -1, 2) - w[0])
    assert np.count_nonzero(div) == 0
    assert np.count_nonzero(w) > 0
