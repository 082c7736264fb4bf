"""GPU parity: the HIP engine through its C ABI against the FP64 CPU oracle on
the same seeded inputs (complex128 to 1e-12 relative, complex64 to the
tolerances stated per test), plus closed forms on the GPU."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402

TOL128 = 1e-12


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def seeded_state(shape, seed, phase=True):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r2 = ((ix - nx / 2) / (nx / 4)) ** 2 + ((iy - ny / 2) / (ny / 4)) ** 2 + ((iz - nz / 2) / (nz / 4)) ** 2
    amp = np.exp(-r2) * (1 + 0.05 * rng.random(shape))
    ph = 0.3 * rng.standard_normal(shape) if phase else 0.0
    return (amp * np.exp(1j * ph)).astype(np.complex128)


def pair(oracle, shape, alpha, beta, dt, npart, interactions, vext, params, precision=E.C128, gpe_for=0):
    g = E.Engine(shape, alpha, beta, dt, npart, interactions, vext, gpe_for, precision)
    o = oracle.OracleEngine(shape, alpha, beta, dt, npart, interactions, vext, gpe_for)
    g.set_user_params(params)
    o.set_user_params(params)
    return g, o


@pytest.mark.parametrize("shape", [(32, 32, 32), (256, 32, 32), (24, 16, 12), (48, 8, 64), (16, 96, 8), (8, 8, 512),
                                   (512, 8, 8), (12, 8, 24), (8, 12, 48)])
@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_fft3d_matches_numpy(shape, prec):
    say("fft3d", shape, prec)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    g = E.Engine(shape, 1.0, 0.0, 0.01, 1.0, precision=prec)
    g.set_psi(x)
    g.fft(-1)
    f, _ = g.get_psi()
    tol = 1e-14 if prec == E.C128 else 3e-6
    assert rel_err(f, np.fft.fftn(x)) < tol
    g.fft(+1)
    b, _ = g.get_psi()
    assert rel_err(b / x.size, x) < tol
    g.close()


@pytest.mark.parametrize("prec,tol", [(E.C128, 1e-13), (E.C64, 2e-6)])
def test_plane_wave_exact_phase(prec, tol):
    say("plane wave", prec)
    shape = (32, 16, 24)
    nx, ny, nz = shape
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    k = (2 * np.pi * 3 / nx, -2 * np.pi * 2 / ny, 2 * np.pi * 1 / nz)
    psi = 0.7 * np.exp(1j * (k[0] * ix + k[1] * iy + k[2] * iz))
    dt, a = 0.05, 0.013
    g = E.Engine(shape, 1.0, 0.0, dt, 1.0, E.CONTACT, E.HARMONIC, 0, prec)
    g.set_user_params([0.0, 0.0, 0.0, a])
    g.set_psi(psi)
    g.evolve_qf(4)
    out, t = g.get_psi()
    ph = -dt * (sum(x * x for x in k) / 2 + 4 * np.pi * a * 0.49)
    assert abs(t - 4 * dt) < 1e-15
    assert rel_err(out, psi * np.exp(4j * ph)) < tol
    g.close()


@pytest.mark.parametrize("mode", ["rte", "ite", "qf"])
def test_512_point_pencils_complex64_vs_complex128(mode):
    """z pencils of 512 points: complex64 runs them in the wave kernels (16
    lanes x 32 points, two-phase exchange), complex128 in the LDS-tile
    kernels, so the two precisions check each other's code paths.  Bound:
    2e-5 of max |psi| after 12 steps (complex64 rounding)."""
    say("512-point pencils", mode)
    shape = (8, 12, 512)
    psi = seeded_state(shape, 21)
    out = {}
    for prec in (E.C64, E.C128):
        alpha, beta = (1.0, 0.0) if mode != "ite" else (0.0, 1.0)
        g = E.Engine(shape, alpha, beta, 0.02, float(np.prod(shape)) * 0.1, E.CONTACT, E.COSINE, 0, prec)
        g.set_user_params([0.05, 4.0, 0.0, 0.01])
        if mode == "qf":
            g.set_quantum_friction_coeff(0.1)
        g.set_psi(psi)
        g.normalize_psi()
        g.evolve_qf(12)
        out[prec] = g.get_psi()[0]
        g.close()
    err = rel_err(out[E.C64], out[E.C128])
    say("512 %s: c64 vs c128 %.2e" % (mode, err))
    assert err < 2e-5


CASES = [
    # name, shape, alpha, beta, interactions, vext, params, gpe_for
    ("ufg_harmonic", (32, 32, 32), 0.1, [0.1, 0.12, 0.15], 0, 0, 0),
    ("compare_lattice", (256, 32, 32), 0.025, [0.01, 0.10, 0.11], 0, 0, 0),
    ("contact_cosine", (24, 16, 12), 0.02, [0.1, 2.0, 0.0, 0.02], -1, 2, 0),
    ("etf_woods_saxon", (48, 8, 64), 0.02, [1.0, 0.5, 6.0, 1.5], 1, 1, 0),
    ("ufg_particles", (16, 96, 8), 0.03, [0.2, 0.05, 0.3], 0, 0, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("mode", ["rte", "ite"])
def test_evolve_matches_oracle(oracle, case, mode):
    name, shape, dt, params, inter, vext, gfor = case
    alpha, beta = (1.0, 0.0) if mode == "rte" else (0.0, 1.0)
    say("evolve", name, mode)
    g, o = pair(oracle, shape, alpha, beta, dt, 800.0, inter, vext, params, gpe_for=gfor)
    psi = seeded_state(shape, 7)
    for x in (g, o):
        x.set_psi(psi)
        x.normalize_psi()
    for nsteps in (1, 9):
        mg = g.evolve_qf(nsteps, want_mu=True)
        mo = o.evolve_qf(nsteps, want_mu=True)
        pg, tg = g.get_psi()
        po, to = o.get_psi()
        assert tg == to
        assert rel_err(pg, po) < TOL128, (name, mode, nsteps)
        if mode == "ite":
            assert abs(mg - mo) <= TOL128 * max(1.0, abs(mo))
    eg, eo = g.energy(), o.energy()
    for a, b in zip(eg[1:], eo[1:]):
        assert abs(a - b) <= TOL128 * max(abs(b), 1e-3 * sum(abs(x) for x in eo[1:]))
    assert abs(g.get_norm() - o.get_norm()) < TOL128 * 800.0
    g.close()


@pytest.mark.parametrize("mode", ["rte", "ite", "complex"])
def test_quantum_friction_matches_oracle(oracle, mode):
    alpha, beta = {"rte": (1.0, 0.0), "ite": (0.0, 1.0), "complex": (1.0, 0.1)}[mode]
    say("qf", mode)
    shape = (32, 24, 16)
    g, o = pair(oracle, shape, alpha, beta, 0.02, 500.0, 0, 0, [0.1, 0.12, 0.15])
    psi = seeded_state(shape, 11)
    for x in (g, o):
        x.set_quantum_friction_coeff(0.3)
        x.set_psi(psi)
        x.normalize_psi()
    g.evolve_qf(6)
    o.evolve_qf(6)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128
    # switching quantum friction off mid-run
    for x in (g, o):
        x.set_quantum_friction_coeff(0.0)
    g.evolve_qf(2)
    o.evolve_qf(2)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128
    g.close()


@pytest.mark.parametrize("mode", ["rte", "ite"])
def test_vortex_evolution_matches_oracle(oracle, mode):
    alpha, beta = (1.0, 0.0) if mode == "rte" else (0.0, 1.0)
    say("vortex", mode)
    shape = (32, 32, 16)
    g, o = pair(oracle, shape, alpha, beta, 0.025, 1000.0, 0, 0, [0.05, 0.05, 0.1])
    psi = seeded_state(shape, 3, phase=False)
    for x in (g, o):
        x.set_psi(psi)
        x.normalize_psi()
        x.set_vortex(0.5, 3.5, 1)
    mg = g.evolve_vortex(7, want_mu=True)
    mo = o.evolve_vortex(7, want_mu=True)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128
    if mode == "ite":
        assert abs(mg - mo) < TOL128 * abs(mo)
    g.close()


def test_vortex_requires_set_vortex():
    g = E.Engine((16, 16, 16), 0.0, 1.0, 0.01, 1.0)
    with pytest.raises(E.GpeError) as ex:
        g.evolve_vortex(1)
    assert ex.value.code == -13
    g.close()


def test_observables_match_oracle(oracle):
    say("observables")
    shape = (24, 32, 16)
    g, o = pair(oracle, shape, 1.0, 0.0, 0.01, 300.0, 0, 0, [0.1, 0.12, 0.15])
    psi = seeded_state(shape, 5)
    for x in (g, o):
        x.set_psi(psi, t=2.5)
    # density and currents (orginal/gpe_engine.cu:1115-1212)
    dg, tg = g.get_density()
    do, to = o.get_density()
    assert tg == to == 2.5
    assert rel_err(dg, do) < 1e-15
    jg, _ = g.get_currents()
    jo, _ = o.get_currents()
    for a, b in zip(jg, jo):
        assert rel_err(a, b) < TOL128
    # energy renormalises in real time (orginal/gpe_engine.cu:1084-1088)
    eg, eo = g.energy(), o.energy()
    for a, b in zip(eg, eo):
        assert abs(a - b) <= TOL128 * max(abs(b), 1.0)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128
    assert abs(g.get_norm() - 300.0) < 1e-9
    g.close()


@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_windings_bit_exact(oracle, prec):
    say("windings", prec)
    rng = np.random.default_rng(8)
    shape = (32, 24, 16)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    g = E.Engine(shape, 1.0, 0.0, 0.01, 1.0, precision=prec)
    g.set_psi(psi)
    wg, cg = g.detect_vortices(0.05)
    on_device = g.get_psi()[0]
    ref_in = on_device.astype(np.complex64) if prec == E.C64 else on_device
    wo, co = oracle.plaquette_windings(ref_in, 0.05)
    assert cg == co and cg > 0
    assert np.array_equal(wg, wo)
    g.close()


@pytest.mark.parametrize("state,steps,tol", [("vortex", 500, 1e-6), ("random", 200, 2e-6)])
def test_c64_energy_error_vs_fp64(oracle, state, steps, tol):
    """complex64 engine vs FP64 oracle in real time on 64^3.  The vortex state
    is the bench workload's (near-uniform density, 20 vortex lines, cosine
    potential); the random-phase state (phase noise of 0.5 rad per site, so
    |k| up to pi everywhere) is the harsh case and carries a looser bound."""
    say("c64 accuracy", state)
    import bench

    shape = (64, 64, 64)
    params = [0.1, 2.0, 0.0, 0.008]
    g, o = pair(oracle, shape, 1.0, 0.0, 0.02, float(np.prod(shape)), E.CONTACT, E.COSINE, params, precision=E.C64)
    if state == "random":
        rng = np.random.default_rng(2)
        psi = np.exp(1j * 0.5 * rng.standard_normal(shape)) * (1 + 0.01 * rng.random(shape))
    else:
        psi = bench.initial_state(shape)
    for x in (g, o):
        x.set_psi(psi)
        x.normalize_psi()
    g.evolve_qf(steps)
    o.evolve_qf(steps)
    ng, no = g.get_norm(), o.get_norm()
    eg, eo = g.energy(), o.energy()
    Eg, Eo = sum(eg[1:]), sum(eo[1:])
    rel = abs(Eg - Eo) / abs(Eo)
    say("c64 %s: rel energy error after %d steps %.3e, norm drift %.2e" % (state, steps, rel, ng / no - 1))
    assert rel < tol
    assert abs(ng / no - 1) < 5e-6
    dg, _ = g.get_density()
    do, _ = o.get_density()
    assert np.max(np.abs(dg - do)) / np.max(do) < 1e-3
    g.close()


@pytest.mark.parametrize("variant", ["vortex2", "vortex3", "phase"])
@pytest.mark.parametrize("mode", ["rte", "ite", "ite_qf"])
def test_pinned_variants_match_oracle(oracle, variant, mode):
    say("pinned", variant, mode)
    alpha, beta = (1.0, 0.0) if mode == "rte" else (0.0, 1.0)
    shape = (32, 24, 16)
    g, o = pair(oracle, shape, alpha, beta, 0.025, 700.0, 0, 0, [0.05, 0.06, 0.1])
    psi = seeded_state(shape, 19)
    psi[3, 5, :] = 0.0  # exact zeros exercise the guarded imprint
    rng = np.random.default_rng(4)
    target = rng.uniform(-np.pi, np.pi, shape)
    for x in (g, o):
        if mode == "ite_qf":
            x.set_quantum_friction_coeff(0.2)
        x.set_psi(psi)
        x.normalize_psi()
        x.set_vortex(0.5, -2.5, 1)
        if variant == "phase":
            x.set_phase(target)
    fn = {"vortex2": "evolve_vortex2", "vortex3": "evolve_vortex3", "phase": "evolve_enforced_phase"}[variant]
    mg = getattr(g, fn)(6, want_mu=True)
    mo = getattr(o, fn)(6, want_mu=True)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128
    if mode != "rte":
        assert abs(mg - mo) < TOL128 * abs(mo)
    # gpe_get_phase: arg psi, kept as the enforced phase
    assert np.max(np.abs(g.get_phase() - o.get_phase())) < 1e-10
    g.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 8, 11, 64, 66])
@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 8, 64), (16, 96, 8), (8, 12, 48), (16, 8, 24), (64, 128, 12),
                                   (256, 8, 8), (8, 512, 8)])
def test_launcher_variants_match_oracle(oracle, variant, shape):
    """Every launcher choice of the autotuner (z passes in wave registers or LDS
    tiles, full or half-width strided tiles, the alternative radix schedules
    of bit 6: 32 = 4.8, 48 = 3.16, 64 = 4.16, 96 = 3.2.16, 128 = 8.16,
    256 = 4.4.16, 512 = 16.32 / 4.8.16) follows the oracle's step, with the
    kinetic factor split over the y and x turn-arounds (the default) and
    unsplit in the x turn-around (bit 3)."""
    say("variant", variant, shape)
    for mode in ("rte", "ite"):
        alpha, beta = (1.0, 0.0) if mode == "rte" else (0.0, 1.0)
        g, o = pair(oracle, shape, alpha, beta, 0.02, 600.0, E.CONTACT, E.COSINE, [0.1, 2.0, 0.0, 0.02])
        g.set_variant(variant)
        psi = seeded_state(shape, 23)
        for x in (g, o):
            x.set_psi(psi)
            x.normalize_psi()
        mg = g.evolve_qf(4, want_mu=True)
        mo = o.evolve_qf(4, want_mu=True)
        assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128, mode
        if mode == "ite":
            assert abs(mg - mo) < TOL128 * abs(mo)
        g.close()


@pytest.mark.parametrize("prec,variant,shape", [
    (E.C64, 2, (256, 64, 8)),    # x rows of 8 complex64 (64 B): pairs, 4 full groups
    (E.C64, 2, (256, 24, 8)),    # 24 tiles: one full group and 8 in plain order
    (E.C128, 0, (512, 16, 8)),   # x rows of 4 complex128 (64 B): pairs
    (E.C128, 2, (512, 16, 8)),   # x rows of 2 complex128 (32 B): groups of 4
    (E.C128, 0, (512, 12, 8)),   # 24 tiles: one full group, 8 in plain order
    (E.C128, 2, (16, 512, 12)),  # y axis, 96 tiles: 3 full groups of 32
])
def test_short_row_tile_order_matches_oracle(oracle, prec, variant, shape):
    """Strided tiles with rows shorter than 128 bytes are handed to workgroups
    in line-sharing groups per XCD (tile_of_block): every tile is still
    transformed exactly once, so the step follows the oracle (complex128 to
    the parity bound, complex64 to its rounding), with full and trailing
    partial groups."""
    say("short-row tiles", prec, variant, shape)
    g = E.Engine(shape, 1.0, 0.0, 0.02, 600.0, E.CONTACT, E.COSINE, 0, prec)
    o = oracle.OracleEngine(shape, 1.0, 0.0, 0.02, 600.0, E.CONTACT, E.COSINE, 0)
    for x in (g, o):
        x.set_user_params([0.1, 2.0, 0.0, 0.02])
    g.set_variant(variant)
    psi = seeded_state(shape, 29)
    for x in (g, o):
        x.set_psi(psi)
        x.normalize_psi()
        x.evolve_qf(4)
    err = rel_err(g.get_psi()[0], o.get_psi()[0])
    assert err < (TOL128 if prec == E.C128 else 2e-5), err
    g.fft(-1)
    o_k = np.fft.fftn(o.get_psi()[0])
    assert rel_err(g.get_psi()[0], o_k) < (1e-12 if prec == E.C128 else 2e-5)
    g.close()
    o.close()


@pytest.mark.parametrize("mode", ["rte", "ite"])
@pytest.mark.parametrize("shape", [(32, 24, 16), (24, 16, 12)])
def test_dipolar_matches_oracle(oracle, mode, shape):
    """gpe_evolve_dipolar and both dipolar energies (dev/gpe_engine.cu:1490-1810)."""
    say("dipolar", mode, shape)
    alpha, beta = (1.0, 0.0) if mode == "rte" else (0.0, 1.0)
    params = [0.05, 0.06, 0.08, 0.03, 0.02]  # trap, a_con, a_dip
    g, o = pair(oracle, shape, alpha, beta, 0.02, 400.0, E.DIPOLAR, E.HARMONIC, params)
    psi = seeded_state(shape, 31)
    for x in (g, o):
        x.set_psi(psi)
        x.normalize_psi()
    for nsteps in (1, 6):
        mg = g.evolve_dipolar(nsteps, want_mu=True)
        mo = o.evolve_dipolar(nsteps, want_mu=True)
        assert rel_err(g.get_psi()[0], o.get_psi()[0]) < TOL128
        if mode == "ite":
            assert abs(mg - mo) < TOL128 * abs(mo)
    for method in (1, 2):
        tg, eg = g.energy_dipolar(method)
        to, eo = o.energy_dipolar(method)
        assert tg == to
        assert abs(eg - eo) < TOL128 * abs(eo)
    g.close()


def test_dipolar_complex64_tracks_oracle(oracle):
    say("dipolar c64")
    shape = (32, 32, 32)
    params = [0.05, 0.05, 0.05, 0.03, 0.02]
    g, o = pair(oracle, shape, 1.0, 0.0, 0.02, 1000.0, E.DIPOLAR, E.HARMONIC, params, precision=E.C64)
    psi = seeded_state(shape, 32)
    for x in (g, o):
        x.set_psi(psi)
        x.normalize_psi()
    g.evolve_dipolar(50)
    o.evolve_dipolar(50)
    assert rel_err(g.get_psi()[0], o.get_psi()[0]) < 2e-5
    _, eg = g.energy_dipolar(2)
    _, eo = o.energy_dipolar(2)
    assert abs(eg - eo) < 1e-5 * abs(eo)
    g.close()


@pytest.mark.parametrize("mode", ["rte", "ite", "qf", "vortex", "c64"])
def test_graph_replays_match_direct_launches(mode):
    """Evolve calls replayed from captured HIP graphs equal direct launches bit
    for bit, over several calls with host work between them (energies, which
    renormalise in real time), a parameter change (which must recapture) and a
    switch of step plan."""
    say("graphs", mode)
    shape = (32, 24, 16)
    prec = E.C64 if mode == "c64" else E.C128
    alpha, beta = (0.0, 1.0) if mode == "ite" else (1.0, 0.0)
    engs = []
    for graph in (True, False):
        g = E.Engine(shape, alpha, beta, 0.02, 500.0, E.CONTACT, E.COSINE, 0, prec)
        g.set_graph(graph)
        g.set_user_params([0.1, 2.0, 0.0, 0.02])
        if mode == "qf":
            g.set_quantum_friction_coeff(0.2)
        g.set_psi(seeded_state(shape, 41))
        g.normalize_psi()
        if mode == "vortex":
            g.set_vortex(0.5, 1.5, 1)
        engs.append(g)
    step = (lambda g, n: g.evolve_vortex(n, want_mu=True)) if mode == "vortex" else (
        lambda g, n: g.evolve_qf(n, want_mu=True))
    for n in (5, 4, 3):
        mus = [step(g, n) for g in engs]
        ps = [g.get_psi()[0] for g in engs]
        assert np.array_equal(ps[0], ps[1]), n
        if mode == "ite":
            assert mus[0] == mus[1]
        es = [g.energy() for g in engs]
        assert es[0] == es[1]
    for g in engs:
        g.set_user_params([0.15, 2.0, 0.0, 0.03])
    for g in engs:
        step(g, 6)
    assert np.array_equal(engs[0].get_psi()[0], engs[1].get_psi()[0])
    for g in engs:
        g.evolve_qf(3)  # another step plan for the vortex mode
    assert np.array_equal(engs[0].get_psi()[0], engs[1].get_psi()[0])
    for g in engs:
        g.close()
