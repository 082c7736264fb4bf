"""The benchmark configurations at their own workloads, against the FP64
oracle: the 32^3 imaginary-time ground state, the 128^3 unitary Fermi gas with
quantum friction and a vortex line, the 512-lattice truncated-Wigner ensemble
and the 64^3 adjoint gradient (the 256^3 headline runs in
test_gpu_fullsize.py).  Inputs come from libcugpe_amd/workloads.py, the same
builders bench.py uses."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402
from libcugpe_amd import workloads as W  # noqa: E402


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / abs(b)


def test_config1_ground_state_32_matches_oracle(oracle):
    """32^3 complex128 imaginary time, contact + harmonic, 2000 steps: E, mu and
    rho within 1e-12 of the FP64 oracle."""
    cfg = W.ground_state_32()
    say("config 1: 32^3 ITE ground state, %d steps" % cfg["steps"])
    g = E.Engine(cfg["shape"], 0.0, 1.0, cfg["dt_lattice"], cfg["npart"], E.CONTACT, E.HARMONIC, 0, E.C128)
    o = oracle.OracleEngine(cfg["shape"], 0.0, 1.0, cfg["dt_lattice"], cfg["npart"], E.CONTACT, E.HARMONIC, 0)
    mus = []
    for x in (g, o):
        x.set_user_params(cfg["params"])
        x.set_psi(cfg["psi0"])
        x.normalize_psi()
        mus.append(x.evolve_qf(cfg["steps"], want_mu=True))
    eg, eo = g.energy(), o.energy()
    rg, _ = g.get_density()
    ro, _ = o.get_density()
    g.close()
    o.close()
    say("E %.15e vs %.15e, mu %.15e vs %.15e" % (sum(eg[1:]), sum(eo[1:]), mus[0], mus[1]))
    assert eg[0] == eo[0]
    assert rel(sum(eg[1:]), sum(eo[1:])) < 1e-12
    for a, b in zip(eg[1:], eo[1:]):
        assert abs(a - b) < 1e-12 * abs(sum(eo[1:]))
    assert rel(mus[0], mus[1]) < 1e-12
    assert np.max(np.abs(rg - ro)) < 1e-12 * np.max(ro)


def _ufg_engine(cfg, prec):
    g = E.Engine(cfg["shape"], 1.0, 0.0, cfg["dt_lattice"], cfg["npart"], E.UFG, E.HARMONIC, 0, prec)
    g.set_user_params(cfg["params"])
    g.set_quantum_friction_coeff(cfg["qf"])
    g.set_psi(cfg["psi0"])
    g.normalize_psi()
    return g


def test_config2_ufg_friction_vortex_c128_matches_oracle(oracle):
    """128^3 complex128 real time with quantum friction 0.1 and a vortex line,
    200 steps: psi and energies within 1e-12 of the oracle, plaquette windings
    identical."""
    cfg = W.ufg_vortex_128()
    steps = 200
    say("config 2: 128^3 UFG + QF + vortex, c128 vs oracle, %d steps" % steps)
    g = _ufg_engine(cfg, E.C128)
    o = oracle.OracleEngine(cfg["shape"], 1.0, 0.0, cfg["dt_lattice"], cfg["npart"], E.UFG, E.HARMONIC, 0)
    o.set_user_params(cfg["params"])
    o.set_quantum_friction_coeff(cfg["qf"])
    o.set_psi(cfg["psi0"])
    o.normalize_psi()
    g.evolve_qf(steps)
    o.evolve_qf(steps)
    pg, _ = g.get_psi()
    po, _ = o.get_psi()
    assert np.max(np.abs(pg - po)) < 1e-12 * np.max(np.abs(po))
    rmin = 1e-2 * float(np.max(np.abs(po) ** 2))
    wg, cg = g.detect_vortices(rmin)
    wo, co = oracle.plaquette_windings(po, rmin)
    say("windings: %d plaquettes (engine) / %d (oracle)" % (cg, co))
    assert cg > 0 and cg == co
    assert np.array_equal(wg, wo)
    eg, eo = g.energy(), o.energy()
    assert rel(sum(eg[1:]), sum(eo[1:])) < 1e-12
    g.close()
    o.close()


def test_config2_windings_c64_follow_c128_over_the_full_run():
    """The same workload for its full 10^4 steps in both precisions: the
    complex64 windings must equal the complex128 ones at every checkpoint (every
    1000 steps); the energy difference is reported."""
    cfg = W.ufg_vortex_128()
    say("config 2: c64 vs c128 windings over %d steps" % cfg["steps"])
    g64, g128 = _ufg_engine(cfg, E.C64), _ufg_engine(cfg, E.C128)
    chunk = 1000
    first_diff = None
    for k in range(cfg["steps"] // chunk):
        g64.evolve_qf(chunk)
        g128.evolve_qf(chunk)
        p128, _ = g128.get_psi()
        rmin = 1e-2 * float(np.max(np.abs(p128) ** 2))
        w64, c64 = g64.detect_vortices(rmin)
        w128, c128 = g128.detect_vortices(rmin)
        if first_diff is None and not np.array_equal(w64, w128):
            where = np.argwhere(w64 != w128)[0]
            first_diff = ((k + 1) * chunk, tuple(int(v) for v in where))
            say("first differing winding after %d steps at plaquette (orientation, ix, iy, iz) %s" % first_diff)
        say("step %d: %d / %d winding plaquettes" % ((k + 1) * chunk, c64, c128))
    e64, e128 = g64.energy(), g128.energy()
    say("energy after %d steps: c64 %.12e c128 %.12e rel diff %.2e"
        % (cfg["steps"], sum(e64[1:]), sum(e128[1:]), rel(sum(e64[1:]), sum(e128[1:]))))
    g64.close()
    g128.close()
    assert first_diff is None, first_diff


def test_config4_full_ensemble_mix():
    """The 512-lattice mix (256 x 24^3, 128 x 32^3, 64 x 48^3, 64 x 64^3),
    complex64, 500 steps in one ensemble: every 48^3 and 64^3 lattice and a
    sample of the others bit for bit equal to a single engine of its shape
    with the same launcher choice, and the mean density of 48^3 and 64^3
    samples against the FP64 oracle."""
    from oracle import pyoracle

    steps = 500
    say("config 4: 512-lattice ensemble, %d steps" % steps)
    shapes, seeds = W.ensemble_shapes()
    gs = {L: W.ensemble_ground_state(E, L) for L, _ in W.ENS_MIX}
    ens = E.Ensemble(shapes, 1.0, 0.0, W.ENS_DT, 1.0, E.CONTACT, E.COSINE, 0, E.C64)
    ens.set_user_params(W.ENS_PARAMS)
    init = {}
    for i, (sh, sd) in enumerate(zip(shapes, seeds)):
        init[i] = W.wigner_sample(gs[sh[0]], sd)
        ens.set_psi(i, init[i])
    ens.evolve(steps)
    variant = {g[0][0]: g[2] for g in ens.groups()}
    check = [i for i, s in enumerate(shapes) if s[0] >= 48]
    check += [i for i, s in enumerate(shapes) if s[0] == 24][:8] + [i for i, s in enumerate(shapes) if s[0] == 32][:8]
    say("comparing %d lattices with single engines" % len(check))
    for i in check:
        L = shapes[i][0]
        g = E.Engine(shapes[i], 1.0, 0.0, W.ENS_DT, 1.0, E.CONTACT, E.COSINE, 0, E.C64)
        g.set_user_params(W.ENS_PARAMS)
        g.set_variant(variant[L])
        g.set_psi(init[i])
        g.evolve_qf(steps)
        got, t = ens.get_psi(i)
        want, tw = g.get_psi()
        g.close()
        assert t == tw
        assert np.array_equal(got, want), (i, shapes[i])
    # mean density of 4 lattices of each large shape against the FP64 oracle
    for L in (48, 64):
        members = [i for i, s in enumerate(shapes) if s[0] == L][:4]
        sub = E.Ensemble([(L,) * 3] * 4, 1.0, 0.0, W.ENS_DT, 1.0, E.CONTACT, E.COSINE, 0, E.C64)
        sub.set_user_params(W.ENS_PARAMS)
        rho_o = np.zeros((L,) * 3)
        for j, i in enumerate(members):
            sub.set_psi(j, init[i])
            o = pyoracle.OracleEngine((L,) * 3, 1.0, 0.0, W.ENS_DT, 1.0, E.CONTACT, E.COSINE, 0)
            o.set_user_params(W.ENS_PARAMS)
            o.set_psi(init[i])
            o.evolve_qf(steps)
            rho_o += np.abs(o.get_psi()[0]) ** 2 / 4
            o.close()
        sub.evolve(steps)
        rho_g = sub.mean_density(0)
        sub.close()
        err = np.linalg.norm(rho_g - rho_o) / np.linalg.norm(rho_o)
        say("%d^3 mean density, complex64 vs FP64 oracle after %d steps: rel L2 %.2e" % (L, steps, err))
        assert err < 2e-5
    ens.close()


def test_config5_adjoint_64_full_run(oracle):
    """64^3, T = 2000 real-time steps, 512 Gaussian bumps: the complex128
    adjoint gradient against central finite differences (h = 1e-4) of the
    complex128 engine on three bumps within 1e-6 of the largest component, and
    dL/dV against the FP64 oracle's adjoint over the full run.  The complex64
    gradient's deviation is reported."""
    import torch

    from libcugpe_amd import adjoint as A

    T = 2000
    say("config 5: 64^3 adjoint, T=%d" % T)
    shape = W.ADJ_SHAPE
    n = int(np.prod(shape))
    inp = W.adjoint_inputs()
    centres, amp0, psi0 = inp["centres"], inp["amp0"], inp["psi0"]
    vt = A.gaussian_bumps(torch.tensor(inp["amp_target"]), inp["centres_target"], W.ADJ_WIDTH, shape).numpy()
    g = E.Engine(shape, 1.0, 0.0, W.ADJ_DT, float(n), E.UFG, E.VARRAY, 0, E.C128)
    g.set_external_potential(vt)
    g.set_psi(psi0)
    g.evolve_qf(T)
    psit, _ = g.get_psi()

    def loss128(a):
        V = A.gaussian_bumps(torch.tensor(a), centres, W.ADJ_WIDTH, shape).numpy()
        g.set_external_potential(V)
        g.set_psi(psi0)
        g.evolve_qf(T)
        p, _ = g.get_psi()
        return 1.0 - abs(np.vdot(psit, p)) ** 2 / (np.vdot(psit, psit).real * np.vdot(psi0, psi0).real)

    grads = {}
    for prec in (E.C128, E.C64):
        eng = E.Engine(shape, 1.0, 0.0, W.ADJ_DT, float(n), E.UFG, E.VARRAY, 0, prec)
        amp = torch.tensor(amp0, requires_grad=True)
        A.overlap_loss(A.gaussian_bumps(amp, centres, W.ADJ_WIDTH, shape), eng, psi0, psit, T).backward()
        grads[prec] = amp.grad.numpy().copy()
        if prec == E.C128:
            gv = eng.adjoint_backward()  # dL/dV of the same forward run
        eng.close()
    scale = np.max(np.abs(grads[E.C128]))
    h = 1e-4
    for j in (0, 101, 307):
        ap, am = amp0.copy(), amp0.copy()
        ap[j] += h
        am[j] -= h
        fd = (loss128(ap) - loss128(am)) / (2 * h)
        say("bump %d: fd %.10e adjoint c128 %.10e c64 %.10e" % (j, fd, grads[E.C128][j], grads[E.C64][j]))
        assert abs(fd - grads[E.C128][j]) < 1e-6 * scale
    g.close()
    c64_err = np.max(np.abs(grads[E.C64] - grads[E.C128])) / scale
    say("complex64 gradient: max deviation from complex128 %.2e of the largest component" % c64_err)
    assert c64_err < 1e-2
    V = A.gaussian_bumps(torch.tensor(amp0), centres, W.ADJ_WIDTH, shape).numpy()
    o = oracle.OracleEngine(shape, 1.0, 0.0, W.ADJ_DT, float(n), E.UFG, E.VARRAY, 0)
    o.set_vext_array(V)
    o.set_psi(psi0)
    Lo, go = o.adjoint_overlap(T, psit)
    o.close()
    err = np.linalg.norm(gv - go) / np.linalg.norm(go)
    say("dL/dV c128 vs oracle adjoint over T=%d: rel L2 %.2e" % (T, err))
    assert err < 1e-9
