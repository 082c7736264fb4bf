"""Headline benchmark: real-time split steps of a 256^3 Bose gas in complex64.

Workload (one "step" = one predictor-corrector split step of the reference's
gpe_evolve_qf, orginal/gpe_engine.cu:948-1013): 256^3 lattice, contact bosons
(gamma = 1, a = 0.008, one particle per site), cosine external potential of
amplitude 0.1 and period L/8 on each axis, dt = 0.02, a seeded random-phase
state with 20 z-line vortices of winding +-1.  The state is resident in HBM
before the timed region starts.

Prints one JSON line with the throughput, the HIP-event roofline of the
dominant kernel (the slower of the y and x turn-around passes, two launches
of each per step), the energy error of complex64 against the complex128
engine on the same steps, and the FP64 CPU oracle timed on this host as the
CPU baseline.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 50]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from libcugpe_amd import workloads as W  # noqa: E402

METRIC = "split-steps/s at 256^3 c64 vs c128 (rel-E err≤1e-6), %HBM peak; ensemble steps/s"
HBM_PEAK_GBS = 8000.0
A_SCAT = W.HEADLINE_A_SCAT
DT = W.HEADLINE_DT
initial_state = W.headline_state


VARIANT = None  # launcher variant (--variant); None keeps the engine's default choice


def make_engine(E, shape, precision, psi):
    n = int(np.prod(shape))
    eng = E.Engine(shape, 1.0, 0.0, DT, float(n), E.CONTACT, E.COSINE, 0, precision)
    if VARIANT is not None:
        eng.set_variant(VARIANT)
    eng.set_user_params(W.HEADLINE_PARAMS)
    eng.set_psi(psi)
    eng.normalize_psi()
    return eng


def timed(eng, steps):
    eng.synchronize()
    t0 = time.perf_counter()
    eng.evolve_qf(steps)
    eng.synchronize()
    return time.perf_counter() - t0


def pmc_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what it
    reads for a plain copy kernel of known size)."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        return json.load(open(p))[kernel]["bytes_per_launch"]
    except (OSError, ValueError, KeyError, TypeError):
        return None


def rounding_energy_error(E, shape, psi, segments=50, seg_steps=20):
    """Energy error that complex64 rounding adds over segments*seg_steps steps,
    free of trajectory divergence: at the start of every segment the complex128
    engine takes the complex64 engine's state, both run seg_steps, and the
    difference of their energy changes is summed.  Segments are short because
    the two trajectories of the turbulent state drift apart within a segment,
    and the split-step energy drift of the drifted trajectory differs (segments
    of 100 steps show a per-segment spread of ~1e-7 from that alone)."""
    e64 = make_engine(E, shape, E.C64, psi)
    e128 = make_engine(E, shape, E.C128, psi)
    total, worst, e_ref = 0.0, 0.0, None
    for _ in range(segments):
        cur, _ = e64.get_psi()
        e128.set_psi(cur)
        _, a1, a2, a3 = e64.energy()
        _, b1, b2, b3 = e128.energy()
        e64.evolve_qf(seg_steps)
        e128.evolve_qf(seg_steps)
        _, c1, c2, c3 = e64.energy()
        _, d1, d2, d3 = e128.energy()
        if e_ref is None:
            e_ref = b1 + b2 + b3
        d = ((c1 + c2 + c3) - (a1 + a2 + a3)) - ((d1 + d2 + d3) - (b1 + b2 + b3))
        total += d
        worst = max(worst, abs(d))
    e64.close()
    e128.close()
    return {"steps": segments * seg_steps, "segments": segments, "segment_steps": seg_steps,
            "rel_err": abs(total) / abs(e_ref), "signed_rel_err": total / abs(e_ref),
            "worst_segment_rel_err": worst / abs(e_ref)}


def host_cpu():
    """(nproc, lscpu model name) of this host."""
    model = None
    try:
        import subprocess

        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return os.cpu_count(), model


def cpu_baseline(E, shape, psi, steps, runs=3):
    """The FP64 oracle (own FFT, OpenMP over every thread it is given) on the
    same 256^3 workload: one warm-up step, then the median of `runs` timed runs
    of `steps` steps.  The complex64 engine then runs the same number of steps
    from the same state, and its energy is compared with the oracle's."""
    from oracle import pyoracle

    n = int(np.prod(shape))
    o = pyoracle.OracleEngine(shape, 1.0, 0.0, DT, float(n), -1, 2, 0)
    o.set_user_params(W.HEADLINE_PARAMS)
    o.set_psi(psi)
    o.normalize_psi()
    o.evolve_qf(1)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        o.evolve_qf(steps)
        times.append(time.perf_counter() - t0)
        print("[bench] oracle run: %d steps in %.2f s" % (steps, times[-1]), file=sys.stderr, flush=True)
    total = 1 + runs * steps
    _, a1, a2, a3 = o.energy()
    threads = pyoracle.lib().gpeo_num_threads()
    o.close()
    g = make_engine(E, shape, E.C64, psi)
    g.evolve_qf(total)
    _, b1, b2, b3 = g.energy()
    g.close()
    nproc, model = host_cpu()
    med = float(np.median(times))
    return {
        "value": steps / med,
        "unit": "split-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "median of %d runs of %d FP64 oracle steps of the same 256^3 workload after one warm-up step "
        "(complex128, own FFT, %d OpenMP threads)" % (runs, steps, threads),
        "run_seconds": times,
        "nproc": nproc,
        "omp_threads": threads,
        "threads_equal_nproc": threads == nproc,
        "cpu_model": model,
        "fft_backend": "own mixed-radix FFT (FFTW absent)",
        "energy_rel_err_c64_vs_fp64": abs((b1 + b2 + b3) - (a1 + a2 + a3)) / abs(a1 + a2 + a3),
        "energy_compare_steps": total,
    }


def ensemble_bench(E, steps, seq_steps, ref_members, ref_steps):
    from time import perf_counter

    shapes, seeds = W.ensemble_shapes()
    gs = {L: W.ensemble_ground_state(E, L) for L, _ in W.ENS_MIX}
    ens = E.Ensemble(shapes, 1.0, 0.0, W.ENS_DT, 1.0, E.CONTACT, E.COSINE, 0, E.C64)
    ens.set_user_params(W.ENS_PARAMS)
    for i, (sh, sd) in enumerate(zip(shapes, seeds)):
        ens.set_psi(i, W.wigner_sample(gs[sh[0]], sd))
    ens.autotune(3)
    ens.evolve(5)  # warm-up (builds the step graphs)
    t0 = perf_counter()
    ens.evolve(steps)
    t_ens = perf_counter() - t0
    points = sum(int(np.prod(s)) for s in shapes)

    # sequential per-lattice launches: one engine per shape, each lattice run on its own
    seq_step_s = 0.0
    for L, cnt in W.ENS_MIX:
        g = E.Engine((L, L, L), 1.0, 0.0, W.ENS_DT, float(L ** 3), E.CONTACT, E.COSINE, 0, E.C64)
        g.set_user_params(W.ENS_PARAMS)
        g.set_psi(W.wigner_sample(gs[L], 1))
        g.evolve_qf(5)
        g.synchronize()
        t0 = perf_counter()
        g.evolve_qf(seq_steps)
        g.synchronize()
        seq_step_s += cnt * (perf_counter() - t0) / seq_steps
        g.close()

    # mean density of the first lattices of the smallest shape against the FP64 oracle
    from oracle import pyoracle

    L0 = W.ENS_MIX[0][0]
    members = list(range(ref_members))
    ens_ref = E.Ensemble([(L0,) * 3] * ref_members, 1.0, 0.0, W.ENS_DT, 1.0, E.CONTACT, E.COSINE, 0, E.C64)
    ens_ref.set_user_params(W.ENS_PARAMS)
    rho_o = np.zeros((L0,) * 3)
    for i in members:
        psi = W.wigner_sample(gs[L0], seeds[i])
        ens_ref.set_psi(i, psi)
        o = pyoracle.OracleEngine((L0,) * 3, 1.0, 0.0, W.ENS_DT, float(L0 ** 3), -1, 2, 0)
        o.set_user_params(W.ENS_PARAMS)
        o.set_psi(psi)
        o.evolve_qf(ref_steps)
        rho_o += np.abs(o.get_psi()[0]) ** 2 / ref_members
        o.close()
    ens_ref.evolve(ref_steps)
    rho_g = ens_ref.mean_density(0)
    ens_ref.close()
    groups = ens.groups()
    ens.close()
    sps = steps / t_ens
    S22 = 22 * 8 * points
    return {
        "workload": "512 truncated-Wigner lattices 24^3:32^3:48^3:64^3 = 256:128:64:64 (%d points), complex64 "
        "real time, contact a=0.008 + cosine V_ext, ground state + variance-1/2 plane-wave noise" % points,
        "ensemble_steps_per_s": sps,
        "lattice_steps_per_s": sps * len(shapes),
        "ms_per_ensemble_step": 1e3 / sps,
        "sequential_ensemble_steps_per_s": 1.0 / seq_step_s,
        "speedup_vs_sequential": sps * seq_step_s,
        "sequential_method": "one engine per shape, %d steps each, per-step time x lattices of that shape" % seq_steps,
        "hbm_frac": S22 * sps / (HBM_PEAK_GBS * 1e9),
        "graph": True,
        "groups": [{"shape": list(g[0]), "lattices": g[1], "variant": g[2]} for g in groups],
        "mean_density_rel_err_vs_fp64": float(np.linalg.norm(rho_g - rho_o) / np.linalg.norm(rho_o)),
        "mean_density_max_rel_err_vs_fp64": float(np.max(np.abs(rho_g - rho_o)) / np.max(rho_o)),
        "mean_density_sample": "%d lattices of %d^3, %d steps, FP64 oracle" % (ref_members, L0, ref_steps),
    }


def adjoint_bench(E, T, oracle_steps, fd_bumps=(0, 101, 307)):
    """64^3 complex64 optimal-control gradient: L = 1 - |<psi_t|psi(T)>|^2 over
    T real-time steps (dt = 1e-3, unitary Fermi gas), dL/dA for 512 seeded
    Gaussian bumps of V_ext (amplitude ~ N(0, 0.05), width 1.5) through the
    torch.autograd wrapper of the engine's adjoint."""
    import torch
    from time import perf_counter

    from libcugpe_amd import adjoint as A

    shape = W.ADJ_SHAPE
    n = int(np.prod(shape))
    dt_a = W.ADJ_DT
    inp = W.adjoint_inputs()
    centres, amp0, psi0 = inp["centres"], inp["amp0"], inp["psi0"]
    # target: psi0 evolved under another seeded bump potential
    vt = A.gaussian_bumps(torch.tensor(inp["amp_target"]), inp["centres_target"], W.ADJ_WIDTH, shape).numpy()
    g = E.Engine(shape, 1.0, 0.0, dt_a, float(n), E.UFG, E.VARRAY, 0, E.C128)
    g.set_external_potential(vt)
    g.set_psi(psi0)
    g.evolve_qf(T)
    psit, _ = g.get_psi()

    def loss128(a):
        V = A.gaussian_bumps(torch.tensor(a), centres, 1.5, shape).numpy()
        g.set_external_potential(V)
        g.set_psi(psi0)
        g.evolve_qf(T)
        p, _ = g.get_psi()
        ov = np.vdot(psit, p)
        return 1.0 - abs(ov) ** 2 / (np.vdot(psit, psit).real * np.vdot(psi0, psi0).real)

    eng = E.Engine(shape, 1.0, 0.0, dt_a, float(n), E.UFG, E.VARRAY, 0, E.C64)
    amp = torch.tensor(amp0, requires_grad=True)
    A.overlap_loss(A.gaussian_bumps(amp, centres, 1.5, shape), eng, psi0, psit, 20).backward()  # warm-up
    amp = torch.tensor(amp0, requires_grad=True)
    t0 = perf_counter()
    L = A.overlap_loss(A.gaussian_bumps(amp, centres, 1.5, shape), eng, psi0, psit, T)
    t1 = perf_counter()
    L.backward()
    t2 = perf_counter()
    grad = amp.grad.numpy().copy()
    # the same gradient from a complex128 engine through the same wrapper
    eng128 = E.Engine(shape, 1.0, 0.0, dt_a, float(n), E.UFG, E.VARRAY, 0, E.C128)
    amp = torch.tensor(amp0, requires_grad=True)
    t4 = perf_counter()
    A.overlap_loss(A.gaussian_bumps(amp, centres, 1.5, shape), eng128, psi0, psit, T).backward()
    t5 = perf_counter()
    grad128 = amp.grad.numpy().copy()
    eng128.close()
    scale = float(np.max(np.abs(grad)))
    h = 1e-4
    fd = []
    for j in fd_bumps:
        ap, am = amp0.copy(), amp0.copy()
        ap[j] += h
        am[j] -= h
        d = (loss128(ap) - loss128(am)) / (2 * h)
        fd.append({"bump": j, "fd_c128": d, "adjoint_c64": float(grad[j]), "adjoint_c128": float(grad128[j]),
                   "abs_err_over_max_grad": abs(d - grad[j]) / scale,
                   "abs_err_c128_over_max_grad": abs(d - grad128[j]) / scale})
    g.close()
    # the oracle's FP64 adjoint on the first oracle_steps steps against the engine's
    from oracle import pyoracle

    V = A.gaussian_bumps(torch.tensor(amp0), centres, 1.5, shape).numpy()
    o = pyoracle.OracleEngine(shape, 1.0, 0.0, dt_a, float(n), 0, 3, 0)
    o.set_vext_array(V)
    o.set_psi(psi0)
    t3 = perf_counter()
    Lo, go = o.adjoint_overlap(oracle_steps, psit)
    t_or = perf_counter() - t3
    o.close()
    eng.set_external_potential(V)
    eng.set_psi(psi0)
    Lg = eng.adjoint_forward(oracle_steps, psit)
    gg = eng.adjoint_backward()
    eng.close()
    return {
        "workload": "64^3 complex64 unitary Fermi gas, T=%d real-time steps dt=1e-3, L = 1-|<psi_t|psi(T)>|^2, "
        "dL/dA for 512 Gaussian bumps of V_ext (A ~ N(0,0.05), width 1.5), sqrt(T) checkpoints" % T,
        "loss": float(L.item()),
        "gradient_wall_s": t2 - t0,
        "forward_s": t1 - t0,
        "backward_s": t2 - t1,
        "fd_check": fd,
        "max_grad": scale,
        "max_abs_err_vs_fd_over_max_grad": max(x["abs_err_over_max_grad"] for x in fd),
        "max_abs_err_vs_fd": max(x["abs_err_over_max_grad"] for x in fd) * scale,
        "c128": {"gradient_wall_s": t5 - t4,
                 "max_abs_err_vs_fd_over_max_grad": max(x["abs_err_c128_over_max_grad"] for x in fd),
                 "max_abs_err_vs_fd": max(x["abs_err_c128_over_max_grad"] for x in fd) * scale},
        "oracle_sample": {
            "steps": oracle_steps,
            "loss_abs_err": abs(Lg - Lo),
            "dLdV_rel_l2_err": float(np.linalg.norm(gg - go) / np.linalg.norm(go)),
            "oracle_s": t_or,
        },
    }


def ite_ground_state_bench(E):
    """Imaginary-time ground state of a trapped contact BEC, 32^3 complex128:
    box L = 20 (oscillator units, dx = 0.625), V = r^2/2, N = 1e4, g = 1.25
    (read as 1.0 N / L^3), Gaussian start of width 2 times (1 + 0.05 U) with U
    seeded uniform, dt = 1e-3, 2000 steps normalised every step.  Lattice
    units: dt_l = dt/dx^2, omega_l = omega dx^2, g_l = g/dx (a_l = g_l/4pi)."""
    from time import perf_counter

    from oracle import pyoracle

    cfg = W.ground_state_32()
    shape, N, dx, steps, up, psi0 = cfg["shape"], cfg["npart"], cfg["dx"], cfg["steps"], cfg["params"], cfg["psi0"]
    dt = cfg["dt_lattice"] * dx ** 2
    g128 = E.Engine(shape, 0.0, 1.0, dt / dx ** 2, N, E.CONTACT, E.HARMONIC, 0, E.C128)
    o = pyoracle.OracleEngine(shape, 0.0, 1.0, dt / dx ** 2, N, -1, 0, 0)
    for x in (g128, o):
        x.set_user_params(up)
        x.set_psi(psi0)
        x.normalize_psi()
    g128.synchronize()
    t0 = perf_counter()
    mu_g = g128.evolve_qf(steps, want_mu=True)
    t_g = perf_counter() - t0
    t0 = perf_counter()
    mu_o = o.evolve_qf(steps, want_mu=True)
    t_o = perf_counter() - t0
    eg, eo = g128.energy(), o.energy()
    Eg, Eo = sum(eg[1:]), sum(eo[1:])
    rg, _ = g128.get_density()
    ro, _ = o.get_density()
    g128.close()
    o.close()
    return {
        "workload": "32^3 complex128 imaginary time, contact BEC N=1e4 g=1.25, V=r^2/2, L=20, dt=1e-3, 2000 steps",
        "steps_per_s": steps / t_g,
        "oracle_steps_per_s": steps / t_o,
        "energy_per_particle_osc_units": Eg / N / dx ** 2,
        "mu_osc_units": mu_g / dx ** 2,
        "energy_rel_err_vs_fp64": abs(Eg - Eo) / abs(Eo),
        "mu_rel_err_vs_fp64": abs(mu_g - mu_o) / abs(mu_o),
        "density_max_rel_err_vs_fp64": float(np.max(np.abs(rg - ro)) / np.max(ro)),
        "density_l2_rel_err_vs_fp64": float(np.linalg.norm(rg - ro) / np.linalg.norm(ro)),
    }


def ufg_vortex_qf_bench(E, steps, oracle_steps):
    """Real-time unitary Fermi gas (dimers, xi = 0.37) with quantum friction
    0.1 in an anisotropic trap omega = (1, 1, 0.5), 128^3, box L = 32 (dx =
    0.25), dt = 5e-4, from a seeded Thomas-Fermi density with one vortex line
    along z imprinted as exp(i theta).  complex128 and complex64 on the full
    run; the FP64 oracle on the first oracle_steps steps for the windings
    (bit-exact) and the energies."""
    from time import perf_counter

    from oracle import pyoracle

    cfg = W.ufg_vortex_128()
    shape, N, dx, up, psi0 = cfg["shape"], cfg["npart"], cfg["dx"], cfg["params"], cfg["psi0"]
    dt = cfg["dt_lattice"] * dx ** 2
    out = {"workload": "128^3 real-time UFG (dimers), QF 0.1, trap (1,1,0.5), L=32, dt=5e-4, %d steps, vortex line" % steps}
    res = {}
    for prec, name in ((E.C128, "c128"), (E.C64, "c64")):
        g = E.Engine(shape, 1.0, 0.0, dt / dx ** 2, N, E.UFG, E.HARMONIC, 0, prec)
        g.set_user_params(up)
        g.set_quantum_friction_coeff(0.1)
        g.set_psi(psi0)
        g.normalize_psi()
        _, a, b, cc = g.energy()
        g.synchronize()
        t0 = perf_counter()
        g.evolve_qf(steps)
        g.synchronize()
        tt = perf_counter() - t0
        _, a2, b2, c2 = g.energy()
        pf, _ = g.get_psi()
        _, cnt = g.detect_vortices(1e-2 * float(np.max(np.abs(pf) ** 2)))
        res[name] = (a + b + cc, a2 + b2 + c2)
        out[name + "_steps_per_s"] = steps / tt
        out[name + "_energy_change"] = (a2 + b2 + c2) / (a + b + cc) - 1
        out[name + "_vortex_plaquettes"] = int(cnt)
        g.close()
    out["energy_rel_err_c64_vs_c128_end"] = abs(res["c64"][1] - res["c128"][1]) / abs(res["c128"][1])
    # the oracle on a bounded sample: windings bit-exact, energy
    g = E.Engine(shape, 1.0, 0.0, dt / dx ** 2, N, E.UFG, E.HARMONIC, 0, E.C128)
    o = pyoracle.OracleEngine(shape, 1.0, 0.0, dt / dx ** 2, N, 0, 0, 0)
    for x in (g, o):
        x.set_user_params(up)
        x.set_quantum_friction_coeff(0.1)
        x.set_psi(psi0)
        x.normalize_psi()
    g.evolve_qf(oracle_steps)
    o.evolve_qf(oracle_steps)
    po, _ = o.get_psi()
    rmin = 1e-2 * float(np.max(np.abs(po) ** 2))
    wg, cg = g.detect_vortices(rmin)
    wo, co = pyoracle.plaquette_windings(po, rmin)
    eg, eo = g.energy(), o.energy()
    g.close()
    o.close()
    out["oracle_sample"] = {
        "steps": oracle_steps,
        "windings_bit_exact": bool(np.array_equal(wg, wo)),
        "vortex_plaquettes": int(cg),
        "energy_rel_err_c128_vs_fp64": abs(sum(eg[1:]) - sum(eo[1:])) / abs(sum(eo[1:])),
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=25)  # x3 runs: about 20 s of the FP64 oracle on 16 threads
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c128", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-extra-workloads", action="store_true")
    ap.add_argument("--ufg-steps", type=int, default=10000)
    ap.add_argument("--ufg-oracle-steps", type=int, default=40)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-oracle-steps", type=int, default=200)
    ap.add_argument("--variant", type=int, default=None, help="launcher variant of the 256^3 engines (bits: 1 z in "
                    "LDS tiles, 2 half-width strided tiles, 4 pipelined strided passes)")
    a = ap.parse_args()
    global VARIANT
    VARIANT = a.variant
    if a.gpus != 1:
        print("bench.py: this engine runs on one GPU; a %d-GPU run is not supported "
              "(one lattice is not sharded across GPUs)" % a.gpus, file=sys.stderr)
        return 2

    from libcugpe_amd import engine as E

    E.lib()
    shape = (a.size,) * 3
    n = int(np.prod(shape))
    print("[bench] building the %s initial state" % (shape,), file=sys.stderr, flush=True)
    psi = initial_state(shape)

    print("[bench] complex64 engine: warmup %d, timed %d" % (a.warmup, a.steps), file=sys.stderr, flush=True)
    eng = make_engine(E, shape, E.C64, psi)
    _, ek, ei, ee = eng.energy()
    e64_0 = ek + ei + ee
    eng.evolve_qf(a.warmup)
    t = timed(eng, a.steps)
    sps = a.steps / t

    # the same number of steps again with HIP events around every launch of
    # each turn-around kernel, y (forward y, K_y K_z, inverse y) and x (forward
    # x, K_x, inverse x): two launches of each per step, together the largest
    # share of step time; the dominant one of the two is the roofline kernel
    turn = {}
    t_prof = None
    for cls, name in ((3, "x"), (2, "y")):
        eng.profile(cls, True)
        tp = timed(eng, a.steps)
        t_prof = tp if t_prof is None else min(t_prof, tp)
        ms, cnt = eng.profile_read()
        eng.profile(cls, False)
        turn[name] = (ms / max(cnt, 1), cnt)
    # the other pass classes on 1/10 of the steps each
    sub = max(a.steps // 10, 1)
    per_class = {}
    for cls, name in ((0, "strided_y_and_x"), (1, "z")):
        eng.profile(cls, True)
        eng.evolve_qf(sub)
        ms, cnt = eng.profile_read()
        eng.profile(cls, False)
        per_class[name] = ms / max(cnt, 1)
    per_class["x"] = turn["x"][0]
    per_class["y"] = turn["y"][0]
    copy_ms = min(eng.copy_benchmark() for _ in range(5))
    total_steps = a.warmup + 3 * a.steps + 2 * sub
    _, ek, ei, ee = eng.energy()
    e64 = ek + ei + ee
    eng.close()

    S = 8 * n  # bytes of one complex64 state array
    # the committed counter summary covers 256^3 with the default (split) kinetic plan
    default_plan = a.size == 256 and (VARIANT is None or not (VARIANT & 8))
    bytes_launch = 2 * S  # one read and one write of the state per axis pass

    def turn_roofline(name):
        avg_ms, cnt = turn[name]
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        what = {"x": "x turn-around (forward x, K_x, inverse x)", "y": "y turn-around (forward y, K_y K_z, inverse y)"}
        return {
            "kernel": "fft_pass_kernel<float, %d> %s" % (a.size, what[name]),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": pmc_traffic(name + "_pass") if default_plan else None,
            "avg_launch_ms": avg_ms,
            "launches_timed": cnt,
            "algorithmic_bytes_per_launch": bytes_launch,
        }

    dominant = max(("x", "y"), key=lambda k: turn[k][0])
    roofline = turn_roofline(dominant)
    other_turn = turn_roofline("y" if dominant == "x" else "x")
    strided_ms = per_class["strided_y_and_x"]
    strided = {
        "kernel": "strided y and x passes (4 of 6 launches per step)" if default_plan else "strided y and x passes",
        "avg_launch_ms": strided_ms,
        "achieved": bytes_launch / (strided_ms * 1e-3) / 1e9,
        "frac": bytes_launch / (strided_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": pmc_traffic("strided") if default_plan else None,
    }

    result = {
        "metric": METRIC,
        "value": sps,
        "unit": "split-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "256^3 real-time contact BEC (a=%g, 1 particle/site), cosine V_ext A=0.1 period L/8, "
            "20 z-line vortices, dt=%g, complex64 state, FP64 reductions" % (A_SCAT, DT),
            "lattice": list(shape),
            "precision": "complex64",
        },
        "step_hbm_frac": 22 * S * sps / (HBM_PEAK_GBS * 1e9),
        "step_algorithmic_bytes": 22 * S,
        "pass_avg_ms": per_class,
        "strided_class": strided,
        "copy_kernel_GBs": 2 * S / (copy_ms * 1e-3) / 1e9,
        "profiled_region_steps_per_s": a.steps / t_prof,
        "roofline": roofline,
        "roofline_other_turnaround": other_turn,
    }

    if not a.no_c128:
        print("[bench] complex128 engine on the same steps", file=sys.stderr, flush=True)
        e128 = make_engine(E, shape, E.C128, psi)
        _, ek, ei, ee = e128.energy()
        e128_0 = ek + ei + ee
        e128.evolve_qf(a.warmup)
        t128 = timed(e128, a.steps)
        e128.evolve_qf(total_steps - a.warmup - a.steps)
        _, ek, ei, ee = e128.energy()
        eref = ek + ei + ee
        e128.close()
        result["c128_steps_per_s"] = a.steps / t128
        result["c128_step_hbm_frac"] = 22 * 2 * S * (a.steps / t128) / (HBM_PEAK_GBS * 1e9)
        result["energy_rel_err_c64_vs_c128"] = abs(e64 - eref) / abs(eref)
        result["energy_steps_compared"] = total_steps
        # each precision's own energy drift over the run: the split step's
        # truncation error follows the (chaotic) trajectory, rounding does not
        result["energy_drift_c64"] = e64 / e64_0 - 1
        result["energy_drift_c128"] = eref / e128_0 - 1
        result["energy_rel_err_c64_vs_c128_step0"] = abs(e64_0 - e128_0) / abs(e128_0)
        print("[bench] complex64 rounding error of the energy (segmented)", file=sys.stderr, flush=True)
        result["energy_rounding_err_c64"] = rounding_energy_error(E, shape, psi)

    if not a.no_ensemble:
        print("[bench] batched ensemble (%d steps)" % a.ensemble_steps, file=sys.stderr, flush=True)
        result["ensemble"] = ensemble_bench(E, a.ensemble_steps, 50, 8, 100)

    if not a.no_extra_workloads:
        print("[bench] 32^3 complex128 ground state", file=sys.stderr, flush=True)
        result["ite_ground_state_32"] = ite_ground_state_bench(E)
        print("[bench] 128^3 UFG with quantum friction (%d steps)" % a.ufg_steps, file=sys.stderr,
              flush=True)
        result["ufg_vortex_qf_128"] = ufg_vortex_qf_bench(E, a.ufg_steps, a.ufg_oracle_steps)

    if not a.no_adjoint:
        print("[bench] adjoint gradient (T=%d)" % a.adjoint_steps, file=sys.stderr, flush=True)
        result["adjoint"] = adjoint_bench(E, a.adjoint_steps, a.adjoint_oracle_steps)

    if not a.no_cpu_baseline:
        print("[bench] FP64 CPU oracle baseline (%d steps)" % a.cpu_steps, file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(E, shape, psi, a.cpu_steps)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
