"""Synthetic inputs of the benchmark configurations, shared by bench.py and the
tests.  Every random number comes from a seeded generator, so the engine and
the FP64 oracle start from bit-identical host states.

Physical units map to the lattice's dx = 1 (the engine's convention) as
dt_l = dt/dx^2, omega_l = omega dx^2, g_l = g/dx (contact: a_l = g_l/(4 pi)),
and energies and chemical potentials divide by dx^2 on the way back.
"""
import numpy as np


def splitmix64_uniform(seed, n):
    """n uniforms in [0, 1) from SplitMix64 (the same stream on any host)."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


# ---------------------------------------------------------------- headline
# 256^3 real-time contact bosons, one particle per site, cosine V_ext of
# amplitude 0.1 and period L/8, dt = 0.02, 20 seeded z-line vortices.
HEADLINE_A_SCAT = 0.008
HEADLINE_DT = 0.02
HEADLINE_PARAMS = [0.1, 8.0, 0.0, HEADLINE_A_SCAT]  # V_AMP, V_PERIODS, -, A_SCAT


def headline_state(shape, seed=3, nvort=20):
    """Seeded random-phase state with nvort z-line vortices of winding +-1."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    ix, iy = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64), indexing="ij")
    ph = np.zeros((nx, ny))
    x0 = rng.uniform(0, nx, nvort) + 0.5
    y0 = rng.uniform(0, ny, nvort) + 0.5
    w = rng.choice([-1.0, 1.0], nvort)
    for j in range(nvort):
        ph += w[j] * np.arctan2(iy - y0[j], ix - x0[j])
    psi = np.empty(shape, dtype=np.complex128)
    noise = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    amp = (1.0 + 0.01 * rng.random((nx, ny, nz))).astype(np.float32)
    psi[...] = np.exp(1j * (ph[:, :, None] + 0.1 * noise)) * amp
    return psi


# ------------------------------------------------- 32^3 imaginary-time ground state
def ground_state_32():
    """Trapped contact BEC, 32^3 complex128 imaginary time: box L = 20
    (oscillator units, dx = 0.625), V = r^2/2, N = 1e4, g = 1.25 (the
    configuration's "g = 1.0 N/L^3" read as 1.25), Gaussian start of width
    2 (1 + 0.05 U) with U from SplitMix64 seed 1, dt = 1e-3, 2000 steps."""
    n1, L, N, g, dt = 32, 20.0, 1.0e4, 1.25, 1e-3
    dx = L / n1
    shape = (n1,) * 3
    c = (np.arange(n1) - n1 // 2) * dx
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    psi0 = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * 2.0 ** 2)) * (1 + 0.05 * splitmix64_uniform(1, n1 ** 3).reshape(shape))
    return {
        "shape": shape,
        "dx": dx,
        "npart": N,
        "dt_lattice": dt / dx ** 2,
        "steps": 2000,
        "params": [dx * dx] * 3 + [g / dx / (4 * np.pi)],  # OMEGA_X/Y/Z, A_SCAT in lattice units
        "psi0": psi0.astype(np.complex128),
    }


# ------------------------------------------------- 128^3 UFG, friction, vortex line
def ufg_vortex_128(n1=128):
    """Real-time unitary Fermi gas (dimers, xi = 0.37) with quantum friction 0.1
    in a trap omega = (1, 1, 0.5), box L = 32 (dx = 0.25 at 128^3), dt = 5e-4,
    from a Thomas-Fermi density with 1% SplitMix64 noise (seed 2) and one vortex
    line along z imprinted as exp(i theta), its core between lattice points."""
    L, N, dt = 32.0, 1.0e4, 5e-4
    dx = L / n1
    shape = (n1,) * 3
    om = (1.0, 1.0, 0.5)
    c = (np.arange(n1) - n1 // 2) * dx
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    V = 0.5 * (om[0] ** 2 * X ** 2 + om[1] ** 2 * Y ** 2 + om[2] ** 2 * Z ** 2)
    rho = np.maximum(1.0 - V / 25.0, 0.0) ** 1.5 * (1 + 0.01 * splitmix64_uniform(2, n1 ** 3).reshape(shape))
    theta = np.arctan2(Y + 0.5 * dx, X + 0.5 * dx)
    return {
        "shape": shape,
        "dx": dx,
        "npart": N,
        "dt_lattice": dt / dx ** 2,
        "steps": 10000,
        "qf": 0.1,
        "params": [w * dx * dx for w in om],
        "psi0": (np.sqrt(rho) * np.exp(1j * theta)).astype(np.complex128),
    }


# ---------------------------------------------------- truncated-Wigner ensemble
ENS_MIX = ((24, 256), (32, 128), (48, 64), (64, 64))  # 512 lattices, 4:2:1:1
ENS_PARAMS = [0.1, 4.0, 0.0, 0.008]                    # cosine lattice A=0.1, P=4; contact a=0.008
ENS_DT = 0.02


def ensemble_shapes():
    """(shapes, seeds) of the 512-lattice mix, seed 5000 + lattice index."""
    shapes, seeds = [], []
    for L, cnt in ENS_MIX:
        for _ in range(cnt):
            shapes.append((L, L, L))
            seeds.append(5000 + len(seeds))
    return shapes, seeds


def wigner_sample(psi0, seed):
    """psi0 plus complex Gaussian noise of variance 1/2 per plane-wave mode."""
    rng = np.random.default_rng(seed)
    shape = psi0.shape
    beta = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 0.5
    return psi0 + np.fft.ifftn(beta) * np.sqrt(psi0.size)


def ensemble_ground_state(E, L, ite_steps=300):
    """Imaginary-time ground state of one L^3 lattice (one particle per site),
    computed by the engine E (libcugpe_amd.engine) in complex128."""
    shape = (L, L, L)
    g = E.Engine(shape, 0.0, 1.0, ENS_DT, float(L ** 3), E.CONTACT, E.COSINE, 0, E.C128)
    g.set_user_params(ENS_PARAMS)
    g.set_psi(np.ones(shape, dtype=np.complex128))
    g.normalize_psi()
    g.evolve_qf(ite_steps)
    psi, _ = g.get_psi()
    g.close()
    return psi


# ---------------------------------------------------------- adjoint (64^3)
ADJ_SHAPE = (64, 64, 64)
ADJ_DT = 1e-3
ADJ_BUMPS = 512
ADJ_WIDTH = 1.5


def adjoint_inputs():
    """psi0 (seed 5), bump centres (seed 5) and amplitudes ~ N(0, 0.05)
    (seed 5), and the target's potential: another seeded bump field (seed 55)."""
    rng = np.random.default_rng(5)
    from . import adjoint as A

    centres = A.bump_centres(ADJ_SHAPE, ADJ_BUMPS, 5)
    amp0 = 0.05 * rng.standard_normal(ADJ_BUMPS)
    psi0 = (1 + 0.05 * rng.random(ADJ_SHAPE)) * np.exp(1j * 0.1 * rng.standard_normal(ADJ_SHAPE))
    rt = np.random.default_rng(55)
    amp_t = 0.05 * rt.standard_normal(ADJ_BUMPS)
    centres_t = A.bump_centres(ADJ_SHAPE, ADJ_BUMPS, 55)
    return {"centres": centres, "amp0": amp0, "psi0": psi0, "amp_target": amp_t, "centres_target": centres_t}
