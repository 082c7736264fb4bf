"""Drop-in check on the GPU: a C++ driver written against include/gpe_engine.h
with the reference's build flags (-DNX -DNY -DNZ) runs the call sequence of
the reference's drivers (test_compare/gpe_imag.c, gpe_real.c,
test_vortex/ite_vortex.c) through libgpe_hip.so; every psi it writes and every
energy it prints is compared with the FP64 oracle replaying the same calls."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "drivers", "bin", "ite_rte_driver")
SHAPE = (32, 16, 24)  # tests/drivers/Makefile NX NY NZ
BLOCKS = 5


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built (make -C tests/drivers)")
def test_driver_matches_oracle(tmp_path, oracle):
    print("[gpu-case] ite_rte_driver", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path), str(BLOCKS)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "driver: OK" in r.stdout
    assert "lattice: 32 x 16 x 24" in r.stdout

    lines = [l.split() for l in open(tmp_path / "energies.txt")]
    load = lambda name: np.fromfile(tmp_path / name, dtype=np.complex128).reshape(SHAPE)

    params = [0.01, 0.10, 0.11]
    o = oracle.OracleEngine(SHAPE, 0.0, 1.0, 0.025, 1000.0, 0, 0)
    o.set_user_params(params)
    o.set_psi(np.ones(SHAPE, dtype=complex))
    o.normalize_psi()
    ref = [("ite",) + o.energy()]
    etot = sum(ref[0][2:])
    for _ in range(BLOCKS):
        mu = o.evolve_qf(20, want_mu=True)
        e = o.energy()
        ref.append(("ite",) + e + (mu,))
        prev, etot = etot, sum(e[1:])
        if abs((prev - etot) / 1000.0) < 1e-9:
            break
    psi_ite, _ = o.get_psi()
    assert rel(load("psi_ite.bin"), psi_ite) < 1e-12

    rte = oracle.OracleEngine(SHAPE, 1.0, 0.0, 0.025, 1000.0, 0, 0)
    rte.set_user_params(params)
    rte.set_psi(psi_ite)
    for _ in range(3):
        rte.evolve_qf(20)
        ref.append(("rte",) + rte.energy())
    assert rel(load("psi_rte.bin"), rte.get_psi()[0]) < 1e-12
    rte.change_alpha_beta(0.0, 1.0)
    rte.set_vortex(0.0, 2.5, 1)
    rte.evolve_vortex(20)
    ref.append(("vortex",) + rte.energy())
    assert rel(load("psi_vortex.bin"), rte.get_psi()[0]) < 1e-12

    assert len(lines) == len(ref)
    for got, want in zip(lines, ref):
        assert got[0] == want[0]
        assert abs(float(got[1]) - want[1]) < 1e-9  # time
        ek, ei, ee = (float(x) for x in got[3:6])
        for a, b in zip((ek, ei, ee), want[2:5]):
            assert abs(a - b) <= 1e-12 * max(abs(b), 1.0)
        if want[0] == "ite" and len(want) > 5:
            assert abs(float(got[6]) - want[5]) <= 1e-12 * max(abs(want[5]), 1.0)

    info = [l.split()[0] for l in open(tmp_path / "psi_ite.bin.info")]
    assert info[:3] == ["32", "16", "24"] and float(info[3]) == 0.025 and float(info[9]) == 1000.0
