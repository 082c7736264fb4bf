"""The reference's differential test (compare_with_orginal.sh: gpe_imag, then
gpe_real from its ground state, on the 256x32x32 lattice) with the
reference's own driver sources, edited only at their CUDA runtime calls
(tests/drivers/ref_drivers.py) and linked against libgpe_hip.so, against the
FP64 oracle replaying the same calls.  The binaries are built where the
reference tree exists; without them this test is skipped."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "drivers", "bin")
IMAG, REAL = os.path.join(BIN, "gpe_imag_dev"), os.path.join(BIN, "gpe_real_dev")
SHAPE = (256, 32, 32)
PARAMS = [0.01, 0.10, 0.11]
NPART, DT = 1000.0, 0.025


def table(path):
    return [[float(x) for x in l.split()] for l in open(path) if not l.startswith("#")]


@pytest.mark.skipif(not (os.path.exists(IMAG) and os.path.exists(REAL)),
                    reason="reference drivers not built (tests/drivers/ref_drivers.py)")
def test_reference_drivers_match_oracle(tmp_path, oracle):
    print("[gpu-case] reference drivers gpe_imag / gpe_real, 256x32x32", flush=True)
    r = subprocess.run([IMAG], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "lattice: 256 x 32 x 32" in r.stdout
    ite = table(tmp_path / "energy_ite_dev.txt")
    psi_imag = np.fromfile(tmp_path / "psi_dev.bin", dtype=np.complex128).reshape(SHAPE)

    # the oracle replays gpe_imag.c: psi = 1, normalise, energy, then blocks of
    # 100 steps with an energy call until the energy per particle moves < 1e-5
    o = oracle.OracleEngine(SHAPE, 0.0, 1.0, DT, NPART, 0, 0, 0)
    o.set_user_params(PARAMS)
    o.set_psi(np.ones(SHAPE, dtype=complex))
    o.normalize_psi()
    rows = []
    t, a, b, c = o.energy()
    rows.append((t, (a + b + c) / NPART))
    while True:
        o.evolve_qf(100)
        t, a, b, c = o.energy()
        prev = rows[-1][1]
        rows.append((t, (a + b + c) / NPART))
        if abs(prev - rows[-1][1]) < 1e-5:
            break
    assert len(rows) == len(ite)
    for (t, e), line in zip(rows, ite):
        assert abs(line[0] - t) < 0.006 and abs(line[1] - e) < 1.5e-8  # printed as %8.2f %12.8f
    po, _ = o.get_psi()
    err = np.max(np.abs(psi_imag - po)) / np.max(np.abs(po))
    print("[gpu-case] gpe_imag: %d blocks, psi rel err %.2e" % (len(rows) - 1, err), flush=True)
    assert err < 1e-12

    r = subprocess.run([REAL], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    rte = table(tmp_path / "energy_rte_dev.txt")
    psi_real = np.fromfile(tmp_path / "psi_dev.dat", dtype=np.complex128).reshape(SHAPE)
    # gpe_real.c: real time from psi_dev.bin, blocks of 100 steps with an
    # energy call (which renormalises in real time) until t > 10
    o.change_alpha_beta(1.0, 0.0)
    o.set_psi(psi_imag)
    rows = []
    t, a, b, c = o.energy()
    rows.append((t, (a + b + c) / NPART))
    while True:
        o.evolve_qf(100)
        t, a, b, c = o.energy()
        rows.append((t, (a + b + c) / NPART))
        if t > 10.0:
            break
    assert len(rows) == len(rte)
    for (t, e), line in zip(rows, rte):
        assert abs(line[0] - t) < 0.006 and abs(line[1] - e) < 1.5e-8
    po, _ = o.get_psi()
    o.close()
    err = np.max(np.abs(psi_real - po)) / np.max(np.abs(po))
    print("[gpu-case] gpe_real: %d blocks, psi rel err %.2e" % (len(rows) - 1, err), flush=True)
    assert err < 1e-12


RTE_VORTEX = os.path.join(BIN, "rte_vortex_dev")


@pytest.mark.skipif(not os.path.exists(RTE_VORTEX), reason="reference drivers not built (tests/drivers/ref_drivers.py)")
def test_reference_rte_vortex_resumes_from_a_dftc_file(tmp_path, oracle):
    """test_vortex/rte_vortex.c on 32^3: it reads <file>.dftc.info with
    get_info, checks the lattice, loads the frame, evolves 10 blocks of 100
    steps (a fixed loop count, rte_vortex.c:322; its snapshot argument is read
    but unused) with an energy call and hands every psi to its writer thread through
    the frame queue (include/gpe_queue.h, OpenMP sections), which appends it to
    psi_frames.dftc.  Every frame must equal the oracle's replay to 1e-12."""
    from libcugpe_amd import dftc

    print("[gpu-case] reference driver rte_vortex from a .dftc file", flush=True)
    shape, aspect, blocks = (32, 32, 32), 4.0, 10
    params = [0.01, aspect * aspect * 0.01, aspect * aspect * 0.01, 0.0]  # rte_vortex.c:271-275
    o = oracle.OracleEngine(shape, 0.0, 1.0, DT, NPART, 0, 0, 0)
    o.set_user_params(params)
    o.set_psi(np.ones(shape, dtype=complex))
    o.normalize_psi()
    o.evolve_qf(200)
    psi_init, _ = o.get_psi()
    path = str(tmp_path / "psi_init.dftc")
    psi_init.astype(np.complex128).tofile(path)
    dftc.write_info(path, *shape, DT, 1, 0.0, aspect, 0.01, 0.0, NPART)
    r = subprocess.run([RTE_VORTEX, path, str(blocks)], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    out_dirs = [d for d in os.listdir(tmp_path) if d.startswith("AspectRatio_4.0_32x32x32_RTE")]
    assert len(out_dirs) == 1, os.listdir(tmp_path)
    frames = np.fromfile(tmp_path / out_dirs[0] / "psi_frames.dftc", dtype=np.complex128)
    assert frames.size == blocks * int(np.prod(shape))
    frames = frames.reshape((blocks,) + shape)
    # the oracle replays rte_vortex.c: set_psi (no normalisation), energy, then
    # blocks of evolve_qf(100) + energy + get_psi
    o.set_psi(psi_init)
    o.energy()
    for k in range(blocks):
        o.evolve_qf(100)
        o.energy()
        want, _ = o.get_psi()
        err = np.max(np.abs(frames[k] - want)) / np.max(np.abs(want))
        assert err < 1e-12, (k, err)
    o.close()
