"""Wave-function files (libcugpe_amd/dftc.py): raw complex128 frames plus the
.info sidecar in both forms of dev/gpe_io.h, read and written in Python and
through include/gpe_io.h compiled with gcc."""
import os
import subprocess

import numpy as np
import pytest

from libcugpe_amd import dftc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_frames_round_trip(tmp_path):
    shape = (6, 4, 5)
    rng = np.random.default_rng(0)
    frames = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    path = str(tmp_path / "psi.dftc")
    with dftc.FrameWriter(path, shape, dt=0.01, omega_x=0.25, npart=100.0, system="fermions", mode="unitary") as w:
        for k, f in enumerate(frames):
            w.append(f, t=0.01 * (k + 1))
    assert os.path.getsize(path) == 3 * 16 * np.prod(shape)
    info = dftc.read_info(path)
    assert (info["nx"], info["ny"], info["nz"], info["nom"]) == (6, 4, 5, 3)
    assert info["system"] == "fermions" and info["mode"] == "unitary"
    assert info["total_time"] == pytest.approx(0.03, rel=1e-14)
    got = dftc.read_frames(path)
    for k in range(3):
        assert np.array_equal(got[k], frames[k])
    # density as the reference's reader: gamma |psi|^2 a_ho^3, gamma 2 for fermions, a_ho = omega_x^-1/2
    rho, ph = dftc.density_phase(path, frame=1)
    assert np.allclose(rho, 2.0 * np.abs(frames[1]) ** 2 * 8.0, rtol=1e-15)
    assert np.array_equal(ph, np.angle(frames[1]))
    # appending keeps the frame count
    with dftc.FrameWriter(path, shape, dt=0.01, omega_x=0.25, system="fermions", mode="unitary", append=True) as w:
        w.append(frames[0])
    assert dftc.read_info(path)["nom"] == 4
    assert np.array_equal(dftc.read_frames(path)[3], frames[0])


@pytest.fixture(scope="module")
def info_tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("io") / "info_io")
    subprocess.check_call(["gcc", "-O1", "-DINTERACTIONS=0", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "drivers", "info_io.c"), "-o", exe])
    return exe


def test_c_writer_python_reader(tmp_path, info_tool):
    base = str(tmp_path / "gs.dftc")
    subprocess.check_call([info_tool, "w", base])
    a = dftc.read_info(base)
    b = dftc.read_info(base + "_h")
    for d in (a, b):
        assert (d["nx"], d["ny"], d["nz"], d["nom"]) == (24, 16, 32, 7)
        assert d["dt"] == pytest.approx(0.0025, abs=1e-6)
        assert d["omega_x"] == pytest.approx(0.390625, abs=1e-6)
        assert d["npart"] == pytest.approx(10000.0)
    assert "system" not in a
    # INTERACTIONS 0 builds the fermion (dimer) header of gpe_save_info
    assert b["system"] == "fermions" and b["mode"] == "unitary"
    assert b["total_time"] == 17.5
    assert b["dt"] == 0.0025  # %.15e keeps the value


def test_python_writer_c_reader(tmp_path, info_tool):
    base = str(tmp_path / "x.dftc")
    dftc.write_info(base, 8, 12, 20, 0.5, 3, 0.75, 1.25, 0.5, 2.0, 321.0)
    out = subprocess.check_output([info_tool, "r", base + ".info"]).decode().split()
    assert [int(v) for v in out[:3]] == [8, 12, 20]
    assert float(out[3]) == 0.5 and int(out[4]) == 3
    assert [float(v) for v in out[5:]] == [0.75, 1.25, 0.5, 2.0, 321.0]


def test_side_files_and_slices(tmp_path):
    """.energy / .trajectory / .currents side files (python/gpe_dftcutils.py:79-145,
    194-218) written and read back, slices and grids."""
    path = str(tmp_path / "run_rte.dftc")
    shape = (4, 3, 5)
    dftc.write_info(path, *shape, 0.01, 2, 0.0, 1.0, 1.0, 0.0, 10.0, total_time=0.02, system="bosons", mode="rte")
    rows = [(0.01 * k, -1.0 - k, 0.5, -1.7, 0.2 + k, 1e-9 * k) for k in range(3)]
    dftc.write_energy(path, rows, mode="rte")
    e = dftc.read_energy(path)  # mode from the file name
    assert set(e) == {"time", "etot", "ekin", "eint", "eext", "comp"}
    assert np.array_equal(e["etot"], np.array([r[1] for r in rows]))
    with pytest.raises(ValueError):
        dftc.write_energy(path, [(0.0, 1.0)], mode="ite")
    traj = [(0.0, 1.0, 2.0, 0.0), (0.1, 1.5, 2.5, 0.0), (0.1, 1.0, 2.0, 3.0)]
    dftc.write_trajectory(path, traj)
    T, X, Y, Z = dftc.read_trajectory(path)
    assert np.array_equal(X, [1.0, 1.5, 1.0])
    t, x, y = dftc.trajectory_in_xy(T, X, Y, Z, 0.0)
    assert np.array_equal(t, [0.0, 0.1]) and np.array_equal(y, [2.0, 2.5])
    x, y, z = dftc.trajectory_at_t(T, X, Y, Z, 0.1)
    assert np.array_equal(z, [0.0, 3.0])
    rng = np.random.default_rng(2)
    js = [[rng.standard_normal(shape) for _ in range(3)] for _ in range(2)]
    with dftc.CurrentsWriter(path, shape) as w:
        for j in js:
            w.append(*j)
    assert os.path.getsize(path + ".currents") == 2 * 3 * 8 * int(np.prod(shape))
    for k in range(2):
        got = dftc.read_currents(path, frame=k)
        for a, b in zip(got, js[k]):
            assert np.array_equal(a, b)
    d = np.arange(np.prod(shape)).reshape(shape)
    assert np.array_equal(dftc.get_slice_x(d, 1, 2), d[:, 1, 2])
    assert np.array_equal(dftc.get_slice_yz(d, 3), d[3])
    x, y, z = dftc.create_grid(4, 3, 5, scale=0.5)
    assert np.array_equal(x, [-1.0, -0.5, 0.0, 0.5]) and np.array_equal(y, [-0.75, -0.25, 0.25])
    X, Y = dftc.create_meshgrid(4, 3, 0)
    assert X.shape == (4, 3) and np.array_equal(X[:, 0], 2 * x) and np.array_equal(Y[0], 2 * y)
