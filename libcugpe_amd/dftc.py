"""Wave-function files of the reference's drivers and analysis scripts.

A ``.dftc`` file is raw complex128 frames in C order ``[nom][nx][ny][nz]``
(written by the drivers with ``fwrite`` of psi, test_vortex/ite_vortex.c:221-230);
``<file>.info`` next to it is a text sidecar, one value per line followed by
``#`` and a name.  Two forms exist (include/gpe_io.h restates both):

- ``save_info``: 10 lines nx ny nz dt nom a_scattering aspect_ratio omega_x r0
  npart (dev/gpe_io.h:16-44);
- ``gpe_save_info``: a system-type line ("bosons"/"fermions") and a mode line
  first, the same 10 values, then the total time (dev/gpe_io.h:46-96).

The reference's Python readers (python/read_dftc_info.py:17-38,
python/gpe_dftcutils.py:43-56) take the header form and return the density
as gamma |psi|^2 a_ho^3 with a_ho = 1/sqrt(omega_x) and gamma = 2 for fermions.

Side files of a run (python/gpe_dftcutils.py:79-145, 194-218), written here
by the ``write_*`` / ``*Writer`` helpers from engine outputs:

- ``<file>.energy``: a header line, then one text row per measurement:
  time, E_tot, E_kin, E_int, E_ext and, in imaginary time, mu and the energy
  change, in real time one more column;
- ``<file>.trajectory``: a header line, then rows t x y z (vortex positions);
- ``<file>.currents``: raw float64 frames ``[nom][3][nx][ny][nz]`` (j_x, j_y, j_z).
"""
import math
import os

import numpy as np

NAMES = ("nx", "ny", "nz", "dt", "nom", "a_scattering", "aspect_ratio", "omega_x", "r0", "npart")


def info_path(dftc):
    return dftc + ".info"


def write_info(dftc, nx, ny, nz, dt, nom, a_scattering, aspect_ratio, omega_x, r0, npart, total_time=None,
               system=None, mode=None):
    """Write ``<dftc>.info``: the header form when ``system`` is given (with
    ``mode`` and ``total_time``), the 10-line form otherwise."""
    lines = []
    if system is not None:
        lines.append("%s   # system type" % system)
        lines.append("%s   # mode" % (mode or ""))
        fmt = "%.15e"
    else:
        fmt = "%f"
    vals = (nx, ny, nz, dt, nom, a_scattering, aspect_ratio, omega_x, r0, npart)
    for name, v in zip(NAMES, vals):
        if name in ("nx", "ny", "nz", "nom"):
            lines.append("%d   # %s" % (int(v), name))
        else:
            lines.append((fmt % float(v)) + "   # " + name)
    if system is not None:
        lines.append((fmt % float(total_time or 0.0)) + "   # total time")
    with open(info_path(dftc), "w") as f:
        f.write("\n".join(lines) + "\n")


def _value(tok):
    return float("inf") if "inf" in tok.lower() else float(tok)


def read_info(dftc):
    """The sidecar as a dict (both forms); ``system``/``mode``/``total_time``
    are present for the header form."""
    with open(info_path(dftc)) as f:
        toks = [ln.split("#")[0].split() for ln in f.read().splitlines() if ln.strip()]
    toks = [t[0] if t else "" for t in toks]
    out = {}
    try:
        float(toks[0])
        start = 0
    except ValueError:
        out["system"], out["mode"] = toks[0], toks[1]
        start = 2
    for i, name in enumerate(NAMES):
        v = _value(toks[start + i])
        out[name] = int(v) if name in ("nx", "ny", "nz", "nom") else v
    if len(toks) > start + len(NAMES):
        out["total_time"] = _value(toks[start + len(NAMES)])
    return out


class FrameWriter:
    """Appends psi frames to a .dftc file and keeps its sidecar's ``nom`` current."""

    def __init__(self, dftc, shape, dt, a_scattering=0.0, aspect_ratio=1.0, omega_x=1.0, r0=0.0, npart=0.0,
                 system=None, mode=None, append=False):
        self.path = dftc
        self.shape = tuple(int(s) for s in shape)
        self.meta = dict(dt=dt, a_scattering=a_scattering, aspect_ratio=aspect_ratio, omega_x=omega_x, r0=r0,
                         npart=npart)
        self.system, self.mode = system, mode
        n = int(np.prod(self.shape))
        self.nom = (os.path.getsize(dftc) // (16 * n)) if (append and os.path.exists(dftc)) else 0
        self.f = open(dftc, "ab" if append else "wb")
        self.total_time = 0.0

    def append(self, psi, t=None):
        a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128))
        if a.shape != self.shape:
            raise ValueError("frame shape %s, file %s" % (a.shape, self.shape))
        a.tofile(self.f)
        self.f.flush()
        self.nom += 1
        if t is not None:
            self.total_time = float(t)
        self._info()

    def _info(self):
        m = self.meta
        write_info(self.path, *self.shape, m["dt"], self.nom, m["a_scattering"], m["aspect_ratio"], m["omega_x"],
                   m["r0"], m["npart"], total_time=self.total_time, system=self.system, mode=self.mode)

    def close(self):
        if self.f:
            self.f.close()
            self.f = None
            self._info()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_frames(dftc):
    """All frames, shape (nom, nx, ny, nz), memory-mapped read-only; nom is
    taken from the file size when the sidecar disagrees with it."""
    info = read_info(dftc)
    n = info["nx"] * info["ny"] * info["nz"]
    data = np.memmap(dftc, dtype=np.complex128, mode="r")
    nom = info["nom"] if data.size == info["nom"] * n else data.size // n
    return data[: nom * n].reshape(nom, info["nx"], info["ny"], info["nz"])


def density_phase(dftc, frame=0):
    """(gamma |psi|^2 a_ho^3, arg psi) of one frame, as gpe_dftcutils.get_data."""
    info = read_info(dftc)
    a_ho = 1.0 / math.sqrt(info["omega_x"])
    gamma = 2.0 if "fermions" in info.get("system", "") else 1.0
    psi = np.asarray(read_frames(dftc)[frame])
    return gamma * np.abs(psi) ** 2 * a_ho ** 3, np.angle(psi)


# ------------------------------------------------------------ side files


def write_energy(dftc, rows, mode="ite"):
    """Write ``<dftc>.energy``: rows of (time, etot, ekin, eint, eext, chem,
    diff) in imaginary time or (time, etot, ekin, eint, eext, comp) in real time."""
    cols = ENERGY_COLUMNS[mode]
    with open(dftc + ".energy", "w") as f:
        f.write("# " + " ".join(cols) + "\n")
        for r in rows:
            if len(r) != len(cols):
                raise ValueError("energy row of %d values, %s mode has %d" % (len(r), mode, len(cols)))
            f.write(" ".join("%.15e" % float(x) for x in r) + "\n")


ENERGY_COLUMNS = {"ite": ("time", "etot", "ekin", "eint", "eext", "chem", "diff"),
                  "rte": ("time", "etot", "ekin", "eint", "eext", "comp")}


def read_energy(dftc, mode=None):
    """``<dftc>.energy`` as a dict of columns, as gpe_dftcutils.get_energy: the
    mode is 'rte' when the file name holds '_rte', 'ite' otherwise."""
    if mode is None:
        mode = "rte" if "_rte" in dftc else "ite"
    data = np.loadtxt(dftc + ".energy", skiprows=1, ndmin=2)
    return {name: data[:, k] for k, name in enumerate(ENERGY_COLUMNS[mode])}


def write_trajectory(dftc, rows):
    """Write ``<dftc>.trajectory``: rows of (t, x, y, z)."""
    with open(dftc + ".trajectory", "w") as f:
        f.write("# t x y z\n")
        for t, x, y, z in rows:
            f.write("%.15e %.15e %.15e %.15e\n" % (t, x, y, z))


def read_trajectory(dftc):
    """(t, x, y, z) columns of ``<dftc>.trajectory``, as gpe_dftcutils.get_trajectory."""
    data = np.loadtxt(dftc + ".trajectory", skiprows=1, ndmin=2)
    return data[:, 0], data[:, 1], data[:, 2], data[:, 3]


def trajectory_in_xy(T, X, Y, Z, slice_z):
    """Points of a trajectory in the plane z = slice_z: (t, x, y)."""
    m = Z == slice_z
    return T[m], X[m], Y[m]


def trajectory_at_t(T, X, Y, Z, slice_t):
    """Points of a trajectory at time slice_t: (x, y, z)."""
    m = T == slice_t
    return X[m], Y[m], Z[m]


class CurrentsWriter:
    """Appends (j_x, j_y, j_z) frames, float64 [3][nx][ny][nz] each, to ``<dftc>.currents``."""

    def __init__(self, dftc, shape, append=False):
        self.shape = tuple(int(s) for s in shape)
        self.f = open(dftc + ".currents", "ab" if append else "wb")

    def append(self, jx, jy, jz):
        for j in (jx, jy, jz):
            a = np.ascontiguousarray(np.asarray(j, dtype=np.float64))
            if a.shape != self.shape:
                raise ValueError("current shape %s, file %s" % (a.shape, self.shape))
            a.tofile(self.f)
        self.f.flush()

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_currents(dftc, frame=0):
    """(j_x, j_y, j_z) of one frame of ``<dftc>.currents``, as gpe_dftcutils.get_currents."""
    info = read_info(dftc)
    n = info["nx"] * info["ny"] * info["nz"]
    data = np.memmap(dftc + ".currents", dtype=np.float64, mode="r")[frame * 3 * n:(frame + 1) * 3 * n]
    v = np.asarray(data).reshape(3, info["nx"], info["ny"], info["nz"])
    return v[0], v[1], v[2]


# ------------------------------------------------------------ slices and grids
# (python/gpe_dftcutils.py:149-191)

def get_slice_x(data, y_slice=32, z_slice=32):
    return data[:, y_slice, z_slice]


def get_slice_y(data, x_slice=32, z_slice=32):
    return data[x_slice, :, z_slice]


def get_slice_z(data, x_slice=32, y_slice=32):
    return data[x_slice, y_slice, :]


def get_slice_xy(data, z_slice=32):
    return data[:, :, z_slice]


def get_slice_xz(data, y_slice=32):
    return data[:, y_slice, :]


def get_slice_yz(data, x_slice=32):
    return data[x_slice, :, :]


def create_grid(Nx=8, Ny=8, Nz=8, scale=1.0):
    """Lattice coordinates i - n/2 per axis, times scale."""
    return tuple((np.arange(n, dtype=np.float64) - n / 2.0) * scale for n in (Nx, Ny, Nz))


def create_meshgrid(Nx=8, Ny=8, Nz=0, scale=1.0):
    """Meshgrid (indexing 'ij') of the axes whose size is not 0."""
    axes = [a for a, n in zip(create_grid(Nx, Ny, Nz, scale), (Nx, Ny, Nz)) if n != 0]
    return tuple(np.meshgrid(*axes, indexing="ij"))
