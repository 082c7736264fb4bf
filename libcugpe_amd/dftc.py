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
