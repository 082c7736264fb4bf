"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package
(libcugpe_amd) never imports this module.

The oracle restates orginal/gpe_engine.cu (and the vortex / chemical-potential
additions of dev/gpe_engine.cu) in FP64 C; see gpe_oracle.c for the per-function
reference citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


def build(force=False):
    """Compile oracle/liboracle.so with the committed Makefile (make's own
    dependency check: the sources and the user hook file).  Where no
    compiler is present, as on a box that only runs the prebuilt library, an
    existing library is used as is."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    else:
        try:
            subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
        except (OSError, subprocess.CalledProcessError):
            pass
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.gpeo_create.restype = _p
        L.gpeo_create.argtypes = [_i, _i, _i, _i, _i, _i, _d, _d, _d, _d]
        L.gpeo_destroy.argtypes = [_p]
        for name, args in [
            ("gpeo_change_alpha_beta", [_p, _d, _d]),
            ("gpeo_set_time", [_p, _d]),
            ("gpeo_set_user_params", [_p, _i, _p]),
            ("gpeo_set_quantum_friction_coeff", [_p, _d]),
            ("gpeo_set_vortex", [_p, _d, _d, _i]),
            ("gpeo_set_vext_array", [_p, _p]),
            ("gpeo_set_psi", [_p, _d, _p]),
            ("gpeo_get_psi", [_p, _dp, _p]),
            ("gpeo_normalize_psi", [_p]),
            ("gpeo_get_norm", [_p, _dp]),
            ("gpeo_evolve_qf", [_p, _i, _dp]),
            ("gpeo_evolve_vortex", [_p, _i, _dp]),
            ("gpeo_evolve_vortex2", [_p, _i, _dp]),
            ("gpeo_evolve_vortex3", [_p, _i, _dp]),
            ("gpeo_evolve_enforced_phase", [_p, _i, _dp]),
            ("gpeo_set_phase", [_p, _p]),
            ("gpeo_get_phase", [_p, _p]),
            ("gpeo_energy", [_p, _dp, _dp, _dp, _dp]),
            ("gpeo_get_density", [_p, _dp, _p]),
            ("gpeo_get_currents", [_p, _dp, _p, _p, _p]),
            ("gpeo_tables", [_p, _p, _p, _p, _p, _p, _p, _p]),
            ("gpeo_vext_value", [_p, _i, _i, _i, _dp]),
            ("gpeo_dedfdn_value", [_p, _d, _dp, _dp]),
            ("gpeo_fft3d", [_i, _i, _i, _p, _i]),
            ("gpeo_fft1d", [_i, _p, _i]),
            ("gpeo_adjoint_overlap", [_p, _i, _p, _i, _dp, _p]),
            ("gpeo_evolve_dipolar", [_p, _i, _dp]),
            ("gpeo_energy_dipolar", [_p, _dp, _dp]),
            ("gpeo_energy_dipolar2", [_p, _dp, _dp]),
            ("gpeo_dipolar_potential", [_p, _p]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _i
        L.gpeo_gamma.argtypes = [_p]
        L.gpeo_gamma.restype = _d
        L.gpeo_unit_root.argtypes = [ctypes.c_long, ctypes.c_long, _dp, _dp]
        L.gpeo_unit_root.restype = None
        L.gpeo_plaquette_windings.argtypes = [_i, _i, _i, _p, _p, _d, _p]
        L.gpeo_plaquette_windings.restype = ctypes.c_long
        L.gpeo_num_threads.restype = _i
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def fft3d(a, sign):
    """Unnormalised 3-D DFT of a complex128 array (copy), sign -1 forward."""
    b = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nx, ny, nz = b.shape
    rc = lib().gpeo_fft3d(nx, ny, nz, _ptr(b), sign)
    assert rc == 0
    return b


def fft1d(a, sign):
    b = np.ascontiguousarray(a, dtype=np.complex128).copy()
    rc = lib().gpeo_fft1d(b.size, _ptr(b), sign)
    assert rc == 0
    return b


def unit_root(m, n):
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().gpeo_unit_root(m, n, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def plaquette_windings(psi, rho_min):
    """int8 windings [3, nx, ny, nz] of psi (complex128 or complex64)."""
    nx, ny, nz = psi.shape
    w = np.zeros((3, nx, ny, nz), dtype=np.int8)
    if psi.dtype == np.complex64:
        a = np.ascontiguousarray(psi)
        cnt = lib().gpeo_plaquette_windings(nx, ny, nz, None, _ptr(a), rho_min, _ptr(w))
    else:
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        cnt = lib().gpeo_plaquette_windings(nx, ny, nz, _ptr(a), None, rho_min, _ptr(w))
    return w, int(cnt)


class OracleEngine:
    """The oracle engine behind the same call surface as the reference's
    gpe_engine.h (dev/gpe_engine.h:299-558), handle-based."""

    def __init__(self, shape, alpha, beta, dt, npart, interactions=0, vext=0, gpe_for=0):
        self.shape = tuple(int(s) for s in shape)
        nx, ny, nz = self.shape
        self._h = lib().gpeo_create(nx, ny, nz, interactions, vext, gpe_for, alpha, beta, dt, npart)
        if not self._h:
            raise ValueError("oracle: invalid lattice %r" % (self.shape,))
        self.n = nx * ny * nz

    def close(self):
        if self._h:
            lib().gpeo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def gamma(self):
        return lib().gpeo_gamma(self._h)

    def change_alpha_beta(self, a, b):
        assert lib().gpeo_change_alpha_beta(self._h, a, b) == 0

    def set_time(self, t0):
        assert lib().gpeo_set_time(self._h, t0) == 0

    def set_user_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        return lib().gpeo_set_user_params(self._h, p.size, _ptr(p))

    def set_quantum_friction_coeff(self, qf):
        assert lib().gpeo_set_quantum_friction_coeff(self._h, qf) == 0

    def set_vortex(self, x0, y0, Q=1):
        assert lib().gpeo_set_vortex(self._h, x0, y0, Q) == 0

    def set_vext_array(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(self.shape)
        assert lib().gpeo_set_vext_array(self._h, _ptr(a)) == 0

    def set_psi(self, psi, t=0.0):
        a = np.ascontiguousarray(psi, dtype=np.complex128).reshape(self.shape)
        assert lib().gpeo_set_psi(self._h, t, _ptr(a)) == 0

    def get_psi(self):
        a = np.empty(self.shape, dtype=np.complex128)
        t = ctypes.c_double()
        assert lib().gpeo_get_psi(self._h, ctypes.byref(t), _ptr(a)) == 0
        return a, t.value

    def normalize_psi(self):
        assert lib().gpeo_normalize_psi(self._h) == 0

    def get_norm(self):
        v = ctypes.c_double()
        assert lib().gpeo_get_norm(self._h, ctypes.byref(v)) == 0
        return v.value

    def evolve_qf(self, nt, want_mu=False):
        mu = ctypes.c_double(float("nan"))
        assert lib().gpeo_evolve_qf(self._h, nt, ctypes.byref(mu) if want_mu else None) == 0
        return mu.value if want_mu else None

    def evolve_vortex(self, nt, want_mu=False):
        mu = ctypes.c_double(float("nan"))
        assert lib().gpeo_evolve_vortex(self._h, nt, ctypes.byref(mu) if want_mu else None) == 0
        return mu.value if want_mu else None

    def _evolve_variant(self, fn, nt, want_mu):
        mu = ctypes.c_double(float("nan"))
        rc = fn(self._h, nt, ctypes.byref(mu) if want_mu else None)
        if rc != 0:
            raise RuntimeError("oracle evolve failed: %d" % rc)
        return mu.value if want_mu else None

    def evolve_vortex2(self, nt, want_mu=False):
        return self._evolve_variant(lib().gpeo_evolve_vortex2, nt, want_mu)

    def evolve_vortex3(self, nt, want_mu=False):
        return self._evolve_variant(lib().gpeo_evolve_vortex3, nt, want_mu)

    def evolve_enforced_phase(self, nt, want_mu=False):
        return self._evolve_variant(lib().gpeo_evolve_enforced_phase, nt, want_mu)

    def set_phase(self, phase):
        a = np.ascontiguousarray(phase, dtype=np.float64).reshape(self.shape)
        assert lib().gpeo_set_phase(self._h, _ptr(a)) == 0

    def get_phase(self):
        a = np.empty(self.shape, dtype=np.float64)
        assert lib().gpeo_get_phase(self._h, _ptr(a)) == 0
        return a

    def energy(self):
        t, ek, ei, ee = (ctypes.c_double() for _ in range(4))
        assert lib().gpeo_energy(self._h, ctypes.byref(t), ctypes.byref(ek), ctypes.byref(ei), ctypes.byref(ee)) == 0
        return t.value, ek.value, ei.value, ee.value

    def get_density(self):
        a = np.empty(self.shape, dtype=np.float64)
        t = ctypes.c_double()
        assert lib().gpeo_get_density(self._h, ctypes.byref(t), _ptr(a)) == 0
        return a, t.value

    def evolve_dipolar(self, nt, want_mu=False):
        mu = ctypes.c_double(float("nan"))
        assert lib().gpeo_evolve_dipolar(self._h, int(nt), ctypes.byref(mu) if want_mu else None) == 0
        return mu.value if want_mu else None

    def energy_dipolar(self, method=1):
        t, ed = ctypes.c_double(), ctypes.c_double()
        fn = lib().gpeo_energy_dipolar if method == 1 else lib().gpeo_energy_dipolar2
        assert fn(self._h, ctypes.byref(t), ctypes.byref(ed)) == 0
        return t.value, ed.value

    def dipolar_potential(self):
        v = np.empty(self.shape, dtype=np.float64)
        assert lib().gpeo_dipolar_potential(self._h, _ptr(v)) == 0
        return v

    def adjoint_overlap(self, nt, psi_target, ckpt=0):
        """(L, dL/dV_ext) of L = 1 - |<psi_t|psi(T)>|^2/(|psi_t|^2 |psi(0)|^2) over nt steps."""
        pt = np.ascontiguousarray(np.asarray(psi_target, dtype=np.complex128))
        grad = np.empty(self.shape, dtype=np.float64)
        loss = ctypes.c_double()
        rc = lib().gpeo_adjoint_overlap(self._h, int(nt), _ptr(pt), int(ckpt), ctypes.byref(loss), _ptr(grad))
        if rc != 0:
            raise ValueError("gpeo_adjoint_overlap: %d" % rc)
        return loss.value, grad

    def get_currents(self):
        js = [np.empty(self.shape, dtype=np.float64) for _ in range(3)]
        t = ctypes.c_double()
        assert lib().gpeo_get_currents(self._h, ctypes.byref(t), *(_ptr(j) for j in js)) == 0
        return js, t.value

    def tables(self):
        nx, ny, nz = self.shape
        kx, ky, kz = np.empty(nx), np.empty(ny), np.empty(nz)
        Kx, Ky, Kz = (np.empty(n, dtype=np.complex128) for n in (nx, ny, nz))
        c = np.empty(2)
        lib().gpeo_tables(self._h, *(_ptr(x) for x in (kx, ky, kz, Kx, Ky, Kz, c)))
        return dict(kx=kx, ky=ky, kz=kz, Kx=Kx, Ky=Ky, Kz=Kz, c=complex(c[0], c[1]))

    def vext(self, ix, iy, iz):
        v = ctypes.c_double()
        lib().gpeo_vext_value(self._h, ix, iy, iz, ctypes.byref(v))
        return v.value

    def dedfdn(self, rho):
        dv, e = ctypes.c_double(), ctypes.c_double()
        lib().gpeo_dedfdn_value(self._h, rho, ctypes.byref(dv), ctypes.byref(e))
        return dv.value, e.value
