"""Python binding of the engine's C ABI (include/gpe_engine.h) over ctypes.

`Engine` mirrors the reference's gpe_engine.h call surface
(dev/gpe_engine.h:299-558) on a handle: same method names without the gpe_
prefix, same argument meaning, and every non-zero return code raised as
`GpeError` carrying the code and gpe_get_error_string().  The compute runs in
libcugpe_amd/libgpe_hip.so; if that library is missing this module raises
instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpe_hip.so")

# physics codes (include/gpe_user_defined.h); USER selects the user hook file
UFG, ETF, CONTACT, DIPOLAR = 0, 1, -1, -2
HARMONIC, WOODS_SAXON, COSINE, VARRAY = 0, 1, 2, 3
USER = 100
PARTICLES, DIMERS = 1, 2
C128, C64 = 128, 64

_lib = None


class GpeError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__("%s failed: error %d (%s)" % (where, code, error_string(code)))


class Config(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nx", "ny", "nz", "interactions", "vext", "gpe_for", "precision")]


_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Load libgpe_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libgpe_hip.so not built: run `python -m libcugpe_amd.build` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "gpe_engine_create": [ctypes.POINTER(_p), ctypes.POINTER(Config), _d, _d, _d, _d],
            "gpe_engine_destroy": [_p],
            "gpe_engine_change_alpha_beta": [_p, _d, _d],
            "gpe_engine_set_time": [_p, _d],
            "gpe_engine_set_user_params": [_p, _i, _p],
            "gpe_engine_set_quantum_friction_coeff": [_p, _d],
            "gpe_engine_set_vortex": [_p, _d, _d, _i],
            "gpe_engine_set_external_potential": [_p, _p],
            "gpe_engine_set_external_potential_device": [_p, _p],
            "gpe_engine_adjoint_backward_device": [_p, _p],
            "gpe_engine_set_psi": [_p, _d, _p],
            "gpe_engine_get_psi": [_p, _dp, _p],
            "gpe_engine_normalize_psi": [_p],
            "gpe_engine_get_norm": [_p, _dp],
            "gpe_engine_evolve_qf": [_p, _i, _dp],
            "gpe_engine_evolve_vortex": [_p, _i, _dp],
            "gpe_engine_evolve_vortex2": [_p, _i, _dp],
            "gpe_engine_evolve_vortex3": [_p, _i, _dp],
            "gpe_engine_evolve_enforced_phase": [_p, _i, _dp],
            "gpe_engine_energy": [_p, _dp, _dp, _dp, _dp],
            "gpe_engine_get_density": [_p, _dp, _p],
            "gpe_engine_get_currents": [_p, _dp, _p, _p, _p],
            "gpe_engine_get_phase": [_p, _p],
            "gpe_engine_set_phase": [_p, _p],
            "gpe_engine_detect_vortices": [_p, _d, _p, ctypes.POINTER(ctypes.c_long)],
            "gpe_engine_cloud_params": [_p, _d, _p],
            "gpe_detect_vortices_noalloc": [_p, _i, _i, _i, _i, _d, _p, _p, _p],
            "gpe_engine_get_time": [_p, _dp],
            "gpe_engine_fft": [_p, _i],
            "gpe_engine_get_config": [_p, ctypes.POINTER(Config)],
            "gpe_engine_device_psi": [_p, ctypes.POINTER(_p)],
            "gpe_engine_stream": [_p, ctypes.POINTER(_p)],
            "gpe_engine_synchronize": [_p],
            "gpe_engine_profile": [_p, _i, _i],
            "gpe_engine_profile_read": [_p, _dp, ctypes.POINTER(ctypes.c_long)],
            "gpe_engine_copy_benchmark": [_p, _dp],
            "gpe_engine_autotune": [_p, _i, ctypes.POINTER(_i)],
            "gpe_engine_evolve_dipolar": [_p, _i, _dp],
            "gpe_engine_energy_dipolar": [_p, _i, _dp, _dp],
            "gpe_engine_adjoint_forward": [_p, _i, _p, _i, _dp],
            "gpe_engine_adjoint_backward": [_p, _p],
            "gpe_engine_set_variant": [_p, _i],
            "gpe_engine_set_graph": [_p, _i],
            "gpe_ensemble_create": [ctypes.POINTER(_p), _i, _p, _i, _i, _i, _i, _d, _d, _d, _d],
            "gpe_ensemble_destroy": [_p],
            "gpe_ensemble_set_user_params": [_p, _i, _p],
            "gpe_ensemble_set_quantum_friction_coeff": [_p, _d],
            "gpe_ensemble_set_psi": [_p, _i, _p],
            "gpe_ensemble_set_external_potential": [_p, _i, _p],
            "gpe_ensemble_set_instance_params": [_p, _i, _i, _p],
            "gpe_ensemble_get_psi": [_p, _i, _dp, _p],
            "gpe_ensemble_evolve": [_p, _i],
            "gpe_ensemble_set_graph": [_p, _i],
            "gpe_ensemble_autotune": [_p, _i],
            "gpe_ensemble_set_variant": [_p, _i, _i],
            "gpe_ensemble_groups": [_p, ctypes.POINTER(_i)],
            "gpe_ensemble_group_info": [_p, _i] + [ctypes.POINTER(_i)] * 5,
            "gpe_ensemble_instance": [_p, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)],
            "gpe_ensemble_mean_density": [_p, _i, _p],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _i
        L.gpe_get_error_string.argtypes = [_i]
        L.gpe_get_error_string.restype = ctypes.c_char_p
        L.gpe_build_info.restype = ctypes.c_char_p
        _lib = L
    return _lib


def error_string(code):
    try:
        return lib().gpe_get_error_string(code).decode()
    except OSError:
        return "library not loaded"


def _check(rc, where):
    if rc != 0:
        raise GpeError(rc, where)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Engine:
    """One engine instance (handle) of a given lattice, physics and precision."""

    def __init__(self, shape, alpha, beta, dt, npart, interactions=UFG, vext=HARMONIC, gpe_for=0, precision=C128):
        self.shape = tuple(int(s) for s in shape)
        self.n = int(np.prod(self.shape))
        self.precision = precision
        cfg = Config(self.shape[0], self.shape[1], self.shape[2], interactions, vext, gpe_for, precision)
        h = ctypes.c_void_p()
        _check(lib().gpe_engine_create(ctypes.byref(h), ctypes.byref(cfg), alpha, beta, dt, npart), "gpe_engine_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().gpe_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------- setters
    def change_alpha_beta(self, alpha, beta):
        _check(lib().gpe_engine_change_alpha_beta(self._h, alpha, beta), "gpe_change_alpha_beta")

    def set_rte_evolution(self):
        self.change_alpha_beta(1.0, 0.0)

    def set_ite_evolution(self):
        self.change_alpha_beta(0.0, 1.0)

    def set_time(self, t0):
        _check(lib().gpe_engine_set_time(self._h, t0), "gpe_set_time")

    def set_user_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        _check(lib().gpe_engine_set_user_params(self._h, p.size, _ptr(p)), "gpe_set_user_params")

    def set_quantum_friction_coeff(self, qf):
        _check(lib().gpe_engine_set_quantum_friction_coeff(self._h, qf), "gpe_set_quantum_friction_coeff")

    def set_vortex(self, x0, y0, Q=1):
        _check(lib().gpe_engine_set_vortex(self._h, x0, y0, int(Q)), "gpe_set_vortex")

    def set_external_potential(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(self.shape)
        _check(lib().gpe_engine_set_external_potential(self._h, _ptr(a)), "gpe_set_external_potential")

    def set_external_potential_device(self, ptr):
        """VEXT 3 array from FP64 device memory at address `ptr` (nx*ny*nz
        doubles, C order); the producer's writes must be complete."""
        _check(lib().gpe_engine_set_external_potential_device(self._h, ctypes.c_void_p(int(ptr))),
               "gpe_engine_set_external_potential_device")

    # ------------------------------------------------------ wave function
    def set_psi(self, psi, t=0.0):
        a = np.ascontiguousarray(psi, dtype=np.complex128).reshape(self.shape)
        _check(lib().gpe_engine_set_psi(self._h, t, _ptr(a)), "gpe_set_psi")

    def get_psi(self):
        a = np.empty(self.shape, dtype=np.complex128)
        t = ctypes.c_double()
        _check(lib().gpe_engine_get_psi(self._h, ctypes.byref(t), _ptr(a)), "gpe_get_psi")
        return a, t.value

    def normalize_psi(self):
        _check(lib().gpe_engine_normalize_psi(self._h), "gpe_normalize_psi")

    def get_norm(self):
        v = ctypes.c_double()
        _check(lib().gpe_engine_get_norm(self._h, ctypes.byref(v)), "gpe_get_norm")
        return v.value

    # ---------------------------------------------------------- evolution
    def evolve_qf(self, nt, want_mu=False):
        mu = ctypes.c_double(float("nan"))
        _check(lib().gpe_engine_evolve_qf(self._h, nt, ctypes.byref(mu) if want_mu else None), "gpe_evolve_qf")
        return mu.value if want_mu else None

    evolve = evolve_qf

    def evolve_vortex(self, nt, want_mu=False):
        mu = ctypes.c_double(float("nan"))
        _check(lib().gpe_engine_evolve_vortex(self._h, nt, ctypes.byref(mu) if want_mu else None), "gpe_evolve_vortex")
        return mu.value if want_mu else None

    def _evolve_variant(self, name, nt, want_mu):
        mu = ctypes.c_double(float("nan"))
        _check(getattr(lib(), name)(self._h, nt, ctypes.byref(mu) if want_mu else None), name)
        return mu.value if want_mu else None

    def evolve_vortex2(self, nt, want_mu=False):
        return self._evolve_variant("gpe_engine_evolve_vortex2", nt, want_mu)

    def evolve_vortex3(self, nt, want_mu=False):
        return self._evolve_variant("gpe_engine_evolve_vortex3", nt, want_mu)

    def evolve_enforced_phase(self, nt, want_mu=False):
        return self._evolve_variant("gpe_engine_evolve_enforced_phase", nt, want_mu)

    # -------------------------------------------------------- observables
    def energy(self):
        t, ek, ei, ee = (ctypes.c_double() for _ in range(4))
        _check(lib().gpe_engine_energy(self._h, ctypes.byref(t), ctypes.byref(ek), ctypes.byref(ei), ctypes.byref(ee)),
               "gpe_energy")
        return t.value, ek.value, ei.value, ee.value

    def get_density(self):
        a = np.empty(self.shape, dtype=np.float64)
        t = ctypes.c_double()
        _check(lib().gpe_engine_get_density(self._h, ctypes.byref(t), _ptr(a)), "gpe_get_density")
        return a, t.value

    def get_currents(self):
        js = [np.empty(self.shape, dtype=np.float64) for _ in range(3)]
        t = ctypes.c_double()
        _check(lib().gpe_engine_get_currents(self._h, ctypes.byref(t), *(_ptr(j) for j in js)), "gpe_get_currents")
        return js, t.value

    def get_phase(self):
        a = np.empty(self.shape, dtype=np.float64)
        _check(lib().gpe_engine_get_phase(self._h, _ptr(a)), "gpe_get_phase")
        return a

    def set_phase(self, phase):
        a = np.ascontiguousarray(phase, dtype=np.float64).reshape(self.shape)
        _check(lib().gpe_engine_set_phase(self._h, _ptr(a)), "gpe_set_phase")

    def detect_vortices(self, rho_min, want_windings=True):
        w = np.zeros((3,) + self.shape, dtype=np.int8) if want_windings else None
        cnt = ctypes.c_long()
        _check(lib().gpe_engine_detect_vortices(self._h, rho_min, _ptr(w) if w is not None else None,
                                                ctypes.byref(cnt)), "gpe_detect_vortices")
        return w, cnt.value

    def cloud_params(self, epsilon=1e-8):
        """(n0, max density, range x, range y, range z) computed on the device
        (dev/gpe_cloud_params.h:60-119: density at the centre, largest density,
        first distance from the centre with density below epsilon, NaN if none)."""
        out = np.empty(5, dtype=np.float64)
        _check(lib().gpe_engine_cloud_params(self._h, float(epsilon), _ptr(out)), "gpe_engine_cloud_params")
        return tuple(float(x) for x in out)

    def fft(self, sign):
        """In-place unnormalised 3-D DFT of psi (sign -1 forward)."""
        _check(lib().gpe_engine_fft(self._h, sign), "gpe_engine_fft")

    def get_time(self):
        t = ctypes.c_double()
        _check(lib().gpe_engine_get_time(self._h, ctypes.byref(t)), "gpe_get_time")
        return t.value

    # -------------------------------------------------------- interop
    def stream(self):
        s = ctypes.c_void_p()
        _check(lib().gpe_engine_stream(self._h, ctypes.byref(s)), "gpe_engine_stream")
        return s.value

    def device_psi(self):
        p = ctypes.c_void_p()
        _check(lib().gpe_engine_device_psi(self._h, ctypes.byref(p)), "gpe_engine_device_psi")
        return p.value

    def synchronize(self):
        _check(lib().gpe_engine_synchronize(self._h), "gpe_engine_synchronize")

    def profile(self, axis_class, enable=True):
        _check(lib().gpe_engine_profile(self._h, axis_class, 1 if enable else 0), "gpe_engine_profile")

    def copy_benchmark(self):
        """Milliseconds of one plain copy kernel of the state array."""
        ms = ctypes.c_double()
        _check(lib().gpe_engine_copy_benchmark(self._h, ctypes.byref(ms)), "gpe_engine_copy_benchmark")
        return ms.value

    def evolve_dipolar(self, nt, want_mu=False):
        mu = ctypes.c_double(float("nan"))
        _check(lib().gpe_engine_evolve_dipolar(self._h, int(nt), ctypes.byref(mu) if want_mu else None),
               "gpe_evolve_dipolar")
        return mu.value if want_mu else None

    def energy_dipolar(self, method=1):
        t, ed = ctypes.c_double(), ctypes.c_double()
        _check(lib().gpe_engine_energy_dipolar(self._h, int(method), ctypes.byref(t), ctypes.byref(ed)),
               "gpe_energy_dipolar")
        return t.value, ed.value

    def adjoint_forward(self, nt, psi_target, ckpt=0):
        """nt real-time steps keeping checkpoints; returns
        L = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2 |psi(0)|^2)."""
        pt = np.ascontiguousarray(np.asarray(psi_target, dtype=np.complex128))
        if pt.shape != self.shape:
            raise ValueError("target shape %s, lattice %s" % (pt.shape, self.shape))
        loss = ctypes.c_double()
        _check(lib().gpe_engine_adjoint_forward(self._h, int(nt), _ptr(pt), int(ckpt), ctypes.byref(loss)),
               "gpe_engine_adjoint_forward")
        self.adjoint_generation = getattr(self, "adjoint_generation", 0) + 1
        return loss.value

    def adjoint_backward(self):
        """dL/dV_ext at every lattice point for the last adjoint_forward."""
        g = np.empty(self.shape, dtype=np.float64)
        _check(lib().gpe_engine_adjoint_backward(self._h, _ptr(g)), "gpe_engine_adjoint_backward")
        return g

    def adjoint_backward_device(self, ptr):
        """dL/dV_ext written to FP64 device memory at address `ptr` (N doubles)."""
        _check(lib().gpe_engine_adjoint_backward_device(self._h, ctypes.c_void_p(int(ptr))),
               "gpe_engine_adjoint_backward_device")

    def autotune(self, steps=3):
        """Time each launcher variant over `steps` real-time steps, keep the fastest; returns it."""
        v = ctypes.c_int()
        _check(lib().gpe_engine_autotune(self._h, int(steps), ctypes.byref(v)), "gpe_engine_autotune")
        return v.value

    def set_graph(self, on):
        _check(lib().gpe_engine_set_graph(self._h, 1 if on else 0), "gpe_engine_set_graph")

    def set_variant(self, variant):
        _check(lib().gpe_engine_set_variant(self._h, int(variant)), "gpe_engine_set_variant")

    def profile_read(self):
        ms = ctypes.c_double()
        cnt = ctypes.c_long()
        _check(lib().gpe_engine_profile_read(self._h, ctypes.byref(ms), ctypes.byref(cnt)), "gpe_engine_profile_read")
        return ms.value, cnt.value


class Ensemble:
    """Many independent lattices advanced together in real time
    (gpe_ensemble_* of include/gpe_engine.h).  `shapes` lists one (nx, ny, nz)
    per lattice; lattices of one shape form a group."""

    def __init__(self, shapes, alpha, beta, dt, npart, interactions=UFG, vext=HARMONIC, gpe_for=0, precision=C64):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        dims = np.ascontiguousarray(np.array(self.shapes, dtype=np.int32).reshape(-1))
        h = ctypes.c_void_p()
        _check(lib().gpe_ensemble_create(ctypes.byref(h), len(self.shapes), _ptr(dims), interactions, vext, gpe_for,
                                         precision, alpha, beta, dt, npart), "gpe_ensemble_create")
        self.h = h

    def close(self):
        if self.h:
            _check(lib().gpe_ensemble_destroy(self.h), "gpe_ensemble_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_user_params(self, params):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        _check(lib().gpe_ensemble_set_user_params(self.h, len(p), _ptr(p)), "gpe_ensemble_set_user_params")

    def set_quantum_friction_coeff(self, qf):
        _check(lib().gpe_ensemble_set_quantum_friction_coeff(self.h, qf), "gpe_ensemble_set_quantum_friction_coeff")

    def set_instance_params(self, inst, params):
        """User parameters of one lattice (read by the user hooks as its d_user_param)."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        _check(lib().gpe_ensemble_set_instance_params(self.h, inst, len(p), _ptr(p)), "gpe_ensemble_set_instance_params")

    def set_external_potential(self, inst, v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        if a.shape != self.shapes[inst]:
            raise ValueError("potential shape %s, lattice %d is %s" % (a.shape, inst, self.shapes[inst]))
        _check(lib().gpe_ensemble_set_external_potential(self.h, inst, _ptr(a)), "gpe_ensemble_set_external_potential")

    def set_psi(self, inst, psi):
        a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128))
        if a.shape != self.shapes[inst]:
            raise ValueError("psi shape %s, lattice %d is %s" % (a.shape, inst, self.shapes[inst]))
        _check(lib().gpe_ensemble_set_psi(self.h, inst, _ptr(a)), "gpe_ensemble_set_psi")

    def get_psi(self, inst):
        a = np.empty(self.shapes[inst], dtype=np.complex128)
        t = ctypes.c_double()
        _check(lib().gpe_ensemble_get_psi(self.h, inst, ctypes.byref(t), _ptr(a)), "gpe_ensemble_get_psi")
        return a, t.value

    def evolve(self, nt):
        _check(lib().gpe_ensemble_evolve(self.h, int(nt)), "gpe_ensemble_evolve")

    def set_graph(self, on):
        _check(lib().gpe_ensemble_set_graph(self.h, 1 if on else 0), "gpe_ensemble_set_graph")

    def autotune(self, steps=3):
        _check(lib().gpe_ensemble_autotune(self.h, int(steps)), "gpe_ensemble_autotune")

    def set_variant(self, group, variant):
        """Launcher variant of one group (bits of Engine.set_variant; 16: fused whole steps, 32: y-z plane plan)."""
        _check(lib().gpe_ensemble_set_variant(self.h, int(group), int(variant)), "gpe_ensemble_set_variant")

    def groups(self):
        """[(shape, count, variant)] per group."""
        n = ctypes.c_int()
        _check(lib().gpe_ensemble_groups(self.h, ctypes.byref(n)), "gpe_ensemble_groups")
        out = []
        for g in range(n.value):
            v = [ctypes.c_int() for _ in range(5)]
            _check(lib().gpe_ensemble_group_info(self.h, g, *[ctypes.byref(x) for x in v]), "gpe_ensemble_group_info")
            out.append(((v[0].value, v[1].value, v[2].value), v[3].value, v[4].value))
        return out

    def instance(self, inst):
        g, s = ctypes.c_int(), ctypes.c_int()
        _check(lib().gpe_ensemble_instance(self.h, inst, ctypes.byref(g), ctypes.byref(s)), "gpe_ensemble_instance")
        return g.value, s.value

    def mean_density(self, group):
        shape = self.groups()[group][0]
        out = np.empty(shape, dtype=np.float64)
        _check(lib().gpe_ensemble_mean_density(self.h, group, _ptr(out)), "gpe_ensemble_mean_density")
        return out
