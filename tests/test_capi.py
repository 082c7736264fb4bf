"""CPU-side checks of the C ABI: the shared library loads, exports every
function include/gpe_engine.h declares, and rejects invalid input before any
device work (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpe_engine.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char\s*\*|gpe_result_t)\s*\*?\s*(gpe_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from libcugpe_amd import engine as E

    if not os.path.exists(E.LIB_PATH):
        from libcugpe_amd import build

        build.build()
    return E.lib()


def test_header_declares_reference_surface():
    names = declared_functions()
    # every call made by the reference's drivers (test_vortex/*.c, test_compare/*.c)
    for f in ["gpe_get_lattice", "gpe_create_engine", "gpe_set_user_params", "gpe_set_psi", "gpe_normalize_psi",
              "gpe_energy", "gpe_evolve_qf", "gpe_get_psi", "gpe_set_vortex", "gpe_evolve_vortex",
              "gpe_destroy_engine"]:
        assert f in names, f
    assert len(names) > 40


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "libcugpe_amd", "libgpe_hip.so")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_invalid_input_rejected_without_device(lib):
    from libcugpe_amd import engine as E

    assert lib.gpe_engine_destroy(None) == -10
    assert lib.gpe_engine_energy(None, None, None, None, None) == -10
    for shape in [(20, 16, 16), (16, 1, 16), (1024, 16, 16), (7, 8, 8)]:
        with pytest.raises(E.GpeError) as ex:
            E.Engine(shape, 1.0, 0.0, 0.01, 1.0)
        assert ex.value.code == -11
    with pytest.raises(E.GpeError) as ex:
        E.Engine((16, 16, 16), 1.0, 0.0, 0.01, 1.0, gpe_for=3)
    assert ex.value.code == -99
    with pytest.raises(E.GpeError):
        E.Engine((16, 16, 16), 1.0, 0.0, 0.01, 1.0, precision=32)
    assert E.error_string(-13).startswith("vortex")


def test_singleton_needs_a_lattice(lib):
    # gpe_create_engine without a configured lattice fails with the lattice code
    assert lib.gpe_destroy_engine() == -10
    nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.gpe_get_lattice(ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz))
    if (nx.value, ny.value, nz.value) == (0, 0, 0):
        f = lib.gpe_create_engine
        f.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int]
        assert f(1.0, 0.0, 0.01, 1.0, 1024) == -11
