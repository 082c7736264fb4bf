"""gpe_detect_vortices_noalloc: vortex detection that allocates nothing,
called through the C ABI on a wave function in device memory and checked bit
for bit against gpe_engine_detect_vortices on the same state.  Each pair of
calls runs a lattice and then a smaller one, so the second call reduces fewer
partial sums than the first left in the library's scratch array."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from libcugpe_amd import engine as E  # noqa: E402


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def vortex_state(shape, seed):
    """two crossing vortex lines (along z and along x) plus noise, which adds
    windings of both signs where the density is small"""
    nx, ny, nz = shape
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    a = ((X - 0.3 * nx) + 1j * (Y - 0.6 * ny)) / nx
    b = ((Y - 0.25 * ny) + 1j * (Z - 0.55 * nz)) / ny
    rng = np.random.default_rng(seed)
    noise = 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return a * b + noise


def engine_with(shape, prec, seed):
    g = E.Engine(shape, 1.0, 0.0, 0.01, float(np.prod(shape)), E.CONTACT, E.HARMONIC, 0, prec)
    g.set_psi(vortex_state(shape, seed))
    return g


def noalloc(g, prec, rho_min, with_count=True):
    nx, ny, nz = g.shape
    n = nx * ny * nz
    w = torch.full((3 * n,), 0x55, dtype=torch.int8, device="cuda")
    c = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.synchronize()
    r = E.lib().gpe_detect_vortices_noalloc(ctypes.c_void_p(g.device_psi()), nx, ny, nz, prec, rho_min,
                                            ctypes.c_void_p(w.data_ptr()),
                                            ctypes.c_void_p(c.data_ptr()) if with_count else None,
                                            ctypes.c_void_p(g.stream()))
    assert r == 0, E.lib().gpe_get_error_string(r)
    g.synchronize()
    return w.cpu().numpy().reshape((3,) + g.shape), float(c.cpu().numpy()[0])


@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_noalloc_matches_engine_detection(prec):
    for shape, rho_min, seed in (((64, 48, 32), 1e-4, 1), ((32, 16, 24), 0.0, 2)):
        say("noalloc vortex detection", shape, prec, rho_min)
        g = engine_with(shape, prec, seed)
        want_w, want_c = g.detect_vortices(rho_min)
        got_w, got_c = noalloc(g, prec, rho_min)
        assert want_c > 0
        assert np.array_equal(got_w, want_w)
        assert got_c == float(want_c)
        # without a count buffer the windings are the same
        got_w2, _ = noalloc(g, prec, rho_min, with_count=False)
        assert np.array_equal(got_w2, want_w)
        g.close()


def test_noalloc_rejects_bad_arguments():
    g = engine_with((16, 16, 16), E.C128, 3)
    p = ctypes.c_void_p(g.device_psi())
    lib = E.lib()
    # checked on the host before anything is launched
    assert lib.gpe_detect_vortices_noalloc(p, 2048, 2048, 2048, E.C128, 0.0, p, None, None) == -16
    assert b"too large" in lib.gpe_get_error_string(-16)
    assert lib.gpe_detect_vortices_noalloc(p, 16, 16, 16, 32, 0.0, p, None, None) == -12
    assert lib.gpe_detect_vortices_noalloc(p, 0, 16, 16, E.C128, 0.0, p, None, None) == -12
    assert lib.gpe_detect_vortices_noalloc(None, 16, 16, 16, E.C128, 0.0, p, None, None) == -12
    g.close()
