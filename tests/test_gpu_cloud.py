"""gpe_engine_cloud_params (the cloud parameters of dev/gpe_cloud_params.h:60-119
computed on the device) against the same quantities from the host copy of psi:
density at the centre, largest density (FP64, to 1e-15 relative) and the
first distance from the centre below epsilon along each axis (exact, NaN when
none)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from libcugpe_amd import engine as E  # noqa: E402


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def expected(psi, gamma, eps):
    nx, ny, nz = psi.shape
    rho = gamma * (psi.real ** 2 + psi.imag ** 2)
    out = [rho[nx // 2, ny // 2, nz // 2], rho.max()]
    for line, L in ((rho[:, ny // 2, nz // 2], nx), (rho[nx // 2, :, nz // 2], ny), (rho[nx // 2, ny // 2, :], nz)):
        r = math.nan
        for i in range(L // 2, 0, -1):
            if line[i] < eps:
                r = float(L // 2 - i)
                break
        out.append(r)
    return out


@pytest.mark.parametrize("prec", [E.C128, E.C64])
@pytest.mark.parametrize("shape,inter", [((32, 24, 48), E.CONTACT), ((64, 64, 64), E.UFG), ((256, 32, 32), E.UFG)])
def test_cloud_params_match_host_copy(prec, shape, inter):
    say("cloud params", prec, shape)
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    r2 = sum(((g - n // 2 + 0.3) / (n / 8)) ** 2 for g, n in zip((ix, iy, iz), shape))
    psi = (np.exp(-r2) * np.exp(0.2j * iy)).astype(np.complex128)
    g = E.Engine(shape, 1.0, 0.0, 0.01, 1000.0, inter, E.HARMONIC, 0, prec)
    g.set_user_params([0.1, 0.1, 0.1, 0.01])
    g.set_psi(psi)
    g.evolve_qf(3)
    host, _ = g.get_psi()
    gamma = 1.0 if inter == E.CONTACT else 2.0
    for eps in (1e-8, 1e-3, 1e-40):
        got = g.cloud_params(eps)
        want = expected(host, gamma, eps)
        assert got[0] == pytest.approx(want[0], rel=1e-15)
        assert got[1] == pytest.approx(want[1], rel=1e-15)
        for a, b in zip(got[2:], want[2:]):
            assert (math.isnan(a) and math.isnan(b)) or a == b, (eps, got, want)
    g.close()
