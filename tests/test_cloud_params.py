"""include/gpe_cloud_params.h (the host-array cloud parameters of
dev/gpe_cloud_params.h:60-119) against closed forms on numpy arrays: the
density at the centre, the largest density, and the first distance from the
centre along each axis where the density falls below epsilon (NaN when it
never does), for gamma 1 (INTERACTIONS -1) and gamma 2 (INTERACTIONS 0)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(tmp, inter):
    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    out = str(tmp / ("cloud_check_%d" % (inter + 10)))
    subprocess.check_call([cc, "-std=c99", "-O1", "-Wall", "-Werror", "-DINTERACTIONS=%d" % inter,
                           "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "drivers", "cloud_check.c"), "-o", out, "-lm"])
    return out


def expected(psi, gamma, eps):
    nx, ny, nz = psi.shape
    rho = gamma * np.abs(psi) ** 2
    out = [rho[nx // 2, ny // 2, nz // 2], rho.max()]
    for line, L in ((rho[:, ny // 2, nz // 2], nx), (rho[nx // 2, :, nz // 2], ny), (rho[nx // 2, ny // 2, :], nz)):
        r = math.nan
        for i in range(L // 2, 0, -1):
            if line[i] < eps:
                r = float(L // 2 - i)
                break
        out.append(r)
    return out


def gaussian(shape, widths, center_shift=(0, 0, 0)):
    ix, iy, iz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    r2 = sum(((g - n // 2 - s) / w) ** 2 for g, n, w, s in zip((ix, iy, iz), shape, widths, center_shift))
    return np.exp(-r2 / 2).astype(np.complex128) * np.exp(0.3j * ix)


@pytest.mark.parametrize("inter,gamma", [(-1, 1.0), (0, 2.0)])
def test_cloud_params_header(tmp_path, inter, gamma):
    exe = build(tmp_path, inter)
    cases = [
        (gaussian((24, 16, 20), (3.0, 2.0, 4.0)), 1e-8),
        (gaussian((17, 9, 12), (2.0, 1.5, 1.0), (1, 0, 2)), 1e-3),
        (gaussian((8, 8, 8), (50.0, 50.0, 50.0)), 1e-8),  # never below epsilon: NaN on every axis
    ]
    for psi, eps in cases:
        f = tmp_path / "psi.bin"
        np.ascontiguousarray(psi).tofile(str(f))
        nx, ny, nz = psi.shape
        r = subprocess.run([exe, str(f), str(nx), str(ny), str(nz), repr(eps)], stdout=subprocess.PIPE, text=True,
                           timeout=60)
        assert r.returncode == 0
        want = expected(psi, gamma, eps)
        for line in r.stdout.splitlines():
            got = [float(x) for x in line.split()[1:]]
            for g, w in zip(got, want):
                if math.isnan(w):
                    assert math.isnan(g), (line, want)
                else:
                    assert g == pytest.approx(w, rel=1e-15, abs=0.0), (line, want)
