"""Pins the oracle's own FFT (it replaces cuFFT, orginal/gpe_engine.cu:519-527,
963-1000) against the DFT definition, numpy golden vectors and identities."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def brute_dft3(x, sign):
    out = x.astype(np.complex128)
    for ax in range(3):
        n = x.shape[ax]
        j = np.arange(n)
        M = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
        out = np.moveaxis(np.tensordot(M, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 8), (24, 3, 4), (2, 5, 3), (10, 2, 7)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft3_matches_brute_force_dft(oracle, shape, sign):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = oracle.fft3d(x, sign)
    ref = brute_dft3(x, sign)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_fft3_matches_numpy_golden(oracle):
    g = np.load(os.path.join(HERE, "golden", "fft_golden.npz"))
    keys = [k[3:] for k in g.files if k.startswith("in_")]
    assert len(keys) >= 4
    for key in keys:
        x = g["in_" + key]
        for sign, name in ((-1, "fwd_"), (1, "inv_")):
            got = oracle.fft3d(x, sign)
            ref = g[name + key]
            err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
            assert err < 1e-14, (key, sign, err)


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024, 7, 11, 100])
def test_fft1d_lengths_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = oracle.fft1d(x, -1)
    assert np.max(np.abs(f - np.fft.fft(x))) <= 2e-15 * np.sqrt(n) * np.max(np.abs(f)) * np.log2(max(n, 2))
    b = oracle.fft1d(x, +1)
    assert np.max(np.abs(b - np.fft.ifft(x) * n)) <= 2e-15 * np.sqrt(n) * np.max(np.abs(b)) * np.log2(max(n, 2))


def test_round_trip_and_parseval(oracle):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((32, 24, 16)) + 1j * rng.standard_normal((32, 24, 16))
    f = oracle.fft3d(x, -1)
    # Parseval for the unnormalised DFT: sum |F|^2 = N sum |x|^2
    assert abs(np.sum(np.abs(f) ** 2) / (x.size * np.sum(np.abs(x) ** 2)) - 1) < 1e-14
    back = oracle.fft3d(f, +1)
    assert np.max(np.abs(back / x.size - x)) < 1e-14


def test_unit_roots_exact_octants(oracle):
    # exp(-2 pi i m/n) against values fixed by symmetry and numpy at large n
    assert oracle.unit_root(0, 8) == 1
    assert oracle.unit_root(2, 8) == complex(0, -1)
    assert oracle.unit_root(4, 8) == -1
    assert oracle.unit_root(6, 8) == complex(0, 1)
    for n in (3, 12, 256, 1 << 20, 3 * (1 << 18)):
        m = np.arange(0, n, max(1, n // 997))
        got = np.array([oracle.unit_root(int(k), n) for k in m])
        # reference in x87 extended precision: pi and the angle carry ~11 more bits
        pi_l = np.arccos(np.longdouble(-1))
        ang = 2 * pi_l * m.astype(np.longdouble) / np.longdouble(n)
        err_re = np.abs(got.real.astype(np.longdouble) - np.cos(ang))
        err_im = np.abs(got.imag.astype(np.longdouble) + np.sin(ang))
        assert float(max(err_re.max(), err_im.max())) < 2.3e-16
