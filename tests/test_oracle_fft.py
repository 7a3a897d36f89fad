"""Pins the oracle's 3-D transforms (restating Calc::r2c_3d_f / c2r_3d_b,
Calc.cu:88-104) against numpy/pocketfft golden vectors, a brute-force DFT,
round trips, Parseval and single-mode closed forms."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, grid, rel_l2


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "fft_golden.npz"))


@pytest.mark.parametrize("nx,nz", [(8, 8), (16, 16), (8, 12), (16, 6)])
def test_r2c_c2r_vs_pocketfft(oracle_mod, golden, nx, nz):
    o = oracle_mod.Oracle(nx, nz)
    key = "%dx%dx%d" % (nx, nx, nz)
    s = o.r2c(golden["u_" + key])
    assert rel_l2(s, golden["s_" + key]) < 1e-14
    u2 = o.c2r(golden["s2_" + key])
    assert rel_l2(u2, golden["u2_" + key]) < 1e-14


def test_brute_force_dft(oracle_mod):
    nx, nz = 4, 6
    rng = np.random.default_rng(7)
    u = rng.standard_normal((nz, nx, nx))
    o = oracle_mod.Oracle(nx, nz)
    s = o.r2c(u)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(nx), indexing="ij")
    for kz in range(nz):
        for ky in range(nx):
            for kx in range(nx // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * x / nx + ky * y / nx + kz * z / nz))
                ref = (u * ph).sum() / u.size
                assert abs(s[kz, ky, kx] - ref) < 1e-15 * 10


@pytest.mark.parametrize("n", [6, 10, 12, 18, 20, 24, 48])
def test_fft1d_mixed_radix(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(oracle_mod.fft1d(x, -1) - np.fft.fft(x)).max() < 1e-13 * n
    assert np.abs(oracle_mod.fft1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-13 * n


@pytest.mark.parametrize("nx,nz", [(16, 16), (32, 32), (16, 24)])
def test_round_trip_and_parseval(oracle_mod, nx, nz):
    rng = np.random.default_rng(nx + nz)
    u = rng.standard_normal((nz, nx, nx))
    o = oracle_mod.Oracle(nx, nz)
    s = o.r2c(u)
    assert np.abs(o.c2r(s) - u).max() < 1e-13
    # Parseval over the Hermitian-completed spectrum (weights 1 on kx = 0, nx/2)
    w = np.full(s.shape[-1], 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    assert abs((w * np.abs(s) ** 2).sum() - (u ** 2).mean()) < 1e-13


def test_single_mode(oracle_mod):
    n = 16
    X, Y, Z = grid(n)
    o = oracle_mod.Oracle(n)
    # cos(2x + 3y - z): amplitude 1/2 at (kx,ky,kz) = (2,3,-1) and at its mirror
    u = np.cos(2 * X + 3 * Y - Z)
    s = o.r2c(u)
    # x_i = i h - pi shifts phases by exp(-i k pi) = (-1)^k
    ph = (-1) ** (2 + 3 - 1)
    assert abs(s[n - 1, 3, 2] - 0.5 * ph) < 1e-15
    s[n - 1, 3, 2] = 0
    assert np.abs(s).max() < 1e-15
