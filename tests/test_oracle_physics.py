"""Pins the oracle's spectral operators, pointwise terms and integrals
(Calc.cu:107-437, phasefield.cpp:140-325, belastic.cpp:68-86) with closed
forms that a dropped term, wrong sign or swapped index would break."""
import numpy as np
import pytest

from conftest import grid, rel, rel_l2


def test_laplace_trig_polynomial(oracle_mod):
    n = 16
    X, Y, Z = grid(n)
    o = oracle_mod.Oracle(n)
    u = np.sin(2 * X) * np.cos(3 * Y) * np.sin(Z) + 0.5 * np.cos(5 * Z)
    lap = o.lap_real(u)
    ref = -(4 + 9 + 1) * np.sin(2 * X) * np.cos(3 * Y) * np.sin(Z) - 0.5 * 25 * np.cos(5 * Z)
    assert np.abs(lap - ref).max() < 1e-12


def test_nyquist_rules(oracle_mod):
    # laplace keeps the Nyquist mode (Calc.cu:117-120), the gradient zeroes it (Calc.cu:311-332)
    n = 8
    X, Y, Z = grid(n)
    o = oracle_mod.Oracle(n)
    for ax, A in enumerate((X, Y, Z)):
        u = np.cos(n / 2 * A)
        assert np.abs(o.lap_real(u) + (n / 2) ** 2 * u).max() < 1e-12
        d = o.fftgradient(o.r2c(u))
        assert np.abs(d[ax]).max() < 1e-13


def test_gradient_components(oracle_mod):
    n = 16
    X, Y, Z = grid(n)
    o = oracle_mod.Oracle(n)
    u = np.sin(X + 2 * Y) + np.cos(3 * Z - Y)
    dx, dy, dz = o.fftgradient(o.r2c(u))
    assert np.abs(dx - np.cos(X + 2 * Y)).max() < 1e-12
    assert np.abs(dy - (2 * np.cos(X + 2 * Y) + np.sin(3 * Z - Y))).max() < 1e-12
    assert np.abs(dz - (-3 * np.sin(3 * Z - Y))).max() < 1e-12


def slab(n, R, eps):
    X, Y, Z = grid(n)
    return np.tanh((R - np.abs(X)) / (np.sqrt(2) * eps))


def test_planar_interface_identities(oracle_mod):
    # phi = tanh(x / (sqrt2 eps)) solves eps phi'' - (phi^3 - phi)/eps = 0, so f -> 0;
    # the area functional counts interface area (factor 3/(2 sqrt2), phasefield.cpp:313)
    n = 128
    h = 2 * np.pi / n
    eps = 2 * h
    o = oracle_mod.Oracle(n, 8, eps=eps)
    X, Y, Z = grid(n, 8)
    R = 1.5
    d = np.tanh((R - np.abs(X)) / (np.sqrt(2) * eps))
    f = o.f(d)
    assert np.abs(f).max() < 2e-3      # spectral error of the profile only
    A = o.surface(d)
    assert rel(A, 2 * (2 * np.pi) ** 2) < 1e-6
    V = o.vol(d)
    assert rel(V, 2 * R * (2 * np.pi) ** 2) < 1e-6
    assert o.elastic_eng(d) < 1e-5    # vs ~23.7 for a sphere


def test_sphere_bending_energy(oracle_mod):
    # E_b of a sphere -> 16 sqrt2 pi / 3, independent of R (eps/R -> 0)
    n = 64
    h = 2 * np.pi / n
    eps = 2 * h
    o = oracle_mod.Oracle(n, eps=eps)
    X, Y, Z = grid(n)
    R = 1.5
    d = np.tanh((R - np.sqrt(X ** 2 + Y ** 2 + Z ** 2)) / (np.sqrt(2) * eps))
    E, V, A, Eb = o.energy(d, constrained=False)
    assert rel(Eb, 16 * np.sqrt(2) * np.pi / 3) < 1e-4
    assert rel(A, 4 * np.pi * R ** 2) < 2e-2
    # (d+1)/2 is a logistic step of width w = sqrt2 eps / 2 -> V = 4pi/3 (R^3 + pi^2 w^2 R)
    w = np.sqrt(2) * eps / 2
    assert rel(V, 4 * np.pi / 3 * (R ** 3 + np.pi ** 2 * w ** 2 * R)) < 1e-3
    assert E == pytest.approx(Eb * 1.0)


def test_energy_penalty_terms(oracle_mod):
    n = 16
    o = oracle_mod.Oracle(n, M1=3.0, M2=5.0, eta=2.0, alpha=1.0, beta=2.0)
    rng = np.random.default_rng(3)
    d = np.tanh(rng.standard_normal((n, n, n)))
    E, V, A, Eb = o.energy(d)
    assert E == pytest.approx(2.0 * Eb + 0.5 * 3.0 * (V - 1.0) ** 2 + 0.5 * 5.0 * (A - 2.0) ** 2, rel=1e-14)


def test_intf_against_fsum(oracle_mod):
    import math
    n = 16
    o = oracle_mod.Oracle(n)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((n, n, n))
    L3 = (2 * np.pi) ** 3
    ref = (math.fsum(u.ravel()) / u.size + 0.75) * 1.25 * L3
    assert rel(o.intf(u, 0.75, 1.25), ref) < 1e-14


def test_nld_term_linear_split(oracle_mod):
    # N = phi_t - L^ phi for eta = 1 (derived from Calc.cu:278-282 and 600):
    # phi_t = -(g0 + lambda1/2 + lambda2 * 3/(2 sqrt2) f)
    n = 16
    o = oracle_mod.Oracle(n, M1=10.0, M2=20.0, alpha=5.0, beta=9.0)
    X, Y, Z = grid(n)
    d = np.tanh((1.8 - np.sqrt(X ** 2 + 1.3 * Y ** 2 + Z ** 2)) / (np.sqrt(2) * o.p.eps))
    nhat, V, A = o.nld_term(d)
    l1, l2 = 10.0 * (V - 5.0), 20.0 * (A - 9.0)
    g0 = o.g0(d)
    f = o.f(d)
    phit = -(g0 + 0.5 * l1 + l2 * 3 / (2 * np.sqrt(2)) * f)
    Het = o.Het()
    lin = o.c2r(Het * o.r2c(d))
    N = o.c2r(nhat)
    assert rel_l2(N, phit - lin) < 1e-11


def test_lhat_closed_form_and_sign(oracle_mod):
    n, nz = 16, 8
    o = oracle_mod.Oracle(n, nz)
    Het = o.Het()
    eps, kappa, D = o.p.eps, 2.0, 1.5
    kz = np.minimum(np.arange(nz), nz - np.arange(nz))[:, None, None]
    ky = np.minimum(np.arange(n), n - np.arange(n))[None, :, None]
    kx = np.arange(n // 2 + 1)[None, None, :]
    m = (kx ** 2 + ky ** 2 + kz ** 2).astype(float)
    ref = -eps * m ** 2 - kappa * m / eps - D * (m + kappa / eps ** 2) / eps
    assert np.abs(Het - ref).max() <= 1e-12 * np.abs(ref).max()
    assert (Het < 0).all()
