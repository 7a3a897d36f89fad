"""Pins the oracle's ETD combines (etd1/etd2, etd41..etd44, Calc.cu:446-577)
and steppers (belastic.cpp:491-552):
 - exactness for a constant nonlinear term (all ETD schemes integrate
   u' = L u + N exactly when N does not change);
 - the convergence order on the scalar test u' = L u + mu u: the reference's
   ETDRK4 uses N^(a) in the first bracket term of etd44 (Calc.cu:559,564)
   instead of N^(u), which makes it first order (Cox-Matthews would be fourth
   order); ETD2RK is second order;
 - the phi-functions against the 40-digit mpmath fixture."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel


def _P(a):
    return a.ctypes.data_as(C.c_void_p)


class SpecOps:
    """The oracle's spectral ETD kernels on a tiny grid whose L^ we control via eps."""

    def __init__(self, oracle_mod, eps=1.0):
        self.m = oracle_mod
        self.o = oracle_mod.Oracle(4, eps=eps)
        self.L = oracle_mod.lib()
        self.H = self.o.Het()

    def call(self, fn, out, *arrs, h):
        getattr(self.L, fn)(self.o.h, _P(out), *[_P(np.ascontiguousarray(a)) for a in arrs], _P(self.H), h)
        return out

    def etd4(self, u, mu, h, Nfun=None):
        N = (lambda x: mu * x) if Nfun is None else Nfun
        z = np.zeros_like(u)
        Nu = N(u)
        a = self.call("vbo_etd41", z.copy(), u, Nu, h=h)
        Na = N(a)
        b = self.call("vbo_etd42", z.copy(), u, Na, h=h)
        Nb = N(b)
        c = a.copy()
        self.L.vbo_etd43(self.o.h, _P(c), _P(Nu), _P(Nb), _P(self.H), h)
        Nc = N(c)
        out = z.copy()
        self.L.vbo_etd44(self.o.h, _P(out), _P(u), _P(Na), _P(Nb), _P(Nc), _P(self.H), h)
        return out

    def etd2(self, u, mu, h):
        Nu = mu * u
        a = self.call("vbo_etd1", np.zeros_like(u), u, Nu, h=h)
        Na = mu * a
        out = np.zeros_like(u)
        self.L.vbo_etd2(self.o.h, _P(out), _P(a), _P(Na), _P(Nu), _P(self.H), h)
        return out


def test_constant_forcing_is_exact(oracle_mod):
    s = SpecOps(oracle_mod)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(s.H.shape) + 1j * rng.standard_normal(s.H.shape)
    Nc = rng.standard_normal(s.H.shape) + 1j * rng.standard_normal(s.H.shape)
    for h in (0.05, 0.3):
        exact = np.exp(h * s.H) * u + np.expm1(h * s.H) / s.H * Nc
        got4 = s.etd4(u, 0.0, h, Nfun=lambda x: Nc)
        assert np.abs(got4 - exact).max() <= 1e-12 * np.abs(exact).max()
        a = s.call("vbo_etd1", np.zeros_like(u), u, Nc, h=h)
        got2 = np.zeros_like(u)
        s.L.vbo_etd2(s.o.h, _P(got2), _P(a), _P(Nc), _P(Nc), _P(s.H), h)
        assert np.abs(got2 - exact).max() <= 1e-12 * np.abs(exact).max()


def _order(step, u0, lam, T=1.0, ks=(3, 4, 5, 6, 7)):
    errs = []
    for k in ks:
        h = 2.0 ** -k
        u = u0.copy()
        for _ in range(int(round(T / h))):
            u = step(u, h)
        errs.append(np.abs(u - np.exp(lam * T) * u0).max())
    return [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


def test_reference_etdrk4_is_first_order(oracle_mod):
    s = SpecOps(oracle_mod)
    mu = -2.0
    u0 = np.zeros(s.H.shape, complex)
    u0[0, 0, 0] = 1.0          # single mode, L^ = -D kappa / eps^3 = -3 at eps = 1
    lam = s.H[0, 0, 0] + mu
    slopes = _order(lambda u, h: s.etd4(u, mu, h), u0, lam)
    assert all(0.85 < q < 1.3 for q in slopes[1:]), slopes


def test_etdrk2_is_second_order(oracle_mod):
    s = SpecOps(oracle_mod)
    mu = -2.0
    u0 = np.zeros(s.H.shape, complex)
    u0[0, 0, 0] = 1.0
    lam = s.H[0, 0, 0] + mu
    slopes = _order(lambda u, h: s.etd2(u, mu, h), u0, lam)
    assert all(1.8 < q < 2.3 for q in slopes[1:]), slopes


def test_oracle_coefficients_vs_mpmath(oracle_mod):
    tab = np.load(os.path.join(GOLDEN, "phi_mpmath.npz"))["table"]
    for z, e1, p1h, ez, f1, f2, f3, p2 in tab:
        H, h = -1.0, -z          # z = h H
        c = oracle_mod.etd_coeffs(H, h)
        assert rel(c[0], e1) < 1e-15 and rel(c[2], ez) < 1e-15
        s = 1.0 / (h * h * H * H * H)
        if abs(z) >= 0.1:       # the reference brackets are well conditioned here
            assert rel(s * c[3] / h, f1) < 1e-11
            assert rel(s * c[4] / h, f2) < 1e-11
            assert rel(s * c[5] / h, f3) < 1e-11
            assert rel(c[6] / h, p2) < 1e-12
