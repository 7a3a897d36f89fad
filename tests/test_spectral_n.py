"""The spectral form of the nonlinear term that the HIP stage combines use.

The reference forms N in real space and transforms it (belastic.cpp:473-488,
Calc.cu:265-297, 428-444):

    nld = -lap(tmp1)/eps + s,   s = D tmp1/eps^3 + (3 phi^2 - D - 1) f/eps^2,
    N   = -lambda1/2 - lambda2 3/(2 sqrt 2) f - eta nld,        N^ = F(N).

The GPU path never takes lap(tmp1) back to real space: it transforms tmp1 and
s and forms

    N^ = -lambda1/2 delta_k0 - lambda2 3/(2 sqrt 2) F(f) - eta F(s) + (eta/eps) lap^(F(tmp1)),
    F(f) = -eps lap^(v^) + (F(tmp1) + kappa v^)/eps,

with lap^ the reference's spectral Laplacian (folded |k|^2, Nyquist kept).
These tests check that identity with the oracle's own operators, so a dropped
term, a wrong sign or a wrong factor in it fails here on the CPU."""
import numpy as np
import pytest

from conftest import grid, rel_l2


def spectral_n(o, d, V, A):
    q = o.params()
    eps, eta, kappa, D = q.eps, q.eta, q.kappa, q.D
    l1, l2 = q.M1 * (V - q.alpha), q.M2 * (A - q.beta)
    vh = o.r2c(d)
    tmp1 = d * (d * d - 1.0 - kappa)
    f = o.f(d)
    s = D * tmp1 / eps ** 3 + (3 * d * d - D - 1) * f / eps ** 2
    th, sh = o.r2c(tmp1), o.r2c(s)
    fh = -eps * o.laplace(vh) + (th + kappa * vh) / eps
    nh = -l2 * 3.0 / (2.0 * np.sqrt(2.0)) * fh - eta * sh + (eta / eps) * o.laplace(th)
    nh[0, 0, 0] -= 0.5 * l1
    return nh


def vesicle(n, seed):
    X, Y, Z = grid(n)
    rng = np.random.default_rng(seed)
    eps = 2.0 * 2 * np.pi / n
    r = np.sqrt((X / 1.3) ** 2 + (Y / 1.0) ** 2 + (Z / 0.8) ** 2)
    d = np.tanh((1.0 - r) / (np.sqrt(2.0) * eps / 1.2))
    return d + 1e-3 * rng.standard_normal(d.shape), eps


@pytest.mark.parametrize("n,eta,M", [(16, 1.0, 1e4), (32, 1.0, 1e3), (32, 0.7, 0.0)])
def test_spectral_n_matches_reference_dataflow(oracle_mod, n, eta, M):
    d, eps = vesicle(n, 7 + n)
    o = oracle_mod.Oracle(n, eps=eps, eta=eta, M1=M, M2=M, alpha=20.0, beta=30.0)
    ref, V, A = o.nld_term(d)
    got = spectral_n(o, d, V, A)
    assert rel_l2(got, ref) < 1e-12


def test_spectral_n_pins_each_term(oracle_mod):
    # a planted mistake in any term of the spectral form must break the match
    n = 16
    d, eps = vesicle(n, 3)
    o = oracle_mod.Oracle(n, eps=eps, eta=1.0, M1=1e4, M2=1e4, alpha=20.0, beta=30.0)
    ref, V, A = o.nld_term(d)
    q = o.params()
    vh, th = o.r2c(d), o.r2c(d * (d * d - 1.0 - q.kappa))
    good = spectral_n(o, d, V, A)
    assert rel_l2(good, ref) < 1e-12
    kappa_dropped = good + q.M2 * (A - q.beta) * 3.0 / (2.0 * np.sqrt(2.0)) * q.kappa * vh / q.eps
    lap_sign = good - 2.0 * (q.eta / q.eps) * o.laplace(th)
    no_lambda1 = good.copy()
    no_lambda1[0, 0, 0] += 0.5 * q.M1 * (V - q.alpha)
    for bad in (kappa_dropped, lap_sign, no_lambda1):
        assert rel_l2(bad, ref) > 1e-6
