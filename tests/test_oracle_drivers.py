"""Pins the oracle's time-stepping drivers (belastic.cpp:174-283, 286-469,
588-746) through properties their control flow fixes: a rejected attempt
leaves no trace in the state, accepted steps never raise the constrained
energy, the penalty drift holds V and A at their targets, and the Euler and
ETD paths agree to first order in h."""
import numpy as np
import pytest

from conftest import grid, rel


def ellipsoid(n, eps):
    X, Y, Z = grid(n)
    r = np.sqrt((X / 1.4) ** 2 + Y ** 2 + (Z / 0.9) ** 2)
    return np.tanh((1.6 - r) / (np.sqrt(2) * eps))


N = 16
EPS = 2 * 2 * np.pi / N


def test_adaptive_rejection_restores_state_exactly(oracle_mod):
    o = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    d0 = ellipsoid(N, EPS)
    d = d0.copy()
    s = o.fEIF_begin(d, 1e-5)
    done, tr = o.fEIF_steps(d, s, 1, trace=True)
    assert done == 1 and s.rejections >= 1 and s.evaluations == s.rejections + 1
    # first attempt at 2e-5 (doubling at step 0), halved once per rejection
    assert tr[0, 0] == 2e-5 / 2 ** s.rejections
    ref = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    ref.set_targets(o.params().alpha, o.params().beta)
    assert np.array_equal(d, ref.etd4rk(d0, tr[0, 0]))
    assert s.time_current == tr[0, 0]


def test_adaptive_energy_non_increasing_and_dt_rule(oracle_mod):
    o = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    d = ellipsoid(N, EPS)
    s = o.fEIF_begin(d, 1e-6)
    E0 = o.energy(d)[0]
    done, tr = o.fEIF_steps(d, s, 12, trace=True)
    assert done == 12
    E = np.concatenate([[E0], tr[:, 1]])
    assert (np.diff(E) <= 0).all()
    # without rejections dt doubles at steps 0, 4, 8 and stays fixed between
    if s.rejections == 0:
        assert list(tr[:, 0]) == [2e-6] * 4 + [4e-6] * 4 + [8e-6] * 4
    assert rel(s.time_current, tr[:, 0].sum()) < 1e-15


def test_penalty_drift_holds_targets(oracle_mod):
    d0 = ellipsoid(N, EPS)
    h = 2e-6
    drift = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    d1 = d0.copy()
    s = drift.fEIF_begin(d1, h, fixed_time_step=h)
    V0, A0 = s.alpha0, s.beta0
    drift.fEIF_steps(d1, s, 20)
    assert s.evaluations == 40
    # targets moved by the accumulated drift, each step twice (belastic.cpp:676-690)
    assert drift.params().alpha != V0 and drift.params().beta != A0
    plain = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    plain.set_targets(V0, A0)
    d2 = d0.copy()
    for _ in range(20):
        d2 = plain.etd4rk(d2, h)
    ev = lambda o, d: (abs(o.vol(d) - V0) / V0, abs(o.surface(d) - A0) / A0)
    dv1, da1 = ev(drift, d1)
    dv2, da2 = ev(plain, d2)
    assert da1 < da2 and dv1 <= max(dv2, 1e-9)


def test_shape_adjust_lowers_unconstrained_energy(oracle_mod):
    o = oracle_mod.Oracle(N, eps=EPS)
    d = ellipsoid(N, EPS)
    rng = np.random.default_rng(5)
    d += 0.05 * rng.standard_normal(d.shape)
    E0 = o.energy(d, constrained=False)[0]
    it = o.shape_adjust(d, max_iter=200)
    assert it > 0
    assert o.energy(d, constrained=False)[0] < E0
    assert np.isfinite(d).all()


def test_euler_adaptive_energy_non_increasing(oracle_mod):
    # run_static (belastic.cpp:367-437): dt grows by 1.4 every 4 steps, halves on a rise
    o = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    d = ellipsoid(N, EPS)
    o.set_targets(1.01 * o.vol(d), 0.99 * o.surface(d))
    s = oracle_mod.Drive()
    s.time_step = 1e-6
    Es = [o.energy(d)[0]]
    for _ in range(10):
        assert o.static_steps(d, s, 1) == 1
        Es.append(o.energy(d)[0])
    assert (np.diff(Es) <= 0).all()
    assert s.evaluations == 10 + s.rejections


def test_reference_etdrk4_small_z_cancellation(oracle_mod):
    """The reference's etd44 brackets (Calc.cu:558-567) are O(z^3) differences of O(1)
    terms divided by h^2 L^3; at |z| = |h L^| ~ 1e-7 the rounding error swamps them.
    The oracle reproduces this (one step moves phi by O(1) where the true change,
    the Euler step, is ~1e-4); the product's accurate tables avoid it."""
    o = oracle_mod.Oracle(N, eps=EPS, M1=1e4, M2=1e4)
    d0 = ellipsoid(N, EPS)
    o.set_targets(1.01 * o.vol(d0), 0.99 * o.surface(d0))
    h = 1e-8
    s = oracle_mod.Drive()
    s.time_step = h
    s.fixed_time_step = h
    de = d0.copy()
    o.static_steps(de, s, 1)
    dE = o.etd4rk(d0, h)
    assert np.abs(de - d0).max() < 1e-3
    assert np.abs(dE - d0).max() > 1.0
    # at h = 1e-6 (|z| >= 6e-6) the same step is within 15% of the Euler step scaled by 100
    step = np.abs(o.etd4rk(d0, 1e-6) - d0).max()
    assert abs(step / (100 * np.abs(de - d0).max()) - 1) < 0.15


def _fd_gradient(o, d, psi, h=1e-5):
    """directional derivative of the oracle's constrained energy along psi"""
    Ep = o.energy(d + h * psi)[0]
    Em = o.energy(d - h * psi)[0]
    return (Ep - Em) / (2 * h)


@pytest.mark.parametrize("M1,M2", [(1e3, 0.0), (0.0, 2e3), (1e3, 2e3)])
def test_euler_rhs_against_the_energy_gradient(oracle_mod, M1, M2):
    """vbo_nld1 through the Euler right-hand side r = lambda1/2 + lambda2 f + eta g0
    (belastic.cpp:52-64, Calc.cu:410-426), pinned per coefficient by the
    functional derivative of E = eta E_b + M1/2 (V - alpha)^2 + M2/2 (A - beta)^2
    taken by finite differences of the oracle's own energy: dE_b/dphi = g0 exactly
    (the spectral Laplacian is symmetric), dV/dphi = 1/2 and dA/dphi = 3/(2 sqrt2) f
    (along a smooth direction, whose Nyquist modes vanish).  So
    <r, psi> = dE(psi) - lambda2 (3/(2 sqrt2) - 1) <f, psi>: the reference's Euler
    step weights the area term by lambda2, not by its gradient's 3/(2 sqrt2) lambda2.
    eta = 1.7, so a dropped eta, a lost 1/2 on lambda1 or a 3/(2 sqrt2) on lambda2
    each break the balance."""
    n = 16
    eta = 1.7
    d = ellipsoid(n, EPS)
    o = oracle_mod.Oracle(n, eps=EPS, eta=eta, M1=M1, M2=M2)
    V0, A0 = o.vol(d), o.surface(d)
    o.set_targets(0.97 * V0, 1.04 * A0)
    X, Y, Z = grid(n)
    psi = np.cos(X) * np.sin(2 * Y) + 0.5 * np.cos(Z - 0.3) + 0.25
    L3 = (2 * np.pi) ** 3
    w = L3 / d.size                        # the discrete integral: sum * L^3 / N
    r = o.sum_fgq(d)
    f = o.f(d)
    lam2 = M2 * (A0 - 1.04 * A0)
    dE = _fd_gradient(o, d, psi)
    lhs = w * np.sum(r * psi)
    rhs = dE - lam2 * (3 / (2 * np.sqrt(2)) - 1) * w * np.sum(f * psi)
    assert rel(lhs, rhs) < 1e-6, (lhs, rhs)
    # the terms one at a time, for the record of their sizes
    g0 = o.g0(d)
    parts = [w * np.sum(eta * g0 * psi), M1 * (V0 - 0.97 * V0) / 2 * w * np.sum(psi), lam2 * w * np.sum(f * psi)]
    assert rel(lhs, sum(parts)) < 1e-12
    # every active term is at least 1e-4 of the total, so a dropped eta (41% of its
    # term), 1/2 (50%) or an extra 3/(2 sqrt2) (6%) moves the balance by at least
    # 6e-6, six times the tolerance (measured: 5e-5 to 1, against at most 2e-8)
    for k, q in enumerate(parts):
        if abs(q) > 0:
            assert abs(q) > 1e-4 * abs(lhs), (k, parts)


def test_nld1_coefficients_one_at_a_time(oracle_mod):
    """vbo_nld1 on constant fields: f = g = 0 leaves lambda1/2; f = 1, g = 0 adds
    lambda2 (no 3/(2 sqrt2)); f = 0, g = 1 adds eta (Calc.cu:417, eta = 2.5)."""
    n = 8
    o = oracle_mod.Oracle(n, eps=EPS, eta=2.5)
    z, one = np.zeros((n, n, n)), np.ones((n, n, n))
    assert np.all(o.nld1(z, z, 3.0, 7.0) == 1.5)
    assert np.all(o.nld1(one, z, 3.0, 7.0) == 1.5 + 7.0)
    assert np.all(o.nld1(z, one, 3.0, 7.0) == 1.5 + 2.5)
