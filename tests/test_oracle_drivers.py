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
