"""GPU parity of the fused solver path (one evaluation of ETD4RK / ETD2RK,
and the run_static_fEIF / run_static / shape_adjust drivers) against the
oracle, which restates belastic.cpp's dataflow with the reference's own
F(F^-1(.)) round trips."""
import numpy as np
import pytest

from conftest import flush_print, rel, rel_l2

pytestmark = pytest.mark.gpu


def make(vb, nx, nz=None, kind=2, seed=20260002, **kw):
    nz = nx if nz is None else nz
    eps = kw.pop("eps", 2.0 * 2 * np.pi / nx)
    phi, _ = vb.shape(kind, seed, nx, nz, 1.0, eps)
    p = vb.params_default(nx=nx, nz=nz, eps=eps, **kw)
    return p, phi


def oracle_for(oracle_mod, p):
    return oracle_mod.Oracle(p.nx, p.nz, eps=p.eps, eta=p.eta, ga=p.ga, M1=p.M1, M2=p.M2,
                             alpha=p.alpha, beta=p.beta)


@pytest.mark.parametrize("nx,nz,scheme", [(32, 32, 4), (64, 64, 4), (64, 32, 4), (48, 48, 4),
                                           (32, 32, 2), (64, 64, 2)])
def test_single_evaluation(vb, gpu, oracle_mod, nx, nz, scheme):
    flush_print("evaluate", nx, nz, scheme)
    p, phi = make(vb, nx, nz, M1=1e4, M2=1e4, scheme=scheme)
    o0 = oracle_mod.Oracle(nx, nz, eps=p.eps)
    # targets 1% off the shape's own V and A, so the multipliers are active but bounded
    p.alpha, p.beta = 1.01 * o0.vol(phi), 0.99 * o0.surface(phi)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    for h in (1e-7, 5e-7):
        s.evaluate(h)
        phi = o.etd4rk(phi, h) if scheme == 4 else o.etd2rk(phi, h)
        assert np.abs(phi).max() < 2.0
        assert rel_l2(s.get_phi(), phi) < 1e-12, h


def test_query_matches_oracle(vb, gpu, oracle_mod):
    p, phi = make(vb, 64, M1=1e4, M2=2e4, alpha=20.0, beta=40.0, eta=1.5)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    E, V, A, Eb = s.query()
    Eo, Vo, Ao, Ebo = o.energy(phi)
    assert rel(V, Vo) < 1e-13 and rel(A, Ao) < 1e-13 and rel(Eb, Ebo) < 1e-12 and rel(E, Eo) < 1e-12
    assert rel_l2(s.get_phi(), phi) < 1e-15


def test_fixed_plain_100_steps(vb, gpu, oracle_mod):
    """ETDRK4, fixed dt, no penalties: 100 steps agree to <= 1e-10."""
    p, phi = make(vb, 32, M1=0.0, M2=0.0, fixed_time_step=1e-6, time_step=1e-6)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    r = s.step(100)
    assert r["steps"] == 100 and r["evaluations"] == 100 and r["status"] == 0
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-6, fixed_time_step=1e-6)
    o.fEIF_steps(d, st, 100)
    assert rel_l2(s.get_phi(), d) < 1e-10
    E, V, A, Eb = s.query()
    Eo, Vo, Ao, Ebo = o.energy(d)
    assert rel(E, Eo) < 1e-12 and rel(V, Vo) < 1e-12 and rel(A, Ao) < 1e-12


def test_fixed_penalty_drift(vb, gpu, oracle_mod):
    """Fixed dt with M1, M2 != 0: two evaluations per step and the alpha/beta drift
    (belastic.cpp:670-691)."""
    p, phi = make(vb, 32, M1=1e4, M2=1e4, fixed_time_step=1e-6, time_step=1e-6, alpha=-1.0, beta=-1.0)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    r = s.step(20)
    assert r["steps"] == 20 and r["evaluations"] == 40
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-6, fixed_time_step=1e-6)
    o.fEIF_steps(d, st, 20)
    assert rel_l2(s.get_phi(), d) < 1e-10
    q = o.params()
    sv = s.state()
    assert rel(sv["alpha"], q.alpha) < 1e-12 and rel(sv["beta"], q.beta) < 1e-12


def test_adaptive_trace(vb, gpu, oracle_mod):
    """Adaptive dt: doubling every 4 steps, rejection when E rises (belastic.cpp:639-669);
    the dt sequence and the rejection count must match exactly."""
    p, phi = make(vb, 32, kind=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-6)
    for k in range(6):
        r = s.step(5)
        o.fEIF_steps(d, st, 5)
        assert r["time_step"] == st.time_step, (k, r["time_step"], st.time_step)
        assert abs(r["time_current"] - st.time_current) <= 1e-15 * st.time_current
    assert rel_l2(s.get_phi(), d) < 1e-10
    E, V, A, Eb = s.query()
    assert rel(E, o.energy(d)[0]) < 1e-11


def test_adaptive_energy_monotone(vb, gpu):
    p, phi = make(vb, 64, kind=3, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-7)
    s = vb.Solver(p, phi)
    last = s.query()[0]
    for _ in range(10):
        r = s.step(2)
        E = s.query()[0]
        assert E <= last * (1 + 1e-14)
        last = E


def test_euler_static_steps(vb, gpu, oracle_mod):
    p, phi = make(vb, 32, scheme=vb.EULER, M1=1e3, M2=1e3, alpha=-1.0, beta=-1.0,
                  time_step=1e-7, fixed_time_step=1e-7)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    s.step(10)
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-7, fixed_time_step=1e-7)
    o.static_steps(d, st, 10)
    assert rel_l2(s.get_phi(), d) < 1e-11


def test_shape_adjust(vb, gpu, oracle_mod):
    p, phi = make(vb, 32, kind=1, shape_adjust=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0,
                  time_step=1e-6, fixed_time_step=1e-6)
    o = oracle_for(oracle_mod, p)
    s = vb.Solver(p, phi)
    s.step(1)   # the driver start runs the Euler pre-relaxation
    d = phi.copy()
    o.shape_adjust(d)
    st = o.fEIF_begin(d, 1e-6, fixed_time_step=1e-6)
    o.fEIF_steps(d, st, 1)
    assert rel_l2(s.get_phi(), d) < 1e-10


def test_checkpoint_resume(vb, gpu, tmp_path):
    p, phi = make(vb, 32, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, fixed_time_step=1e-6, time_step=1e-6)
    a = vb.Solver(p, phi)
    a.step(5)
    a.save(str(tmp_path))
    b = vb.Solver(p, phi)
    b.load(str(tmp_path))
    assert rel_l2(b.get_phi(), a.get_phi()) < 1e-15
    sa, sb = a.state(), b.state()
    assert sa["alpha"] == sb["alpha"] and sa["beta"] == sb["beta"]
    info = np.fromfile(tmp_path / "32x32x32.info", dtype=np.float64)
    assert info.shape == (7,) and info[1] == 1e-6
    out = np.fromfile(tmp_path / "d32x32x32.out", dtype=np.float64)
    assert out.size == 32 ** 3
