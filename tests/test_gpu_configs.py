"""GPU parity at the sizes and settings of the benchmark configurations.

 - the CPU-runnable case: 64^3 FP64 ETDRK2, seeded ellipsoid, fixed dt = 1e-7,
   M1 = M2 = 1e4, alpha/beta from t = 0, 200 driver steps;
 - the drop-in case: 128^3 FP64 ETDRK4, two seeded ellipsoids, adaptive dt from
   1e-7 doubling every 10 steps (a shortened run);
 - one ETDRK4 evaluation at every large supported size;
 - the precision case: 256^3 FP32-mixed against FP64 on the same GPU.
The oracle restates the reference's dataflow (belastic.cpp:491-552)."""
import numpy as np
import pytest

from conftest import flush_print, rel, rel_l2

pytestmark = pytest.mark.gpu


def oracle_for(oracle_mod, p):
    return oracle_mod.Oracle(p.nx, p.nz, eps=p.eps, eta=p.eta, ga=p.ga, M1=p.M1, M2=p.M2,
                             alpha=p.alpha, beta=p.beta)


def test_cpu_case_64_etdrk2_200_steps(vb, gpu, oracle_mod):
    n = 64
    eps = 2.0 * 2 * np.pi / n
    phi, _ = vb.shape(vb.SHAPE_ELLIPSOID, 20260001, n, n, 1.0, eps)
    p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=0.0, beta=0.0, scheme=vb.ETDRK2,
                          time_step=1e-7, fixed_time_step=1e-7)
    flush_print("64^3 ETDRK2, 200 steps")
    s = vb.Solver(p, phi)
    r = s.step(200)
    assert r["status"] == 0 and r["steps"] == 200 and r["evaluations"] == 400
    o = oracle_for(oracle_mod, p)
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-7, fixed_time_step=1e-7)
    assert o.fEIF_steps(d, st, 200, scheme=2) == 200
    assert rel_l2(s.get_phi(), d) < 1e-10
    E, V, A, Eb = s.query()
    Eo, Vo, Ao, Ebo = o.energy(d)
    assert rel(V, Vo) < 1e-12 and rel(A, Ao) < 1e-12 and rel(E, Eo) < 1e-10
    sv, q = s.state(), o.params()
    assert rel(sv["alpha"], q.alpha) < 1e-12 and rel(sv["beta"], q.beta) < 1e-12


def test_dropin_case_128_adaptive(vb, gpu, oracle_mod):
    n = 128
    eps = 2.0 * 2 * np.pi / n
    phi, _ = vb.shape(vb.SHAPE_TWO_ELLIPSOIDS, 20260002, n, n, 1.0, eps)
    p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=0.0, beta=0.0, scheme=vb.ETDRK4,
                          time_step=1e-7, fixed_time_step=0.0, dt_double_every=10)
    flush_print("128^3 ETDRK4 adaptive")
    s = vb.Solver(p, phi)
    o = oracle_for(oracle_mod, p)
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-7, dt_double_every=10)
    for k in range(3):
        r = s.step(8)
        o.fEIF_steps(d, st, 8)
        assert r["time_step"] == st.time_step, k
        assert r["rejections"] >= 0 and abs(r["time_current"] - st.time_current) <= 1e-15 * st.time_current
    assert rel_l2(s.get_phi(), d) < 1e-10
    assert rel(s.query()[0], o.energy(d)[0]) < 1e-11


@pytest.mark.parametrize("n", [128, 192, 256, 384, 512])
def test_one_evaluation_large(vb, gpu, oracle_mod, n):
    eps = 2.0 * 2 * np.pi / n
    phi, _ = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps, noise=1e-3)
    o0 = oracle_mod.Oracle(n, eps=eps)
    a, b = 1.01 * o0.vol(phi), 0.99 * o0.surface(phi)
    del o0
    p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=a, beta=b, scheme=vb.ETDRK4)
    flush_print("one evaluation", n)
    s = vb.Solver(p, phi)
    s.evaluate(5e-7)
    got = s.get_phi()
    s.close()
    ref = oracle_for(oracle_mod, p).etd4rk(phi, 5e-7)
    assert rel_l2(got, ref) < 1e-12


def test_precision_case_256(vb, gpu):
    n = 256
    eps = 2.0 * 2 * np.pi / n
    phi, _ = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps, noise=1e-3)
    out = {}
    for prec in (vb.FP64, vb.FP32_MIXED):
        p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=0.0, beta=0.0, precision=prec,
                              phi_lut_mode=vb.LUT_ACCURATE, time_step=5e-7, fixed_time_step=5e-7)
        s = vb.Solver(p, phi)
        r = s.step(10)
        assert r["status"] == 0
        out[prec] = (s.get_phi(), s.query(), s.state())
        s.close()
    (p64, q64, s64), (p32, q32, s32) = out[vb.FP64], out[vb.FP32_MIXED]
    assert rel_l2(p32, p64) < 1e-5            # phi tolerance of the precision form
    assert rel(q32[1], q64[1]) < 1e-7         # V
    assert rel(q32[2], q64[2]) < 1e-7         # A
    # E_b weights the high modes, where the FP32 transforms' rounding lives, by
    # about eps k^2 (up to ~800 here): over the first 100 steps of this case its
    # difference ranges from 5e-8 to 2e-5 while phi stays within 3e-7; E also
    # carries the drifted targets of each run
    assert rel(q32[3], q64[3]) < 3e-5 and rel(q32[0], q64[0]) < 1e-3


@pytest.mark.parametrize("n", [384, 512])
def test_fp32_mixed_large(vb, gpu, n):
    # the FP32-mixed pipeline at the large sizes against the FP64 one on the same
    # field: three penalty-drift steps (six evaluations, the probe stage included)
    eps = 2.0 * 2 * np.pi / n
    phi, _ = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps, noise=1e-3)
    out = {}
    for prec in (vb.FP64, vb.FP32_MIXED):
        flush_print("fp32-mixed large", n, prec)
        p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=0.0, beta=0.0, precision=prec,
                              phi_lut_mode=vb.LUT_ACCURATE, time_step=5e-7, fixed_time_step=5e-7)
        s = vb.Solver(p, phi)
        r = s.step(3)
        assert r["status"] == 0 and r["evaluations"] == 6
        out[prec] = (s.get_phi(), s.query())
        s.close()
    (p64, q64), (p32, q32) = out[vb.FP64], out[vb.FP32_MIXED]
    assert rel_l2(p32, p64) < 1e-5                                  # the FP32-mixed bound on phi
    assert rel(q32[1], q64[1]) < 1e-6 and rel(q32[2], q64[2]) < 1e-6   # V, A
