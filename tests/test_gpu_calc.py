"""GPU parity of every reference Calc operation (Calc.cu:107-606) through the
C ABI (vb_calc_*), against the oracle on the same seeded inputs."""
import numpy as np
import pytest

from conftest import flush_print, grid, rel, rel_l2

pytestmark = pytest.mark.gpu

NX, NZ = 32, 16


@pytest.fixture(scope="module")
def setup(vb, gpu, oracle_mod):
    import torch
    o = oracle_mod.Oracle(NX, NZ, M1=1e4, M2=1e4)
    c = vb.Calc(NX, NZ)
    X, Y, Z = grid(NX, NZ)
    d = np.tanh((1.6 - np.sqrt(X ** 2 + 1.4 * Y ** 2 + 0.8 * Z ** 2)) / (np.sqrt(2) * o.p.eps))
    d += 0.01 * np.random.default_rng(1).standard_normal(d.shape)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    R = lambda: torch.empty((NZ, NX, NX), dtype=torch.float64, device=gpu)
    Cx = lambda: torch.empty((NZ, NX, NX // 2 + 1), dtype=torch.complex128, device=gpu)
    return dict(o=o, c=c, d=d, T=T, R=R, C=Cx, eps=o.p.eps)


def H(t):
    return t.cpu().numpy()


def test_laplace_and_lap(setup):
    s = setup
    o, c, d = s["o"], s["c"], s["d"]
    fd = o.r2c(d)
    x = s["T"](fd)
    c.laplace(x)
    assert rel_l2(H(x), o.laplace(fd)) < 1e-15
    out = s["R"]()
    c.lap_spec(s["T"](fd), out)
    ref = o.c2r(o.laplace(fd))
    assert rel_l2(H(out), ref) < 1e-13
    c.lap_real(s["T"](d), out)
    assert rel_l2(H(out), o.lap_real(d)) < 1e-13


def test_fftgradient(setup):
    s = setup
    o, c, d = s["o"], s["c"], s["d"]
    fd = o.r2c(d)
    dx, dy, dz = s["R"](), s["R"](), s["R"]()
    c.fftgradient(s["T"](fd), dx, dy, dz)
    for a, b in zip((dx, dy, dz), o.fftgradient(fd)):
        assert rel_l2(H(a), b) < 1e-13


def test_pointwise_kernels(setup):
    s = setup
    o, c, d, eps = s["o"], s["c"], s["d"], s["eps"]
    lap = o.lap_real(d)
    f = s["R"]()
    c.f(f, s["T"](lap), s["T"](d), eps)
    fref = o.f_kernel(lap, d)
    assert rel_l2(H(f), fref) < 1e-15
    # g0_kernel works in place on lap f
    g0 = s["T"](o.lap_real(fref))
    c.g0(g0, s["T"](d), s["T"](fref), eps)
    assert rel_l2(H(g0), o.g0(d)) < 1e-13
    nl = s["R"]()
    tmp1 = s["R"]()
    c.nld(nl, s["T"](d), s["T"](fref), tmp1, eps)
    assert rel_l2(H(nl), o.nld(d, fref)) < 1e-13
    dx, dy, dz = o.fftgradient(o.r2c(d))
    st = s["R"]()
    c.s_term(st, s["T"](d), s["T"](dx), s["T"](dy), s["T"](dz), eps)
    assert rel_l2(H(st), o.s_term(d, dx, dy, dz)) < 1e-15
    m = s["R"]()
    c.multi(s["T"](d), s["T"](fref), m)
    assert rel_l2(H(m), d * fref) < 1e-16
    u = s["T"](d)
    c.plus(u, s["T"](fref), -3e-3)
    assert rel_l2(H(u), d + (-3e-3) * fref) < 1e-15
    n1 = s["R"]()
    c.nld1(n1, s["T"](fref), s["T"](d), 2.5, -1.5, 1.0)
    assert rel_l2(H(n1), o.nld1(fref, d, 2.5, -1.5)) < 1e-15
    n2 = s["R"]()
    c.nld2(n2, s["T"](fref), s["T"](d), 2.5, -1.5, 1.0)
    assert rel_l2(H(n2), o.nld2(fref, d, 2.5, -1.5)) < 1e-15


def test_intf(setup):
    s = setup
    o, c, d = s["o"], s["c"], s["d"]
    for add, mul in [(1.0, 0.5), (0.0, 3 / (2 * np.sqrt(2))), (0.0, 1.0)]:
        assert rel(c.intf(s["T"](d), add, mul), o.intf(d, add, mul)) < 1e-13


def test_init_fEIF(setup, gpu):
    import torch
    s = setup
    o, c = s["o"], s["c"]
    Het = torch.empty((NZ, NX, NX // 2 + 1), dtype=torch.float64, device=gpu)
    c.init_fEIF(Het, s["eps"])
    assert np.array_equal(H(Het), o.Het())


@pytest.mark.parametrize("h", [1e-7, 5e-6, 1e-4])
def test_etd_combines(setup, oracle_mod, h):
    import ctypes as C
    from oracle.pyoracle import lib as olib
    s = setup
    o, c, d, eps = s["o"], s["c"], s["d"], s["eps"]
    rng = np.random.default_rng(9)
    fd = o.r2c(d)
    spec = lambda: o.r2c(rng.standard_normal(d.shape))
    n1, n2, n3 = spec(), spec(), spec()
    Het = o.Het()
    P = lambda a: a.ctypes.data_as(C.c_void_p)

    def ref(fn, *arrs):
        out = np.zeros_like(fd)
        getattr(olib(), fn)(o.h, P(out), *[P(np.ascontiguousarray(a)) for a in arrs], P(Het), h)
        return out
    out = s["C"]()
    c.etd1(out, s["T"](fd), s["T"](n1), eps, h)
    assert rel_l2(H(out), ref("vbo_etd1", fd, n1)) < 1e-14
    c.etd2(out, s["T"](fd), s["T"](n1), s["T"](n2), eps, h)
    assert rel_l2(H(out), ref("vbo_etd2", fd, n1, n2)) < 1e-14
    c.etd42(out, s["T"](fd), s["T"](n1), eps, h)
    assert rel_l2(H(out), ref("vbo_etd42", fd, n1)) < 1e-14
    cc = s["T"](fd.copy())
    c.etd43(cc, s["T"](n1), s["T"](n2), eps, h)
    ref43 = fd.copy()
    olib().vbo_etd43(o.h, P(ref43), P(n1), P(n2), P(Het), h)
    assert rel_l2(H(cc), ref43) < 1e-14
    dr = s["R"]()
    c.etd41(dr, s["T"](fd), s["T"](n1), eps, h)
    assert rel_l2(H(dr), o.c2r(ref("vbo_etd41", fd, n1))) < 1e-13
    c.etd44(dr, s["T"](fd), s["T"](n1), s["T"](n2), s["T"](n3), eps, h)
    r44 = np.zeros_like(fd)
    olib().vbo_etd44(o.h, P(r44), P(fd), P(n1), P(n2), P(n3), P(Het), h)
    assert rel_l2(H(dr), o.c2r(r44)) < 1e-12
