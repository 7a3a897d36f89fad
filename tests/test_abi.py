"""CPU checks of the C ABI (include/vesicle.h): the library loads and exports
every declared entry point, and the host-only functions (parameters, ini
loading, eps table, seeded shapes, coefficient tables) behave as the
reference's host code does.  No GPU computation is started here."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, grid, rel

HEADER = os.path.join(ROOT, "include", "vesicle.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vb_[A-Za-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(vb):
    import ctypes
    lib = vb.lib()
    names = declared_symbols()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    nm = os.popen("nm -D --defined-only %s" % vb.LIB_PATH).read()
    for n in names:
        assert re.search(r"\bT %s\b" % n, nm), n


def test_version_and_strerror(vb):
    assert "gfx950" in vb.version()
    for code in range(10):
        assert vb.lib().vb_strerror(code)


def test_params_default_matches_reference_defaults(vb):
    p = vb.params_default()
    assert (p.kappa, p.D) == (2.0, 1.5)              # Calc.cu:581-582
    assert p.dt_cap == 1e-3 and p.dt_double_every == 4   # belastic.cpp:644-651
    assert p.penalty_drift == 1 and p.scheme == vb.ETDRK4
    assert p.accept_chain == 0                       # the host decides each attempt, as the reference


@pytest.mark.parametrize("code,mult", [(0, 2.0 * 2), (1, 2.5 * np.sqrt(2.0)), (2, 2.5 * 2), (3, 1.5 * 2),
                                       (4, 1.25 * 2), (5, 1.0 * 2), (6, 0.7 * 2), (7, 4.0 * 2),
                                       (8, 2.5 * np.sqrt(2.0) * 1.22), (150, 1.5 * 2), (200, 2.0 * 2),
                                       (300, 3.0 * 2)])
def test_eps_table(vb, code, mult):
    # fluidbase.cpp:68-108 with L = 2 pi, n = 128: eps = mult * L / 2 / n
    L, n = 2 * np.pi, 128
    assert rel(vb.eps_from_code(code, n, L), mult * L / 2.0 / n) < 1e-15


def _ini(tmp_path, name):
    for line in open(os.path.join(GOLDEN, "ini_lines.txt")):
        k, v = line.rstrip("\n").split("\t", 1)
        if k == name:
            p = tmp_path / name
            p.write_text(v + "\n")
            return str(p)
    raise KeyError(name)


def test_params_from_ini0(vb, tmp_path):
    # ini0.dat: 128^3, shape adjust on, dt 1e-7, fixed 1, epssize 200, M1 = M2 = 1e4, alpha/beta auto
    p = vb.params_from_ini(_ini(tmp_path, "ini0.dat"))
    assert (p.nx, p.ny, p.nz) == (128, 128, 128)
    assert p.shape_adjust == 1 and p.time_step == 1e-7 and p.fixed_time_step == 1.0
    assert rel(p.eps, 2.0 * 2 * np.pi / 128) < 1e-15
    assert p.M1 == 1e4 and p.M2 == 1e4
    assert p.alpha == 0.0 and p.beta == 0.0          # beta < 0 -> 0 (fluidbase.cpp:143)
    assert p.eta == 1.0 and p.ga == 1.0 and p.tmax == 100.0


def test_params_from_ini_odd_grid_rejected(vb, tmp_path):
    # ini1.dat has nx = 33; the reference's Calc rejects odd nx (Calc.cu:52-55)
    p = vb.params_from_ini(_ini(tmp_path, "ini1.dat"))
    assert p.nx == 33 and p.nz == 32
    phi = np.zeros((p.nz, p.nx, p.nx))
    with pytest.raises(vb.VbError) as e:
        vb.Solver(p, phi)
    assert e.value.status == vb.VB_ERR_ODD_NX


def test_params_from_ini_eta_default(vb, tmp_path):
    # ini3.dat: eta = 5 as given; ini12.dat: M1 = M2 = 1e5, epssize 125
    p = vb.params_from_ini(_ini(tmp_path, "ini3.dat"))
    assert p.eta == 5.0 and p.epssize == 3
    p = vb.params_from_ini(_ini(tmp_path, "ini12.dat"))
    assert p.M1 == 1e5 and rel(p.eps, 1.25 * 2 * np.pi / 65) < 1e-15


def test_shapes_seeded_and_geometric(vb, oracle_mod):
    n = 64
    eps = 2 * 2 * np.pi / n
    a, sa = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps)
    b, sb = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    c, sc = vb.shape(vb.SHAPE_TORUS, 20260004, n, n, 1.0, eps)
    assert not np.array_equal(sa, sc)
    R, r = sa[0], sa[1]
    assert 1.2 <= R <= 1.8 and 0.4 <= r <= 0.7
    assert np.abs(a).max() <= 1.0
    # (phi+1)/2 is a logistic step of width w = sqrt2 eps / 2 across the tube wall, so
    # the tube cross-section is pi (r^2 + pi^2 w^2 / 3) and V = 2 pi R times that
    o = oracle_mod.Oracle(n, eps=eps)
    w = np.sqrt(2) * eps / 2
    assert rel(o.vol(a), 2 * np.pi ** 2 * R * (r ** 2 + np.pi ** 2 * w ** 2 / 3)) < 5e-3
    noisy, _ = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps, noise=1e-3)
    d = noisy - a
    assert abs(d.std() - 1e-3) < 2e-5 and abs(d.mean()) < 2e-5
    for kind in range(5):
        phi, sp = vb.shape(kind, 7, 32, 32, 1.0, 2 * 2 * np.pi / 32)
        assert np.isfinite(phi).all() and phi.max() > 0.5 and phi.min() < -0.5


def test_lut_reference_mode_matches_oracle_bitwise(vb, oracle_mod):
    # the product's table and the oracle's per-mode reference expressions are the same
    # doubles: same L^ (Calc.cu:599-600), same exp, same operation order (Calc.cu:495, 556-567)
    n = 64
    eps = 2 * 2 * np.pi / n
    mmax = 3 * (n // 2) ** 2
    for h in (1e-7, 5e-7, 3e-5):
        t = vb.lut(vb.ETDRK4, vb.LUT_REFERENCE, h, eps, mmax=mmax)
        o = oracle_mod.Oracle(n, eps=eps)
        Het = o.Het()
        ref = {}
        kz = np.minimum(np.arange(n), n - np.arange(n))[:, None, None]
        ky = np.minimum(np.arange(n), n - np.arange(n))[None, :, None]
        kx = np.arange(n // 2 + 1)[None, None, :]
        m = (kx ** 2 + ky ** 2 + kz ** 2)
        for mm, H in zip(m.ravel()[::97], Het.ravel()[::97]):
            assert t[mm, 7] == H
            c = oracle_mod.etd_coeffs(H, h)
            assert t[mm, 0] == c[0] and t[mm, 1] == c[1] and t[mm, 2] == c[2]
            s = 1.0 / (h * h * H * H * H)
            assert t[mm, 3] == s * c[3] and t[mm, 4] == s * (2 * c[4]) and t[mm, 5] == s * c[5]
        t2 = vb.lut(vb.ETDRK2, vb.LUT_REFERENCE, h, eps, mmax=mmax)
        for mm, H in zip(m.ravel()[::131], Het.ravel()[::131]):
            c = oracle_mod.etd_coeffs(H, h)
            assert t2[mm, 2] == c[6]


def test_lut_accurate_mode_vs_mpmath(vb):
    """phi-function values against 40-digit mpmath; the reference closed forms lose
    accuracy for small |z| (cancellation), the accurate tables do not."""
    tab = np.load(os.path.join(GOLDEN, "phi_mpmath.npz"))["table"]
    eps = 2 * 2 * np.pi / 64
    H0 = -1.5 * 2.0 / eps ** 3      # L^ at m = 0 (Calc.cu:600)
    worst_ref = 0.0
    for z, e1, p1h, ez, f1, f2, f3, p2 in tab:
        h = z / H0
        a = vb.lut(vb.ETDRK4, vb.LUT_ACCURATE, h, eps, mmax=0)[0]
        zz = h * a[7]
        assert rel(zz, z) < 1e-15
        # the table's z = h L^ carries its own rounding; it moves exp(z) by ~|z| ulps
        tol = 2e-15 + 1e-15 * abs(z)
        assert rel(a[0], e1) < tol and rel(a[2], ez) < tol
        assert rel(a[1] / h, p1h) < 1e-14
        assert rel(a[3] / h, f1) < 1e-13 and rel(a[4] / (2 * h), f2) < 1e-13 and rel(a[5] / h, f3) < 1e-13
        b = vb.lut(vb.ETDRK2, vb.LUT_ACCURATE, h, eps, mmax=0)[0]
        assert rel(b[2] / h, p2) < 1e-13
        r = vb.lut(vb.ETDRK4, vb.LUT_REFERENCE, h, eps, mmax=0)[0]
        worst_ref = max(worst_ref, rel(r[3] / h, f1))
    assert worst_ref > 1e-6   # the reference expressions are ill-conditioned near z = 0


def test_footprint(vb):
    f64 = vb.footprint(vb.params_default(nx=256))
    f32 = vb.footprint(vb.params_default(nx=256, precision=vb.FP32_MIXED))
    f128 = vb.footprint(vb.params_default(nx=128))
    assert f64 > f128 > 0 and f32 > 0
    assert f64 < 4 << 30          # well under 4 GiB per 256^3 solve


def test_create_without_gpu_reports_no_device(vb):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    p = vb.params_default(nx=32)
    with pytest.raises(vb.VbError) as e:
        vb.Solver(p, np.zeros((32, 32, 32)))
    assert e.value.status == vb.VB_ERR_NO_DEVICE


def test_schedules_without_gpu(vb):
    # kernel schedules per grid length (vb_schedule_*): host-side tables only
    assert vb.schedule_count(256, vb.FP32_MIXED) == 4
    assert vb.schedule_count(128, vb.FP32_MIXED) == 4
    assert vb.schedule_count(192, vb.FP32_MIXED) == 4
    # at 192 the default x plan is 24 x 8 (one LDS exchange), the other 4 x 4 x 4 x 3
    assert vb.schedule_describe(192, vb.FP32_MIXED, 0) == {"z_tile": 16, "x_passes": 2, "x_radix": 24}
    assert vb.schedule_describe(192, vb.FP32_MIXED, 2) == {"z_tile": 16, "x_passes": 4, "x_radix": 4}
    assert vb.schedule_count(256, vb.FP64) == 1
    assert vb.schedule_count(100, vb.FP32_MIXED) == 0 and vb.schedule_get(100, vb.FP32_MIXED) == -1
    d = [vb.schedule_describe(256, vb.FP32_MIXED, s) for s in range(4)]
    # the default: 8-column z tiles and the 16 x 16 plan of the ETD x kernel
    assert d[0] == {"z_tile": 8, "x_passes": 2, "x_radix": 16}
    assert len({tuple(x.values()) for x in d}) == 4
    saved = vb.schedule_get(256, vb.FP32_MIXED)
    try:
        vb.schedule_set(256, vb.FP32_MIXED, 3)
        assert vb.schedule_get(256, vb.FP32_MIXED) == 3
        assert vb.schedule_get(256, vb.FP64) == 0
        for bad in ((256, vb.FP64, 1), (256, vb.FP32_MIXED, 4), (100, vb.FP32_MIXED, 0), (256, 7, 0)):
            with pytest.raises(vb.VbError):
                vb.schedule_set(*bad)
    finally:
        vb.schedule_set(256, vb.FP32_MIXED, saved)


def test_autotune_arguments_and_no_device(vb):
    # invalid lengths, precisions and step counts are rejected before any device work
    for args in ((100, vb.FP32_MIXED), (64, 7)):
        with pytest.raises(vb.VbError) as e:
            vb.autotune(*args)
        assert e.value.status == vb.VB_ERR_PARAM
    with pytest.raises(vb.VbError) as e:
        vb.autotune(64, vb.FP32_MIXED, steps=0)
    assert e.value.status == vb.VB_ERR_PARAM
    import torch
    if torch.cuda.is_available():
        return
    saved = vb.schedule_get(32, vb.FP32_MIXED)
    with pytest.raises(vb.VbError) as e:
        vb.autotune(32, vb.FP32_MIXED, steps=1)
    assert e.value.status == vb.VB_ERR_NO_DEVICE
    assert vb.schedule_get(32, vb.FP32_MIXED) == saved   # a failed tuning selects nothing


@pytest.mark.parametrize("field,value,status", [
    ("dt_double_every", 0, "VB_ERR_PARAM"), ("time_step", 0.0, "VB_ERR_PARAM"), ("dt_cap", -1.0, "VB_ERR_PARAM"),
    ("scheme", 3, "VB_ERR_PARAM"), ("precision", 2, "VB_ERR_PARAM"), ("phi_lut_mode", 5, "VB_ERR_PARAM"),
    ("eps", 0.0, "VB_ERR_PARAM"), ("nx", 33, "VB_ERR_ODD_NX"), ("nx", 100, "VB_ERR_PARAM"),
    ("accept_chain", -1, "VB_ERR_PARAM"), ("accept_chain", 17, "VB_ERR_PARAM"), ("graphs", 3, "VB_ERR_PARAM")])
def test_invalid_params_rejected_by_every_entry_point(vb, field, value, status):
    # vb_create, vb_footprint and vb_batch_create share one check, made before
    # any device work (a zero doubling period would otherwise divide by zero in
    # the adaptive driver, belastic.cpp:644)
    import ctypes as C
    p = vb.params_default(nx=32)
    setattr(p, field, value)
    code = getattr(vb, status)
    phi = np.zeros((p.nz, p.nx, p.nx))
    with pytest.raises(vb.VbError) as e:
        vb.Solver(p, phi)
    assert e.value.status == code
    assert vb.footprint(p) == 0
    arr = (vb.Params * 1)(p)
    ptrs = (C.c_void_p * 1)(phi.ctypes.data)
    h = C.c_void_p()
    assert vb.lib().vb_batch_create(1, arr, ptrs, C.byref(h)) == code
    zero = vb.Params()
    assert vb.lib().vb_batch_create(1, (vb.Params * 1)(zero), ptrs, C.byref(h)) == vb.VB_ERR_PARAM
