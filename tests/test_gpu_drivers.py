"""GPU checks of the driver exits (belastic.cpp:713-746, fluidbase.cpp:2885-2903),
of checkpoint resume against an uninterrupted run (fluidbase.cpp:2776-2825), of
the adaptive branch of the Euler driver (belastic.cpp:367-437) and of the
FP32-mixed adaptive driver, whose energy-decrease test (belastic.cpp:653-668)
compares FP64 energies of the FP64 state."""
import subprocess
import time

import numpy as np
import pytest

from conftest import flush_print, rel, rel_l2
from test_host import INI

pytestmark = pytest.mark.gpu


def make(vb, n, kind=1, seed=20260002, **kw):
    eps = kw.pop("eps", 2.0 * 2 * np.pi / n)
    phi, _ = vb.shape(kind, seed, n, n, 1.0, eps)
    return vb.params_default(nx=n, nz=n, eps=eps, **kw), phi


@pytest.mark.parametrize("fixed", [1e-6, 0.0])
@pytest.mark.parametrize("prec", ["FP64", "FP32_MIXED"])
def test_nonfinite_field_is_a_blow_up(vb, gpu, fixed, prec):
    # one NaN grid point spreads through every transform: the driver's blow-up
    # exit (belastic.cpp:733-740) ends the call with VB_ERR_NONFINITE
    flush_print("nonfinite", fixed, prec)
    p, phi = make(vb, 32, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6, fixed_time_step=fixed,
                  precision=getattr(vb, prec))
    phi[16, 16, 16] = np.nan
    s = vb.Solver(p, phi)
    r = s.step(4)
    assert r["status"] == vb.VB_ERR_NONFINITE, r
    assert r["steps"] <= 4


@pytest.mark.parametrize("fixed", [1e-24, 0.0])
def test_dt_underflow_exits(vb, gpu, fixed):
    # time_step < 1e-23 (belastic.cpp:667, 713-718): the adaptive start doubles
    # 1e-24 to 2e-24, still below the limit; the fixed step is min(fixed, dt)
    p, phi = make(vb, 32, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-24, fixed_time_step=fixed)
    s = vb.Solver(p, phi)
    r = s.step(5)
    assert r["status"] == vb.VB_ERR_DT_UNDERFLOW and r["steps"] == 0, r
    assert r["time_current"] == 0.0


@pytest.mark.parametrize("prec", ["FP64", "FP32_MIXED"])
def test_small_energy_finishes(vb, gpu, prec):
    # no interface, only a 1e-5 ripple on phi = -1: E is about 4e-7 < 1e-3 and
    # the ripple decays, so the first iteration is accepted and ends the run
    # before time and steps advance (belastic.cpp:727-732)
    n = 32
    p = vb.params_default(nx=n, M1=0.0, M2=0.0, time_step=1e-6, precision=getattr(vb, prec))
    x = np.arange(n) * 2 * np.pi / n - np.pi
    phi = -1.0 + 1e-5 * np.cos(x)[None, None, :] * np.cos(2 * x)[None, :, None] * np.ones((n, 1, 1))
    s = vb.Solver(p, phi)
    E0 = s.query()[0]
    assert 0.0 < E0 < 1e-3
    r = s.step(10)
    assert r["status"] == vb.VB_FINISHED and r["steps"] == 0 and r["evaluations"] == 1, r
    assert 0.0 <= r["energy"] < E0 and r["time_current"] == 0.0


@pytest.mark.parametrize("prec,fixed", [("FP64", 0.0), ("FP32_MIXED", 0.0), ("FP64", 1e-6)])
def test_resume_continues_exactly(vb, gpu, tmp_path, prec, fixed):
    # save after 8 iterations (a multiple of the doubling period: the reference
    # restarts its step counter on resume, belastic.cpp:629), resume in a new
    # solver and continue: every later iteration is the uninterrupted run's, bit
    # for bit (the state spectrum sidecar makes the round trip exact)
    flush_print("resume", prec, fixed)
    p, phi = make(vb, 32, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6, fixed_time_step=fixed,
                  precision=getattr(vb, prec), penalty_drift=0)
    a = vb.Solver(p, phi)
    assert a.step(8)["status"] == 0
    a.save(str(tmp_path))
    assert (tmp_path / "d32x32x32.out").exists() and (tmp_path / "d32x32x32.spec").exists()
    ra = a.step(12)
    b = vb.Solver(p, np.zeros((32, 32, 32)))
    b.load(str(tmp_path))
    rb = b.step(12)
    for k in ("steps", "evaluations", "rejections", "status", "time_step", "time_current"):
        assert ra[k] == rb[k], (k, ra[k], rb[k])
    assert np.array_equal(a.get_phi(), b.get_phi())
    assert a.query() == b.query()
    # without the sidecar the field file alone is loaded: the same run to the
    # rounding of the transform round trip (1e-16), which in FP32-mixed mode
    # moves FP32 roundings of the stage fields (1e-8)
    (tmp_path / "d32x32x32.spec").unlink()
    c = vb.Solver(p, np.zeros((32, 32, 32)))
    c.load(str(tmp_path))
    rc = c.step(12)
    tol = 1e-13 if prec == "FP64" else 1e-6
    err = rel_l2(c.get_phi(), a.get_phi())
    assert rc["time_step"] == ra["time_step"] and err < tol, err


def test_resume_small_grid_stamped_name(vb, gpu, tmp_path):
    # 16^3 fields are at most 100000 bytes: the checkpoint field carries the
    # output stamp round(t / 1e-6) in its name (fluidbase.cpp:259-275)
    p, phi = make(vb, 16, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6, fixed_time_step=1e-6)
    a = vb.Solver(p, phi)
    a.step(3)
    a.save(str(tmp_path))
    f = tmp_path / "d16x16x16_00000003.out"
    assert f.exists() and not (tmp_path / "d16x16x16.out").exists()
    assert np.array_equal(np.fromfile(f, dtype="<f8").reshape(16, 16, 16), a.get_phi())
    b = vb.Solver(p, np.zeros((16, 16, 16)))
    b.load(str(tmp_path))
    assert np.array_equal(b.get_phi(), a.get_phi())


def test_euler_adaptive_branch_matches_oracle(vb, gpu, oracle_mod):
    # run_static with adaptive dt: x1.4 every 4 steps, halve when E rises
    # (belastic.cpp:367-395); eta != 1 so the eta g0 term is weighted
    p, phi = make(vb, 32, kind=2, scheme=vb.EULER, M1=1e3, M2=1e3, alpha=-1.0, beta=-1.0, eta=1.7,
                  time_step=1e-7, fixed_time_step=0.0)
    o = oracle_mod.Oracle(32, eps=p.eps, eta=1.7, M1=1e3, M2=1e3, alpha=-1.0, beta=-1.0)
    s = vb.Solver(p, phi)
    d = phi.copy()
    st = o.fEIF_begin(d, 1e-7)
    for k in range(4):
        r = s.step(5)
        o.static_steps(d, st, 5)
        assert r["time_step"] == st.time_step, (k, r["time_step"], st.time_step)
    assert st.rejections > 0 or st.time_step > 1e-7    # the branch moved dt
    err = rel_l2(s.get_phi(), d)
    assert err < 1e-10, err
    assert rel(s.query()[0], o.energy(d)[0]) < 1e-11


def test_fp32_mixed_adaptive_energy_test_is_fp64(vb, gpu):
    # The accept test of the FP32-mixed driver compares FP64 energies: the energy
    # it reports for each accepted state is the FP64 query of that state, bit for
    # bit, and never exceeds the previous one.  Its accepted dt sequence is then
    # the FP64 run's wherever the energy change of a decision is far above the
    # 1e-7-level difference of the two states: here the first 20 iterations;
    # later, at dt ~ 6e-6 (beyond the explicit penalties' stability, sigma dt ~ 4),
    # candidates change E by about that much and the two runs may decide apart.
    n = 64
    out = {}
    for prec in (vb.FP64, vb.FP32_MIXED):
        p, phi = make(vb, n, kind=3, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-7, precision=prec,
                      phi_lut_mode=vb.LUT_ACCURATE)
        s = vb.Solver(p, phi)
        s.step(0)                       # the driver start: alpha and beta from the field
        seq, last = [], s.query()[0]
        for _ in range(40):
            r = s.step(1)
            assert r["status"] == 0
            E = s.query()[0]
            assert r["energy"] == E and E <= last, (prec, len(seq), r["energy"], E, last)
            last = E
            seq.append((r["time_step"], r["rejections"], r["energy"]))
        out[prec] = seq
        s.close()
    s64, s32 = out[vb.FP64], out[vb.FP32_MIXED]
    first = next((i for i, (a, b) in enumerate(zip(s32, s64)) if a[:2] != b[:2]), len(s64))
    flush_print("FP32-mixed and FP64 dt sequences agree for", first, "iterations")
    assert first >= 20
    # E of the two states: the penalty terms amplify their 1e-8-level differences
    # in V and A (M (A - beta) A / E up to a few hundred), measured up to 4e-6
    for (_, _, e32), (_, _, e64) in zip(s32[:first], s64[:first]):
        assert rel(e32, e64) < 2e-5


def test_cli_folder_edit_stops_the_run(vb, gpu, tmp_path):
    # check_folder (fluidbase.cpp:2885-2903): every 16 steps the driver reads
    # the first field of its ini file; another folder number ends the run
    # (belastic.cpp:774-777)
    ini = tmp_path / "ini0.dat"
    ini.write_text(INI)
    proc = subprocess.Popen([vb.HOST_PROGRAM, "0", "--results", str(tmp_path), "--max-steps", "100000000",
                             "--output-every", "1e9"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = []
    try:
        t0 = time.time()
        for line in proc.stdout:
            lines.append(line)
            if "Start STATIC Running" in line:
                break
            assert time.time() - t0 < 60, "no start line"
        ini.write_text("102" + INI[3:])
        out, _ = proc.communicate(timeout=60)
    finally:
        if proc.poll() is None:
            proc.kill()
    text = "".join(lines) + out
    flush_print(text[-300:])
    assert proc.returncode == 0
    assert "User terminate!" in text and "End Running" in text
