"""GPU checks of the C++ host program (the belastic-shaped driver over the C
ABI) and of the residual behind its "error" output."""
import os
import subprocess

import numpy as np
import pytest

from conftest import flush_print, rel
from test_host import INI

pytestmark = pytest.mark.gpu


def test_residual_matches_oracle(vb, gpu, oracle_mod):
    n = 32
    eps = 2 * 2 * np.pi / n
    phi = vb.shape_ninit(2, n, eps=eps)
    o0 = oracle_mod.Oracle(n, eps=eps)
    a, b = 1.01 * o0.vol(phi), 0.99 * o0.surface(phi)
    for prec in (vb.FP64, vb.FP32_MIXED):
        p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=2e4, alpha=a, beta=b, eta=1.5, precision=prec)
        s = vb.Solver(p, phi)
        o = oracle_mod.Oracle(n, eps=eps, eta=1.5, M1=1e4, M2=2e4, alpha=a, beta=b)
        ref = np.mean(o.sum_fgq(phi) ** 2) * (2 * np.pi) ** 3     # Calc.cu:250,260
        got = s.residual()
        assert rel(got, ref) < (1e-11 if prec == vb.FP64 else 1e-4), (prec, got, ref)


def run_cli(vb, results, *args):
    r = subprocess.run([vb.HOST_PROGRAM, "0", "--results", str(results)] + list(args),
                       capture_output=True, text=True, timeout=120)
    flush_print(r.stdout[-400:], r.stderr[-400:])
    return r


def test_host_program_run_and_resume(vb, gpu, tmp_path):
    (tmp_path / "ini0.dat").write_text(INI)
    r = run_cli(vb, tmp_path, "--max-steps", "20")
    assert r.returncode == 0
    fold = tmp_path / "101"
    for f in ("output32x32x32.txt", "energy32x32x32.txt", "vol_sur32x32x32.txt", "lamb32x32x32.txt",
              "error32x32x32.txt", "time_step32x32x32.txt", "alpha_beta32x32x32.txt", "32x32x32.info",
              "d32x32x32.out"):
        assert (fold / f).exists(), f
    # the same run through the Python front end: ini -> params, ninit 2 shape, 20 steps
    p = vb.params_from_ini(str(tmp_path / "ini0.dat"))
    s = vb.Solver(p, vb.shape_ninit(2, p.nx, p.nz, eps=p.eps))
    rep = s.step(20)
    assert rep["status"] == 0
    cli_phi = np.fromfile(fold / "d32x32x32.out", dtype="<f8").reshape(32, 32, 32)
    assert np.array_equal(cli_phi, s.get_phi())
    info = np.fromfile(fold / "32x32x32.info", dtype="<f8")
    st = s.state()
    assert info[0] == st["time_current"] and info[2] == st["alpha"] and info[3] == st["beta"]
    # the energy log: first line at t = 0, last at t = 20 dt, stamped in units of 1e-6
    rows = [l.split(",") for l in (fold / "energy32x32x32.txt").read_text().splitlines()]
    assert int(rows[0][0]) == 0 and int(rows[-1][0]) == round(20 * 1e-7 / 1e-6)
    E, V, A, Eb = s.query()
    assert abs(float(rows[-1][7]) - E) < 1e-5 * max(1.0, abs(E))
    # resume: the .info makes the next run continue from t = 20 dt; the same
    # resume through the Python front end loads the checkpoint the CLI wrote
    s2 = vb.Solver(p, np.zeros((32, 32, 32)))
    s2.load(str(fold))
    s2.step(10)
    r = run_cli(vb, tmp_path, "--max-steps", "10")
    assert r.returncode == 0 and "Continue the existing experiment" in r.stdout
    info2 = np.fromfile(fold / "32x32x32.info", dtype="<f8")
    assert abs(info2[0] - 30 * 1e-7) < 1e-13 and info2[0] == s2.state()["time_current"]
    cli_phi = np.fromfile(fold / "d32x32x32.out", dtype="<f8").reshape(32, 32, 32)
    assert np.array_equal(cli_phi, s2.get_phi())
