"""GPU checks of the kernel schedules and the autotuner (vb_schedule_*,
vb_autotune).  The schedules of one grid length differ only in the radix plan
of the ETD x kernel and in the column tile of the z kernels, so every one of
them must compute the same solve to FP32 rounding and stay within the FP32
precision bound of the FP64 run; the autotuner must select the schedule it
timed fastest; and solves created under a selection, alone or batched, must
agree bit for bit (the batched form replays the launches of the solo form)."""
import numpy as np
import pytest

from conftest import flush_print, rel, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore(vb):
    saved = {(n, pr): vb.schedule_get(n, pr) for n in (32, 64, 128, 192, 256) for pr in (0, 1)}
    yield
    for (n, pr), sc in saved.items():
        vb.schedule_set(n, pr, sc)


def params(vb, n, precision, **kw):
    eps = 2.0 * 2 * np.pi / n
    phi, _ = vb.shape(vb.SHAPE_TORUS, 20260005, n, n, 1.0, eps, noise=1e-3)
    kw.setdefault("M1", 1e4)
    kw.setdefault("M2", 1e4)
    p = vb.params_default(nx=n, nz=n, eps=eps, alpha=0.0, beta=0.0, precision=precision,
                          phi_lut_mode=vb.LUT_ACCURATE, time_step=5e-7, fixed_time_step=5e-7, **kw)
    return p, phi


def solve(vb, p, phi, steps, graphs=1):
    p.graphs = graphs
    s = vb.Solver(p, phi)
    r = s.step(steps)
    out = (s.get_phi(), s.query(), r)
    s.close()
    return out


@pytest.mark.parametrize("n", [32, 64, 128, 192, 256])
def test_every_schedule_computes_the_same_solve(vb, gpu, restore, n):
    steps = 5
    p, phi = params(vb, n, vb.FP64)
    phi64, q64, _ = solve(vb, p, phi, steps)
    ref = None
    for sc in range(vb.schedule_count(n, vb.FP32_MIXED)):
        flush_print("schedule", n, sc, vb.schedule_describe(n, vb.FP32_MIXED, sc))
        vb.schedule_set(n, vb.FP32_MIXED, sc)
        p, phi = params(vb, n, vb.FP32_MIXED)
        got = solve(vb, p, phi, steps)
        assert got[2]["status"] == 0 and got[2]["evaluations"] == 2 * steps
        assert rel_l2(got[0], phi64) < 1e-5          # the FP32-mixed bound against FP64
        assert rel(got[1][1], q64[1]) < 1e-6 and rel(got[1][2], q64[2]) < 1e-6   # V, A
        if ref is None:
            ref = got
        else:
            assert rel_l2(got[0], ref[0]) < 1e-6     # schedules differ in rounding only


def test_autotune_selects_the_fastest(vb, gpu, restore):
    flush_print("autotune 64 fp32")
    rep = vb.autotune(64, vb.FP32_MIXED, steps=3)
    assert rep["candidates"] == vb.schedule_count(64, vb.FP32_MIXED) == 2
    t = rep["ms_per_step"]
    assert all(np.isfinite(x) and x > 0 for x in t)
    assert t[rep["chosen"]] == min(t)
    assert vb.schedule_get(64, vb.FP32_MIXED) == rep["chosen"]
    flush_print("autotune 32 fp64")
    rep = vb.autotune(32, vb.FP64, steps=2)
    assert rep["candidates"] == 1 and rep["chosen"] == 0


def test_batch_under_a_selection_matches_solo(vb, gpu, restore):
    for n in (32, 64):
        vb.schedule_set(n, vb.FP32_MIXED, 1)
    cases = [params(vb, 32, vb.FP32_MIXED, M1=m1, M2=m2) for m1, m2 in ((1e4, 1e4), (3e3, 2e4))]
    cases.append(params(vb, 64, vb.FP32_MIXED))
    flush_print("batch under schedule 1 of", len(cases))
    b = vb.Batch([c[0] for c in cases], [c[1] for c in cases])
    b.step(4)
    got = [(b.get_phi(i), b.query(i)) for i in range(len(cases))]
    b.close()
    for i, (p, phi) in enumerate(cases):
        ref_phi, ref_q, _ = solve(vb, p, phi, 4, graphs=0)
        assert np.array_equal(got[i][0], ref_phi), i
        assert got[i][1] == ref_q, i
