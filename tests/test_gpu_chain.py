"""GPU checks of the adaptive ETD driver with its accept/reject decisions taken
on the device (vb_params.accept_chain > 0): a run in chains of attempts must be
the host-decided run (belastic.cpp:639-669, 713-746) bit for bit -- the same
reports after every call, the same field and the same integrals -- through
rejections, dt values that leave the tables a chain was given, calls shorter
than a chain, and each driver exit."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu

KEYS = ("steps", "evaluations", "rejections", "status", "time_step", "time_current", "energy")


def make(vb, n, kind=3, seed=20260002, **kw):
    eps = kw.pop("eps", 2.0 * 2 * np.pi / n)
    phi, _ = vb.shape(kind, seed, n, n, 1.0, eps)
    return vb.params_default(nx=n, nz=n, eps=eps, **kw), phi


def same(x, y):
    return x == y or (isinstance(x, float) and np.isnan(x) and np.isnan(y))


def same_report(a, b, keys=KEYS):
    return all(same(a[k], b[k]) for k in keys)


def run(vb, p, phi, calls):
    s = vb.Solver(p, phi)
    reps = [s.step(k) for k in calls]
    out = reps, s.get_phi(), s.query(), s.state()
    s.close()
    return out


@pytest.mark.parametrize("n,prec,chain,scheme,cap", [
    (32, "FP64", 4, "ETDRK4", 1e-3), (64, "FP32_MIXED", 16, "ETDRK4", 1e-3), (64, "FP64", 3, "ETDRK4", 1e-3),
    (32, "FP32_MIXED", 5, "ETDRK2", 1e-3), (32, "FP32_MIXED", 8, "ETDRK4", 5e-7)])
def test_chain_matches_host_decisions(vb, gpu, n, prec, chain, scheme, cap):
    # dt starts at 1e-4, far above what the penalties allow, so the first
    # iterations reject several times and halve dt below the tables a chain
    # starts with; the calls of 1, 7, 13 and 19 steps end chains early.  With
    # the 5e-7 cap the doubling rule sets dt = cap instead of 2 dt
    # (belastic.cpp:644-651) at the first iteration and every fourth after.
    calls = [1, 7, 13, 19]
    flush_print("chain", n, prec, chain, scheme, cap)
    reps = {}
    for ac in (0, chain):
        p, phi = make(vb, n, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-4, dt_cap=cap,
                      precision=getattr(vb, prec), scheme=getattr(vb, scheme), accept_chain=ac)
        reps[ac] = run(vb, p, phi, calls)
    (r0, phi0, q0, s0), (r1, phi1, q1, s1) = reps[0], reps[chain]
    if cap < 1e-3:
        assert all(r["time_step"] <= cap for r in r0) and any(r["time_step"] == cap for r in r0), r0
    else:
        assert sum(r["rejections"] for r in r0) >= 3, r0
    for a, b in zip(r0, r1):
        assert same_report(a, b), (a, b)
    assert np.array_equal(phi0, phi1)
    assert q0 == q1 and s0 == s1


@pytest.mark.parametrize("prec", ["FP64", "FP32_MIXED"])
def test_chain_exits_match_host(vb, gpu, prec):
    n = 32
    cases = []
    # blow-up: one NaN grid point (belastic.cpp:733-740)
    p, phi = make(vb, n, kind=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6, precision=getattr(vb, prec))
    phi[16, 16, 16] = np.nan
    cases.append(("nonfinite", p, phi, vb.VB_ERR_NONFINITE))
    # dt underflow (belastic.cpp:713-718)
    p, phi = make(vb, n, kind=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-24, precision=getattr(vb, prec))
    cases.append(("underflow", p, phi, vb.VB_ERR_DT_UNDERFLOW))
    # E < 1e-3 (belastic.cpp:727-732): a 1e-5 ripple on phi = -1
    x = np.arange(n) * 2 * np.pi / n - np.pi
    phi = -1.0 + 1e-5 * np.cos(x)[None, None, :] * np.cos(2 * x)[None, :, None] * np.ones((n, 1, 1))
    p = vb.params_default(nx=n, M1=0.0, M2=0.0, time_step=1e-6, precision=getattr(vb, prec))
    cases.append(("finished", p, phi, vb.VB_FINISHED))
    for name, p, phi, status in cases:
        flush_print("chain exit", name, prec)
        out = {}
        for ac in (0, 8):
            p.accept_chain = ac
            s = vb.Solver(p, phi)
            r = s.step(6)
            r2 = s.step(6)   # the call after an exit, as the host driver would make it
            out[ac] = (r, r2, s.state())
            s.close()
        assert out[0][0]["status"] == status, out[0]
        for a, b in zip(out[0][:2], out[8][:2]):
            assert same_report(a, b), (name, a, b)
        assert same_report(out[0][2], out[8][2], out[0][2].keys()), (name, out[0][2], out[8][2])


@pytest.mark.parametrize("prec", ["FP64", "FP32_MIXED"])
def test_euler_attempt_graphs_match_eager(vb, gpu, prec):
    # The shape-adjust pre-pass (belastic.cpp:174-283) and the adaptive Euler
    # driver (belastic.cpp:367-437) replay one captured attempt for every dt,
    # the update reading its dt from device memory: bit for bit the eager run.
    n = 32
    out = {}
    for graphs in (0, 2):
        flush_print("euler attempts, graphs", graphs, prec)
        p, phi = make(vb, n, kind=1, shape_adjust=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6,
                      fixed_time_step=1e-6, precision=getattr(vb, prec), graphs=graphs)
        s = vb.Solver(p, phi)
        s.step(1)   # the driver start runs the shape adjustment
        shaped = s.get_phi()
        s.close()
        # from dt = 1e-5 the explicit update first raises E: rejections halve dt
        p, phi = make(vb, n, kind=2, scheme=vb.EULER, M1=1e3, M2=1e3, alpha=-1.0, beta=-1.0, eta=1.7,
                      time_step=1e-5, fixed_time_step=0.0, precision=getattr(vb, prec), graphs=graphs)
        s = vb.Solver(p, phi)
        reps = [s.step(5) for _ in range(4)]
        out[graphs] = (shaped, reps, s.get_phi(), s.query())
        s.close()
    assert np.array_equal(out[0][0], out[2][0])
    for a, b in zip(out[0][1], out[2][1]):
        assert same_report(a, b), (a, b)
    flush_print("Euler attempts", sum(r["evaluations"] for r in out[0][1]),
                "rejections", sum(r["rejections"] for r in out[0][1]))
    assert sum(r["rejections"] for r in out[0][1]) > 0   # the rejection path ran in graph form too
    assert np.array_equal(out[0][2], out[2][2]) and out[0][3] == out[2][3]


@pytest.mark.parametrize("n,prec,kind", [(32, "FP64", 1), (32, "FP32_MIXED", 1), (64, "FP32_MIXED", 3)])
def test_shape_adjust_chain_matches_host(vb, gpu, n, prec, kind):
    # The shape-adjust pre-pass (belastic.cpp:174-283) decided on the device in
    # chains of attempts: the same field as the host loop, bit for bit, and the
    # same state for the steps after it (its pipeline state is the host's).
    out = {}
    for ac in (0, 8):
        flush_print("shape adjust, accept_chain", ac, n, prec)
        p, phi = make(vb, n, kind=kind, shape_adjust=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-6,
                      fixed_time_step=1e-6, precision=getattr(vb, prec), accept_chain=ac)
        s = vb.Solver(p, phi)
        s.step(0)    # the driver start runs the shape adjustment
        shaped, q = s.get_phi(), s.query()
        r = s.step(3)
        out[ac] = (shaped, q, r, s.get_phi(), s.state())
        s.close()
    a, b = out[0], out[8]
    assert not np.array_equal(a[0], make(vb, n, kind=kind)[1])   # the pre-pass moved the field
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    assert same_report(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert same_report(a[4], b[4], a[4].keys())
