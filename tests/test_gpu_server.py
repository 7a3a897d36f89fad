"""GPU checks of the request loop: each request computes exactly what vb_step
computes for it alone (bit-identical whatever runs beside it), admission
respects the memory budget, and the energy-drop exit fires after the window."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def request(vb, nx, kind, seed, precision, nz=None, **kw):
    nz = nx if nz is None else nz
    eps = 2.0 * 2 * np.pi / nx
    phi, _ = vb.shape(kind, seed, nx, nz, 1.0, eps)
    p = vb.params_default(nx=nx, nz=nz, eps=eps, M1=1e4, M2=1e4, time_step=1e-7, fixed_time_step=0.0,
                          precision=precision, phi_lut_mode=vb.LUT_ACCURATE, tmax=10.0, **kw)
    return p, phi


def mix(vb):
    return [request(vb, 32, 1, 31, 1), request(vb, 64, 3, 32, 1), request(vb, 48, 2, 33, 0, nz=32),
            request(vb, 32, 4, 34, 0), request(vb, 64, 1, 35, 1)]


def run(vb, reqs, **cfg):
    srv = vb.Server(keep_phi=True, **cfg)
    for i, (p, phi) in enumerate(reqs):
        srv.submit(i, p, phi)
    res = {r["id"]: r for r in srv.drain()}
    phis = {i: srv.take_phi(i) for i in res}
    srv.close()
    return res, phis


KEYS = ("status", "steps", "evaluations", "rejections", "E", "V", "A", "Eb", "time_step", "time_current")


def test_alone_and_in_the_mix_identical(vb, gpu):
    reqs = mix(vb)
    flush_print("server: in the mix")
    mixed, mphi = run(vb, reqs, max_active=5, streams=4, max_steps=25)
    flush_print("server: one at a time")
    alone, aphi = run(vb, reqs, max_active=1, streams=1, max_steps=25)
    assert sorted(mixed) == sorted(alone) == list(range(len(reqs)))
    for i in range(len(reqs)):
        assert mixed[i]["status"] == 0 and mixed[i]["steps"] == 25, mixed[i]
        for k in KEYS:
            assert mixed[i][k] == alone[i][k], (i, k)
        assert np.array_equal(mphi[i], aphi[i]), i
        assert np.isfinite(mphi[i]).all()


def test_server_equals_vb_step(vb, gpu):
    reqs = mix(vb)[:3]
    res, phis = run(vb, reqs, max_active=3, streams=2, max_steps=20)
    for i, (p, phi) in enumerate(reqs):
        s = vb.Solver(p, phi)
        r = s.step(20)
        assert (r["steps"], r["evaluations"], r["rejections"]) == \
            (res[i]["steps"], res[i]["evaluations"], res[i]["rejections"]), i
        assert np.array_equal(s.get_phi(), phis[i]), i
        E, V, A, Eb = s.query()
        assert (E, V, A, Eb) == (res[i]["E"], res[i]["V"], res[i]["A"], res[i]["Eb"]), i


def test_admission_by_memory_budget(vb, gpu):
    reqs = [request(vb, 64, 1, 40 + i, 1) for i in range(5)]
    foot = vb.footprint(reqs[0][0])
    res, _ = run(vb, reqs, max_active=16, streams=4, max_steps=6, memory_budget=int(2.5 * foot))
    assert all(r["status"] == 0 and r["steps"] == 6 for r in res.values())
    # at no time were more than two requests resident
    ev = sorted([(r["t_admit"], 1) for r in res.values()] + [(r["t_done"], -1) for r in res.values()],
                key=lambda e: (e[0], e[1]))
    live = peak = 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak <= 2
    # arrival order is kept
    order = sorted(res, key=lambda i: res[i]["t_admit"])
    assert order == sorted(order)


def test_energy_drop_exit(vb, gpu):
    p, phi = request(vb, 32, 1, 50, 1)
    res, _ = run(vb, [(p, phi)], converge_window=5, converge_tol=1.0)
    assert res[0]["status"] == 0 and res[0]["steps"] == 6


def test_slot_pool_reuse_matches_fresh_solves(vb, gpu):
    # one slot: every request after the first reuses the previous request's
    # solver (same grid and precision) or, when its kind differs and the budget
    # holds one solver only, replaces it; each result is what a fresh vb_step
    # computes for that request
    reqs = [request(vb, 32, 1, 61, 1), request(vb, 32, 3, 62, 1, eta=1.5),
            request(vb, 64, 4, 63, 1), request(vb, 32, 2, 64, 1)]
    reqs[1][0].eps *= 1.5           # other eps: other tables and constants in the reused solver
    reqs[3][0].M1 = 3e3
    budget = vb.footprint(reqs[2][0]) + vb.footprint(reqs[0][0]) // 2   # a 64^3 solver or a 32^3 one
    res, phis = run(vb, reqs, max_active=1, streams=1, max_steps=12, memory_budget=budget)
    for i, (p, phi) in enumerate(reqs):
        flush_print("slot pool: fresh solve", i)
        s = vb.Solver(p, phi)
        r = s.step(12)
        assert (r["steps"], r["evaluations"], r["rejections"]) == \
            (res[i]["steps"], res[i]["evaluations"], res[i]["rejections"]), i
        assert np.array_equal(s.get_phi(), phis[i]), i
        assert s.query() == (res[i]["E"], res[i]["V"], res[i]["A"], res[i]["Eb"]), i
        s.close()


def test_shared_launch_rounds_match_solo_solves(vb, gpu):
    # seven requests of one kind (32^3 FP32-mixed ETDRK4) and three of another
    # (32^3 FP64): the attempts of a kind's running solves go out together as
    # shared launches, each solve with its own dt, tables, constants and state;
    # two 96^3 requests beside them go out one attempt at a time; every result
    # is what a fresh vb_step computes for that request alone
    reqs = [request(vb, 32, 1 + k % 4, 70 + k, 1) for k in range(7)] + \
           [request(vb, 32, 1 + k % 4, 80 + k, 0) for k in range(3)] + \
           [request(vb, 96, 1 + k, 90 + k, 1) for k in range(2)]
    for k, (p, _) in enumerate(reqs):
        p.M1 = 1e3 * (1 + k)
        p.eps *= 1.0 + 0.1 * (k % 3)
    flush_print("server: shared rounds")
    res, phis = run(vb, reqs, max_active=16, streams=4, max_steps=15)
    for i, (p, phi) in enumerate(reqs):
        s = vb.Solver(p, phi)
        r = s.step(15)
        assert res[i]["status"] == 0 and res[i]["steps"] == 15, res[i]
        assert (r["steps"], r["evaluations"], r["rejections"]) == \
            (res[i]["steps"], res[i]["evaluations"], res[i]["rejections"]), i
        assert np.array_equal(s.get_phi(), phis[i]), i
        assert s.query() == (res[i]["E"], res[i]["V"], res[i]["A"], res[i]["Eb"]), i
        s.close()
