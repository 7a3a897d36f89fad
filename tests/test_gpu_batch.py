"""GPU checks of graph-replayed fixed-dt steps and of batched solves: a graph
replays exactly the launches the eager driver issues, so every result must be
bit-identical to the eager, one-solve-at-a-time run of the same parameters."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def case(vb, nx, nz=None, kind=1, seed=7, **kw):
    nz = nx if nz is None else nz
    eps = kw.pop("eps", 2.0 * 2 * np.pi / nx)
    phi, _ = vb.shape(kind, seed, nx, nz, 1.0, eps)
    kw.setdefault("fixed_time_step", 5e-7)
    kw.setdefault("time_step", 5e-7)
    p = vb.params_default(nx=nx, nz=nz, eps=eps, **kw)
    return p, phi


def solo(vb, p, phi, chunks, graphs):
    p.graphs = graphs
    s = vb.Solver(p, phi)
    reps = [s.step(c) for c in chunks]
    return s.get_phi(), s.state(), s.query(), reps


@pytest.mark.parametrize("prec,drift", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_graph_replay_matches_eager(vb, gpu, prec, drift):
    flush_print("graph vs eager", prec, drift)
    M = 1e4 if drift else 0.0
    p, phi = case(vb, 32, precision=prec, M1=M, M2=M)
    chunks = [9, 1, 4, 7]
    a = solo(vb, p, phi, chunks, graphs=1)
    b = solo(vb, p, phi, chunks, graphs=0)
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1]
    assert a[2] == b[2]
    assert [r["evaluations"] for r in a[3]] == [r["evaluations"] for r in b[3]]
    assert a[1]["time_current"] == b[1]["time_current"]


def test_graph_replay_with_queries_between(vb, gpu):
    # a query between calls changes the host-side pipeline state (plain mode)
    p, phi = case(vb, 32, M1=0.0, M2=0.0)
    outs = []
    for g in (1, 0):
        p.graphs = g
        s = vb.Solver(p, phi)
        q = []
        for c in (4, 5, 6):
            s.step(c)
            q.append(s.query())
        outs.append((s.get_phi(), q))
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]


def test_batch_matches_solo(vb, gpu):
    specs = [
        dict(nx=64, kind=3, seed=11, precision=1, M1=1e3, M2=3e4),
        dict(nx=32, kind=1, seed=12, precision=1, M1=3e4, M2=1e3, eps=1.5 * 2 * np.pi / 32, phi_lut_mode=1),
        dict(nx=48, nz=32, kind=2, seed=13, precision=0, M1=1e4, M2=1e4),
        dict(nx=32, kind=4, seed=14, precision=1, M1=0.0, M2=0.0, eps=3.0 * 2 * np.pi / 32),
        dict(nx=64, kind=1, seed=15, precision=0, M1=2e4, M2=5e3),
        # adaptive dt: stepped with host decisions inside the batch
        dict(nx=32, kind=1, seed=16, precision=0, M1=1e4, M2=1e4, fixed_time_step=0.0, time_step=1e-7),
    ]
    ps, phis = [], []
    for sp in specs:
        sp = dict(sp)
        p, phi = case(vb, sp.pop("nx"), sp.pop("nz", None), sp.pop("kind"), sp.pop("seed"), **sp)
        ps.append(p)
        phis.append(phi)
    flush_print("batch of", len(ps))
    b = vb.Batch(ps, phis)
    reps = b.step(5) + b.step(6)
    got = [(b.get_phi(i), b.query(i)) for i in range(len(ps))]
    b.close()
    for i, (p, phi) in enumerate(zip(ps, phis)):
        assert np.isfinite(got[i][0]).all() and reps[i]["status"] == 0, i
        ref_phi, _, ref_q, ref_reps = solo(vb, p, phi, [5, 6], graphs=0)
        assert np.array_equal(got[i][0], ref_phi), i
        assert got[i][1] == ref_q, i
        assert reps[i]["evaluations"] == ref_reps[0]["evaluations"], i
        assert reps[len(ps) + i]["steps"] == ref_reps[1]["steps"], i


@pytest.mark.parametrize("prec", [0, 1])
def test_adaptive_attempt_graphs_match_eager(vb, gpu, prec):
    # adaptive attempts replay one graph per (state, flags, dt): rejections and the
    # doubling ladder revisit dt values, so graphs are reused and re-keyed
    p, phi = case(vb, 32, precision=prec, M1=1e4, M2=1e4, fixed_time_step=0.0, time_step=1e-6)
    a = solo(vb, p, phi, [12, 1, 9], graphs=1)
    b = solo(vb, p, phi, [12, 1, 9], graphs=0)
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1] and a[2] == b[2]
    assert [(r["steps"], r["evaluations"], r["rejections"]) for r in a[3]] == \
        [(r["steps"], r["evaluations"], r["rejections"]) for r in b[3]]
    assert sum(r["rejections"] for r in a[3]) > 0      # the reject path was exercised


def test_size_group_launches_match_solo(vb, gpu):
    # solves of one size group share every launch (grid.z over the solves);
    # each must still compute exactly what it computes alone
    specs = []
    for j, (e, m1, m2) in enumerate([(1.5, 1e3, 2e4), (2.0, 1e4, 1e4), (3.0, 3e4, 1e3), (2.0, 5e3, 5e4)]):
        specs.append(dict(nx=32, kind=(1, 3, 4, 1)[j], seed=60 + j, precision=1, M1=m1, M2=m2,
                          eps=e * 2 * np.pi / 32, phi_lut_mode=1, time_step=2e-7, fixed_time_step=2e-7))
    for j in range(3):
        specs.append(dict(nx=48, nz=32, kind=2, seed=70 + j, precision=0, M1=1e4 * (j + 1), M2=1e4))
    for j in range(2):
        specs.append(dict(nx=32, kind=1, seed=80 + j, precision=1, M1=0.0, M2=0.0))
    ps, phis = [], []
    for sp in specs:
        sp = dict(sp)
        p, phi = case(vb, sp.pop("nx"), sp.pop("nz", None), sp.pop("kind"), sp.pop("seed"), **sp)
        ps.append(p)
        phis.append(phi)
    flush_print("grouped batch of", len(ps))
    b = vb.Batch(ps, phis)
    reps = b.step(7) + b.step(6)
    got = [(b.get_phi(i), b.query(i), b.report(i)) for i in range(len(ps))]
    b.close()
    for i, (p, phi) in enumerate(zip(ps, phis)):
        assert np.isfinite(got[i][0]).all() and reps[i]["status"] == 0, i
        ref_phi, ref_state, ref_q, ref_reps = solo(vb, p, phi, [7, 6], graphs=0)
        assert np.array_equal(got[i][0], ref_phi), i
        assert got[i][1] == ref_q, i
        assert got[i][2]["time_current"] == ref_state["time_current"], i
        assert reps[i]["evaluations"] == ref_reps[0]["evaluations"], i


def test_batch_large_sizes_match_solo(vb, gpu):
    # the 128^3 and 192^3 size groups of the batched configuration: each solve of
    # a shared launch computes exactly what it computes alone
    specs = [dict(nx=128, kind=1, seed=90, M1=1e4, M2=3e4), dict(nx=128, kind=3, seed=91, M1=3e3, M2=1e4),
             dict(nx=192, kind=4, seed=92, M1=1e4, M2=1e4)]
    ps, phis = [], []
    for sp in specs:
        sp = dict(sp)
        p, phi = case(vb, sp.pop("nx"), None, sp.pop("kind"), sp.pop("seed"), precision=1, phi_lut_mode=1,
                      time_step=1e-7, fixed_time_step=1e-7, **sp)
        ps.append(p)
        phis.append(phi)
    flush_print("large batch of", len(ps))
    b = vb.Batch(ps, phis)
    reps = b.step(4)
    got = [(b.get_phi(i), b.query(i)) for i in range(len(ps))]
    b.close()
    for i, (p, phi) in enumerate(zip(ps, phis)):
        assert reps[i]["status"] == 0 and reps[i]["evaluations"] == 8, i
        ref_phi, _, ref_q, _ = solo(vb, p, phi, [4], graphs=0)
        assert np.array_equal(got[i][0], ref_phi), i
        assert got[i][1] == ref_q, i


def test_batched_mix_48_solves(vb, gpu):
    # the batched configuration as stated: 16 x 32^3, 16 x 64^3, 12 x 128^3 and
    # 4 x 192^3 FP32-mixed ETDRK4 solves with the penalty drift, 300 steps each in
    # one batch; every solve ends without a driver exit, and one solve of every
    # size group matches its own solo run bit for bit
    from cudabubble_amd.workloads import batch_mix
    ps, phis = batch_mix(vb)
    assert len(ps) == 48
    flush_print("48-solve mix: building")
    b = vb.Batch(ps, phis)
    reps = b.step(300)
    assert all(r["status"] == 0 and r["steps"] == 300 and r["evaluations"] == 600 for r in reps), \
        [(i, r) for i, r in enumerate(reps) if r["status"] != 0 or r["steps"] != 300][:4]
    first = {}
    for i, p in enumerate(ps):
        first.setdefault(p.nx, i)
    got = {n: (b.get_phi(i), b.query(i)) for n, i in first.items()}
    b.close()
    for n, i in sorted(first.items()):
        flush_print("48-solve mix: solo run of solve", i, "n =", n)
        s = vb.Solver(ps[i], phis[i])
        assert s.step(300)["status"] == 0
        assert np.array_equal(s.get_phi(), got[n][0]), n
        assert s.query() == got[n][1], n
        s.close()


def test_batched_mix_48_solves_shared_stages(vb, gpu):
    # the same mix with the stage combines of the 32^3, 64^3 and 128^3 groups as
    # one launch each (default kernel schedules), 60 steps in uneven chunks:
    # every solve ends without a driver exit, and the first solve of every
    # sharing size matches its solo run bit for bit
    from cudabubble_amd.workloads import batch_mix
    ps, phis = batch_mix(vb)
    saved = {n: vb.schedule_get(n, vb.FP32_MIXED) for n in (32, 64, 128)}
    try:
        for n in saved:
            vb.schedule_set(n, vb.FP32_MIXED, 0)
        flush_print("48-solve mix, shared stage launches: building")
        b = vb.Batch(ps, phis)
        assert b.set_shared_stages(True) == 3
        chunks = [7, 30, 23]
        reps = [b.step(c) for c in chunks][-1]
        statuses = [b.report(i)["status"] for i in range(len(ps))]
        assert all(st == 0 for st in statuses), [(i, st) for i, st in enumerate(statuses) if st != 0][:6]
        first = {}
        for i, p in enumerate(ps):
            if p.nx in saved:
                first.setdefault(p.nx, i)
        got = {n: (b.get_phi(i), b.query(i)) for n, i in first.items()}
        b.close()
        for n, i in sorted(first.items()):
            flush_print("48-solve mix, shared stage launches: solo run of solve", i, "n =", n)
            s = vb.Solver(ps[i], phis[i])
            for c in chunks:
                assert s.step(c)["status"] == 0
            assert np.array_equal(s.get_phi(), got[n][0]), n
            assert s.query() == got[n][1], n
            s.close()
    finally:
        for n, sc in saved.items():
            vb.schedule_set(n, vb.FP32_MIXED, sc)


def test_shared_stage_launches_match_solo(vb, gpu):
    # solves of four lengths whose ETD stage combines go out as one launch per
    # stage over all of them (vb_batch_set_shared_stages), beside a size that
    # a shared launch does not carry (48 x 32): every solve computes exactly
    # what it computes alone, through the first steps (per-solve launches until
    # the pipelines recur) and the captured shared graph
    specs = [dict(nx=32, kind=1, seed=100, M1=1e4, M2=3e4), dict(nx=32, kind=3, seed=101, M1=3e3, M2=1e4),
             dict(nx=64, kind=4, seed=102, M1=1e4, M2=1e4), dict(nx=64, kind=1, seed=103, M1=2e4, M2=5e3),
             dict(nx=128, kind=2, seed=104, M1=1e4, M2=1e4), dict(nx=256, kind=3, seed=105, M1=1e4, M2=1e4),
             dict(nx=48, nz=32, kind=1, seed=106, M1=1e4, M2=1e4)]
    ps, phis = [], []
    for sp in specs:
        sp = dict(sp)
        p, phi = case(vb, sp.pop("nx"), sp.pop("nz", None), sp.pop("kind"), sp.pop("seed"), precision=1,
                      phi_lut_mode=1, time_step=1e-7, fixed_time_step=1e-7, **sp)
        ps.append(p)
        phis.append(phi)
    flush_print("shared stage launches: batch of", len(ps))
    b = vb.Batch(ps, phis)
    assert b.set_shared_stages(True) == 4          # the 32, 64, 128 and 256 groups
    # odd chunks: per-solve steps on the groups' streams before and after the
    # shared graph's launches on the first member's stream
    reps = b.step(5) + b.step(3) + b.step(1)
    got = [(b.get_phi(i), b.query(i), b.report(i)) for i in range(len(ps))]
    b.close()
    for i, (p, phi) in enumerate(zip(ps, phis)):
        flush_print("shared stage launches: solo run", i)
        assert reps[i]["status"] == 0 and reps[i]["steps"] == 5, i
        ref_phi, ref_state, ref_q, _ = solo(vb, p, phi, [5, 3, 1], graphs=0)
        assert np.array_equal(got[i][0], ref_phi), i
        assert got[i][1] == ref_q, i
        assert got[i][2]["time_current"] == ref_state["time_current"], i
