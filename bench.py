"""Benchmark: ETDRK4 grid-point-steps/s at 256^3, FP32-mixed vs FP64.

Workload (the precision configuration): 256^3 grid on [-pi, pi]^3, eps = 2h,
seeded torus (R ~ U[1.2, 1.8], r ~ U[0.4, 0.7]) with a tanh profile plus
N(0, 1e-3) noise, ETDRK4 with fixed dt = 5e-7, M1 = M2 = 1e4, alpha/beta from
t = 0.  With fixed dt and penalties on, one driver step of the reference is two
ETDRK4 evaluations with the alpha/beta drift correction (belastic.cpp:670-691).

A "step" below is one driver step; the unit of work is one ETDRK4 evaluation
on one grid point, so value = 256^3 * evaluations / second (FP32-mixed: FP32
FFTs and pointwise stages, FP64 coefficient tables, reductions and penalty
logic).  The FP64 rate, the FP32-vs-FP64 error after the same steps, the
per-step latency percentiles, the roofline of the dominant kernel and a CPU
baseline (the FP64 oracle on the host cores, a bounded sample) are reported
in the same JSON line.  Before the FP32-mixed solves of a grid size are built,
vb_autotune times the kernel schedules compiled for that size and selects the
fastest (--no-tune keeps schedule 0); the line reports its timings and choice.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ETDRK4 grid-pt-steps/s at 256³ (FP32-mixed vs FP64) + p50/p99 sweep latency"
HBM_PEAK = 8.0e12    # MI355X HBM3E, bytes/s


def alg_bytes(name, n, prec_bytes):
    """Algorithmic bytes of one launch of a pipeline kernel: reads + writes of
    its arrays (coefficient tables and twiddles excluded as cache-resident).
    Work spectra are complex of the work precision (cs bytes per mode); the
    state spectrum is complex FP64 (16 bytes per mode) in both precisions."""
    cs = 2 * prec_bytes
    modes = (n // 2 + 1) * n * n
    C = modes * cs
    R = n ** 3 * prec_bytes
    table = {
        "y_inv2": 4 * C, "y_fwd2": 4 * C, "y_fwd": 2 * C, "y_inv": 2 * C, "z_lap": 2 * C, "z_lap_fin": 2 * C,
        "z_fwd": 2 * C, "z_inv": 2 * C,
        # xa: P, Q in; spectra of tmp1 and s out (Euler: tmp1, f and the real v, f)
        "xa": 4 * C, "xa_f": 4 * C + 2 * R, "xa_sums": C,
        "zprep": modes * (16 + 2 * cs),
        # stage combines: spectra of tmp1 and s in, the stage field's spectrum
        # (state or Vh), stage/state spectra; next-stage field (Vh), P, Q out
        "zstage1": modes * (6 * cs + 16), "zstage2": modes * (7 * cs + 16),
        "zstage3": modes * 8 * cs, "zstage4": modes * (7 * cs + 32),
        # stage 4 of a drift step's first evaluation: no state out, P alone
        "zstage4p": modes * (6 * cs + 16),
    }
    return table.get(name)


def config_params(vb, n, precision, lut_mode):
    from cudabubble_amd.workloads import precision_case
    return precision_case(vb, n, precision, lut_mode)


def sync():
    import torch
    torch.cuda.synchronize()


def run_timed(vb, p, phi, warmup, steps):
    s = vb.Solver(p, phi)
    r = s.step(warmup) if warmup else None
    sync()
    t0 = time.perf_counter()
    r = s.step(steps)
    sync()
    t1 = time.perf_counter()
    assert r["status"] == 0, r
    return s, (t1 - t0), r["evaluations"]


def latency(s, nsteps):
    out = []
    for _ in range(nsteps):
        sync()
        t0 = time.perf_counter()
        s.step(1)
        sync()
        out.append((time.perf_counter() - t0) * 1e3)
    return out


def kernel_profile(vb, s, n, prec_bytes, steps=2):
    s.profile(True)
    s.step(steps)
    kt = s.kernel_times()
    s.profile(False)
    rows = []
    for name, (cnt, ms) in kt.items():
        b = alg_bytes(name, n, prec_bytes)
        rows.append((ms, name, cnt, b))
    rows.sort(reverse=True)
    return rows


def cpu_baseline(n, seconds=10.0):
    """The FP64 oracle (own FFT, OpenMP) on this host: ETDRK4 evaluations of the
    same n^3 field, repeated until about `seconds` of CPU work have been timed."""
    from oracle import pyoracle
    import cudabubble_amd as vb
    eps = 2.0 * (2 * np.pi) / n
    phi, _ = vb.shape(vb.SHAPE_TORUS, 20260003, n, n, 1.0, eps, noise=1e-3)
    o = pyoracle.Oracle(n, eps=eps, M1=1e4, M2=1e4)
    o.set_targets(o.vol(phi), o.surface(phi))
    d = o.etd4rk(phi, 5e-7)          # untimed: first touch of the oracle's buffers
    evals, dt = 0, 0.0
    while dt < seconds:
        t0 = time.perf_counter()
        d = o.etd4rk(d, 5e-7)
        dt += time.perf_counter() - t0
        evals += 1
    # and one evaluation on one core (the oracle's sums do not depend on the thread count)
    lib = pyoracle.lib()
    cores = lib.vbo_threads()
    lib.vbo_set_threads(1)
    t0 = time.perf_counter()
    d = o.etd4rk(d, 5e-7)
    t1 = time.perf_counter() - t0
    lib.vbo_set_threads(cores)
    return {"value": n ** 3 * evals / dt, "unit": "grid-point-steps/s", "cores": cores,
            "kind": "port",
            "sample": "%d ETDRK4 evaluations (reference dataflow, 36 FFTs each) of the same %d^3 torus in FP64 "
                      "by the CPU oracle, %.1f s timed" % (evals, n, dt),
            "single_core": {"value": n ** 3 / t1, "cores": 1,
                            "sample": "1 ETDRK4 evaluation of the same field on one thread, %.1f s" % t1}}


def batched(vb, steps, seed=20260004):
    """The batched form: 48 FP32-mixed ETDRK4 solves of uneven size resident
    together on the GPU (16 x 32^3, 16 x 64^3, 12 x 128^3, 4 x 192^3), fixed dt
    with the penalty drift; fields, parameters and per-solve dt as in
    cudabubble_amd.workloads.batch_mix."""
    from cudabubble_amd.workloads import BATCH_GROUPS, batch_mix
    ps, phis = batch_mix(vb, seed)

    def run(shared):
        b = vb.Batch(ps, phis)
        members = b.set_shared_stages(shared)
        b.step(4)     # start-up: alpha/beta, tables, graph capture
        sync()
        t0 = time.perf_counter()
        reps = b.step(steps)
        sync()
        dt = time.perf_counter() - t0
        b.close()
        return sum(p.nx * p.nx * p.nz * r["evaluations"] for p, r in zip(ps, reps)) / dt, dt, reps, members

    value, dt, reps, _ = run(False)
    # the ETD stage combines of the 32^3, 64^3 and 128^3 groups as one launch
    # each; a shared launch runs every size's default kernel schedule, so the
    # tuned schedules of those sizes are set aside for this run
    tuned = {n: vb.schedule_get(n, vb.FP32_MIXED) for n in (32, 64, 128)}
    for n in tuned:
        vb.schedule_set(n, vb.FP32_MIXED, 0)
    shared, _, reps_s, members = run(True)
    for n, sc in tuned.items():
        vb.schedule_set(n, vb.FP32_MIXED, sc)
    return {"value": value, "unit": "grid-point-steps/s", "solves": len(ps), "steps": steps,
            "not_ok": [i for i, r in enumerate(reps) if r["status"] != 0],
            "dt_range": [min(p.fixed_time_step for p in ps), max(p.fixed_time_step for p in ps)],
            "seconds": dt, "grid_points": int(sum(p.nx * p.nx * p.nz for p in ps)),
            "groups": ["%d x %d^3" % (c, n) for n, c in BATCH_GROUPS],
            "shared_stage_launches": {"value": shared, "size_groups_sharing": members,
                                      "not_ok": [i for i, r in enumerate(reps_s) if r["status"] != 0]}}


def group_alone(n, steps):
    """One size group of the batched form alone (its solves as in
    cudabubble_amd.workloads.batch_mix), on the default kernel schedules: the
    driver of the groups' kernel traces.  Prints one JSON line."""
    import cudabubble_amd as vb
    from cudabubble_amd.workloads import BATCH_GROUPS, batch_mix
    if n not in [m for m, _ in BATCH_GROUPS]:
        print("the batched form has no %d^3 solves (its sizes: %s)" % (n, ", ".join(str(m) for m, _ in BATCH_GROUPS)),
              file=sys.stderr)
        sys.exit(2)
    ps, phis = batch_mix(vb)
    sel = [i for i, p in enumerate(ps) if p.nx == n]
    b = vb.Batch([ps[i] for i in sel], [phis[i] for i in sel])
    b.step(4)     # start-up: alpha/beta, tables, graph capture
    sync()
    t0 = time.perf_counter()
    reps = b.step(steps)
    sync()
    dt = time.perf_counter() - t0
    b.close()
    print(json.dumps({"group": "%d x %d^3" % (len(sel), n), "steps": steps, "ms_per_step": 1e3 * dt / steps,
                      "gp_steps_per_s": sum(n ** 3 * r["evaluations"] for r in reps) / dt,
                      "not_ok": [i for i, r in zip(sel, reps) if r["status"] != 0]}), flush=True)


def converge_run(vb, p, phi, window=50, tol=1e-6):
    """One request alone through vb_step, with the server's stopping rule:
    relative energy drop over `window` accepted steps below tol, t >= tmax or a
    driver exit.  Returns the solver and the accepted dt of every step."""
    s = vb.Solver(p, phi)
    hist, dts = [], []
    while True:
        r = s.step(1)
        if r["status"] != 0:
            break
        dts.append(r["time_step"])
        hist.append(r["energy"])
        if r["time_current"] >= p.tmax:
            break
        if len(hist) > window and abs(hist[-window - 1] - hist[-1]) <= tol * abs(hist[-window - 1]):
            break
    return s, dts


def serving(vb, seconds, rate=20.0, seed=20260005, fp64_sample=0.05, sample_cap=None):
    """The serving form: Poisson arrivals at `rate` requests/s for `seconds`,
    grid sizes drawn from {64^3: 50%, 128^3: 35%, 256^3: 14%, 512^3: 1%},
    shapes and parameters drawn as in the batched form, FP32-mixed adaptive
    ETDRK4 from dt = 1e-7 until the relative energy drop over 50 accepted steps
    is below 1e-6 or t >= tmax = 1e-3.  Latency is submit -> result.

    A 5% sample is re-run in FP64, each request alone, two ways: on the FP32-mixed
    run's own accepted dt sequence (vb_evaluate at each accepted dt from the
    same start: the two differ by rounding only), and with its own adaptive dt
    (the FP64 path can accept and reject differently near an equilibrium)."""
    from cudabubble_amd.workloads import serve_request, serve_stream
    reqs, rng = serve_stream(vb, seconds, rate, seed)

    srv = vb.Server(max_active=16, streams=4, converge_window=50, converge_tol=1e-6)
    results, t0 = {}, time.perf_counter()
    t_note = [t0]

    def progress(submitted):
        # a line on stderr every 20 s, so that a long stream shows it is alive
        if time.perf_counter() - t_note[0] >= 20.0:
            t_note[0] = time.perf_counter()
            print("serving: %.0f s, %d of %d submitted, %d done" % (t_note[0] - t0, submitted, len(reqs),
                                                                     len(results)), file=sys.stderr, flush=True)

    for j, req in enumerate(reqs):
        p, phi = serve_request(vb, req, j, vb.FP32_MIXED, seed)     # prepared before its arrival time
        while time.perf_counter() - t0 < req[0]:
            for r in srv.poll():
                results[r["id"]] = r
            progress(j)
            time.sleep(0.0005)
        srv.submit(j, p, phi)
    while True:
        for r in srv.poll():
            results[r["id"]] = r
        st = srv.pending()
        if st["queued"] == 0 and st["running"] == 0:
            break
        progress(len(reqs))
        time.sleep(0.002)
    for r in srv.poll():
        results[r["id"]] = r
    wall = time.perf_counter() - t0
    srv.close()
    lat = np.array([r["t_done"] - r["t_submit"] for r in results.values()])
    wait = np.array([r["t_admit"] - r["t_submit"] for r in results.values()])
    gp_evals = sum(reqs[i][1] ** 3 * r["evaluations"] for i, r in results.items())
    span = max(r["t_done"] for r in results.values()) - min(r["t_submit"] for r in results.values())
    sample = sorted(rng.choice(len(reqs), size=max(1, int(round(fp64_sample * len(reqs)))), replace=False))
    if sample_cap:
        sample = sample[:sample_cap]
    errs = []
    for j in sample:
        j = int(j)
        q = results[j]
        print("serving: FP64 re-run of request %d (%d^3, %d steps)" % (j, reqs[j][1], q["steps"]),
              file=sys.stderr, flush=True)
        p32, phi = serve_request(vb, reqs[j], j, vb.FP32_MIXED, seed)
        s32, dts = converge_run(vb, p32, phi)
        alone = s32.query()
        s32.close()
        p64, _ = serve_request(vb, reqs[j], j, vb.FP64, seed)
        m = vb.Solver(p64, phi)
        m.step(0)                          # the driver start: alpha/beta from the field
        for h in dts:
            m.evaluate(h)
        Em, Vm, Am, Ebm = m.query()
        m.close()
        s64, dts64 = converge_run(vb, p64, phi)
        E, V, A, Eb = s64.query()
        s64.close()
        errs.append({"id": j, "n": reqs[j][1], "steps_fp32": q["steps"], "steps_fp64": len(dts64),
                     "served_equals_alone": [q["E"], q["V"], q["A"], q["Eb"]] == list(alone),
                     "matched_path": {"E_rel": abs(q["E"] - Em) / abs(Em), "V_rel": abs(q["V"] - Vm) / abs(Vm),
                                      "A_rel": abs(q["A"] - Am) / abs(Am), "Eb_rel": abs(q["Eb"] - Ebm) / abs(Ebm)},
                     "own_path": {"E_rel": abs(q["E"] - E) / abs(E), "V_rel": abs(q["V"] - V) / abs(V),
                                  "A_rel": abs(q["A"] - A) / abs(A)}})
    sizes = {}
    for i in results:
        sizes[reqs[i][1]] = sizes.get(reqs[i][1], 0) + 1
    status = {}
    for r in results.values():
        status[str(r["status"])] = status.get(str(r["status"]), 0) + 1
    worst = {k: max(e["matched_path"][k] for e in errs) for k in ("E_rel", "V_rel", "A_rel", "Eb_rel")} if errs else {}
    return {"requests": len(reqs), "completed": sum(1 for r in results.values() if r["status"] == 0),
            "status_counts": status, "arrival_seconds": seconds, "rate_per_s": rate,
            "sizes": {str(k): v for k, v in sorted(sizes.items())},
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
            "queue_wait_s": {"p50": float(np.percentile(wait, 50)), "p99": float(np.percentile(wait, 99))},
            "aggregate_gp_steps_per_s": gp_evals / span, "wall_s": wall,
            "mean_steps": float(np.mean([r["steps"] for r in results.values()])),
            "fp64_sample_worst_matched_path": worst, "fp64_sample": errs}


def copy_calibration():
    """A plain device copy of known size (256 MiB read + 256 MiB written), the
    reference point for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "repeats": 5}))


def copy_bandwidth(repeats=20):
    """Read + write bytes per second of a plain 1 GiB device copy (torch's copy
    kernel): the practical HBM ceiling of a streaming kernel on this device,
    reported beside the 8 TB/s peak."""
    import torch
    n = 1024 * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(repeats):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / repeats
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 4 / sec


def driver_forms():
    """Wall time per attempt of the driver forms on one solve: adaptive ETD with
    host decisions vs chains decided on the device (accept_chain 4 and 16), and
    Euler attempts eager vs in graph form (graphs = 2); FP32-mixed, one JSON
    line per grid.  Each pair is also checked to end in the same field."""
    import time
    import numpy as np
    import cudabubble_amd as vb

    def adaptive(n, ac, steps):
        eps = 2.0 * 2 * np.pi / n
        phi, _ = vb.shape(3, 20260002, n, n, 1.0, eps)
        p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0, time_step=1e-7,
                              precision=vb.FP32_MIXED, accept_chain=ac)
        s = vb.Solver(p, phi)
        s.step(steps)                 # graphs and tables in place
        t0 = time.perf_counter()
        r = s.step(steps)
        sec = time.perf_counter() - t0
        out = s.get_phi()
        s.close()
        return 1e6 * sec / r["evaluations"], r, out

    def euler(n, graphs):
        eps = 2.0 * 2 * np.pi / n
        phi, _ = vb.shape(2, 20260002, n, n, 1.0, eps)
        p = vb.params_default(nx=n, nz=n, eps=eps, scheme=vb.EULER, M1=1e3, M2=1e3, alpha=-1.0, beta=-1.0,
                              time_step=1e-7, fixed_time_step=0.0, precision=vb.FP32_MIXED, graphs=graphs)
        s = vb.Solver(p, phi)
        s.step(20)
        t0 = time.perf_counter()
        r = s.step(200)
        sec = time.perf_counter() - t0
        out = s.get_phi()
        s.close()
        return 1e6 * sec / r["evaluations"], out

    def shape(n, graphs, ac):
        # the shape-adjust pre-pass alone (the driver start of a fixed-dt run)
        eps = 2.0 * 2 * np.pi / n
        phi, _ = vb.shape(1, 20260002, n, n, 1.0, eps)
        p = vb.params_default(nx=n, nz=n, eps=eps, shape_adjust=1, M1=1e4, M2=1e4, alpha=-1.0, beta=-1.0,
                              time_step=1e-6, fixed_time_step=1e-6, precision=vb.FP32_MIXED, graphs=graphs,
                              accept_chain=ac)
        s = vb.Solver(p, phi)
        t0 = time.perf_counter()
        s.step(0)
        sec = time.perf_counter() - t0
        out = s.get_phi()
        s.close()
        return sec, out

    for n in (32, 64, 128, 256):
        steps = 64 if n <= 128 else 24
        line = {"grid": n, "precision": "fp32_mixed", "adaptive_iterations": steps}
        base = None
        for ac in (0, 4, 16):
            us, r, out = adaptive(n, ac, steps)
            line["us_per_attempt_chain%d" % ac] = round(us, 1)
            if base is None:
                base = (r, out)
                line["attempts"], line["rejections"] = r["evaluations"], r["rejections"]
            else:
                line["same_chain%d" % ac] = bool(np.array_equal(out, base[1]) and r == base[0])
        if n <= 128:
            e0, o0 = euler(n, 0)
            e2, o2 = euler(n, 2)
            line["euler_us_per_attempt_eager"] = round(e0, 1)
            line["euler_us_per_attempt_graph"] = round(e2, 1)
            line["euler_same"] = bool(np.array_equal(o0, o2))
            shape(n, 1, 8)   # warm: the kernels' first launches in the process (each solve captures its own graphs)
            se, so = shape(n, 1, 0)
            sg, sgo = shape(n, 2, 0)
            sc, sco = shape(n, 1, 8)
            line["shape_adjust_ms_eager"] = round(1e3 * se, 2)
            line["shape_adjust_ms_graph"] = round(1e3 * sg, 2)
            line["shape_adjust_ms_device_decisions"] = round(1e3 * sc, 2)
            line["shape_same"] = bool(np.array_equal(so, sgo) and np.array_equal(so, sco))
        print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--latency-steps", type=int, default=20)
    ap.add_argument("--batch-steps", type=int, default=300,
                    help="steps of the 48-solve batched form (0: skip it)")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="arrival window of the serving form at 20 requests/s (0: skip it); the "
                         "configuration states 120 s, the default is a reduced window")
    ap.add_argument("--serve-sample-cap", type=int, default=0,
                    help="at most this many requests of the 5%% FP64 sample (0: all of them)")
    ap.add_argument("--precision-steps", type=int, default=500,
                    help="driver steps after which the FP32-mixed state is compared with the FP64 one")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep schedule 0 instead of autotuning the kernel schedules of each grid size")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (counter calibration)")
    ap.add_argument("--group-alone", type=int, default=0, metavar="N",
                    help="run only the batched form's N^3 size group alone for --steps steps "
                         "(default schedules; the driver of the groups' kernel traces)")
    ap.add_argument("--driver-forms", action="store_true",
                    help="run only the per-attempt timing of the driver forms (host vs device "
                         "decisions, eager vs graph Euler attempts)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.driver_forms:
        driver_forms()
        return
    if args.group_alone:
        group_alone(args.group_alone, args.steps)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver path runs on a single device and "
              "independent solves shard trivially; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import cudabubble_amd as vb
    vb.lib()
    n = args.n
    tuned = {}

    def tune(size):
        # vb_autotune: time every compiled kernel schedule of this size and keep the fastest
        if not args.no_tune and "%d" % size not in tuned:
            r = vb.autotune(size, vb.FP32_MIXED, steps=4)
            tuned["%d" % size] = {"chosen": r["chosen"], "ms_per_step": [round(x, 4) for x in r["ms_per_step"]],
                                  "schedules": r["schedules"]}

    # FP32-mixed (headline)
    tune(n)
    p32, phi, sp = config_params(vb, n, vb.FP32_MIXED, vb.LUT_ACCURATE)
    s32, t32, ev32 = run_timed(vb, p32, phi, args.warmup, args.steps)
    value = n ** 3 * ev32 / t32
    ms_step = t32 * 1e3 / args.steps
    lat = latency(s32, args.latency_steps)
    prof32 = kernel_profile(vb, s32, n, 4)

    out_extra = {}
    if not args.no_fp64:
        p64, _, _ = config_params(vb, n, vb.FP64, vb.LUT_ACCURATE)
        s64, t64, ev64 = run_timed(vb, p64, phi, args.warmup, args.steps)
        v64 = n ** 3 * ev64 / t64
        # both solvers continue (untimed) to the precision configuration's length
        done32 = args.warmup + args.steps + args.latency_steps + 2
        done64 = args.warmup + args.steps
        done = max(done32, args.precision_steps)
        s32.step(done - done32)
        s64.step(done - done64)
        done32 = done
        phi32, phi64 = s32.get_phi(), s64.get_phi()
        E32, V32, A32, Eb32 = s32.query()
        E64, V64, A64, Eb64 = s64.query()
        st64 = s64.state()
        rl2 = float(np.linalg.norm(phi32 - phi64) / np.linalg.norm(phi64))
        prof64 = kernel_profile(vb, s64, n, 8)
        out_extra["fp64"] = {"value": v64, "ms_per_step": t64 * 1e3 / args.steps,
                             "dominant_kernel": prof64[0][1]}
        out_extra["speedup_fp32_vs_fp64"] = value / v64
        # E includes the penalty targets, which the drift correction integrates from V and A;
        # E_same_targets evaluates the FP32-mixed state with the FP64 run's targets
        E32s = p64.eta * Eb32 + 0.5 * p64.M1 * (V32 - st64["alpha"]) ** 2 + 0.5 * p64.M2 * (A32 - st64["beta"]) ** 2
        out_extra["error_vs_fp64"] = {"steps": done32, "phi_rel_l2": rl2,
                                      "E_rel": abs(E32 - E64) / abs(E64),
                                      "E_rel_same_targets": abs(E32s - E64) / abs(E64),
                                      "V_rel": abs(V32 - V64) / abs(V64),
                                      "A_rel": abs(A32 - A64) / abs(A64), "Eb_rel": abs(Eb32 - Eb64) / abs(Eb64)}

    ms, name, cnt, b = prof32[0]
    avg = ms / cnt * 1e-3
    achieved = b / avg if b else None
    roof = {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9 if achieved else None,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None,
            "traffic": None, "avg_launch_us": avg * 1e6, "alg_bytes_per_launch": b}
    # the streaming ceiling: the library's own 16-byte copy kernels (best variant)
    # and torch's copy, each over 1 GiB
    copy_bw, copy_each = vb.copy_bandwidth(1 << 30, repeats=20)
    copy_bw *= 1e9
    roof["copy_GBps"] = copy_bw / 1e9
    roof["copy_variants_GBps"] = copy_each
    roof["copy_GBps_torch"] = copy_bandwidth() / 1e9
    roof["frac_of_copy"] = achieved / copy_bw if achieved else None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%d_fp32.json" % (name, n))
    if os.path.exists(tfile):
        with open(tfile) as f:
            roof["traffic"] = json.load(f).get("bytes_per_launch")
    # whole-evaluation roofline against the canonical fused dataflow (35 C + 44 R bytes)
    C4 = (n // 2 + 1) * n * n * 8
    R4 = n ** 3 * 4
    B_alg = 35 * C4 + 44 * R4
    t_eval = t32 / ev32
    # A penalty-drift step is two evaluations, and `value` counts both; the bytes
    # it needs are fewer: the second evaluation reuses the first stage's prefix
    # of the state (one N-evaluation, 5 C + 11 R) and the first evaluation's
    # result is used only for V and A (no state write, C)
    B_step = 2 * B_alg - (5 * C4 + 11 * R4) - C4
    t_step = t32 / args.steps
    out_extra["step_roofline"] = {
        "B_alg_per_evaluation": B_alg, "achieved_GBps": B_alg / t_eval / 1e9,
        "frac": B_alg / t_eval / HBM_PEAK, "frac_of_copy": B_alg / t_eval / copy_bw,
        "drift_step": {"B_needed": B_step, "achieved_GBps": B_step / t_step / 1e9, "frac": B_step / t_step / HBM_PEAK,
                       "note": "2 evaluations per drift step less the reused stage-1 prefix and the probe's "
                               "unwritten state"}}
    # the plain fixed-dt ETDRK4 step (targets fixed, no drift): one evaluation per step
    pp, _, _ = config_params(vb, n, vb.FP32_MIXED, vb.LUT_ACCURATE)
    pp.penalty_drift = 0
    sp_, tp, evp = run_timed(vb, pp, phi, args.warmup, args.steps)
    sp_.close()
    out_extra["plain_etdrk4_fp32"] = {"value": n ** 3 * evp / tp, "ms_per_step": tp * 1e3 / args.steps,
                                      "evaluations_per_step": evp / args.steps,
                                      "frac": B_alg / (tp / evp) / HBM_PEAK}
    out_extra["latency_ms_per_step"] = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                        "n": len(lat)}
    out_extra["kernels_fp32"] = [{"name": nm, "launches": c, "ms": round(m, 3),
                                  "GBps": (bb / (m / c * 1e-3) / 1e9) if bb else None}
                                 for m, nm, c, bb in prof32]
    # the single-solve measurements are done: free their HBM before the other forms
    s32.close()
    if not args.no_fp64:
        s64.close()
    if args.batch_steps > 0:
        from cudabubble_amd.workloads import BATCH_GROUPS
        for size, _ in BATCH_GROUPS:
            tune(size)
        bt = batched(vb, args.batch_steps)
        if "fp64" in out_extra:
            bt["ratio_vs_fp64_single_%d" % n] = bt["value"] / out_extra["fp64"]["value"]
        out_extra["batched"] = bt
    if args.serve_seconds > 0:
        sv = serving(vb, args.serve_seconds, sample_cap=args.serve_sample_cap or None)
        if args.serve_seconds < 120:
            sv["reduced"] = "%g s of arrivals instead of the configuration's 120 s" % args.serve_seconds
        out_extra["serving"] = sv
    cpu = None if args.no_cpu else cpu_baseline(n)

    line = {"metric": METRIC, "value": value, "unit": "grid-point-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "%d^3 torus (seed 20260003, noise 1e-3), eps=2h, ETDRK4 fixed dt=5e-7, "
                                   "M1=M2=1e4 penalty drift (2 evaluations per step, both counted in value: "
                                   "the first only feeds the drift of the targets), FP32-mixed" % n,
                       "grid": n, "evaluations_per_step": ev32 / args.steps, "shape": [round(x, 6) for x in sp[:2]],
                       "lut": "accurate", "precision": "fp32-mixed (f64 tables, reductions, penalty/dt logic)",
                       "parallelism": "1 GPU"},
            "roofline": roof, "cpu_baseline": cpu}
    line.update(out_extra)
    if tuned:
        line["autotune_fp32"] = tuned
    print(json.dumps(line))


if __name__ == "__main__":
    main()
