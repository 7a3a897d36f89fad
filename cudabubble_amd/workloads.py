"""Synthetic workloads of the benchmark configurations, shared by bench.py and
the GPU tests: the seeded fields and parameters of the precision form (one
256^3 torus), of the batched form (48 solves of uneven size) and of the
serving form (a Poisson request stream).  Host-side set-up only."""
import numpy as np

BATCH_GROUPS = ((32, 16), (64, 16), (128, 12), (192, 4))    # (n, solves)
SERVE_SIZES = ((64, 0.50), (128, 0.35), (256, 0.14), (512, 0.01))


def precision_case(vb, n, precision, lut_mode, seed=20260003):
    """n^3 torus (R ~ U[1.2, 1.8], r ~ U[0.4, 0.7]) with a tanh profile plus
    N(0, 1e-3) noise, eps = 2h, ETDRK4 with fixed dt = 5e-7, M1 = M2 = 1e4,
    alpha/beta from t = 0 (the penalty drift: two evaluations per step)."""
    eps = 2.0 * (2 * np.pi) / n
    phi, sp = vb.shape(vb.SHAPE_TORUS, seed, n, n, 1.0, eps, noise=1e-3)
    p = vb.params_default(nx=n, nz=n, eps=eps, M1=1e4, M2=1e4, alpha=0.0, beta=0.0,
                          time_step=5e-7, fixed_time_step=5e-7, precision=precision,
                          phi_lut_mode=lut_mode, scheme=vb.ETDRK4)
    return p, phi, sp


def bending_energy(phi, eps, L):
    """E_b = (1/2eps) int f^2, f = -eps lap phi + (phi^2-1) phi / eps, by numpy FFT."""
    n = phi.shape[-1]
    k = np.fft.fftfreq(n, 1.0 / n) * (2 * np.pi / L)
    kz = np.fft.fftfreq(phi.shape[0], 1.0 / phi.shape[0]) * (2 * np.pi / L)
    k2 = kz[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    lap = np.real(np.fft.ifftn(-k2 * np.fft.fftn(phi)))
    f = -eps * lap + (phi * phi - 1) * phi / eps
    return float(np.mean(f * f)) * L ** 3 / (2 * eps)


def batch_mix(vb, seed=20260004, groups=BATCH_GROUPS):
    """The batched form: FP32-mixed ETDRK4 solves of uneven size (16 x 32^3,
    16 x 64^3, 12 x 128^3, 4 x 192^3).  Solve j is an ellipsoid, torus or
    biconcave disk (in turn) drawn by vb_shape from seed + j; eps in
    {1.5h, 2h, 3h} and M1, M2 log-uniform in [1e3, 1e5] come from numpy's
    PCG64(seed); fixed dt with the penalty drift (two evaluations per step).
    The penalties enter N explicitly, so their stiffness bounds dt: the volume
    mode relaxes at rate M1 L^3 / 4 and the area mode at about
    M2 (9/8) 2 eps E_b, and the reference scheme diverges once dt times their
    sum exceeds about 1 (torus and disk shapes at M2 ~ 1e5 have E_b ~ 1e2-1e3).
    Each solve takes dt = min(1e-7, 0.5 / that sum) from its initial field.
    Returns the parameter structs and the initial fields."""
    rng = np.random.default_rng(seed)
    ps, phis = [], []
    j = 0
    for n, count in groups:
        for _ in range(count):
            kind = (vb.SHAPE_ELLIPSOID, vb.SHAPE_TORUS, vb.SHAPE_DISK)[j % 3]
            eps = (1.5, 2.0, 3.0)[int(rng.integers(3))] * (2 * np.pi) / n
            M1, M2 = 10.0 ** rng.uniform(3, 5, size=2)
            phi, _ = vb.shape(kind, seed + j, n, n, 1.0, eps)
            L = 2 * np.pi
            stiff = M1 * L ** 3 / 4 + M2 * 9.0 / 8.0 * 2 * eps * bending_energy(phi, eps, L)
            dt = min(1e-7, 0.5 / stiff)
            ps.append(vb.params_default(nx=n, nz=n, eps=eps, M1=M1, M2=M2, alpha=0.0, beta=0.0,
                                        time_step=dt, fixed_time_step=dt, precision=vb.FP32_MIXED,
                                        phi_lut_mode=vb.LUT_ACCURATE, scheme=vb.ETDRK4))
            phis.append(phi)
            j += 1
    return ps, phis


def serve_stream(vb, seconds, rate=20.0, seed=20260005):
    """The serving form's arrivals: Poisson at `rate` requests/s for `seconds`,
    grid sizes drawn from {64^3: 50%, 128^3: 35%, 256^3: 14%, 512^3: 1%},
    shapes and parameters drawn as in the batched form.  Returns the arrival
    list [(t, n, kind, eps, M1, M2)] and the numpy generator (for the sample)."""
    rng = np.random.default_rng(seed)
    t_arr, reqs = 0.0, []
    while True:
        t_arr += rng.exponential(1.0 / rate)
        if t_arr > seconds:
            break
        n = int(rng.choice([s for s, _ in SERVE_SIZES], p=[q for _, q in SERVE_SIZES]))
        j = len(reqs)
        kind = (vb.SHAPE_ELLIPSOID, vb.SHAPE_TORUS, vb.SHAPE_DISK)[j % 3]
        eps = (1.5, 2.0, 3.0)[int(rng.integers(3))] * (2 * np.pi) / n
        M1, M2 = 10.0 ** rng.uniform(3, 5, size=2)
        reqs.append((t_arr, n, kind, eps, M1, M2))
    return reqs, rng


def serve_request(vb, req, j, precision, seed=20260005):
    """Parameters and field of request j: adaptive-dt ETDRK4 from dt = 1e-7,
    tmax = 1e-3 (its stopping rule is the server's)."""
    _, n, kind, eps, M1, M2 = req
    phi, _ = vb.shape(kind, seed + j, n, n, 1.0, eps)
    p = vb.params_default(nx=n, nz=n, eps=eps, M1=M1, M2=M2, alpha=0.0, beta=0.0, time_step=1e-7,
                          fixed_time_step=0.0, precision=precision, phi_lut_mode=vb.LUT_ACCURATE,
                          scheme=vb.ETDRK4, tmax=1e-3)
    return p, phi
