"""cudabubble_amd -- MI355X-native pseudo-spectral vesicle solver.

Python front end of libvesicle.so (C ABI in include/vesicle.h).  The solver
path is hand-written HIP for gfx950; this module only loads the library and
converts arguments.  There is no fallback: if the library is missing or no
GPU is present the calls raise.

Mirrors the reference's call surface (belastic / phasefield / Calc):
  Solver  -- belastic + phasefield: create from parameters and phi0, step the
             run_static_fEIF driver, query E/V/A, get/set phi, checkpoint.
  Calc    -- class Calc: individual transforms and kernels on device tensors.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvesicle.so")

VB_OK, VB_ERR_PARAM, VB_ERR_ODD_NX, VB_ERR_OOM, VB_ERR_HIP = 0, 1, 2, 3, 4
VB_ERR_NONFINITE, VB_ERR_DT_UNDERFLOW, VB_ERR_IO, VB_ERR_NO_DEVICE, VB_FINISHED = 5, 6, 7, 8, 9
EULER, ETDRK2, ETDRK4 = 1, 2, 4
FP64, FP32_MIXED = 0, 1
LUT_REFERENCE, LUT_ACCURATE = 0, 1
SHAPE_SPHERE, SHAPE_ELLIPSOID, SHAPE_TWO_ELLIPSOIDS, SHAPE_TORUS, SHAPE_DISK = 0, 1, 2, 3, 4
SUPPORTED_N = (16, 32, 48, 64, 96, 128, 192, 256, 384, 512)


class VbError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().vb_strerror(status).decode() if _LIB is not None else str(status)
        super().__init__("%s: %s (status %d)" % (what, msg, status))


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("scheme", C.c_int),
                ("len", C.c_double), ("eps", C.c_double), ("epssize", C.c_int),
                ("precision", C.c_int), ("eta", C.c_double), ("ga", C.c_double),
                ("M1", C.c_double), ("M2", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("kappa", C.c_double), ("D", C.c_double), ("time_step", C.c_double),
                ("fixed_time_step", C.c_double), ("dt_cap", C.c_double),
                ("dt_double_every", C.c_int), ("penalty_drift", C.c_int),
                ("phi_lut_mode", C.c_int), ("shape_adjust", C.c_int), ("tmax", C.c_double),
                ("device", C.c_int), ("graphs", C.c_int), ("accept_chain", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepReport(C.Structure):
    _fields_ = [("steps", C.c_int), ("evaluations", C.c_int), ("rejections", C.c_int),
                ("status", C.c_int), ("time_step", C.c_double), ("time_current", C.c_double),
                ("energy", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("max_active", C.c_int), ("streams", C.c_int), ("keep_phi", C.c_int),
                ("memory_budget", C.c_size_t), ("converge_window", C.c_int), ("max_steps", C.c_int),
                ("converge_tol", C.c_double)]


class Result(C.Structure):
    _fields_ = [("id", C.c_longlong), ("status", C.c_int), ("steps", C.c_int), ("evaluations", C.c_int),
                ("rejections", C.c_int), ("E", C.c_double), ("V", C.c_double), ("A", C.c_double),
                ("Eb", C.c_double), ("time_step", C.c_double), ("time_current", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


MAX_SCHEDULES = 4


class TuneReport(C.Structure):
    _fields_ = [("n", C.c_int), ("precision", C.c_int), ("steps", C.c_int), ("candidates", C.c_int),
                ("chosen", C.c_int), ("ms_per_step", C.c_double * MAX_SCHEDULES)]


_LIB = None


def lib():
    """Load libvesicle.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libvesicle.so is not built (run __graft_entry__.build())")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and
    # loads it by path, so a process that uses both must load torch first; the
    # library's DT_NEEDED libamdhip64.so.7 then binds to that same runtime
    # instead of starting a second one that would not see the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    I, D, P, S = C.c_int, C.c_double, C.c_void_p, C.c_size_t
    L.vb_version.restype = C.c_char_p
    L.vb_strerror.restype = C.c_char_p
    L.vb_strerror.argtypes = [I]
    L.vb_params_default.argtypes = [C.POINTER(Params)]
    L.vb_params_from_ini.argtypes = [C.c_char_p, C.POINTER(Params)]
    L.vb_eps_from_code.restype = D
    L.vb_eps_from_code.argtypes = [I, I, D]
    L.vb_create.argtypes = [C.POINTER(Params), P, C.POINTER(P)]
    L.vb_destroy.argtypes = [P]
    L.vb_step.argtypes = [P, I, C.POINTER(StepReport)]
    L.vb_query.argtypes = [P, C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(D)]
    L.vb_state.argtypes = [P, C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(D)]
    L.vb_get_phi.argtypes = [P, P]
    L.vb_set_phi.argtypes = [P, P]
    L.vb_evaluate.argtypes = [P, D]
    L.vb_residual.argtypes = [P, C.POINTER(D)]
    L.vb_shape_ninit.argtypes = [I, I, I, I, D, D, P]
    L.vb_read_field.argtypes = [C.c_char_p, I, I, I, P]
    L.vb_save.argtypes = [P, C.c_char_p]
    L.vb_load.argtypes = [P, C.c_char_p]
    L.vb_stream.restype = P
    L.vb_stream.argtypes = [P]
    L.vb_profile.argtypes = [P, I]
    L.vb_kernel_times.argtypes = [P, C.c_char_p, I, C.POINTER(I), C.POINTER(D), I]
    L.vb_footprint.restype = S
    L.vb_footprint.argtypes = [C.POINTER(Params)]
    L.vb_calc_create.argtypes = [I, I, D, I, C.POINTER(P)]
    L.vb_calc_destroy.argtypes = [P]
    L.vb_shape.argtypes = [I, C.c_ulonglong, I, I, D, D, D, P, P]
    L.vb_lut.argtypes = [I, I, D, D, D, D, D, I, P]
    L.vb_batch_create.argtypes = [I, C.POINTER(Params), C.POINTER(P), C.POINTER(P)]
    L.vb_batch_step.argtypes = [P, I]
    L.vb_batch_report.argtypes = [P, I, C.POINTER(StepReport)]
    L.vb_batch_query.argtypes = [P, I, C.POINTER(D), C.POINTER(D), C.POINTER(D), C.POINTER(D)]
    L.vb_batch_get_phi.argtypes = [P, I, P]
    L.vb_batch_set_shared_stages.argtypes = [P, I, C.POINTER(I)]
    L.vb_batch_destroy.argtypes = [P]
    L.vb_server_config_default.argtypes = [C.POINTER(ServerConfig)]
    L.vb_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(P)]
    L.vb_server_submit.argtypes = [P, C.c_longlong, C.POINTER(Params), P]
    L.vb_server_poll.argtypes = [P, C.POINTER(Result), I, C.POINTER(I)]
    L.vb_server_pending.argtypes = [P, C.POINTER(I), C.POINTER(I), C.POINTER(S)]
    L.vb_server_take_phi.argtypes = [P, C.c_longlong, P]
    L.vb_server_clock.restype = D
    L.vb_server_clock.argtypes = [P]
    L.vb_server_destroy.argtypes = [P]
    for name, nargs in [("vb_calc_r2c", 2), ("vb_calc_c2r", 2), ("vb_calc_laplace", 1),
                        ("vb_calc_lap_spec", 2), ("vb_calc_lap_real", 2),
                        ("vb_calc_fftgradient", 4), ("vb_calc_multi", 3), ("vb_calc_sync", 0)]:
        getattr(L, name).argtypes = [P] + [P] * nargs
    L.vb_calc_f.argtypes = [P, P, P, P, D]
    L.vb_calc_g0.argtypes = [P, P, P, P, D]
    L.vb_calc_intf.argtypes = [P, P, D, D, C.POINTER(D)]
    L.vb_calc_nld.argtypes = [P, P, P, P, P, D, D, D]
    L.vb_calc_s_term.argtypes = [P, P, P, P, P, P, D]
    L.vb_calc_plus.argtypes = [P, P, P, D]
    L.vb_calc_nld1.argtypes = [P, P, P, P, D, D, D]
    L.vb_calc_nld2.argtypes = [P, P, P, P, D, D, D]
    L.vb_calc_init_fEIF.argtypes = [P, P, D]
    L.vb_calc_etd1.argtypes = [P, P, P, P, D, D]
    L.vb_calc_etd2.argtypes = [P, P, P, P, P, D, D]
    L.vb_calc_etd41.argtypes = [P, P, P, P, D, D]
    L.vb_calc_etd42.argtypes = [P, P, P, P, D, D]
    L.vb_calc_etd43.argtypes = [P, P, P, P, D, D]
    L.vb_calc_etd44.argtypes = [P, P, P, P, P, P, D, D]
    L.vb_schedule_count.argtypes = [I, I]
    L.vb_schedule_get.argtypes = [I, I]
    L.vb_schedule_set.argtypes = [I, I, I]
    L.vb_schedule_describe.argtypes = [I, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
    L.vb_autotune.argtypes = [I, I, I, I, C.POINTER(TuneReport)]
    L.vb_copy_bandwidth.argtypes = [I, C.c_size_t, I, C.POINTER(D), C.POINTER(D), I]
    L.vb_intf_static.argtypes = [P, C.c_longlong, D, D, D, P, P]
    _LIB = L
    return L


def _check(rc, what):
    if rc != VB_OK:
        raise VbError(rc, what)


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def params_default(**kw):
    p = Params()
    lib().vb_params_default(C.byref(p))
    nx = kw.get("nx")
    for k, v in kw.items():
        setattr(p, k, v)
    if nx is not None:
        p.ny = nx
        if "nz" not in kw:
            p.nz = nx
        if "eps" not in kw:
            p.eps = eps_from_code(p.epssize, nx, p.len * 2 * np.pi)
    return p


def params_from_ini(path):
    p = Params()
    _check(lib().vb_params_from_ini(path.encode(), C.byref(p)), "vb_params_from_ini")
    return p


def eps_from_code(epssize, nx, L):
    return lib().vb_eps_from_code(int(epssize), int(nx), float(L))


def lut(scheme, lut_mode, h, eps, length=1.0, kappa=2.0, D=1.5, mmax=0):
    """Host-built ETD coefficient table [mmax+1][8] (see vb_lut in vesicle.h)."""
    out = np.zeros((mmax + 1, 8))
    _check(lib().vb_lut(int(scheme), int(lut_mode), float(h), float(eps), float(length), float(kappa),
                        float(D), int(mmax), _np_ptr(out)), "vb_lut")
    return out


def footprint(p):
    return int(lib().vb_footprint(C.byref(p)))


def shape(kind, seed, nx, nz=None, length=1.0, eps=None, noise=0.0):
    """Seeded initial field [nz][nx][nx] (float64) and its shape parameters."""
    nz = nx if nz is None else nz
    if eps is None:
        eps = 2.0 * (2 * np.pi * length) / nx
    phi = np.zeros((nz, nx, nx))
    sp = np.zeros(8)
    _check(lib().vb_shape(int(kind), C.c_ulonglong(seed), nx, nz, float(length), float(eps), float(noise),
                          _np_ptr(phi), _np_ptr(sp)), "vb_shape")
    return phi, sp


def shape_ninit(ninit, nx, nz=None, length=1.0, eps=None):
    """The reference's deterministic initial shape of ini field ninit (0, 1, 2, 4, 5, 6, 8,
    11, 50, 51)."""
    nz = nx if nz is None else nz
    if eps is None:
        eps = 2.0 * (2 * np.pi * length) / nx
    phi = np.zeros((nz, nx, nx))
    _check(lib().vb_shape_ninit(int(ninit), nx, nx, nz, float(length), float(eps), _np_ptr(phi)), "vb_shape_ninit")
    return phi


def read_field(path, nx, nz=None):
    """A raw [nz][ny][nx] field file; (nx+1) x (ny+1) x nz files are trimmed."""
    nz = nx if nz is None else nz
    phi = np.zeros((nz, nx, nx))
    _check(lib().vb_read_field(os.fsencode(path), nx, nx, nz, _np_ptr(phi)), "vb_read_field")
    return phi


HOST_PROGRAM = os.path.join(HERE, "cudabubble")


class Solver:
    """One vesicle solve (belastic + phasefield, belastic.cpp:15-23)."""

    def __init__(self, params, phi0):
        self.p = params
        phi0 = np.ascontiguousarray(phi0, dtype=np.float64)
        if phi0.shape != (params.nz, params.nx, params.nx):
            raise ValueError("phi0 shape %s != (nz, ny, nx) = %s" % (phi0.shape, (params.nz, params.nx, params.nx)))
        h = C.c_void_p()
        _check(lib().vb_create(C.byref(params), _np_ptr(phi0), C.byref(h)), "vb_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().vb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        r = StepReport()
        _check(lib().vb_step(self.h, int(n), C.byref(r)), "vb_step")
        return r.as_dict()

    def evaluate(self, h):
        _check(lib().vb_evaluate(self.h, float(h)), "vb_evaluate")

    def residual(self):
        """LSerror of the Euler right-hand side (belastic::geterror)."""
        e = C.c_double()
        _check(lib().vb_residual(self.h, C.byref(e)), "vb_residual")
        return e.value

    def query(self):
        E, V, A, Eb = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().vb_query(self.h, C.byref(E), C.byref(V), C.byref(A), C.byref(Eb)), "vb_query")
        return E.value, V.value, A.value, Eb.value

    def state(self):
        a, b, dt, t = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().vb_state(self.h, C.byref(a), C.byref(b), C.byref(dt), C.byref(t)), "vb_state")
        return {"alpha": a.value, "beta": b.value, "time_step": dt.value, "time_current": t.value}

    def get_phi(self):
        out = np.zeros((self.p.nz, self.p.nx, self.p.nx))
        _check(lib().vb_get_phi(self.h, _np_ptr(out)), "vb_get_phi")
        return out

    def set_phi(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        _check(lib().vb_set_phi(self.h, _np_ptr(phi)), "vb_set_phi")

    def save(self, d):
        _check(lib().vb_save(self.h, d.encode()), "vb_save")

    def load(self, d):
        _check(lib().vb_load(self.h, d.encode()), "vb_load")

    def stream(self):
        return lib().vb_stream(self.h)

    def profile(self, on=True):
        _check(lib().vb_profile(self.h, 1 if on else 0), "vb_profile")

    def kernel_times(self):
        cap, nl = 64, 32
        names = C.create_string_buffer(cap * nl)
        counts = (C.c_int * cap)()
        ms = (C.c_double * cap)()
        n = lib().vb_kernel_times(self.h, names, nl, counts, ms, cap)
        if n < 0:
            raise VbError(-n, "vb_kernel_times")
        out = {}
        for i in range(min(n, cap)):
            nm = names.raw[i * nl:(i + 1) * nl].split(b"\0")[0].decode()
            out[nm] = (counts[i], ms[i])
        return out


class Batch:
    """Many independent solves of any supported sizes on one GPU (vb_batch_*):
    resident per-solve state, a few streams, graph replays of fixed-dt steps."""

    def __init__(self, params_list, phi_list):
        n = len(params_list)
        self.params = list(params_list)
        arr = (Params * n)(*params_list)
        self._phis = [np.ascontiguousarray(p, dtype=np.float64) for p in phi_list]
        ptrs = (C.c_void_p * n)(*[p.ctypes.data for p in self._phis])
        h = C.c_void_p()
        _check(lib().vb_batch_create(n, arr, ptrs, C.byref(h)), "vb_batch_create")
        self.h = h

    def step(self, n=1):
        """n driver iterations of every solve; returns their step reports."""
        _check(lib().vb_batch_step(self.h, int(n)), "vb_batch_step")
        return [self.report(i) for i in range(len(self.params))]

    def report(self, i):
        r = StepReport()
        _check(lib().vb_batch_report(self.h, i, C.byref(r)), "vb_batch_report")
        return r.as_dict()

    def set_shared_stages(self, on=True):
        """vb_batch_set_shared_stages: the ETD stage combines of the size groups
        of several lengths as one launch each; returns how many groups share."""
        m = C.c_int()
        _check(lib().vb_batch_set_shared_stages(self.h, int(on), C.byref(m)), "vb_batch_set_shared_stages")
        return m.value

    def query(self, i):
        E, V, A, Eb = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().vb_batch_query(self.h, i, C.byref(E), C.byref(V), C.byref(A), C.byref(Eb)), "vb_batch_query")
        return E.value, V.value, A.value, Eb.value

    def get_phi(self, i):
        p = self.params[i]
        out = np.zeros((p.nz, p.nx, p.nx))
        _check(lib().vb_batch_get_phi(self.h, i, _np_ptr(out)), "vb_batch_get_phi")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().vb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop on one GPU (vb_server_*): submit solves at any time, poll
    finished results.  Requests use adaptive dt and an ETD scheme."""

    def __init__(self, max_active=16, streams=4, keep_phi=False, memory_budget=0, converge_window=50,
                 converge_tol=1e-6, max_steps=0, device=0):
        c = ServerConfig()
        lib().vb_server_config_default(C.byref(c))
        c.device, c.max_active, c.streams, c.keep_phi = device, max_active, streams, int(keep_phi)
        c.memory_budget, c.converge_window, c.max_steps, c.converge_tol = memory_budget, converge_window, \
            max_steps, converge_tol
        h = C.c_void_p()
        _check(lib().vb_server_create(C.byref(c), C.byref(h)), "vb_server_create")
        self.h, self.shapes = h, {}

    def submit(self, rid, params, phi0):
        phi0 = np.ascontiguousarray(phi0, dtype=np.float64)
        if phi0.shape != (params.nz, params.nx, params.nx):
            raise ValueError("phi0 shape %s != (nz, ny, nx)" % (phi0.shape,))
        _check(lib().vb_server_submit(self.h, int(rid), C.byref(params), _np_ptr(phi0)), "vb_server_submit")
        self.shapes[int(rid)] = phi0.shape

    def poll(self, cap=256):
        out = (Result * cap)()
        n = C.c_int()
        _check(lib().vb_server_poll(self.h, out, cap, C.byref(n)), "vb_server_poll")
        return [out[i].as_dict() for i in range(n.value)]

    def pending(self):
        q, r, b = C.c_int(), C.c_int(), C.c_size_t()
        _check(lib().vb_server_pending(self.h, C.byref(q), C.byref(r), C.byref(b)), "vb_server_pending")
        return {"queued": q.value, "running": r.value, "bytes": b.value}

    def take_phi(self, rid):
        out = np.zeros(self.shapes[int(rid)])
        _check(lib().vb_server_take_phi(self.h, int(rid), _np_ptr(out)), "vb_server_take_phi")
        return out

    def clock(self):
        return lib().vb_server_clock(self.h)

    def drain(self, poll_s=0.002):
        """Waits until nothing is queued or running; returns the results."""
        import time
        res = []
        while True:
            res += self.poll()
            st = self.pending()
            if st["queued"] == 0 and st["running"] == 0:
                return res + self.poll()
            time.sleep(poll_s)

    def close(self):
        if getattr(self, "h", None):
            lib().vb_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dptr(t):
    """Device address of a torch tensor (plumbing only)."""
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return C.c_void_p(t.data_ptr())


class Calc:
    """The reference's class Calc (Calc.h:8-119) on device tensors.

    Real fields: float64 [nz][nx][nx]; half spectra: complex128 [nz][nx][nx/2+1]
    (float32 / complex64 for FP32 transforms)."""

    def __init__(self, nx, nz=None, lenx=2 * np.pi, precision=FP64):
        nz = nx if nz is None else nz
        h = C.c_void_p()
        _check(lib().vb_calc_create(int(nx), int(nz), float(lenx), int(precision), C.byref(h)), "vb_calc_create")
        self.h, self.nx, self.nz, self.precision = h, nx, nz, precision

    def close(self):
        if getattr(self, "h", None):
            lib().vb_calc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        conv = [(_dptr(a) if hasattr(a, "data_ptr") else a) for a in args]
        _check(getattr(lib(), name)(self.h, *conv), name)

    def r2c(self, u, s): self._call("vb_calc_r2c", u, s)
    def c2r(self, s, u): self._call("vb_calc_c2r", s, u)
    def laplace(self, s): self._call("vb_calc_laplace", s)
    def lap_spec(self, fd, out): self._call("vb_calc_lap_spec", fd, out)
    def lap_real(self, d, out): self._call("vb_calc_lap_real", d, out)
    def fftgradient(self, fd, dx, dy, dz): self._call("vb_calc_fftgradient", fd, dx, dy, dz)
    def f(self, f, lap, d, eps): self._call("vb_calc_f", f, lap, d, float(eps))
    def g0(self, g0, d, f, eps): self._call("vb_calc_g0", g0, d, f, float(eps))
    def nld(self, nldp, d, f, tmp1, eps, kappa=2.0, D=1.5):
        self._call("vb_calc_nld", nldp, d, f, tmp1, float(eps), float(kappa), float(D))
    def s_term(self, s, d, dx, dy, dz, eps): self._call("vb_calc_s_term", s, d, dx, dy, dz, float(eps))
    def multi(self, a, b, out): self._call("vb_calc_multi", a, b, out)
    def plus(self, u, nld1, t): self._call("vb_calc_plus", u, nld1, float(t))
    def nld1(self, nld, f, g, l1, l2, eta): self._call("vb_calc_nld1", nld, f, g, float(l1), float(l2), float(eta))
    def nld2(self, nld, f, g, l1, l2, eta): self._call("vb_calc_nld2", nld, f, g, float(l1), float(l2), float(eta))
    def init_fEIF(self, Het, eps): self._call("vb_calc_init_fEIF", Het, float(eps))
    def etd1(self, a, fd, nld, eps, h): self._call("vb_calc_etd1", a, fd, nld, float(eps), float(h))
    def etd2(self, a, fd, nnld, nld, eps, h): self._call("vb_calc_etd2", a, fd, nnld, nld, float(eps), float(h))
    def etd41(self, d, fd, nld, eps, h): self._call("vb_calc_etd41", d, fd, nld, float(eps), float(h))
    def etd42(self, b, fd, nlda, eps, h): self._call("vb_calc_etd42", b, fd, nlda, float(eps), float(h))
    def etd43(self, c, nld, nldb, eps, h): self._call("vb_calc_etd43", c, nld, nldb, float(eps), float(h))
    def etd44(self, d, fd, nlda, nldb, nldc, eps, h):
        self._call("vb_calc_etd44", d, fd, nlda, nldb, nldc, float(eps), float(h))

    def intf(self, u, addition, multiply):
        v = C.c_double()
        _check(lib().vb_calc_intf(self.h, _dptr(u), float(addition), float(multiply), C.byref(v)), "vb_calc_intf")
        return v.value


def version():
    return lib().vb_version().decode()


# ---- kernel schedules and the autotuner (vb_schedule_*, vb_autotune) --------
def schedule_count(n, precision):
    return int(lib().vb_schedule_count(int(n), int(precision)))


def schedule_get(n, precision):
    return int(lib().vb_schedule_get(int(n), int(precision)))


def schedule_set(n, precision, sched):
    _check(lib().vb_schedule_set(int(n), int(precision), int(sched)), "vb_schedule_set")


def schedule_describe(n, precision, sched):
    """{z_tile, x_passes, x_radix} of one compiled schedule."""
    zt, xp, xr = C.c_int(), C.c_int(), C.c_int()
    _check(lib().vb_schedule_describe(int(n), int(precision), int(sched), C.byref(zt), C.byref(xp), C.byref(xr)),
           "vb_schedule_describe")
    return {"z_tile": zt.value, "x_passes": xp.value, "x_radix": xr.value}


def autotune(n, precision, steps=4, device=0):
    """Times every schedule of (n, precision) on the device, selects the fastest
    and returns the report as a dict."""
    r = TuneReport()
    _check(lib().vb_autotune(int(n), int(precision), int(steps), int(device), C.byref(r)), "vb_autotune")
    return {"n": r.n, "precision": r.precision, "steps": r.steps, "candidates": r.candidates,
            "chosen": r.chosen, "ms_per_step": [r.ms_per_step[i] for i in range(r.candidates)],
            "schedules": [schedule_describe(n, precision, i) for i in range(r.candidates)]}


COPY_VARIANTS = ("1x16B", "4x16B", "1x16B nontemporal", "4x16B nontemporal")


def copy_bandwidth(nbytes=1 << 30, repeats=20, device=0):
    """Read + write GB/s of the library's own 16-byte copy kernels over an
    nbytes buffer (vb_copy_bandwidth): the best and every variant's."""
    best = C.c_double()
    each = (C.c_double * len(COPY_VARIANTS))()
    _check(lib().vb_copy_bandwidth(int(device), C.c_size_t(nbytes), int(repeats), C.byref(best), each,
                                   len(COPY_VARIANTS)), "vb_copy_bandwidth")
    return best.value, dict(zip(COPY_VARIANTS, list(each)))


def intf_static(u, lenx, addition, multiply, val, stream=None):
    """vb_intf_static on torch device tensors: u (float64, contiguous) and the
    one-element float64 output val; launched on `stream` (a torch stream, or the
    current one) without waiting and without allocating."""
    import torch
    assert u.dtype == torch.float64 and u.is_contiguous() and val.dtype == torch.float64
    st = (stream or torch.cuda.current_stream()).cuda_stream
    _check(lib().vb_intf_static(C.c_void_p(u.data_ptr()), u.numel(), float(lenx), float(addition),
                                float(multiply), C.c_void_p(val.data_ptr()), C.c_void_p(st)), "vb_intf_static")
