"""numpy front end of the CPU FP64 oracle (liboracle.so).

Test infrastructure only: the tests, the smoke check and the benchmark's
CPU-baseline leg use it as the checker.  The product never imports it.

Arrays follow the reference layouts (Calc.h:4, Calc.cu:57-58): real fields are
float64 [nz][ny][nx]; half spectra are complex128 [nz][ny][nx/2+1].
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib_path():
    return os.path.join(_HERE, "liboracle.so")


def build(force=False):
    """Compile liboracle.so from vb_oracle.c (gcc, OpenMP, no FMA contraction)."""
    import subprocess
    out = lib_path()
    src = os.path.join(_HERE, "vb_oracle.c")
    if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
                               "-o", out, src, "-lm"])
    return out


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("pad0", C.c_int),
                ("xlen", C.c_double), ("eps", C.c_double), ("eta", C.c_double), ("ga", C.c_double),
                ("M1", C.c_double), ("M2", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("kappa", C.c_double), ("D", C.c_double)]


class Drive(C.Structure):
    _fields_ = [("time_step", C.c_double), ("fixed_time_step", C.c_double),
                ("time_current", C.c_double), ("alpha0", C.c_double), ("beta0", C.c_double),
                ("eng0", C.c_double), ("eng1", C.c_double), ("dt_cap", C.c_double),
                ("steps", C.c_int), ("dt_double_every", C.c_int), ("rejections", C.c_int),
                ("evaluations", C.c_int), ("status", C.c_int), ("pad", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        D = C.c_double
        I = C.c_int
        L.vbo_new.restype = P
        L.vbo_new.argtypes = [C.POINTER(Params)]
        L.vbo_free.argtypes = [P]
        L.vbo_set_targets.argtypes = [P, D, D]
        L.vbo_get_params.argtypes = [P, C.POINTER(Params)]
        L.vbo_threads.restype = I
        L.vbo_set_threads.argtypes = [I]
        for name, n in [("vbo_r2c", 2), ("vbo_c2r", 2), ("vbo_lap_spec", 2), ("vbo_lap_real", 2),
                        ("vbo_fpf", 2), ("vbo_g0", 2), ("vbo_sum_fgq", 2)]:
            getattr(L, name).argtypes = [P] + [P] * n
        L.vbo_fft1d.argtypes = [I, P, I]
        L.vbo_laplace.argtypes = [P, P]
        L.vbo_gradient.argtypes = [P, P, P, P, P]
        L.vbo_fftgradient.argtypes = [P, P, P, P, P]
        L.vbo_f.argtypes = [P, P, P, P]
        L.vbo_g0k.argtypes = [P, P, P, P]
        L.vbo_intf.restype = D
        L.vbo_intf.argtypes = [P, P, D, D]
        L.vbo_nld.argtypes = [P, P, P, P]
        L.vbo_s_term.argtypes = [P, P, P, P, P, P]
        L.vbo_nld1.argtypes = [P, P, P, P, D, D]
        L.vbo_nld2.argtypes = [P, P, P, P, D, D]
        L.vbo_init_fEIF.argtypes = [P, P]
        L.vbo_etd1.argtypes = [P, P, P, P, P, D]
        L.vbo_etd2.argtypes = [P, P, P, P, P, P, D]
        L.vbo_etd41.argtypes = [P, P, P, P, P, D]
        L.vbo_etd42.argtypes = [P, P, P, P, P, D]
        L.vbo_etd43.argtypes = [P, P, P, P, P, D]
        L.vbo_etd44.argtypes = [P, P, P, P, P, P, P, D]
        for name in ["vbo_vol", "vbo_surface", "vbo_elastic_eng"]:
            getattr(L, name).restype = D
            getattr(L, name).argtypes = [P, P]
        L.vbo_energy.restype = D
        L.vbo_energy.argtypes = [P, P, I, C.POINTER(D), C.POINTER(D), C.POINTER(D)]
        L.vbo_nld_term.argtypes = [P, P, P, C.POINTER(D), C.POINTER(D)]
        L.vbo_etd4rk.argtypes = [P, P, D]
        L.vbo_etd2rk.argtypes = [P, P, D]
        L.vbo_fEIF_begin.argtypes = [P, P, C.POINTER(Drive)]
        L.vbo_fEIF_steps.restype = I
        L.vbo_fEIF_steps.argtypes = [P, P, C.POINTER(Drive), I, I, P]
        L.vbo_shape_adjust.restype = I
        L.vbo_shape_adjust.argtypes = [P, P, I]
        L.vbo_static_steps.restype = I
        L.vbo_static_steps.argtypes = [P, P, C.POINTER(Drive), I]
        L.vbo_etd_coeffs.argtypes = [D, D, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def fft1d(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().vbo_fft1d(x.size, _p(x), int(sign))
    return x


def etd_coeffs(H, h):
    out = np.zeros(7)
    lib().vbo_etd_coeffs(float(H), float(h), _p(out))
    return out


class Oracle:
    """One grid + parameter set, mirroring the reference's Calc/phasefield/belastic."""

    def __init__(self, nx, nz=None, xlen=2 * np.pi, eps=None, eta=1.0, ga=1.0, M1=0.0, M2=0.0,
                 alpha=-1.0, beta=-1.0, kappa=2.0, D=1.5):
        nz = nx if nz is None else nz
        if eps is None:
            eps = 2.0 * xlen / nx  # 2h (fluidbase.cpp:105-107)
        self.p = Params(nx, nx, nz, 0, xlen, eps, eta, ga, M1, M2, alpha, beta, kappa, D)
        self.h = lib().vbo_new(C.byref(self.p))
        if not self.h:
            raise ValueError("invalid grid (nx must be even)")
        self.nx, self.ny, self.nz = nx, nx, nz
        self.nh = nx // 2 + 1
        self.rshape = (nz, nx, nx)
        self.cshape = (nz, nx, self.nh)

    def __del__(self):
        try:
            if self.h:
                lib().vbo_free(self.h)
        except Exception:
            pass

    def params(self):
        q = Params()
        lib().vbo_get_params(self.h, C.byref(q))
        return q

    def set_targets(self, alpha, beta):
        lib().vbo_set_targets(self.h, float(alpha), float(beta))

    def _r(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.rshape, (a.shape, self.rshape)
        return a

    def _c(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == self.cshape, (a.shape, self.cshape)
        return a

    def newr(self):
        return np.zeros(self.rshape)

    def newc(self):
        return np.zeros(self.cshape, dtype=np.complex128)

    def r2c(self, u):
        u = self._r(u); s = self.newc()
        lib().vbo_r2c(self.h, _p(u), _p(s)); return s

    def c2r(self, s):
        s = self._c(s); u = self.newr()
        lib().vbo_c2r(self.h, _p(s), _p(u)); return u

    def laplace(self, s):
        s = self._c(s).copy(); lib().vbo_laplace(self.h, _p(s)); return s

    def gradient(self, s):
        s = self._c(s); a, b, c = self.newc(), self.newc(), self.newc()
        lib().vbo_gradient(self.h, _p(s), _p(a), _p(b), _p(c)); return a, b, c

    def lap_real(self, d):
        d = self._r(d); o = self.newr(); lib().vbo_lap_real(self.h, _p(d), _p(o)); return o

    def fftgradient(self, fd):
        fd = self._c(fd); a, b, c = self.newr(), self.newr(), self.newr()
        lib().vbo_fftgradient(self.h, _p(fd), _p(a), _p(b), _p(c)); return a, b, c

    def f(self, d):
        d = self._r(d); o = self.newr(); lib().vbo_fpf(self.h, _p(d), _p(o)); return o

    def g0(self, d):
        d = self._r(d); o = self.newr(); lib().vbo_g0(self.h, _p(d), _p(o)); return o

    def f_kernel(self, lap, d):
        lap = self._r(lap); d = self._r(d); o = self.newr()
        lib().vbo_f(self.h, _p(o), _p(lap), _p(d)); return o

    def nld(self, d, f):
        d = self._r(d); f = self._r(f); o = self.newr()
        lib().vbo_nld(self.h, _p(o), _p(d), _p(f)); return o

    def s_term(self, d, dx, dy, dz):
        o = self.newr()
        lib().vbo_s_term(self.h, _p(o), _p(self._r(d)), _p(self._r(dx)), _p(self._r(dy)), _p(self._r(dz)))
        return o

    def nld2(self, f, g, l1, l2):
        o = self.newr(); lib().vbo_nld2(self.h, _p(o), _p(self._r(f)), _p(self._r(g)), l1, l2); return o

    def nld1(self, f, g, l1, l2):
        o = self.newr(); lib().vbo_nld1(self.h, _p(o), _p(self._r(f)), _p(self._r(g)), l1, l2); return o

    def intf(self, u, addition, multiply):
        return lib().vbo_intf(self.h, _p(self._r(u)), addition, multiply)

    def Het(self):
        o = np.zeros(self.cshape); lib().vbo_init_fEIF(self.h, _p(o)); return o

    def vol(self, d):
        return lib().vbo_vol(self.h, _p(self._r(d)))

    def surface(self, d):
        return lib().vbo_surface(self.h, _p(self._r(d)))

    def elastic_eng(self, d):
        return lib().vbo_elastic_eng(self.h, _p(self._r(d)))

    def energy(self, d, constrained=True):
        V, A, Eb = C.c_double(), C.c_double(), C.c_double()
        E = lib().vbo_energy(self.h, _p(self._r(d)), int(constrained), C.byref(V), C.byref(A), C.byref(Eb))
        return E, V.value, A.value, Eb.value

    def nld_term(self, d):
        o = self.newc(); V, A = C.c_double(), C.c_double()
        lib().vbo_nld_term(self.h, _p(self._r(d)), _p(o), C.byref(V), C.byref(A))
        return o, V.value, A.value

    def sum_fgq(self, d):
        o = self.newr(); lib().vbo_sum_fgq(self.h, _p(self._r(d)), _p(o)); return o

    def etd4rk(self, d, h):
        d = self._r(d).copy(); lib().vbo_etd4rk(self.h, _p(d), h); return d

    def etd2rk(self, d, h):
        d = self._r(d).copy(); lib().vbo_etd2rk(self.h, _p(d), h); return d

    def fEIF_begin(self, d, time_step, fixed_time_step=0.0, dt_double_every=4):
        s = Drive()
        s.time_step = time_step
        s.fixed_time_step = fixed_time_step
        s.dt_double_every = dt_double_every
        s.dt_cap = 1e-3
        lib().vbo_fEIF_begin(self.h, _p(self._r(d)), C.byref(s))
        return s

    def fEIF_steps(self, d, s, n, scheme=4, trace=False):
        """Runs n loop iterations in place on d (float64 C-contiguous)."""
        assert d.dtype == np.float64 and d.flags.c_contiguous and d.shape == self.rshape
        tr = np.zeros((n, 4)) if trace else None
        done = lib().vbo_fEIF_steps(self.h, _p(d), C.byref(s), n, scheme,
                                    _p(tr) if trace else None)
        return (done, tr[:done]) if trace else done

    def shape_adjust(self, d, max_iter=100000):
        assert d.dtype == np.float64 and d.flags.c_contiguous and d.shape == self.rshape
        return lib().vbo_shape_adjust(self.h, _p(d), max_iter)

    def static_steps(self, d, s, n):
        assert d.dtype == np.float64 and d.flags.c_contiguous and d.shape == self.rshape
        return lib().vbo_static_steps(self.h, _p(d), C.byref(s), n)
