"""Writes the golden fixtures under tests/golden/ (run from the repo root:
`python tests/golden/gen_golden.py`).  Each fixture is data produced by an
implementation independent of this repository's oracle and product:

fft_golden.npz   numpy.fft (pocketfft) forward R2C / backward C2R of seeded
                 fields, with the reference's conventions: forward divided by
                 nx*ny*nz (Calc.cu:88-96), backward unnormalised (Calc.cu:99-104).
phi_mpmath.npz   40-digit mpmath values of the ETD coefficient functions the
                 reference combines in etd41..etd44 / etd1, etd2 (Calc.cu:446-577):
                 e1 = exp(z/2), (e1-1)/z, e^z, f1 = [-4-z+e^z(4-3z+z^2)]/z^3,
                 f2 = [2+z+e^z(z-2)]/z^3, f3 = [-4-3z-z^2+e^z(4-z)]/z^3,
                 phi2 = (e^z-1-z)/z^2.
ini_lines.txt    first lines of reference parameter files (data, verbatim).
"""
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_fixtures():
    rng = np.random.default_rng(20260001)
    out = {}
    for (nz, nx) in [(8, 8), (16, 16), (12, 8), (6, 16)]:
        u = rng.standard_normal((nz, nx, nx))
        s = np.fft.rfftn(u, axes=(0, 1, 2)) / u.size
        # a Hermitian spectrum for the backward transform
        s2 = np.fft.rfftn(rng.standard_normal((nz, nx, nx)), axes=(0, 1, 2)) / u.size
        u2 = np.fft.irfftn(s2, s=(nz, nx, nx), axes=(0, 1, 2)) * u.size
        key = "%dx%dx%d" % (nx, nx, nz)
        out["u_" + key] = u
        out["s_" + key] = s
        out["s2_" + key] = s2
        out["u2_" + key] = u2
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def phi_fixtures():
    mp.mp.dps = 40
    zs = [-1e-8, -1e-6, -1e-5, -1e-4, -3e-4, -1e-3, -3e-3, -1e-2, -3e-2, -0.1, -0.3, -0.7, -1.0, -1.5,
          -3.0, -7.0, -15.0, -30.0, -61.0, -100.0]
    rows = []
    for z in zs:
        zz = mp.mpf(z)
        ez = mp.e ** zz
        e1 = mp.e ** (zz / 2)
        f1 = (-4 - zz + ez * (4 - 3 * zz + zz * zz)) / zz ** 3
        f2 = (2 + zz + ez * (zz - 2)) / zz ** 3
        f3 = (-4 - 3 * zz - zz * zz + ez * (4 - zz)) / zz ** 3
        p1h = (e1 - 1) / zz
        p2 = (ez - 1 - zz) / zz ** 2
        rows.append([z, float(e1), float(p1h), float(ez), float(f1), float(f2), float(f3), float(p2)])
    np.savez_compressed(os.path.join(HERE, "phi_mpmath.npz"), table=np.array(rows),
                        columns=np.array(["z", "e1", "(e1-1)/z", "ez", "f1", "f2", "f3", "phi2"]))


def ini_fixtures():
    ref = "/root/reference/results"
    lines = []
    for i in [0, 1, 2, 3, 10, 12, 33]:
        p = os.path.join(ref, "ini%d.dat" % i)
        with open(p) as f:
            lines.append("ini%d.dat\t%s" % (i, f.readline().rstrip("\n")))
    with open(os.path.join(HERE, "ini_lines.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    fft_fixtures()
    phi_fixtures()
    if os.path.isdir("/root/reference/results"):
        ini_fixtures()
