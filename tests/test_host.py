"""CPU checks of the host-side pieces that need no GPU: the reference's
deterministic initial shapes (closed-form volumes), the field-file reader with
its 65x65x64 trimming, and the C++ host program's command line."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, grid, rel

L = 2 * np.pi


def test_sphere_and_spheroid_volumes(vb, oracle_mod):
    n = 64
    eps = 2 * L / 2 / n
    o = oracle_mod.Oracle(n, eps=eps)
    # (phi+1)/2 = logistic((R - r)/w), w = sqrt2 eps / 2, so a ball holds
    # 4/3 pi (R^3 + pi^2 w^2 R) (the second moment of the logistic density)
    w = np.sqrt(2) * eps / 2
    ball = lambda R: 4 / 3 * np.pi * (R ** 3 + np.pi ** 2 * w ** 2 * R)
    # ninit 0: sphere of radius L/4 (fluidbase.cpp:1080-1099)
    assert rel(o.vol(vb.shape_ninit(0, n, eps=eps)), ball(L / 4)) < 1e-3
    # ninit 1: two disjoint spheres of radii 0.3 L/2, 0.2 L/2 (fluidbase.cpp:1197-1215)
    assert rel(o.vol(vb.shape_ninit(1, n, eps=eps)), ball(0.3 * L / 2) + ball(0.2 * L / 2)) < 2e-3
    # ninit 2: spheroid, semi-axes a, a, 0.7 a with a = 0.7 L/2 (fluidbase.cpp:1862-1877)
    a = 0.7 * L / 2
    assert rel(o.vol(vb.shape_ninit(2, n, eps=eps)), 4 / 3 * np.pi * a * a * 0.7 * a) < 5e-3


def test_gtorus_shape(vb, oracle_mod):
    n = 64
    eps = 2 * L / 2 / n
    phi = vb.shape_ninit(4, n, eps=eps)
    assert np.abs(phi).max() <= 1.0 and phi.max() > 0.9 and phi.min() < -0.9
    # centre line radius 0.5 L/2; the tube is thickest (0.3 L/2) at theta = 0 (x > 0)
    assert phi[n // 2, n // 2, n // 2] < -0.99      # the hole
    i = int(round((0.5 * L / 2 + L / 2) / (L / n)))
    assert phi[n // 2, n // 2, i] > 0.99            # inside the tube on +x
    with pytest.raises(vb.VbError):
        vb.shape_ninit(3, n, eps=eps)


def test_read_field_trims_repeated_planes(vb, tmp_path):
    n, nz = 8, 6
    rng = np.random.default_rng(0)
    core = rng.standard_normal((nz, n, n))
    padded = np.concatenate([core, core[:, :1, :]], axis=1)
    padded = np.concatenate([padded, padded[:, :, :1]], axis=2)
    f = tmp_path / "d9x9x6.out"
    padded.astype("<f8").tofile(f)
    assert np.array_equal(vb.read_field(str(f), n, nz), core)
    g = tmp_path / "d8x8x6.out"
    core.astype("<f8").tofile(g)
    assert np.array_equal(vb.read_field(str(g), n, nz), core)
    h = tmp_path / "bad.out"
    core[:, :, :5].astype("<f8").tofile(h)
    with pytest.raises(vb.VbError):
        vb.read_field(str(h), n, nz)


INI = ("101  32  32  32   1  1  1   0   1e-7  5e-6  1e-7   2  -1  1  200  0  30  1.0  1  1  -1  -1  "
       "1.0e4  1.0e4  0  0  100.0  30  0.05  0  30.0  70  10  -1\n")


def test_host_program_without_gpu(vb, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    exe = vb.HOST_PROGRAM
    assert os.path.exists(exe), "the C++ host program is not built"
    assert subprocess.run([exe, "--help"], capture_output=True).returncode == 0
    (tmp_path / "ini0.dat").write_text(INI)
    r = subprocess.run([exe, "0", "--results", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "GPU" in r.stderr
    # the folder was made and the ini copied, as the reference does
    assert (tmp_path / "101" / "ini0.dat").read_text().split() == INI.split()


@pytest.mark.parametrize("ninit,a,q", [(5, 0.83, 2.0), (6, 0.8, 5.0), (11, 0.8, 3.0)])
def test_oblate_spheroid_volumes(vb, oracle_mod, ninit, a, q):
    # ninit 5, 6, 11 (fluidbase.cpp:1880-1932): tanh((a L/2 - r) / (eps/sqrt2)) with
    # r = sqrt(x^2 + y^2 + q^2 z^2).  In z' = q z the field is a ball whose
    # (phi+1)/2 is a logistic step of scale s = sqrt2 eps / 4, so
    # V = 4/3 pi (A^3 + pi^2 s^2 A) / q with A = a L/2
    n = 128
    eps = 2 * L / 2 / n
    phi = vb.shape_ninit(ninit, n, eps=eps)
    s = np.sqrt(2) * eps / 4
    A = a * L / 2
    o = oracle_mod.Oracle(n, eps=eps)
    # (in z' the grid spacing is q h against s ~ 0.7 h, so the sampled sum of the
    # step is accurate to a few 1e-4 only)
    assert rel(o.vol(phi), 4 / 3 * np.pi * (A ** 3 + np.pi ** 2 * s ** 2 * A) / q) < 1e-3
    # the zero level set crosses the z axis at A / q and the x axis at A
    X, Y, Z = grid(n)
    assert np.all(phi[(np.abs(Z) < 0.95 * A / q) & (np.abs(X) < 1e-12) & (np.abs(Y) < 1e-12)] > 0)
    assert np.all(phi[(np.abs(Z) > 1.05 * A / q) & (np.abs(X) < 1e-12) & (np.abs(Y) < 1e-12)] < 0)


def test_vtorus_volume(vb, oracle_mod):
    # ninit 8 (fluidbase.cpp:2121-2143): a torus about the y axis with centre
    # radius R = 0.5 L/2 and tube radius r = 0.25 L/2, a true distance profile:
    # V = 2 pi R pi (r^2 + pi^2 w^2 / 3), w = sqrt2 eps / 2
    n = 128
    eps = 2 * L / 2 / n
    phi = vb.shape_ninit(8, n, eps=eps)
    R, r, w = 0.5 * L / 2, 0.25 * L / 2, np.sqrt(2) * eps / 2
    o = oracle_mod.Oracle(n, eps=eps)
    assert rel(o.vol(phi), 2 * np.pi ** 2 * R * (r ** 2 + np.pi ** 2 * w ** 2 / 3)) < 2e-4
    # the hole is on the y axis, the tube crosses the x axis at R
    c = n // 2
    assert phi[c, :, c].max() < -0.99
    i = int(round((R + L / 2) / (L / n)))
    assert phi[c, c, i] > 0.99 and phi[i, c, c] > 0.99 and phi[c, i, c] < -0.99


def test_tian_frame(vb):
    # ninit 50 (fluidbase.cpp:1825-1858): 27 rods of radius r = L/12, along each
    # axis through the offset pairs {0, +-4r}^2, cut at |x|, |y|, |z| = 5r.  Each of
    # the 27 lattice points is met by three rods: the union loses the overlap of
    # three cylinders, 8 sqrt2 r^3 (two-cylinder Steinmetz solids 16 r^3 / 3, the
    # three-cylinder one 8 (2 - sqrt2) r^3), so V = 27 r^3 (10 pi - 8 sqrt2)
    n = 130   # r = L/12 is then no whole number of grid spacings: no grid point sits on a rod wall
    phi = vb.shape_ninit(50, n, eps=2 * L / 2 / n)
    assert set(np.unique(phi)) == {-1.0, 1.0}
    r = L / 12
    vol = np.count_nonzero(phi > 0) * (L / n) ** 3
    assert rel(vol, 27 * r ** 3 * (10 * np.pi - 8 * np.sqrt(2))) < 2e-2
    # symmetric under axis permutations (exactly) and reflections (x_i -> x_{n-i})
    assert np.array_equal(phi, phi.transpose(2, 1, 0)) and np.array_equal(phi, phi.transpose(1, 0, 2))
    refl = np.roll(phi[::-1, :, :], 1, axis=0)
    assert np.mean(refl != phi) < 1e-4
    c = n // 2
    i4 = int(round((4 * r + L / 2) / (L / n)))
    i2 = int(round((2 * r + L / 2) / (L / n)))
    i6 = int(round((5.5 * r + L / 2) / (L / n)))
    assert phi[c, c, c] == 1 and phi[c, i4, i4] == 1 and phi[i2, i2, i2] == -1 and phi[c, c, i6] == -1


def test_gyroid(vb):
    # ninit 51 (fluidbase.cpp:2255-2270): sign of cos x sin y + cos y sin z + cos z sin x.
    # The function is odd and invariant under the shift (pi, pi, pi): the two
    # labyrinths have equal volume and the field repeats at half the box
    n = 64
    phi = vb.shape_ninit(51, n, eps=2 * L / 2 / n)
    assert set(np.unique(phi)) == {-1.0, 1.0}
    assert abs(phi.mean()) < 2e-2
    shifted = np.roll(phi, (n // 2, n // 2, n // 2), axis=(0, 1, 2))
    assert np.mean(shifted != phi) < 1e-3
    X, Y, Z = grid(n)
    g = np.cos(X) * np.sin(Y) + np.cos(Y) * np.sin(Z) + np.cos(Z) * np.sin(X)
    far = np.abs(g) > 1e-6
    assert np.array_equal(phi[far], np.where(g[far] > 0, 1.0, -1.0))
