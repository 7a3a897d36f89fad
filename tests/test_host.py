"""CPU checks of the host-side pieces that need no GPU: the reference's
deterministic initial shapes (closed-form volumes), the field-file reader with
its 65x65x64 trimming, and the C++ host program's command line."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel

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
