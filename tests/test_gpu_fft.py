"""GPU parity of the hand-written 3-D R2C / C2R (Calc::r2c_3d_f /
c2r_3d_b, Calc.cu:88-104) against the oracle and, at sizes the oracle is slow
for, against the vendor FFT (rocFFT through torch.fft) as an independent check."""
import numpy as np
import pytest

from conftest import flush_print, rel_l2

pytestmark = pytest.mark.gpu

SIZES = [(16, 16), (32, 32), (48, 48), (64, 64), (96, 96), (128, 128), (192, 192),
         (32, 64), (64, 16), (128, 32), (16, 48)]


def _seeded(nx, nz, seed=11):
    return np.random.default_rng(seed + nx * 1000 + nz).standard_normal((nz, nx, nx))


@pytest.mark.parametrize("nx,nz", SIZES)
def test_r2c_c2r_fp64(vb, gpu, oracle_mod, nx, nz):
    import torch
    flush_print("fft fp64", nx, nz)
    u = _seeded(nx, nz)
    o = oracle_mod.Oracle(nx, nz)
    c = vb.Calc(nx, nz)
    ud = torch.from_numpy(u).to(gpu)
    sd = torch.empty((nz, nx, nx // 2 + 1), dtype=torch.complex128, device=gpu)
    c.r2c(ud, sd)
    ref = o.r2c(u)
    assert rel_l2(sd.cpu().numpy(), ref) < 1e-14
    u2 = torch.empty_like(ud)
    c.c2r(sd, u2)
    assert rel_l2(u2.cpu().numpy(), o.c2r(ref)) < 1e-14
    assert np.abs(u2.cpu().numpy() - u).max() < 1e-12


@pytest.mark.parametrize("nx,nz", [(32, 32), (64, 64), (128, 128), (192, 192), (64, 16)])
def test_r2c_c2r_fp32(vb, gpu, oracle_mod, nx, nz):
    import torch
    flush_print("fft fp32", nx, nz)
    u = _seeded(nx, nz)
    o = oracle_mod.Oracle(nx, nz)
    c = vb.Calc(nx, nz, precision=vb.FP32_MIXED)
    ud = torch.from_numpy(u.astype(np.float32)).to(gpu)
    sd = torch.empty((nz, nx, nx // 2 + 1), dtype=torch.complex64, device=gpu)
    c.r2c(ud, sd)
    ref = o.r2c(u)
    assert rel_l2(sd.cpu().numpy().astype(np.complex128), ref) < 2e-6
    u2 = torch.empty_like(ud)
    c.c2r(sd, u2)
    assert rel_l2(u2.cpu().numpy().astype(np.float64), u) < 2e-6


@pytest.mark.parametrize("n,prec", [(256, "fp64"), (256, "fp32"), (384, "fp32"), (512, "fp32")])
def test_large_vs_vendor_fft(vb, gpu, n, prec):
    """Full-size grids: compare with rocFFT (torch.fft) and check the round trip."""
    import torch
    flush_print("fft large", n, prec)
    fp32 = prec == "fp32"
    rdt = torch.float32 if fp32 else torch.float64
    cdt = torch.complex64 if fp32 else torch.complex128
    g = torch.Generator(device=gpu)
    g.manual_seed(5)
    u = torch.randn((n, n, n), generator=g, device=gpu, dtype=rdt)
    c = vb.Calc(n, n, precision=vb.FP32_MIXED if fp32 else vb.FP64)
    s = torch.empty((n, n, n // 2 + 1), dtype=cdt, device=gpu)
    c.r2c(u, s)
    ref = torch.fft.rfftn(u.double(), dim=(0, 1, 2)) / u.numel()
    err = (torch.linalg.vector_norm(s.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)).item()
    tol = 2e-6 if fp32 else 1e-14
    assert err < tol, err
    del ref
    u2 = torch.empty_like(u)
    c.c2r(s, u2)
    rt = (torch.linalg.vector_norm((u2 - u).double()) / torch.linalg.vector_norm(u.double())).item()
    assert rt < tol, rt


def test_c2r_ignores_self_conjugate_imaginary(vb, gpu, oracle_mod):
    """Imaginary parts on self-conjugate modes are dropped, as a real-output inverse does."""
    import torch
    nx = nz = 16
    o = oracle_mod.Oracle(nx, nz)
    rng = np.random.default_rng(3)
    s = o.r2c(rng.standard_normal((nz, nx, nx)))
    s[0, 0, 0] += 0.25j
    s[0, 0, nx // 2] += 0.5j
    c = vb.Calc(nx, nz)
    sd = torch.from_numpy(s).to(gpu)
    u = torch.empty((nz, nx, nx), dtype=torch.float64, device=gpu)
    c.c2r(sd, u)
    assert rel_l2(u.cpu().numpy(), o.c2r(s)) < 1e-14
