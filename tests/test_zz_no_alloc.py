"""GPU check of the allocation-free integral vb_intf_static (its partial sums
in a static device array) against vb_calc_intf on the same inputs: two calls
in a row on one GPU, the second with a smaller input, each bit for bit."""
import ctypes as C

import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def test_intf_static_matches_calc_intf(vb, gpu):
    import torch
    L = 2 * np.pi
    rng = np.random.default_rng(11)
    prev = None
    for n, nz, a, m in ((128, 96, 0.0, 0.5), (48, 32, 1.0, 1.5 / np.sqrt(2.0))):
        flush_print("intf_static", n, nz)
        u = torch.from_numpy(rng.standard_normal((nz, n, n)) + 0.25).to(gpu)
        out = torch.full((1,), np.nan, dtype=torch.float64, device=gpu)
        vb.intf_static(u, L, a, m, out)
        torch.cuda.synchronize()
        got = float(out.item())
        calc = vb.Calc(n, nz, L, vb.FP64)
        ref = calc.intf(u, a, m)
        calc.close()
        assert got == ref, (n, got, ref)
        assert abs(got - (float(u.double().mean()) + a) * m * L ** 3) <= 1e-12 * abs(got)
        assert got != prev
        prev = got
    # beyond the static array's maximum: refused before any launch
    rc = vb.lib().vb_intf_static(C.c_void_p(u.data_ptr()), (1 << 30) + 1, L, 0.0, 1.0, C.c_void_p(out.data_ptr()),
                                 None)
    assert rc == vb.VB_ERR_PARAM
