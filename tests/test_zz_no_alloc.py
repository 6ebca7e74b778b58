"""plbm_reduce_static (the reduction without any allocation: partial sums in a static
device array) against plbm_reduce_dev, through the C ABI: two calls in a row, the second
on a smaller lattice, each with and without the analytic velocity; the results must be
bit-identical (same kernels, same workgroup partition)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(1 + 0.01 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))),
            np.asfortranarray(0.05 * rng.standard_normal((ny, nx))))


@pytest.mark.parametrize("storage", ["f64", "f32s"])
def test_reduce_static_equals_reduce_dev(P, gpu, storage):
    import torch
    L = P._abi.lib()
    # 1300 x 900: more cells than one sweep of the 1024 workgroups; then 70 x 45
    for nx, ny, seed in ((1300, 900, 3), (70, 45, 4)):
        print(f"case reduce_static {storage} {nx}x{ny}", flush=True)
        lat = P.Lattice(nx, ny, 2, storage)
        lat.set_params(1.5, 0.25)
        lat.equilibrium_init(0, *_fields(nx, ny, seed))
        lat.run("trt", 3, 0, 1)
        _, uxa, uya = _fields(nx, ny, seed + 10)
        d_uxa = torch.from_numpy(np.ravel(uxa, order="F")).cuda()
        d_uya = torch.from_numpy(np.ravel(uya, order="F")).cuda()
        for ua in (None, (d_uxa, d_uya)):
            want = np.zeros(P._abi.NRED)
            pa = ua[0].data_ptr() if ua else None
            pb = ua[1].data_ptr() if ua else None
            P._abi.check(L.plbm_reduce_dev(lat.h, 1, pa, pb,
                                           want.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            out = torch.full((P._abi.NRED,), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            P._abi.check(L.plbm_reduce_static(lat.h, 1, pa, pb, out.data_ptr()))
            lat.sync()
            got = out.cpu().numpy()
            np.testing.assert_array_equal(got, want)
            assert want[P._abi.RED_MASS] > 0
        lat.close()
    with pytest.raises(P.PlbmError):
        P._abi.check(L.plbm_reduce_static(None, 0, None, None, None))
