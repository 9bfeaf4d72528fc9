"""hy_spec_mean_noalloc (no device allocation; partial sums in a fixed device
array per GPU) against hy_spec_mean through the C ABI, on two calls in a row,
the second on a smaller grid: the same kernels, so the same mean bit for bit,
and the array left by the first call does not leak into the second."""
import ctypes as C

import numpy as np
import pytest
import torch

from hydft_amd import abi

pytestmark = pytest.mark.gpu


def _grid(nx, ny, nz):
    h = C.c_void_p()
    abi.call("hy_grid_create", nx, ny, nz, 7.0, 5.0, 3.0, 1, abi.HY_PREC_FP64, C.byref(h))
    return h


def _means(h, u):
    a, b = C.c_double(), C.c_double()
    abi.call("hy_spec_mean", h, u.data_ptr(), C.byref(a), None)
    abi.call("hy_spec_mean_noalloc", h, u.data_ptr(), C.byref(b), None)
    return a.value, b.value


def test_spec_mean_noalloc_matches_spec_mean(gpu):
    gen = torch.Generator(device="cuda").manual_seed(5)
    big = _grid(64, 48, 32)
    small = _grid(16, 12, 8)
    u_big = torch.rand((32, 48, 64), dtype=torch.float64, device="cuda", generator=gen) + 0.25
    u_small = torch.rand((8, 12, 16), dtype=torch.float64, device="cuda", generator=gen) - 0.75
    for h, u in ((big, u_big), (small, u_small)):
        a, b = _means(h, u)
        ref = float(u.double().mean())
        print(f"mean {a!r} noalloc {b!r} torch {ref!r}")
        assert a == b
        assert abs(b - ref) <= 1e-14 * abs(ref)
    # a second pass in the other order: the array keeps only the last call's sums
    a, b = _means(big, u_big)
    assert a == b
    for h in (big, small):
        abi.call("hy_grid_destroy", h)
