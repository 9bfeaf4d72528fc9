"""Host mirror of the reference's spec_t / fft_t / stress_t (spectral.f90,
fft.f90, stress.f90) over libhydft.

Arrays are numpy (C order, shape (nz, ny, nx)) or torch CUDA tensors of the
same shape; spectra are complex128 (nz, ny, nx//2+1).  Every operation runs on
the GPU through the C ABI; numpy inputs are copied to the device and results
copied back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi


def _dev(x, dtype):
    if isinstance(x, torch.Tensor):
        t = x.to(device="cuda", dtype=dtype)
        return t.contiguous(), True
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype), False


def _out(t, as_torch):
    if as_torch:
        return t
    torch.cuda.synchronize()
    return t.cpu().numpy()


class Spec:
    """spec_t: grid + transforms + operators (spectral.f90:19-188)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, dealias=True):
        abi.lib()
        h = C.c_void_p()
        abi.call("hy_grid_create", nx, ny, nz, float(lx), float(ly), float(lz), int(bool(dealias)),
                 abi.HY_PREC_FP64, C.byref(h))
        self.h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nkx = nx // 2 + 1
        nkx, ndim, kept = C.c_int(), C.c_int(), (C.c_int * 3)()
        vol, kmax = C.c_double(), C.c_double()
        abi.call("hy_grid_query", h, C.byref(nkx), C.byref(ndim), kept, C.byref(vol), C.byref(kmax))
        self.ndim = ndim.value
        self.kept = tuple(kept)
        self.volume = vol.value
        self.kmax_dealiased = kmax.value
        self.shape = (nz, ny, nx)
        self.sshape = (nz, ny, self.nkx)

    def close(self):
        if self.h:
            abi.call("hy_grid_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mask(self):
        m = np.zeros(self.sshape)
        abi.call("hy_grid_mask", self.h, m.ctypes.data)
        return m

    # fft.f90:52-69
    def fwd(self, u):
        d, t = _dev(u, torch.float64)
        out = torch.empty(self.sshape, dtype=torch.complex128, device="cuda")
        abi.call("hy_fft_r2c", self.h, d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def bwd(self, uhat):
        d, t = _dev(uhat, torch.complex128)
        out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_fft_c2r", self.h, d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def grad_hat(self, uhat, idir):
        """idir is 1-based as in spectral.f90:67."""
        d, t = _dev(uhat, torch.complex128)
        out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_spec_grad_hat", self.h, d.data_ptr(), int(idir), out.data_ptr(), None)
        return _out(out, t)

    def grad(self, u):
        d, t = _dev(u, torch.float64)
        out = torch.empty((3,) + self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_spec_grad", self.h, d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def lap(self, u):
        d, t = _dev(u, torch.float64)
        out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_spec_lap", self.h, d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def div(self, v):
        """v: (3, nz, ny, nx)."""
        d, t = _dev(np.asarray(v) if not isinstance(v, torch.Tensor) else v, torch.float64)
        out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_spec_div", self.h, d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def convolve(self, kernel, uhat):
        dk, _ = _dev(kernel, torch.float64)
        d, t = _dev(uhat, torch.complex128)
        out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_spec_convolve", self.h, dk.data_ptr(), d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def dealias(self, u):
        d, t = _dev(u, torch.float64)
        d = d.clone()
        abi.call("hy_spec_dealias", self.h, d.data_ptr(), None)
        return _out(d, t)

    def mean(self, u):
        d, _ = _dev(u, torch.float64)
        m = C.c_double()
        abi.call("hy_spec_mean", self.h, d.data_ptr(), C.byref(m), None)
        return m.value

    def mode(self, u, mx, my, mz):
        d, _ = _dev(u, torch.float64)
        out = (C.c_double * 2)()
        abi.call("hy_spec_mode", self.h, d.data_ptr(), int(mx), int(my), int(mz), out, None)
        return complex(out[0], out[1])

    # stress.f90:60-105 (eta, xi in code units)
    def stress_pi0(self, eta, xi, u3):
        d, t = _dev(np.asarray(u3) if not isinstance(u3, torch.Tensor) else u3, torch.float64)
        out = torch.empty((6,) + self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_stress_pi0", self.h, float(eta), float(xi), d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def stress_div(self, pi6):
        d, t = _dev(np.asarray(pi6) if not isinstance(pi6, torch.Tensor) else pi6, torch.float64)
        out = torch.empty((3,) + self.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_stress_div", self.h, d.data_ptr(), out.data_ptr(), None)
        return _out(out, t)
