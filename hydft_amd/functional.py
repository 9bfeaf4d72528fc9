"""Host mirror of the reference's functional_t and fe_term_t (functional.f90,
free_energy.f90) over libhydft: the terms of a species on a drop-in grid,
evaluated on the GPU term by term.

Fields are numpy arrays or torch CUDA tensors in the layouts of
hydft_amd.spectral (real (nz, ny, nx), half spectra (nz, ny, nx//2+1)
complex128).  Terms are addressed by their name or their position in the
list, as in the reference's sp%fun%terms(it)%t.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi
from .spectral import _dev, _out


class Functional:
    """functional_t on a Spec's grid; `species` from hydft_amd.model.build_species."""

    def __init__(self, spec, species):
        self.spec = spec
        self.names = [name for name, _ in species.terms]
        n = len(species.terms)
        arr = (abi.TermDesc * max(n, 1))()
        self._keep = []
        for i, (name, d) in enumerate(species.terms):
            if name == "ry":
                k = np.ascontiguousarray(species.structure.k, dtype=float)
                c = np.ascontiguousarray(species.structure.c, dtype=float)
                self._keep += [k, c]
                d.ntab = k.size
                d.k_tab = k.ctypes.data_as(C.POINTER(C.c_double))
                d.c_tab = c.ctypes.data_as(C.POINTER(C.c_double))
            arr[i] = d
        h = C.c_void_p()
        abi.call("hy_fun_create", spec.h, n, arr, C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            abi.call("hy_fun_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def index(self, term):
        return self.names.index(term) if isinstance(term, str) else int(term)

    def add_mu(self, term, n, nhat, mu):
        """fe_term_t%add_mu: returns mu + dF_term/dn."""
        dn, _ = _dev(n, torch.float64)
        dh, _ = _dev(nhat, torch.complex128)
        dm, t = _dev(mu, torch.float64)
        dm = dm.clone()
        abi.call("hy_fun_add_mu", self.h, self.index(term), dn.data_ptr(), dh.data_ptr(), dm.data_ptr(), None)
        return _out(dm, t)

    def mu(self, n, nhat):
        """functional_mu."""
        dn, t = _dev(n, torch.float64)
        dh, _ = _dev(nhat, torch.complex128)
        out = torch.empty(self.spec.shape, dtype=torch.float64, device="cuda")
        abi.call("hy_fun_mu", self.h, dn.data_ptr(), dh.data_ptr(), out.data_ptr(), None)
        return _out(out, t)

    def kernel(self, k, term=None):
        """fe_term_t%kernel (term) or functional_kernel (term None)."""
        out = C.c_double()
        abi.call("hy_fun_kernel", self.h, -1 if term is None else self.index(term), float(k), C.byref(out))
        return out.value

    def energy(self, n, nhat, term=None):
        """fe_term_t%energy (term) or functional_energy (term None)."""
        dn, _ = _dev(n, torch.float64)
        dh, _ = _dev(nhat, torch.complex128)
        out = C.c_double()
        abi.call("hy_fun_energy", self.h, -1 if term is None else self.index(term), dn.data_ptr(),
                 dh.data_ptr(), C.byref(out), None)
        return out.value
