"""Host mirror of the reference's species factory and model_t (species.f90,
units.f90, model.f90) over libhydft.

The host builds everything that is one-time setup (units, the HNC c(k) table,
transport scalars, term descriptors) and hands it to the C ABI; the GPU owns
the state and runs the RHS / RK3 loop.  Arrays are numpy (nz, ny, nx) C-order,
which is the reference's Fortran (nx, ny, nz) memory layout.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import abi
from .host import hnc as _hnc
from .host import transport as _tr
from .host.params import Params

PI = math.pi
SQRT2 = math.sqrt(2.0)


def fd_eval(which, x):
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=float)))
    y = np.empty_like(x)
    abi.call("hy_fd_eval", int(which), x.size, x.ctypes.data, y.ctypes.data)
    return y


FD_IM12, FD_IP12, FD_IP32, FD_DIM12, FD_D2IM12, FD_XI, FD_DXI, FD_ALPHA = range(8)


@dataclass
class Units:
    """units_t (units.f90:41-70)."""
    kind: str
    gamma: float
    kappa: float
    rs: float
    mass: float = 1.0
    n0: float = 3.0 / (4.0 * PI)
    e2: float = 1.0 / 3.0
    beta: float = 3.0
    hbar: float = 0.0
    cn: float = 0.0
    alpha0: float = 0.0
    theta: float = 1.7976931348623157e308
    lambda_qsp: float = 0.0
    lambda_tf: float = 0.0
    quantum: bool = False

    def __post_init__(self):
        self.beta = 3.0 * self.gamma
        if self.kind == "electron":
            self.quantum = True
            self.kappa = 0.0
            self.hbar = 1.0 / math.sqrt(3.0 * self.rs)
            self.cn = SQRT2 / (PI * PI * self.hbar ** 3 * self.beta ** 1.5) * self.mass ** 1.5
            self.alpha0 = float(fd_eval(FD_ALPHA, self.n0 / self.cn)[0])
            tf = self.hbar ** 2 * (3.0 * PI * PI * self.n0) ** (2.0 / 3.0) / (2.0 * self.mass)
            self.theta = (1.0 / self.beta) / tf
            self.lambda_qsp = 2.0 * math.sqrt(PI * self.gamma / self.rs)
            self.lambda_tf = 1.0 / math.sqrt(4.0 * PI * self.e2 * self.beta * self.cn * 0.5
                                             * float(fd_eval(FD_IM12, self.alpha0)[0]))
        else:
            self.lambda_tf = 1.0 / math.sqrt(3.0 * self.gamma)


@dataclass
class Species:
    units: Units
    structure: object
    terms: list          # list of (name, TermDesc)
    transport: object
    closure: str
    eta_code: float
    xi_code: float
    tau: float
    evolve: bool


def hnc_request(p: Params):
    """The HNC solve species_build asks for: (potential, n0, nr, rmax, tol, mix,
    maxiter, with_derivative), the arguments of hnc.structure_from_hnc."""
    un = Units(p.kind, p.gamma, p.kappa, p.rs, p.mass)
    if un.quantum:
        pot = _hnc.PairPotential("hansen_mcdonald", un.gamma, 0.0, lam=un.lambda_qsp)
    else:
        pot = _hnc.PairPotential("yukawa", un.gamma, un.kappa)
    need_deriv = p.closure == "maxwell" and p.tau_model == "ichimaru"
    return (pot, un.n0, p.hnc_nr, p.hnc_rmax, p.hnc_tol, p.hnc_mix, p.hnc_maxiter, need_deriv)


def build_species(p: Params, structure=None, cache=None) -> Species:
    """species_build (species.f90:49-153).  `cache`: an hnc.StructureCache
    that supplies (and keeps) the HNC c(k)."""
    un = Units(p.kind, p.gamma, p.kappa, p.rs, p.mass)
    st = structure
    if st is None:
        if p.structure == "hnc":
            req = hnc_request(p)
            st = cache.get(*req) if cache is not None else _hnc.structure_from_hnc(*req)
        elif p.structure == "file":
            if not p.structure_file:
                raise abi.HydftError(abi.HY_ERR_INVALID, "structure = file needs structure_file")
            st = _hnc.structure_from_file(p.structure_file, p.structure_kind, un.n0)
        elif p.structure == "none":
            st = None
        else:
            raise abi.HydftError(abi.HY_ERR_UNKNOWN_KIND, "unknown structure source: " + p.structure)
    terms = p.terms if p.terms.strip() else ("tf,gradient,hartree,ry" if un.quantum else "ideal,hartree,ry")
    has_hartree = "hartree" in terms
    descs = []
    for name in [s.strip() for s in terms.split(",") if s.strip()]:
        d = abi.TermDesc()
        d.n0, d.beta = un.n0, un.beta
        d.e2, d.kappa = un.e2, un.kappa
        if name == "ideal":
            d.kind = abi.HY_TERM_IDEAL
        elif name == "tf":
            if not un.quantum:
                raise abi.HydftError(abi.HY_ERR_INVALID, "tf term needs kind = electron")
            d.kind, d.cn, d.alpha0 = abi.HY_TERM_TF, un.cn, un.alpha0
        elif name == "gradient":
            if not un.quantum:
                raise abi.HydftError(abi.HY_ERR_INVALID, "gradient term needs kind = electron")
            d.kind, d.cn, d.alpha0 = abi.HY_TERM_GRADIENT, un.cn, un.alpha0
            d.cg = p.tfk_gamma * (3.0 * SQRT2 * PI * PI / 8.0) * un.hbar ** 5 * un.beta ** 1.5 / un.mass ** 2.5
        elif name == "hartree":
            d.kind = abi.HY_TERM_HARTREE
        elif name == "ry":
            if st is None:
                raise abi.HydftError(abi.HY_ERR_INVALID, "ry term needs a structure source (hnc or file)")
            d.kind = abi.HY_TERM_RY
            d.subtract_mean_field = int(has_hartree)
        else:
            raise abi.HydftError(abi.HY_ERR_UNKNOWN_KIND, "unknown free-energy term: " + name)
        descs.append((name, d))
    tr = _tr.transport_setup(p.transport, p.tau_model, p.eta, p.xi, p.tau, un, p.adiabatic_index, st,
                             p.closure == "maxwell")
    closure = p.closure.lower()
    if closure not in ("newtonian", "maxwell"):
        raise abi.HydftError(abi.HY_ERR_UNKNOWN_KIND, "unknown stress closure: " + closure)
    evolve = closure == "maxwell" and tr.tau > 0.0
    return Species(un, st, descs, tr, closure, tr.eta * un.n0, tr.xi * un.n0,
                   tr.tau if closure == "maxwell" else 0.0, evolve)


_VEXT = {"none": abi.HY_VEXT_NONE, "mode": abi.HY_VEXT_MODE, "driven": abi.HY_VEXT_DRIVEN,
         "gaussian": abi.HY_VEXT_GAUSSIAN}


class Model:
    """model_t on the GPU.  `precision`: 'fp64' or 'mixed' (FP32 transforms and
    functional kernels, FP64 RK state)."""

    def __init__(self, p: Params, structure=None, precision=None, species=None):
        abi.lib()
        self.p = p
        self.sp = species or build_species(p, structure)
        prec = (precision or p.precision).lower()
        if prec not in ("fp64", "mixed"):
            raise abi.HydftError(abi.HY_ERR_INVALID, "precision must be fp64 or mixed")
        self.precision = prec
        sp = self.sp
        n = len(sp.terms)
        arr = (abi.TermDesc * max(n, 1))()
        self._keep = []
        for i, (name, d) in enumerate(sp.terms):
            if name == "ry":
                k = np.ascontiguousarray(sp.structure.k, dtype=float)
                c = np.ascontiguousarray(sp.structure.c, dtype=float)
                self._keep += [k, c]
                d.ntab = k.size
                d.k_tab = k.ctypes.data_as(C.POINTER(C.c_double))
                d.c_tab = c.ctypes.data_as(C.POINTER(C.c_double))
            arr[i] = d
        desc = abi.ModelDesc()
        desc.nx, desc.ny, desc.nz = p.nx, p.ny, p.nz
        desc.lx, desc.ly, desc.lz = p.lx, p.ly, p.lz
        desc.dealias = int(bool(p.dealias))
        desc.precision = abi.HY_PREC_MIXED if prec == "mixed" else abi.HY_PREC_FP64
        desc.mass, desc.n0 = p.mass, sp.units.n0
        desc.nterms, desc.terms = n, arr
        desc.closure = abi.HY_CLOSURE_MAXWELL if sp.closure == "maxwell" else abi.HY_CLOSURE_NEWTONIAN
        desc.eta, desc.xi, desc.tau = sp.eta_code, sp.xi_code, sp.tau
        desc.eta_l = sp.transport.eta_l
        vk = p.vext.lower()
        if vk not in _VEXT:
            raise abi.HydftError(abi.HY_ERR_UNKNOWN_KIND, "unknown external potential: " + vk)
        desc.vext_kind = _VEXT[vk]
        desc.vext_amp = p.vext_amplitude
        for i in range(3):
            desc.vext_mode[i] = int(p.vext_mode[i])
        desc.vext_omega, desc.vext_width, desc.vext_ramp = p.vext_omega, p.vext_width, p.vext_ramp
        desc.dt, desc.cfl = p.dt, p.cfl
        self._desc = desc
        h = C.c_void_p()
        abi.call("hy_model_create", C.byref(desc), C.byref(h))
        self.h = h
        self.shape = (p.nz, p.ny, p.nx)
        self.n0 = sp.units.n0
        self.evolve = sp.evolve

    def close(self):
        if getattr(self, "h", None):
            abi.call("hy_model_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state
    def set_state(self, n, mom=None, pi=None):
        n = np.ascontiguousarray(n, dtype=float).reshape(self.shape)
        mp = None
        if mom is not None:
            mom = np.ascontiguousarray(mom, dtype=float).reshape((3,) + self.shape)
            mp = mom.ctypes.data
        pp = None
        if pi is not None:
            pi = np.ascontiguousarray(pi, dtype=float).reshape((6,) + self.shape)
            pp = pi.ctypes.data
        abi.call("hy_model_set_state", self.h, n.ctypes.data, mp, pp)

    def get_state(self):
        n = np.empty(self.shape)
        mom = np.empty((3,) + self.shape)
        pi = np.empty((6,) + self.shape) if self.evolve else None
        abi.call("hy_model_get_state", self.h, n.ctypes.data, mom.ctypes.data,
                 pi.ctypes.data if pi is not None else None)
        return n, mom, pi

    # ---- time
    @property
    def time(self):
        t, dt, st = C.c_double(), C.c_double(), C.c_int()
        abi.call("hy_model_get_time", self.h, C.byref(t), C.byref(dt), C.byref(st))
        return t.value, dt.value, st.value

    @property
    def t(self):
        return self.time[0]

    @property
    def dt(self):
        return self.time[1]

    @property
    def step(self):
        return self.time[2]

    def set_time(self, t, step=0):
        abi.call("hy_model_set_time", self.h, float(t), int(step))

    def set_dt(self, dt):
        abi.call("hy_model_set_dt", self.h, float(dt))

    def compute_dt(self):
        dt = C.c_double()
        abi.call("hy_model_compute_dt", self.h, C.byref(dt))
        return dt.value

    def advance(self, nsteps=1, dt_every=0, tend=0.0, probe=None):
        """nsteps RK3 steps; returns (taken, probe rows (t, Re, Im dn_k/n0, Re, Im mom_k, dt))."""
        taken = C.c_int()
        buf = None
        pm = None
        if probe is not None:
            buf = np.zeros((max(nsteps, 1), 6))
            pm = (C.c_int * 3)(*[int(v) for v in probe])
        abi.call("hy_model_advance", self.h, int(nsteps), int(dt_every), float(tend), pm,
                 buf.ctypes.data if buf is not None else None, C.byref(taken))
        return taken.value, (buf[:taken.value] if buf is not None else None)

    # ---- captured steps (batched and served runs)
    def graph_build(self, nsteps, dt_every=0, probe=None, ring=None):
        """Capture nsteps RK3 steps (dt recomputed every dt_every) into a graph;
        with `probe`, step s writes its probe row to ring slot (s-1) % ring."""
        pm = (C.c_int * 3)(*[int(v) for v in probe]) if probe is not None else None
        ring = int(ring or nsteps)
        abi.call("hy_model_graph_build", self.h, int(nsteps), int(dt_every), pm, ring)
        self._g_ring = ring

    def graph_launch(self, nlaunch=1):
        """Enqueue nlaunch replays on the model's stream (returns at once)."""
        abi.call("hy_model_graph_launch", self.h, int(nlaunch))

    def graph_probe(self):
        """The probe ring: rows (t, Re, Im dn_k/n0, Re, Im mom_k, dt), slot (s-1) % ring."""
        buf = np.zeros((self._g_ring, 6))
        abi.call("hy_model_graph_probe", self.h, buf.ctypes.data)
        return buf

    def idle(self):
        v = C.c_int()
        abi.call("hy_model_idle", self.h, C.byref(v))
        return bool(v.value)

    def check(self):
        abi.call("hy_model_check", self.h)

    def device_bytes(self):
        v = C.c_ulonglong()
        abi.call("hy_model_device_bytes", self.h, C.byref(v))
        return int(v.value)

    def kept(self):
        """The kept box (Kx, Ky, Kz) of the model's grid."""
        k = (C.c_int * 3)()
        abi.call("hy_model_kept", self.h, k)
        return tuple(int(v) for v in k)

    def probe_ring(self, ring):
        """`ring` rows of the device probe ring (slot (s-1) % ring for step s)."""
        buf = np.zeros((int(ring), 6))
        abi.call("hy_model_probe_ring", self.h, int(ring), buf.ctypes.data)
        return buf

    def rhs(self, t):
        dn = np.empty(self.shape)
        dmom = np.empty((3,) + self.shape)
        dpi = np.empty((6,) + self.shape) if self.evolve else None
        abi.call("hy_model_rhs", self.h, float(t), dn.ctypes.data, dmom.ctypes.data,
                 dpi.ctypes.data if dpi is not None else None)
        return dn, dmom, dpi

    def mu(self, n):
        n = np.ascontiguousarray(n, dtype=float).reshape(self.shape)
        out = np.empty(self.shape)
        abi.call("hy_model_mu", self.h, n.ctypes.data, out.ctypes.data)
        return out

    def budgets(self):
        out = np.zeros(6)
        abi.call("hy_model_budgets", self.h, out.ctypes.data)
        return out[0], out[1:4].copy(), out[4], out[5]

    def probe(self, mode=(1, 0, 0)):
        out = (C.c_double * 4)()
        abi.call("hy_model_probe", self.h, int(mode[0]), int(mode[1]), int(mode[2]), out)
        return complex(out[0], out[1]), complex(out[2], out[3])

    def kernel(self, k):
        out = C.c_double()
        abi.call("hy_model_kernel", self.h, float(k), C.byref(out))
        return out.value


class Batch:
    """Same-shape models stepped together: each kernel of a step is launched
    once for all of them (hy_batch_*).  The models keep their own state and
    probe rings; do not use them while the batch is in flight."""

    def __init__(self, models, nsteps, dt_every=0, probe=None, ring=None):
        self.models = list(models)
        hs = (C.c_void_p * len(self.models))(*[m.h for m in self.models])
        pm = (C.c_int * 3)(*[int(v) for v in probe]) if probe is not None else None
        self.ring = int(ring or nsteps)
        self.nsteps = int(nsteps)
        h = C.c_void_p()
        abi.call("hy_batch_create", hs, len(self.models), int(nsteps), int(dt_every), pm, self.ring,
                 C.byref(h))
        self.h = h.value

    def launch(self, nlaunch=1):
        abi.call("hy_batch_launch", self.h, int(nlaunch))

    def set_stops(self, stops):
        """Per-model step limits (multiples of nsteps; 0: none): a run stops there
        while the others go on."""
        arr = (C.c_int * len(self.models))(*[int(v) for v in stops])
        abi.call("hy_batch_set_stops", self.h, arr)

    def idle(self):
        v = C.c_int()
        abi.call("hy_batch_idle", self.h, C.byref(v))
        return bool(v.value)

    def sync(self):
        abi.call("hy_batch_sync", self.h)

    def probe(self, i):
        return self.models[i].probe_ring(self.ring)

    def close(self):
        if getattr(self, "h", None):
            abi.call("hy_batch_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    """Batches of different shapes (uneven grids) advanced by one graph launch
    per pass (hy_group_*): each batch is a branch of the graph.  Close the
    group before its batches."""

    def __init__(self, batches):
        self.batches = list(batches)
        hs = (C.c_void_p * len(self.batches))(*[b.h for b in self.batches])
        h = C.c_void_p()
        abi.call("hy_group_create", hs, len(self.batches), C.byref(h))
        self.h = h.value

    def launch(self, nlaunch=1):
        abi.call("hy_group_launch", self.h, int(nlaunch))

    def idle(self):
        v = C.c_int()
        abi.call("hy_group_idle", self.h, C.byref(v))
        return bool(v.value)

    def sync(self):
        abi.call("hy_group_sync", self.h)

    def close(self):
        if getattr(self, "h", None):
            abi.call("hy_group_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
