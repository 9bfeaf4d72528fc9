"""Host-side static structure: pair potentials and the OZ-HNC solver that
produce the radial c(k) table the RY correlation kernel interpolates.

One-time FP64 host setup (not on the time-stepping path).  Follows
pair_potential.f90 (potentials and the long/short split) and oz_hnc.f90
(Picard + Ng iteration on gamma_s = h - c_s with a DST-I Fourier-Bessel pair,
restarts with halved mixing, continuation in Gamma).  The DST-I is
scipy.fft.dst(type=1), the same unnormalised transform as FFTW's RODFT00.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.fft

PI = math.pi
LN2 = 0.693147180559945309417232121458


class HncError(RuntimeError):
    pass


@dataclass
class PairPotential:
    """pair_potential.f90:16-139: yukawa | coulomb | hansen_mcdonald."""
    name: str
    gamma: float
    kappa: float = 0.0
    ksplit: float = 1.0
    lam: float = 1.0
    pauli: bool = True

    def beta_v(self, r):
        r = np.asarray(r, float)
        if self.name == "yukawa":
            return self.gamma * np.exp(-self.kappa * r) / r
        if self.name == "coulomb":
            return self.gamma / r
        if self.name in ("hansen_mcdonald", "hm", "qsp"):
            rr = np.where(r > 0.0, r, 1.0)
            v = np.where(r > 0.0, self.gamma * (1.0 - np.exp(-2.0 * PI * r / self.lam)) / rr,
                         self.gamma * 2.0 * PI / self.lam)
            if self.pauli:
                v = v + LN2 * np.exp(-4.0 * PI * r * r / (self.lam ** 2 * LN2))
            return v
        raise HncError("unknown pair potential: " + self.name)

    def beta_v_long_r(self, r):
        r = np.asarray(r, float)
        kp = self.kappa + self.ksplit
        rr = np.where(r > 1.0e-12, r, 1.0)
        return np.where(r > 1.0e-12, self.gamma * (np.exp(-self.kappa * r) - np.exp(-kp * r)) / rr,
                        self.gamma * (kp - self.kappa))

    def beta_v_long_k(self, k):
        kp = self.kappa + self.ksplit
        return 4.0 * PI * self.gamma * (1.0 / (k * k + self.kappa ** 2) - 1.0 / (k * k + kp * kp))

    def beta_v_short(self, r):
        return self.beta_v(r) - self.beta_v_long_r(r)

    def beta_v_k(self, k):
        return 4.0 * PI * self.gamma / (k * k + self.kappa ** 2)


class FourierBessel:
    """Radial transform pair on r_j = j dr, k_m = m dk, dk = pi/((n+1) dr) (oz_hnc.f90:57-96)."""

    def __init__(self, n, rmax):
        self.n = n
        self.dr = rmax / (n + 1)
        self.dk = PI / ((n + 1) * self.dr)
        self.r = np.arange(1, n + 1) * self.dr
        self.k = np.arange(1, n + 1) * self.dk

    def r2k(self, fr):
        out = scipy.fft.dst(self.r * fr, type=1)
        return 2.0 * PI * self.dr * out / self.k

    def k2r(self, fk):
        out = scipy.fft.dst(self.k * fk, type=1)
        return self.dk * out / (4.0 * PI * PI * self.r)


@dataclass
class HncResult:
    r: np.ndarray
    k: np.ndarray
    g: np.ndarray
    cr: np.ndarray
    ck: np.ndarray
    hk: np.ndarray
    sk: np.ndarray
    u_ex: float
    converged: bool
    niter: int
    residual: float


def _solve_direct(pot, n0, nr, rmax, tol, mix, maxiter, gamma_s_init=None):
    expcap = 40.0
    npicard = 5
    fb = FourierBessel(nr, rmax)
    bvs = pot.beta_v_short(fb.r)
    bvl_r = pot.beta_v_long_r(fb.r)
    bvl_k = pot.beta_v_long_k(fb.k)
    ck = -pot.beta_v_k(fb.k)
    gk_mf = n0 * ck * ck / (1.0 - n0 * ck) - bvl_k
    gs = fb.k2r(gk_mf)
    if gamma_s_init is not None and len(gamma_s_init) == nr:
        gs = gamma_s_init.copy()
    mixc = mix
    converged = False
    unstable = False
    resid = 0.0
    it_total = 0
    for _ in range(4):
        hist_in = np.zeros((3, nr))
        hist_out = np.zeros((3, nr))
        nhist = 0
        ng_cooldown = 0
        resid_prev = 1.0e100
        unstable = False
        for it in range(1, maxiter + 1):
            it_total = it
            cs = np.exp(np.minimum(-bvs + gs, expcap)) - 1.0 - gs
            ck = fb.r2k(cs) - bvl_k
            if np.any(n0 * ck >= 1.0):
                unstable = True
                break
            gk = n0 * ck * ck / (1.0 - n0 * ck) - bvl_k
            gnew = fb.k2r(gk)
            resid = float(np.max(np.abs(gnew - gs)))
            if resid < tol:
                gs = gnew
                converged = True
                break
            if resid > 1.0e6:
                unstable = True
                break
            hist_in[2] = hist_in[1]
            hist_in[1] = hist_in[0]
            hist_in[0] = gs
            hist_out[2] = hist_out[1]
            hist_out[1] = hist_out[0]
            hist_out[0] = gnew
            nhist = min(nhist + 1, 3)
            if resid > 10.0 * resid_prev:
                ng_cooldown = 10
            if nhist == 3 and it > npicard and ng_cooldown == 0:
                d0 = hist_out[0] - hist_in[0]
                d01 = d0 - (hist_out[1] - hist_in[1])
                d02 = d0 - (hist_out[2] - hist_in[2])
                a11, a12, a22 = d01 @ d01, d01 @ d02, d02 @ d02
                b1, b2 = d0 @ d01, d0 @ d02
                det = a11 * a22 - a12 * a12
                accepted = False
                if abs(det) > 1.0e-14 * a11 * a22 and a11 > 0.0 and a22 > 0.0:
                    c1 = (b1 * a22 - b2 * a12) / det
                    c2 = (a11 * b2 - a12 * b1) / det
                    if abs(c1) < 20.0 and abs(c2) < 20.0:
                        gtry = (1.0 - c1 - c2) * hist_out[0] + c1 * hist_out[1] + c2 * hist_out[2]
                        if np.max(gtry - bvs) < expcap:
                            gs = gtry
                            accepted = True
                if not accepted:
                    gs = (1.0 - mixc) * gs + mixc * gnew
            else:
                gs = (1.0 - mixc) * gs + mixc * gnew
                if ng_cooldown > 0:
                    ng_cooldown -= 1
            resid_prev = resid
        if converged or not unstable:
            break
        mixc *= 0.5
        gs = fb.k2r(gk_mf)
        if gamma_s_init is not None and len(gamma_s_init) == nr:
            gs = gamma_s_init.copy()
    if unstable:
        return None, gs
    return _finish(pot, n0, fb, bvs, bvl_r, bvl_k, gs, converged, min(it_total, maxiter), resid), gs


def _finish(pot, n0, fb, bvs, bvl_r, bvl_k, gs, converged, niter, resid, expcap=40.0):
    """g, c(r), c(k), h(k), S(k) and the excess energy from the final gamma_s."""
    g = np.exp(np.minimum(-bvs + gs, expcap))
    cs = g - 1.0 - gs
    ck = fb.r2k(cs) - bvl_k
    hk = ck / (1.0 - n0 * ck)
    sk = 1.0 + n0 * hk
    cr = cs - bvl_r
    u_ex = 2.0 * PI * n0 * (fb.dr * float(np.sum(fb.r ** 2 * (g - 1.0) * pot.beta_v(fb.r))))
    return HncResult(fb.r, fb.k, g, cr, ck, hk, sk, u_ex, converged, niter, resid)


def hnc_solve(pot: PairPotential, n0, nr=8192, rmax=80.0, tol=1.0e-9, mix=0.2, maxiter=20000):
    """hnc_solve (oz_hnc.f90:111-140): direct solve, else continuation in Gamma."""
    res, gs = _solve_direct(pot, n0, nr, rmax, tol, mix, maxiter)
    if res is not None:
        return res
    gtarget = pot.gamma
    g = min(gtarget, 2.0)
    gs_init = np.zeros(nr)
    p = PairPotential(pot.name, g, pot.kappa, pot.ksplit, pot.lam, pot.pauli)
    for _ in range(200):
        p.gamma = g
        res, gs = _solve_direct(p, n0, nr, rmax, tol, mix, maxiter, gamma_s_init=gs_init)
        if res is None:
            raise HncError("hnc: continuation in Gamma failed")
        gs_init = gs
        if g >= gtarget:
            break
        g = min(gtarget, g * 1.3)
    return res


def hnc_solve_many_device(pots, n0s, nr=8192, rmax=80.0, tol=1.0e-9, mix=0.2, maxiter=20000):
    """hnc_solve for many state points on one radial grid: the direct solves
    run together on the GPU (hy_hnc_iterate: the same Picard + Ng iteration,
    restarts and rules as _solve_direct, with the DST-I as a matrix product
    on the device); a state point whose direct solve is unstable in every
    attempt continues in Gamma on the host (hnc_solve).  Returns HncResults
    in input order."""
    import ctypes as C
    from .. import abi
    fb = FourierBessel(nr, rmax)
    m = len(pots)
    bvs = np.empty((m, nr)); bvl_r = np.empty((m, nr)); bvl_k = np.empty((m, nr)); gs0 = np.empty((m, nr))
    for i, (pot, n0) in enumerate(zip(pots, n0s)):
        bvs[i] = pot.beta_v_short(fb.r)
        bvl_r[i] = pot.beta_v_long_r(fb.r)
        bvl_k[i] = pot.beta_v_long_k(fb.k)
        ck = -pot.beta_v_k(fb.k)
        gs0[i] = fb.k2r(n0 * ck * ck / (1.0 - n0 * ck) - bvl_k[i])
    n0a = np.ascontiguousarray(n0s, dtype=float)
    gs = np.empty((m, nr))
    status = np.empty(m, dtype=np.int32)
    niter = np.empty(m, dtype=np.int32)
    resid = np.empty(m)
    abi.call("hy_hnc_iterate", m, nr, fb.dr, fb.dk, n0a.ctypes.data, bvs.ctypes.data, bvl_k.ctypes.data,
             gs0.ctypes.data, float(tol), float(mix), int(maxiter), gs.ctypes.data,
             status.ctypes.data_as(C.POINTER(C.c_int)), niter.ctypes.data_as(C.POINTER(C.c_int)), resid.ctypes.data)
    out = []
    for i, (pot, n0) in enumerate(zip(pots, n0s)):
        if status[i] < 0:
            out.append(hnc_solve(pot, n0, nr, rmax, tol, mix, maxiter))
        else:
            out.append(_finish(pot, n0, fb, bvs[i], bvl_r[i], bvl_k[i], gs[i], bool(status[i] == 1),
                               int(niter[i]), float(resid[i])))
    return out


@dataclass
class Structure:
    """structure_t (structure_source.f90:18-41): radial c(k) table (+ HNC energy)."""
    k: np.ndarray
    c: np.ndarray
    n0: float
    u_ex: float = 0.0
    du_dgamma: float = 0.0
    has_energy: bool = False
    source: str = "hnc"


def structure_from_hnc(pot: PairPotential, n0, nr=8192, rmax=80.0, tol=1.0e-9, mix=0.2,
                       maxiter=20000, with_derivative=False) -> Structure:
    """structure_from_hnc (structure_source.f90:50-83)."""
    res = hnc_solve(pot, n0, nr, rmax, tol, mix, maxiter)
    if not res.converged:
        raise HncError("structure: HNC did not converge")
    st = Structure(k=res.k.copy(), c=res.ck.copy(), n0=n0, u_ex=res.u_ex, has_energy=True)
    if with_derivative:
        g0 = pot.gamma
        dg = 0.02 * g0
        pp = PairPotential(pot.name, g0 + dg, pot.kappa, pot.ksplit, pot.lam, pot.pauli)
        rp1 = hnc_solve(pp, n0, nr, rmax, tol, mix, maxiter)
        pm = PairPotential(pot.name, g0 - dg, pot.kappa, pot.ksplit, pot.lam, pot.pauli)
        rm1 = hnc_solve(pm, n0, nr, rmax, tol, mix, maxiter)
        if not (rp1.converged and rm1.converged):
            raise HncError("structure: HNC derivative runs did not converge")
        st.du_dgamma = (rp1.u_ex - rm1.u_ex) / (2.0 * dg)
    return st


def structure_from_file(path, kind, n0) -> Structure:
    """structure_from_file (structure_source.f90:87-138): two columns (k, S) or (k, c)."""
    kk, vv = [], []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            parts = s.replace(",", " ").split()
            try:
                a, b = float(parts[0]), float(parts[1])
            except (ValueError, IndexError):
                continue
            kk.append(a)
            vv.append(b)
    if len(kk) < 2:
        raise HncError("structure file has fewer than two data rows: " + path)
    k = np.array(kk)
    v = np.array(vv)
    kind = kind[:1].lower()
    if kind == "s":
        c = (1.0 - 1.0 / v) / n0
    elif kind == "c":
        c = v
    else:
        raise HncError("structure file kind must be s or c")
    if np.any(np.diff(k) <= 0.0):
        raise HncError("structure file: k must be strictly increasing")
    if k[0] <= 0.0:
        raise HncError("structure file: k must be positive")
    return Structure(k=k, c=c, n0=n0, source="file")


# ------------------------------------------------------------ c(k) supply
def _structure_key(pot: PairPotential, n0, nr, rmax, tol, mix, maxiter, with_derivative) -> tuple:
    return (pot.name, float(pot.gamma), float(pot.kappa), float(pot.ksplit), float(pot.lam), bool(pot.pauli),
            float(n0), int(nr), float(rmax), float(tol), float(mix), int(maxiter), bool(with_derivative))


class StructureCache:
    """c(k) tables keyed by the potential, the state point and the solver
    settings, in memory and optionally on disk (one .npz per key, written and
    read by this class only).  Batched and served runs reuse a solve whenever
    the key repeats; `solve_many` fills the cache for a batch of distinct
    keys in parallel processes."""

    def __init__(self, directory=None, maxsize=1024):
        self.directory = directory
        self.maxsize = maxsize
        self._mem = {}
        self.hits = self.misses = 0
        if directory:
            import os
            os.makedirs(directory, exist_ok=True)

    def _path(self, key):
        import hashlib
        import os
        h = hashlib.sha1(repr(key).encode()).hexdigest()[:24]
        return os.path.join(self.directory, f"ck_{h}.npz")

    def _load(self, key):
        import os
        if not self.directory:
            return None
        path = self._path(key)
        if not os.path.exists(path):
            return None
        z = np.load(path, allow_pickle=False)
        if tuple(z["key"].tolist()) != tuple(str(v) for v in key):
            return None
        return Structure(k=z["k"], c=z["c"], n0=float(z["n0"]), u_ex=float(z["u_ex"]),
                         du_dgamma=float(z["du_dgamma"]), has_energy=True)

    def _store(self, key, st):
        self._mem[key] = st
        if len(self._mem) > self.maxsize:
            self._mem.pop(next(iter(self._mem)))
        if self.directory:
            np.savez(self._path(key), key=np.array([str(v) for v in key]), k=st.k, c=st.c, n0=st.n0,
                     u_ex=st.u_ex, du_dgamma=st.du_dgamma)

    def get(self, pot: PairPotential, n0, nr=8192, rmax=80.0, tol=1.0e-9, mix=0.2, maxiter=20000,
            with_derivative=False) -> Structure:
        key = _structure_key(pot, n0, nr, rmax, tol, mix, maxiter, with_derivative)
        st = self._mem.get(key)
        if st is None:
            st = self._load(key)
            if st is not None:
                self._mem[key] = st
        if st is not None:
            self.hits += 1
            return st
        self.misses += 1
        st = structure_from_hnc(pot, n0, nr, rmax, tol, mix, maxiter, with_derivative)
        self._store(key, st)
        return st

    def solve_many(self, requests, processes=None, device=False):
        """requests: (pot, n0, nr, rmax, tol, mix, maxiter, with_derivative)
        tuples.  Solves the keys not yet cached, in a process pool or, with
        `device`, together on the GPU (hnc_solve_many_device), and returns the
        Structures in request order."""
        import concurrent.futures as cf
        keys = [_structure_key(*r) for r in requests]
        todo = {}
        for key, r in zip(keys, requests):
            if key not in self._mem and self._load(key) is None and key not in todo:
                todo[key] = r
        if todo and device:
            self._solve_device(todo)
        elif todo:
            workers = processes or min(len(todo), 8)
            with cf.ProcessPoolExecutor(max_workers=workers) as ex:
                futs = {key: ex.submit(structure_from_hnc, *r) for key, r in todo.items()}
                for key, f in futs.items():
                    self._store(key, f.result())
                    self.misses += 1
        return [self.get(*r) for r in requests]

    def _solve_device(self, todo):
        """structure_from_hnc for every key, grouped by solver settings; the
        derivative runs (Gamma +- 2 %) join the same device batch."""
        groups = {}
        for key, r in todo.items():
            groups.setdefault(tuple(r[2:7]), []).append((key, r))
        for (nr, rmax, tol, mix, maxiter), items in groups.items():
            pots, n0s, slots = [], [], []
            for key, (pot, n0, *_rest, with_derivative) in items:
                idx = [len(pots)]
                pots.append(pot); n0s.append(n0)
                if with_derivative:
                    dg = 0.02 * pot.gamma
                    for gg in (pot.gamma + dg, pot.gamma - dg):
                        idx.append(len(pots))
                        pots.append(PairPotential(pot.name, gg, pot.kappa, pot.ksplit, pot.lam, pot.pauli))
                        n0s.append(n0)
                slots.append((key, pot, n0, with_derivative, idx))
            res = hnc_solve_many_device(pots, n0s, nr, rmax, tol, mix, maxiter)
            for key, pot, n0, with_derivative, idx in slots:
                r0 = res[idx[0]]
                if not r0.converged:
                    raise HncError("structure: HNC did not converge")
                st = Structure(k=r0.k.copy(), c=r0.ck.copy(), n0=n0, u_ex=r0.u_ex, has_energy=True)
                if with_derivative:
                    rp1, rm1 = res[idx[1]], res[idx[2]]
                    if not (rp1.converged and rm1.converged):
                        raise HncError("structure: HNC derivative runs did not converge")
                    st.du_dgamma = (rp1.u_ex - rm1.u_ex) / (2.0 * 0.02 * pot.gamma)
                self._store(key, st)
                self.misses += 1
