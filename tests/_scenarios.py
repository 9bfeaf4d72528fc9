"""The reference's integration tests (test_conservation, test_mode_newtonian,
test_mode_maxwell, test_driven_mode, test_ground_state) written once against a
small backend interface, run on the CPU oracle and on the GPU model."""
from __future__ import annotations

import math

import numpy as np

from hydft_amd.host.analysis import fit_damped_mode, fit_driven_response
from hydft_amd.host.linear import Linear
from oracle import hydft_oracle as O
from tests import _pairs as P

TWOPI = O.TWOPI


class OracleBackend:
    def __init__(self, p, st):
        self.p = p
        self.m = P.oracle_model(p, st)
        self.n0 = self.m.n0

    def set_state(self, n, mom=None):
        self.m.n = n.copy()
        if mom is not None:
            self.m.mom = [mom[d].copy() for d in range(3)]

    def run(self, tend, mode):
        """do while (t < tend): compute_dt; advance; record mode of n."""
        ts, zs = [], []
        while self.m.t < tend:
            self.m.compute_dt()
            self.m.advance()
            ts.append(self.m.t)
            zs.append(self.m.spc.mode(self.m.n, *mode))
        return np.array(ts), np.array(zs)

    def steps(self, nsteps):
        for _ in range(nsteps):
            self.m.compute_dt()
            self.m.advance()

    def budgets(self):
        mass, mom, ekin, fe = self.m.budgets()
        return mass, np.array(mom), ekin, fe

    def kernel(self, k):
        return float(self.m.fun.kernel(np.array([k]))[0])

    def mu_plus_v(self):
        nhat = self.m.spc.fwd(self.m.n)
        mu = self.m.fun.mu(self.m.spc, self.m.n, nhat)
        return mu + self.m.ext.v

    def mode_n(self, mode):
        return self.m.spc.mode(self.m.n, *mode)

    @property
    def t(self):
        return self.m.t


class GpuBackend:
    def __init__(self, p, st, precision="fp64"):
        from hydft_amd.model import Model
        self.p = p
        self.m = Model(p, structure=st, precision=precision)
        self.n0 = self.m.n0
        self.om = None

    def set_state(self, n, mom=None):
        self.m.set_state(n, mom)

    def run(self, tend, mode):
        taken, pr = self.m.advance(10 ** 7, dt_every=1, tend=-tend, probe=mode)
        return pr[:, 0], (pr[:, 1] + 1j * pr[:, 2]) * self.n0

    def steps(self, nsteps):
        self.m.advance(nsteps, dt_every=1)

    def budgets(self):
        return self.m.budgets()

    def kernel(self, k):
        return self.m.kernel(k)

    def mu_plus_v(self):
        n, _, _ = self.m.get_state()
        mu = self.m.mu(n)
        o = P.oracle_model(self.p, self.m.sp.structure)  # static v_ext on the grid
        return mu + o.ext.v

    def mode_n(self, mode):
        a, _ = self.m.probe(mode)
        return a * self.n0

    @property
    def t(self):
        return self.m.t


def linear_for(backend, p, st):
    un = O.Units(kind=p.kind, gamma=p.gamma, kappa=p.kappa, rs=p.rs, mass=p.mass)
    eta_l = 4.0 * p.eta / 3.0 + p.xi
    tau = p.tau if p.closure == "maxwell" else 0.0
    return Linear(un.n0, lambda k: np.array([backend.kernel(float(np.atleast_1d(k)[0]))]), p.mass, eta_l, tau,
                  un.beta, un.hbar, un.quantum)


# ------------------------------------------------------------------ scenarios
def conservation(make, closure, eta, tau):
    """test_conservation.f90:27-54."""
    p = P.params(kind="ion", gamma=2.0, kappa=0.1, nx=32, ny=4, nz=1, lx=20.0, ly=5.0, transport="constant",
                 eta=eta, closure=closure, tau_model="constant", tau=tau, cfl=0.3)
    st = P.structure("ion", 2.0, 0.1, nr=4096, rmax=60.0)
    b = make(p, st)
    spc = O.Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz)
    rng = np.random.default_rng(7)
    n = O.ic_random_field(spc, P.N0, 0.02, rng.random(spc.g.shape))
    b.set_state(n)
    mass0, mom0, ekin0, fe0 = b.budgets()
    b.steps(100)
    mass, mom, ekin, fe = b.budgets()
    out = {"mass": abs(mass / mass0 - 1.0),
           "mom": float(np.max(np.abs(np.asarray(mom) - np.asarray(mom0)))) / (mass0 * math.sqrt(ekin / mass0)),
           "ekin": ekin, "energy": (ekin + fe - ekin0 - fe0) / abs(ekin0 + fe0)}
    return out


def single_mode(make, kind, closure="newtonian", eta=0.1, tau=0.0, tend=100.0, tfit=5.0):
    """test_mode_newtonian.f90:26-61 / test_mode_maxwell.f90:103-138."""
    gamma = 1.0 if kind == "electron" else 2.0
    p = P.params(kind=kind, gamma=gamma, kappa=0.1, rs=1.86, nx=32, lx=20.0, transport="constant", eta=eta,
                 closure=closure, tau_model="constant", tau=tau, cfl=0.1)
    st = P.structure(kind, gamma, 0.1, 1.86, nr=4096, rmax=60.0)
    b = make(p, st)
    mode = 2
    spc = O.Spec(p.nx, 1, 1, p.lx, 1.0, 1.0)
    b.set_state(O.ic_mode(spc, P.N0, 1.0e-4, (mode, 0, 0)))
    lin = linear_for(b, p, st)
    q = TWOPI * mode / p.lx
    w, _ = lin.dispersion(q)
    ts, zs = b.run(tend, (mode, 0, 0))
    om, ga, nz = fit_damped_mode(ts, zs.real, tfit)
    return {"w_lin": w, "om": om, "ga": ga, "nz": nz, "steps": len(ts)}


def driven(make, closure, tau, wd_over_w0=0.7):
    """test_driven_mode.f90:37-68."""
    mode, amp = 2, 1.0e-5
    p = P.params(kind="ion", gamma=2.0, kappa=0.1, nx=32, lx=20.0, transport="constant", eta=0.3,
                 closure=closure, tau_model="constant", tau=tau, cfl=0.1, vext="driven",
                 vext_amplitude=amp, vext_mode=(mode, 0, 0), vext_ramp=10.0)
    st = P.structure("ion", 2.0, 0.1, nr=4096, rmax=60.0)
    q = TWOPI * mode / p.lx
    # w_drive from the linear kernel of the same functional
    probe = make(p, st)
    lin = linear_for(probe, p, st)
    wd = wd_over_w0 * q * math.sqrt(lin.c2(q))
    p.vext_omega = wd
    b = make(p, st)
    b.set_state(np.full((1, 1, p.nx), P.N0))
    ts, zs = b.run(300.0, (mode, 0, 0))
    a, scatter = fit_driven_response(ts, zs, wd, 250.0)
    chi_sim = 2.0 * a / amp
    chi_lin = lin.chi(q, wd)
    return {"chi_sim": chi_sim, "chi_lin": chi_lin, "scatter": scatter, "err": abs(chi_sim - chi_lin) / abs(chi_lin)}


def ground_state(make, kind):
    """test_ground_state.f90:36-62."""
    gamma = 1.0 if kind == "electron" else 2.0
    amp, mode = 1.0e-3, 1
    p = P.params(kind=kind, gamma=gamma, kappa=0.1, rs=1.86, nx=32, lx=20.0, transport="constant", eta=1.0,
                 closure="newtonian", cfl=0.2, vext="mode", vext_amplitude=amp, vext_mode=(mode, 0, 0))
    st = P.structure(kind, gamma, 0.1, 1.86, nr=4096, rmax=60.0)
    b = make(p, st)
    b.set_state(np.full((1, 1, p.nx), P.N0))
    b.run(150.0, (mode, 0, 0))
    tot = b.mu_plus_v()
    spread = (tot.max() - tot.min()) / amp
    q = TWOPI * mode / p.lx
    dnk = b.mode_n((mode, 0, 0))
    pred = -0.5 * amp / b.kernel(q)
    return {"spread": spread, "err": abs(dnk - pred) / abs(pred)}
