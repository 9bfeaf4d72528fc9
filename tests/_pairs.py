"""Shared test helpers: build a GPU model and the CPU oracle on identical
inputs (parameters, c(k) table, initial fields)."""
from __future__ import annotations

import functools

import numpy as np

from hydft_amd.host import hnc
from hydft_amd.host.params import Params
from oracle import hydft_oracle as O

N0 = 3.0 / (4.0 * np.pi)


@functools.lru_cache(maxsize=None)
def structure(kind, gamma, kappa, rs=1.86, nr=2048, rmax=40.0):
    """HNC c(k) table shared by the GPU model and the oracle."""
    if kind == "electron":
        un = O.Units(kind="electron", gamma=gamma, rs=rs)
        pot = hnc.PairPotential("hansen_mcdonald", gamma, 0.0, lam=un.lambda_qsp)
    else:
        pot = hnc.PairPotential("yukawa", gamma, kappa)
    return hnc.structure_from_hnc(pot, N0, nr=nr, rmax=rmax, tol=1e-10)


def params(**kw) -> Params:
    p = Params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p.finalize()


def oracle_model(p: Params, st):
    return O.make_model(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, kind=p.kind, gamma=p.gamma, kappa=p.kappa,
                        rs=p.rs, mass=p.mass, terms=p.terms, k_tab=st.k, c_tab=st.c,
                        tfk_gamma=p.tfk_gamma, closure=p.closure, eta=p.eta, xi=p.xi, tau=p.tau,
                        dt=p.dt, cfl=p.cfl, dealias=p.dealias, vext=p.vext,
                        vext_amplitude=p.vext_amplitude, vext_mode=p.vext_mode,
                        vext_omega=p.vext_omega, vext_width=p.vext_width, vext_ramp=p.vext_ramp)


def oracle_from_species(p: Params, sp):
    """The oracle built from a resolved species (hydft_amd.model.build_species):
    the same c(k) table, the transport scalars of its fit (eta_bar, xi_bar,
    tau, eta_l) and the same time-step settings."""
    st = sp.structure
    return O.make_model(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, kind=p.kind, gamma=p.gamma, kappa=p.kappa,
                        rs=p.rs, mass=p.mass, terms=p.terms,
                        k_tab=None if st is None else st.k, c_tab=None if st is None else st.c,
                        tfk_gamma=p.tfk_gamma, closure=sp.closure, eta=sp.transport.eta, xi=sp.transport.xi,
                        tau=sp.tau, eta_l=sp.transport.eta_l, dt=p.dt, cfl=p.cfl, dealias=p.dealias,
                        vext=p.vext, vext_amplitude=p.vext_amplitude, vext_mode=p.vext_mode,
                        vext_omega=p.vext_omega, vext_width=p.vext_width, vext_ramp=p.vext_ramp)


def smooth_fields(p: Params, seed, amp_n=0.02, amp_u=0.0, with_pi=False):
    """Band-limited random n, mom (and Pi): seeded uniform deviates, dealiased."""
    spc = O.Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.dealias)
    rng = np.random.default_rng(seed)
    n = O.ic_random_field(spc, N0, amp_n, rng.random(spc.g.shape))
    mom = np.zeros((3,) + spc.g.shape)
    for d in range(spc.g.ndim):
        if amp_u > 0:
            u = spc.dealias(rng.random(spc.g.shape) - 0.5)
            mom[d] = spc.dealias(n * amp_u * u / np.abs(u).max())
    pi = None
    if with_pi:
        pi = np.stack([spc.dealias(1e-3 * (rng.random(spc.g.shape) - 0.5)) for _ in range(6)])
    return n, mom, pi


def set_oracle_state(om, n, mom, pi=None):
    om.n = n.copy()
    om.mom = [mom[d].copy() for d in range(3)]
    if om.stress.evolve:
        om.pi = [pi[c].copy() for c in range(6)] if pi is not None else [np.zeros_like(n) for _ in range(6)]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
