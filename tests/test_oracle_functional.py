"""Pins the oracle's functional terms, stress and stepping (CPU).

- test_kernels.f90:49-69: every term's K_i(k) equals the finite difference of
  its mu_i[n] about n0 (relative 2e-4), for the three reference cases.
- Stress: div Pi0 = eta lap u + (eta/3 + xi) grad(div u) for constant eta, xi.
- RHS invariants: a uniform state has zero RHS; sum dn = 0; sum dmom =
  sum n F_ext.
- One RK3 step of a small single mode equals the RK3 stability polynomial
  P(M) = I + M + M^2/2 + M^3/6 of the linearised 2x2 system."""
import numpy as np
import pytest

from oracle import hydft_oracle as O
from tests import _pairs as P

TWOPI = O.TWOPI


@pytest.mark.parametrize("kind,terms", [("ion", "ideal,hartree,ry"), ("electron", "tf,gradient,hartree,ry"),
                                        ("electron", "tf,gradient,ry")])
def test_kernel_vs_finite_difference(kind, terms):
    gamma = 1.0 if kind == "electron" else 2.0
    st = P.structure(kind, gamma, 0.1, 1.86, nr=4096, rmax=60.0)
    m = O.make_model(64, 1, 1, 20.0, 1.0, 1.0, kind=kind, gamma=gamma, kappa=0.1, rs=1.86, terms=terms,
                     k_tab=st.k, c_tab=st.c)
    spc, n0, eps = m.spc, m.n0, 1.0e-6
    n0a = np.full(spc.g.shape, n0)
    n0hat = spc.fwd(n0a)
    for mm in range(1, 7):
        k = TWOPI * mm / 20.0
        n = n0 * (1.0 + eps * np.cos(k * spc.g.x)).reshape(spc.g.shape)
        nhat = spc.fwd(n)
        for t in m.fun.terms:
            mu = t.add_mu(spc, n, nhat, np.zeros(spc.g.shape))
            mu0 = t.add_mu(spc, n0a, n0hat, np.zeros(spc.g.shape))
            kfd = 2.0 * spc.mode(mu - mu0, mm, 0, 0).real / (eps * n0)
            kan = float(t.kernel(np.array([k]))[0])
            assert abs(kfd - kan) / max(abs(kan), 1e-3) < 2e-4, (t.name, mm)
        mu = m.fun.mu(spc, n, nhat)
        mu0 = m.fun.mu(spc, n0a, n0hat)
        kfd = 2.0 * spc.mode(mu - mu0, mm, 0, 0).real / (eps * n0)
        kan = float(m.fun.kernel(np.array([k]))[0])
        assert abs(kfd - kan) / abs(kan) < 2e-4


def test_stress_divergence_identity():
    spc = O.Spec(16, 12, 8, 10.0, 7.0, 5.0)
    rng = np.random.default_rng(5)
    u = [spc.dealias(rng.standard_normal(spc.g.shape)) for _ in range(3)]
    eta, xi = 0.7, 0.25
    st = O.Stress("newtonian", eta, xi, 0.0, 1.0)
    f = st.divergence(spc, st.pi0(spc, u))
    divu = spc.div(u)
    gdiv = spc.grad(divu)
    for i in range(3):
        ex = eta * spc.lap(u[i]) + (eta / 3.0 + xi) * gdiv[i]
        assert np.abs(f[i] - ex).max() < 1e-11 * np.abs(ex).max()


def test_rhs_invariants():
    st = P.structure("ion", 2.0, 0.1)
    m = O.make_model(16, 8, 8, 20.0, 10.0, 10.0, gamma=2.0, kappa=0.1, k_tab=st.k, c_tab=st.c, eta=0.3,
                     xi=0.1)
    dn, dmom, _ = m.rhs(0.0)
    assert np.abs(dn).max() < 1e-15 and max(np.abs(x).max() for x in dmom) < 1e-15
    p = P.params(nx=16, ny=8, nz=8, lx=20.0, ly=10.0, lz=10.0, eta=0.3, xi=0.1, vext="driven",
                 vext_amplitude=1e-3, vext_mode=(1, 1, 0), vext_omega=0.5, vext_ramp=0.0)
    m = P.oracle_model(p, st)
    n, mom, _ = P.smooth_fields(p, 11, amp_n=0.05, amp_u=0.02)
    P.set_oracle_state(m, n, mom)
    dn, dmom, _ = m.rhs(0.37)
    assert abs(dn.sum()) < 1e-14 * np.abs(dn).sum()
    m.ext.eval(m.spc, 0.37)
    for i in range(3):
        ref = np.sum(m.n * m.ext.f[i])
        assert abs(dmom[i].sum() - ref) < 1e-12 * (np.abs(dmom[i]).sum() + abs(ref)) + 1e-14


def test_rk3_polynomial_single_mode():
    st = P.structure("ion", 2.0, 0.1)
    m = O.make_model(32, 1, 1, 20.0, 1.0, 1.0, gamma=2.0, kappa=0.1, k_tab=st.k, c_tab=st.c, dt=0.3)
    amp = 1e-8
    mm = 2
    k = TWOPI * mm / 20.0
    m.n = O.ic_mode(m.spc, m.n0, amp, (mm, 0, 0))
    nk0 = m.spc.mode(m.n, mm, 0, 0)
    mk0 = m.spc.mode(m.mom[0], mm, 0, 0)
    m.advance()
    nk1 = m.spc.mode(m.n, mm, 0, 0)
    mk1 = m.spc.mode(m.mom[0], mm, 0, 0)
    K = float(m.fun.kernel(np.array([k]))[0])
    # linearised: dn_k/dt = -i k m_k / mass, dm_k/dt = -i k n0 K n_k
    A = np.array([[0.0, -1j * k], [-1j * k * m.n0 * K, 0.0]]) * m.dt
    Pm = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6
    pred = Pm @ np.array([nk0, mk0])
    assert abs(nk1 - pred[0]) < 1e-6 * abs(nk0)
    assert abs(mk1 - pred[1]) < 1e-6 * abs(nk0) * np.sqrt(m.n0 * K)
    # a first-order (Euler) or second-order update would miss by O(z^3) or more
    z = k * np.sqrt(m.n0 * K) * m.dt
    assert z ** 3 / 6 > 1e-4
