"""Closed forms that pin the oracle parts no reference test fixes.

1. The Kirzhnits gradient constant C = gamma (3 sqrt2 pi^2/8) hbar^5
   beta^(3/2) / m^(5/2) (species.f90:127).  In the degenerate limit
   theta -> 0 the gradient kernel 2 C xi(alpha0) k^2 (fe_kinetic_tfk.f90:
   140-145) must tend to the von Weizsaecker form scaled by gamma,
   gamma hbar^2 k^2 / (4 m n0): with I_{-1/2}(a) ~ 2 sqrt(a) and
   n = cn I_{1/2}(a) ~ cn (2/3) a^(3/2), xi = I'_{-1/2}/I_{-1/2}^2 ~ cn/(6 n0),
   and C cn = (3/4) gamma hbar^2/m.  A wrong power of hbar, beta or m, or a
   wrong numerical factor, moves the limit.  Checked on the oracle and on the
   descriptor that the host factory (hydft_amd.model.build_species) hands to
   the library, for two masses.
2. model_compute_dt (model.f90:150-183) on a 1-D ideal-gas ion model, where
   every branch has a closed form: K(k) = 1/(beta n0), so w_max =
   k_max sqrt(1/(beta m)) over the kept modes; nu_max = eta_l/m k_max^2;
   dx/u_max; and the Maxwell wave speed with eta_l/(m tau) added.
"""
import math

import numpy as np
import pytest

from oracle import hydft_oracle as O
from tests import _pairs as P

TWOPI = 2.0 * math.pi


def _gamma_of_theta(rs, theta):
    """params.f90:170: Gamma from r_s and theta for electrons."""
    return 2.0 * rs / (theta * (9.0 * math.pi / 4.0) ** (2.0 / 3.0))


def _vw(un, k, gam=1.0 / 9.0):
    return gam * un.hbar ** 2 * k * k / (4.0 * un.mass * un.n0)


@pytest.mark.parametrize("mass", [1.0, 2.0])
def test_gradient_kernel_degenerate_limit_oracle(mass):
    k = np.array([0.3, 1.0, 2.5])
    errs = []
    for theta in (0.04, 0.02, 0.01):
        un = O.Units(kind="electron", gamma=_gamma_of_theta(1.0, theta), rs=1.0, mass=mass)
        g = O.Gradient(un, 1.0 / 9.0)
        rel = np.abs(g.kernel(k) / _vw(un, k) - 1.0).max()
        errs.append(rel)
    # the Sommerfeld corrections vanish like theta^2
    assert errs[-1] < 2e-3, errs
    assert errs[1] < 0.3 * errs[0] and errs[2] < 0.3 * errs[1], errs


@pytest.mark.parametrize("mass", [1.0, 2.0])
def test_gradient_kernel_degenerate_limit_host_descriptor(mass):
    """The cg and alpha0 the host factory passes to libhydft give the same limit
    (xi evaluated by the library's own Fermi-Dirac routine)."""
    from hydft_amd.model import FD_XI, build_species, fd_eval
    k = np.array([0.3, 1.0, 2.5])
    errs = []
    for theta in (0.04, 0.02, 0.01):
        p = P.params(kind="electron", rs=1.0, theta=theta, mass=mass, terms="tf,gradient",
                     structure="none")
        sp = build_species(p)
        d = dict(sp.terms)["gradient"]
        xi0 = float(fd_eval(FD_XI, d.alpha0)[0])
        kg = 2.0 * d.cg * xi0 * k * k
        un = O.Units(kind="electron", gamma=p.gamma, rs=1.0, mass=mass)
        assert abs(d.cg / O.Gradient(un, 1.0 / 9.0).cg - 1.0) <= 1e-14
        errs.append(np.abs(kg / _vw(un, k) - 1.0).max())
    assert errs[-1] < 2e-3 and errs[2] < 0.3 * errs[1] < 0.09 * errs[0], errs


def _ideal_1d(eta=0.0, closure="newtonian", tau=0.0, mass=1.0, cfl=0.4):
    return O.make_model(32, 1, 1, 20.0, 1.0, 1.0, kind="ion", gamma=3.0, kappa=0.0, mass=mass,
                        terms="ideal", closure=closure, eta=eta, tau=tau, cfl=cfl)


def _kmax_kept(n, length):
    """Largest kept |k| of the 2/3 rule: |m| (2 pi/L) < (2/3)(2 pi/L)(n/2) (grid.f90:90-102)."""
    m = max(j for j in range(n // 2 + 1) if j < n / 3.0)
    return m * TWOPI / length


@pytest.mark.parametrize("mass", [1.0, 2.5])
def test_compute_dt_wave_branch_closed_form(mass):
    om = _ideal_1d(mass=mass)
    beta = 9.0
    wmax = _kmax_kept(32, 20.0) * math.sqrt(1.0 / (beta * mass))
    assert abs(om.compute_dt() / (0.4 * 1.7 / wmax) - 1.0) <= 1e-14


def test_compute_dt_viscous_branch_closed_form():
    eta = 50.0
    om = _ideal_1d(eta=eta)
    eta_l = 4.0 * eta / 3.0
    kmax = _kmax_kept(32, 20.0)
    assert abs(om.spc.g.kmax_dealiased - kmax) <= 1e-14 * kmax
    want = 0.4 * 2.5 / (eta_l * kmax ** 2)
    wave = 0.4 * 1.7 / (kmax * math.sqrt(1.0 / 9.0))
    assert want < wave
    assert abs(om.compute_dt() / want - 1.0) <= 1e-14


def test_compute_dt_advective_branch_closed_form():
    om = _ideal_1d()
    u0 = 40.0
    om.mom = [np.full(om.spc.g.shape, u0 * om.n0), np.zeros(om.spc.g.shape), np.zeros(om.spc.g.shape)]
    want = 0.4 * (20.0 / 32) / u0
    assert abs(om.compute_dt() / want - 1.0) <= 1e-14


def test_compute_dt_maxwell_closed_form():
    eta, tau = 2.0, 0.5
    om = _ideal_1d(eta=eta, closure="maxwell", tau=tau)
    eta_l = 4.0 * eta / 3.0
    c2 = 1.0 / 9.0 + eta_l / tau
    want = 0.4 * 1.7 / (_kmax_kept(32, 20.0) * math.sqrt(c2))
    assert abs(om.compute_dt() / want - 1.0) <= 1e-14


def test_compute_dt_fixed_dt_wins():
    om = _ideal_1d(eta=50.0)
    om.dt_fixed = 0.0123
    assert om.compute_dt() == 0.0123
