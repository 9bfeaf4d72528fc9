"""Host transport coefficients (one-time scalars): shear/bulk viscosity in units
of m n0 wp a^2 and the Maxwell time in 1/wp (transport.f90, stanton_murillo.f90).

Models: constant | yukawa_fit (Bastea MD fit at the effective coupling) |
stanton_murillo (effective-Boltzmann K_22 fit) | electron_fit (Conti-Vignale
interpolated with Stanton-Murillo), and tau_model constant | ichimaru.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

# Stanton & Murillo PRE 93 043203 Table IV fit coefficients of K_nm(g): weak
# coupling (g < 1) a_1..a_5 and strong coupling (g > 1) b_0..b_4, for
# (n,m) = (1,1), (1,2), (1,3), (2,2).  These are published constants.
_A = {
    (1, 1): (1.4660, -1.7836, 1.4313, -0.55833, 0.061162),
    (1, 2): (0.52094, 0.25153, -1.1337, 1.2155, -0.43784),
    (1, 3): (0.30346, 0.23739, -0.62167, 0.56110, -0.18046),
    (2, 2): (0.85401, -0.22898, -0.60059, 0.80591, -0.30555),
}
_B = {
    (1, 1): (0.081033, -0.091336, 0.051760, -0.50026, 0.17044),
    (1, 2): (0.20572, -0.16536, 0.061572, -0.12770, 0.066993),
    (1, 3): (0.68375, -0.38459, 0.10711, 0.10649, 0.028760),
    (2, 2): (0.43475, -0.21147, 0.11116, 0.19665, 0.15195),
}
SQRT3PI = math.sqrt(3.0 * math.pi)


class TransportError(RuntimeError):
    pass


def sm_knm(n, m, g):
    """Reduced collision integral K_nm(g) (stanton_murillo.f90:52-72)."""
    fact = 1.0
    for i in range(2, m):
        fact *= i
    if g < 1.0:
        a = _A[(n, m)]
        s = 0.0
        for i in range(4, -1, -1):
            s = (s + a[i]) * max(g, 2.2250738585072014e-308)
        return -0.25 * n * fact * math.log(s)
    b = _B[(n, m)]
    lg = math.log(g)
    return (b[0] + b[1] * lg + b[2] * lg * lg) / (1.0 + b[3] * g + b[4] * g * g)


def sm_g(gamma, kappa):
    return gamma * math.sqrt(kappa * kappa + 3.0 * gamma / (1.0 + 3.0 * gamma))


def sm_viscosity(gamma, kappa):
    """eta* = 5 sqrt(3 pi)/(36 Gamma^{5/2} K_22(g)) (stanton_murillo.f90:120-124)."""
    return 5.0 * SQRT3PI / (36.0 * gamma ** 2.5 * sm_knm(2, 2, sm_g(gamma, kappa)))


def sm_electron_viscosity(rs, gamma, theta, lambda_tf):
    """eta_l* of the Sci. Rep. supplement (stanton_murillo.f90:134-146)."""
    cv = 1.0 / math.sqrt(3.0 * rs) / (60.0 * rs ** (-1.5) + 80.0 / rs - 40.0 * rs ** (-2.0 / 3.0)
                                      + 62.0 * rs ** (-1.0 / 3.0))
    sm = 5.0 * SQRT3PI / (36.0 * gamma ** 2.5 * sm_knm(2, 2, gamma / lambda_tf))
    return 4.0 / 3.0 * (cv + theta * sm) / (1.0 + theta), cv, sm


@dataclass
class Transport:
    eta: float = 0.0
    xi: float = 0.0
    tau: float = 0.0
    eta_l: float = 0.0
    elastic_modulus: float = 0.0


def transport_setup(model, tau_model, eta_in, xi_in, tau_in, un, gamma_ad, st, maxwell) -> Transport:
    """transport_setup (transport.f90:61-111).  `un` needs gamma, kappa, rs,
    theta, lambda_tf, quantum; `st` needs u_ex, du_dgamma, has_energy."""
    model = model.lower()
    tau_model = tau_model.lower()
    gamma, kappa = un.gamma, un.kappa
    tr = Transport()
    if model == "constant":
        tr.eta, tr.xi = eta_in, xi_in
    elif model == "yukawa_fit":
        geff = gamma * (1.0 + kappa + 0.5 * kappa * kappa) * math.exp(-kappa)
        tr.eta = 0.482 / geff ** 2 + 0.629 / geff ** 0.878 + 1.88e-3 * geff
        tr.xi = xi_in
    elif model == "stanton_murillo":
        if un.quantum:
            raise TransportError("transport = stanton_murillo is the ion (Yukawa) fit; use electron_fit")
        tr.eta = sm_viscosity(gamma, kappa)
        tr.xi = xi_in
    elif model == "electron_fit":
        if not un.quantum:
            raise TransportError("transport = electron_fit needs kind = electron")
        eta_l, _, _ = sm_electron_viscosity(un.rs, gamma, un.theta, un.lambda_tf)
        tr.eta = 0.75 * eta_l
        tr.xi = 0.0
    else:
        raise TransportError("unknown transport model: " + model)
    tr.eta_l = 4.0 * tr.eta / 3.0 + tr.xi
    if maxwell:
        if tau_model == "constant":
            tr.tau = tau_in
        elif tau_model == "ichimaru":
            if st is None or not st.has_energy:
                raise TransportError("tau_model = ichimaru needs the HNC structure source (for E_c)")
            u = st.u_ex
            du = st.du_dgamma
            muc = u / 3.0 + gamma * du / 9.0
            tr.elastic_modulus = 1.0 - gamma_ad * muc + 4.0 * u / 15.0
            if tr.elastic_modulus <= 0.0:
                raise TransportError("ichimaru tau: 1 - gamma mu + 4u/15 <= 0; use tau_model = constant")
            tr.tau = 3.0 * gamma * tr.eta_l / tr.elastic_modulus
        else:
            raise TransportError("unknown tau model: " + tau_model)
    return tr
