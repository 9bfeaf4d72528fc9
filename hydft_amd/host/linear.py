"""Closed-form linear response (host): chi(k,w), dispersion roots and the
dynamic structure factor, built from the same functional kernel K(k),
viscosity and Maxwell time as the time-stepping model (linear_response.f90).

  chi(k,w) = -k^2 n0 / D,  D = -w^2 + k^2 n0 K(k)/m - i k^2 (eta_bar/m) w/(1 - i w tau)

Used as the per-request validator for chi(k,w) / dispersion runs and as the
closed form the driven-mode and single-mode checks compare against.
"""
from __future__ import annotations

import math

import numpy as np


class Linear:
    def __init__(self, n0, kernel, mass=1.0, eta_bar=0.0, tau=0.0, beta=3.0, hbar=0.0,
                 quantum=False):
        """kernel: callable k -> K(k) (functional_kernel, functional.f90:64-72)."""
        self.n0, self.kernel, self.mass = n0, kernel, mass
        self.eta_bar, self.tau, self.beta, self.hbar, self.quantum = eta_bar, tau, beta, hbar, quantum

    def c2(self, k):
        """linear_response.f90:59-63."""
        return self.n0 * float(np.asarray(self.kernel(np.array([k])))[0]) / self.mass

    def denominator(self, k, w):
        """linear_response.f90:65-70."""
        w = complex(w)
        return (-w * w + k * k * self.c2(k)
                - 1j * k * k * (self.eta_bar / self.mass) * w / (1.0 - 1j * w * self.tau))

    def chi(self, k, w):
        """linear_response.f90:72-76."""
        return -k * k * self.n0 / self.denominator(k, complex(w, 0.0))

    def dispersion(self, k):
        """Root of D = 0 with the largest real part (linear_response.f90:82-147);
        the cubic/quadratic is solved with numpy's companion-matrix roots and
        polished by Newton."""
        c2k2 = k * k * self.c2(k)
        nu = k * k * self.eta_bar / self.mass
        if self.tau > 0.0:
            a = [-1j * self.tau, 1.0, 1j * (nu + c2k2 * self.tau), -c2k2]   # highest degree first
        else:
            a = [1.0, 1j * nu, -c2k2]
        roots = np.roots(a).astype(complex)
        p = np.poly1d(a)
        dp = p.deriv()
        for i in range(len(roots)):
            z = roots[i]
            for _ in range(5):
                d = dp(z)
                if d == 0:
                    break
                z = z - p(z) / d
            roots[i] = z
        w = roots[np.argmax(roots.real)]
        return complex(w), roots

    def dsf(self, k, w):
        """S(k,w)/n0 (linear_response.f90:152-171)."""
        if abs(w) < 1.0e-6:
            ratio = self.n0 * self.eta_bar / (self.mass * self.c2(k) ** 2)
        else:
            ratio = -self.chi(k, w).imag / w
        if self.quantum:
            x = self.beta * self.hbar * w
            if abs(x) < 1.0e-6:
                return ratio / (math.pi * self.n0 * self.beta * self.hbar) * (1.0 + 0.5 * x)
            return ratio / (math.pi * self.n0 * self.beta * self.hbar) * x / (1.0 - math.exp(-x))
        return ratio / (math.pi * self.n0 * self.beta)
