"""Pins the Fermi-Dirac functions used by the electron functional (CPU).

The checks of test_fermi_dirac.f90:16-52 (limits, known value I_1/2(0),
derivative relations, inverse), an independent scipy.integrate.quad of
int_0^inf x^p/(1+e^(x-alpha)) dx, and agreement between the oracle's numpy
restatement and libhydft's host implementation (the one the GPU kernels use)."""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import hydft_oracle as O

fd = O.fermi_dirac()
SQRTPI = math.sqrt(math.pi)


def v(f, a):
    return float(f(np.array([a]))[0])


def test_limits_and_known_value():
    a = -20.0
    assert abs(v(fd.ip12, a) / (0.5 * SQRTPI * math.exp(a)) - 1) < 1e-8
    assert abs(v(fd.im12, a) / (SQRTPI * math.exp(a)) - 1) < 1e-8
    assert abs(v(fd.ip32, a) / (0.75 * SQRTPI * math.exp(a)) - 1) < 1e-8
    a = 250.0
    ex = (2.0 / 3.0) * a ** 1.5 * (1.0 + math.pi ** 2 / (8.0 * a * a))
    assert abs(v(fd.ip12, a) / ex - 1) < 1e-8
    ex = 2.0 * math.sqrt(a) * (1.0 - math.pi ** 2 / (24.0 * a * a))
    assert abs(v(fd.im12, a) / ex - 1) < 1e-8
    assert abs(v(fd.ip12, 300 - 1e-9) / v(fd.ip12, 300 + 1e-9) - 1) < 1e-8
    assert abs(v(fd.ip12, -30 - 1e-9) / v(fd.ip12, -30 + 1e-9) - 1) < 1e-8
    assert abs(v(fd.ip12, 0.0) - 0.678093895) < 1e-8


def test_derivative_relations():
    hh = 1e-4
    for ia in range(10):
        a = -5.0 + 4.7 * ia
        d = (v(fd.ip12, a + hh) - v(fd.ip12, a - hh)) / (2 * hh)
        assert abs(d / (0.5 * v(fd.im12, a)) - 1) < 1e-7
        d = (v(fd.ip32, a + hh) - v(fd.ip32, a - hh)) / (2 * hh)
        assert abs(d / (1.5 * v(fd.ip12, a)) - 1) < 1e-7
        d = (v(fd.im12, a + hh) - v(fd.im12, a - hh)) / (2 * hh)
        assert abs(d / v(fd.dim12, a) - 1) < 1e-6
        d = (v(fd.dim12, a + hh) - v(fd.dim12, a - hh)) / (2 * hh)
        assert abs(d / v(fd.d2im12, a) - 1) < 1e-5
        d = (v(fd.xi, a + hh) - v(fd.xi, a - hh)) / (2 * hh)
        assert abs(d / v(fd.dxi, a) - 1) < 1e-5


def test_inverse():
    for ia in range(23):
        a = -25.0 + 13.3 * ia
        y = v(fd.ip12, a)
        assert abs(v(fd.alpha_from_ip12, y) - a) < 1e-9 * max(1.0, abs(a))


@pytest.mark.parametrize("a", [-3.0, 0.0, 1.7, 8.0, 40.0])
def test_against_quadrature(a):
    def I(p):
        f = lambda x: x ** p / (1.0 + math.exp(x - a)) if x - a < 700 else 0.0
        lim = max(a, 0) + 60
        r1 = integrate.quad(f, 0, 1, limit=200, epsabs=0, epsrel=1e-13)[0]
        r2 = integrate.quad(f, 1, lim, limit=400, epsabs=0, epsrel=1e-13)[0]
        return r1 + r2
    assert abs(v(fd.ip12, a) / I(0.5) - 1) < 1e-9
    assert abs(v(fd.ip32, a) / I(1.5) - 1) < 1e-9
    # I_{-1/2} by parts: I_{-1/2} = 2 dI_{1/2}/dalpha; quad of x^-1/2 with the endpoint singularity
    im = integrate.quad(lambda t: 2.0 / (1.0 + math.exp(t * t - a)), 0, math.sqrt(max(a, 0) + 60),
                        limit=400, epsrel=1e-13)[0]
    assert abs(v(fd.im12, a) / im - 1) < 1e-9


def test_library_host_fd_matches_oracle():
    from hydft_amd.model import fd_eval
    xs = np.array([-35.0, -29.99, -12.3, -0.7, 0.0, 0.33, 4.4, 77.7, 299.9, 301.5])
    for which, f in enumerate([fd.im12, fd.ip12, fd.ip32, fd.dim12, fd.d2im12, fd.xi, fd.dxi]):
        np.testing.assert_allclose(fd_eval(which, xs), f(xs), rtol=1e-13, atol=0)
    ys = fd.ip12(xs)
    np.testing.assert_allclose(fd_eval(7, ys), fd.alpha_from_ip12(ys), rtol=1e-12, atol=1e-12)


def test_golden_known_answers():
    """The known answers of test_fermi_dirac.f90 as committed data
    (tests/golden/reference_known_answers.json), against the oracle and
    libhydft's host implementation (hy_fd_eval, the code the kernels use)."""
    import json
    import os
    from hydft_amd.model import fd_eval, FD_IM12, FD_IP12, FD_ALPHA
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))["fermi_dirac"]
    kv = g["known_value"]
    assert abs(v(fd.ip12, kv["alpha"]) - kv["value"]) < kv["abs_tol"]
    assert abs(float(fd_eval(FD_IP12, kv["alpha"])[0]) - kv["value"]) < kv["abs_tol"]
    cl = g["classical_limit"]
    for name, f, code in (("im12", fd.im12, FD_IM12), ("ip12", fd.ip12, FD_IP12)):
        ex = cl["gamma_p_plus_1"][name] * math.exp(cl["alpha"])
        assert abs(v(f, cl["alpha"]) / ex - 1) < cl["rel_tol"]
        assert abs(float(fd_eval(code, cl["alpha"])[0]) / ex - 1) < cl["rel_tol"]
    dl = g["degenerate_limit"]
    assert abs(v(fd.ip12, dl["alpha"]) / dl["ip12"] - 1) < dl["rel_tol"]
    assert abs(v(fd.im12, dl["alpha"]) / dl["im12"] - 1) < dl["rel_tol"]
    assert abs(float(fd_eval(FD_IP12, dl["alpha"])[0]) / dl["ip12"] - 1) < dl["rel_tol"]
    for a in g["inverse"]["alphas"]:
        y = v(fd.ip12, a)
        tol = g["inverse"]["rel_tol"] * max(1.0, abs(a))
        assert abs(v(fd.alpha_from_ip12, y) - a) < tol
        assert abs(float(fd_eval(FD_ALPHA, y)[0]) - a) < tol


def test_field_kernel_forms_match_reference_forms():
    """The Fermi-Dirac code the field kernels run (table gathers; inline
    Sommerfeld and classical expansions outside [-30, 300] with powers from
    one square root; the Halley-started inverse) against the reference forms
    (fermi_dirac.f90:150-267, pow-based tails), compiled from the same source
    for the host (hy_fd_eval 8-12).  Covers both tails and the table."""
    from hydft_amd.model import fd_eval
    xs = np.concatenate([np.linspace(-80.0, -30.05, 57), np.linspace(-30.0, 300.0, 661),
                         np.linspace(300.05, 5000.0, 53), [-30.0, 300.0, -29.999, 299.999]])
    for kernel, ref in ((8, 0), (9, 3), (10, 4), (11, 1)):
        got, want = fd_eval(kernel, xs), fd_eval(ref, xs)
        np.testing.assert_allclose(got, want, rtol=2e-14, atol=0, err_msg=f"form {kernel} vs {ref}")
    # inverse: y over the table and both tails, each back to the reference inverse
    ys = fd_eval(1, np.concatenate([np.linspace(-45.0, -30.5, 8), np.linspace(-29.0, 299.0, 200),
                                    np.linspace(301.0, 900.0, 8)]))
    a_ref = fd_eval(7, ys)
    a_ker = fd_eval(12, ys)
    np.testing.assert_allclose(a_ker, a_ref, rtol=1e-13, atol=1e-13)


def test_density_table_scheme():
    """The mixed electron functional's density table: nodes x_i = -0.2 + i h
    (h = 0.4/2048) of x = (n - n0)/n0 hold alpha - alpha0, xi and d xi/d n
    from the reference evaluation (fermi_dirac.f90:225-267); 4-point Lagrange
    interpolation at any |x| < 0.2 - 2h reproduces that evaluation to 1e-12
    relative, across degenerate and classical alpha0."""
    from hydft_amd.model import fd_eval
    N, R = 2048, 0.2
    h = 2 * R / N
    rng = np.random.default_rng(7)

    def factors(alpha, cn):
        im, d1, d2 = fd_eval(0, alpha), fd_eval(3, alpha), fd_eval(4, alpha)
        return d1 / im ** 2, (d2 * im - 2 * d1 * d1) / im ** 3 / im * (2.0 / cn)

    for a0 in (-6.0, -0.5, 0.4, 3.0, 25.0):
        cn = 0.37
        n0 = cn * float(fd_eval(1, np.array([a0]))[0])
        xn = -R + h * np.arange(N + 1)
        an = fd_eval(7, n0 * (1 + xn) / cn)
        tab = np.stack([an - a0, *factors(an, cn)])
        xs = rng.uniform(-(R - 2 * h), R - 2 * h, 4000)
        ar = fd_eval(7, n0 * (1 + xs) / cn)
        ref = np.stack([ar - a0, *factors(ar, cn)])
        u = (xs + R) / h
        i = u.astype(int) - 1
        t = u - (i + 1)
        w = [-t * (t - 1) * (t - 2) / 6, (t + 1) * (t - 1) * (t - 2) / 2, -(t + 1) * t * (t - 2) / 2,
             (t + 1) * t * (t - 1) / 6]
        got = sum(w[k] * tab[:, i + k] for k in range(4))
        scale = np.abs(ref).max(axis=1, keepdims=True)
        assert np.abs(got - ref).max(axis=1).max() <= 1e-12 * scale.max(), a0
        for c in (1, 2):
            assert np.max(np.abs(got[c] - ref[c]) / np.abs(ref[c])) <= 1e-12, (a0, c)
