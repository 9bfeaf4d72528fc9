"""Host-side setup (CPU): namelist input, HNC structure, transport, linear
response and probe analysis.  Checks follow test_hnc.f90, test_transport.f90,
test_dispersion.f90 and closed forms; nothing here touches a GPU."""
import math

import numpy as np
import pytest

from hydft_amd.host import hnc, transport
from hydft_amd.host.analysis import fit_damped_mode, fit_driven_response
from hydft_amd.host.linear import Linear
from hydft_amd.host.params import Params, parse_namelists, read_params
from oracle import hydft_oracle as O

N0 = 3.0 / (4.0 * math.pi)


def test_namelist_defaults_and_overrides(tmp_path):
    f = tmp_path / "run.in"
    f.write_text("""! a comment
&grid nx = 32, ny = 16, nz = 8, lx = 10.0d0, dealias = .false. /
&species kind = 'Electron', rs = 1.0, theta = 1.0, terms = 'tf,gradient,hartree,ry',
         transport = 'electron_fit' /
&time dt = 0.005, tend = 2.5 /
&initial type = 'random', amplitude = 5e-3, mode = 2, 0, 0, seed = 11 /
&external type = 'driven', amplitude = 1e-4, mode(2) = 3, omega = 0.7 /
&gpu precision = 'mixed' /
""")
    p = read_params(str(f))
    assert (p.nx, p.ny, p.nz, p.lx, p.dealias) == (32, 16, 8, 10.0, False)
    assert p.kind == "electron" and p.transport == "electron_fit"
    # theta > 0 overrides Gamma for electrons and kappa = 0 (params.f90:168-172)
    assert abs(p.gamma - 2.0 / (1.0 * (9.0 * math.pi / 4.0) ** (2.0 / 3.0))) < 1e-15
    assert abs(p.gamma - 0.54301) < 1e-5
    assert p.kappa == 0.0
    assert p.ic == "random" and p.ic_mode == (2, 0, 0) and p.ic_seed == 11
    assert p.vext_mode == (1, 3, 0) and p.vext_omega == 0.7
    assert p.precision == "mixed"
    assert p.cfl == 0.4 and p.hnc_nr == 8192  # untouched defaults (params.f90:42-51)
    assert Params().finalize().terms == "ideal,hartree,ry"


def test_namelist_unknown_variable_is_an_error(tmp_path):
    f = tmp_path / "bad.in"
    f.write_text("&grid nq = 3 /\n")
    with pytest.raises(Exception):
        read_params(str(f))
    assert parse_namelists("&probe mode = 1, 2, 3 /")["probe"]["mode"] == [1, 2, 3]


def test_fourier_bessel_pair():
    """test_hnc.f90:25-35."""
    fb = hnc.FourierBessel(4096, 60.0)
    fr = np.exp(-fb.r)
    fk = fb.r2k(fr)
    sel = fb.k < 20.0
    assert np.max(np.abs(fk[sel] * (fb.k[sel] ** 2 + 1.0) ** 2 / (8.0 * math.pi) - 1.0)) < 1e-4
    assert np.max(np.abs(fb.k2r(fk) - fr)) < 1e-12


@pytest.mark.parametrize("gam", [10.0, 50.0])
def test_ocp_hnc(gam):
    """test_hnc.f90:37-49: OCP energy within 3% of the SDD fit, g(0)=0, perfect screening."""
    res = hnc.hnc_solve(hnc.PairPotential("coulomb", gam), N0, 8192, 80.0, 1e-9, 0.2, 5000)
    ufit = -0.897744 * gam + 0.95043 * gam ** 0.25 + 0.18956 * gam ** (-0.25) - 0.81487
    assert res.converged
    assert abs(res.u_ex / ufit - 1.0) < 0.03
    assert abs(res.g[0]) < 1e-6 and abs(res.g[-1] - 1.0) < 1e-6
    assert abs(res.sk[0] / (res.k[0] ** 2 / (3.0 * gam)) - 1.0) < 0.05


def test_yukawa_and_qsp_hnc():
    """test_hnc.f90:51-70."""
    gam, kap = 5.0, 1.0
    st = hnc.structure_from_hnc(hnc.PairPotential("yukawa", gam, kap), N0, 8192, 80.0, 1e-9, 0.2, 5000,
                                with_derivative=True)
    q = 0.5
    smf = (q * q + kap * kap) / (q * q + kap * kap + 3.0 * gam)
    c = O.interp_linear(st.k, st.c, min(q, st.k[-1]))
    assert abs(1.0 / (1.0 - N0 * c) / smf - 1.0) < 0.25
    assert st.u_ex < 0.0 and st.du_dgamma < 0.0
    res = hnc.hnc_solve(hnc.PairPotential("hansen_mcdonald", 1.0, 0.0, lam=2.0 * math.sqrt(math.pi / 1.86)),
                        N0, 8192, 80.0, 1e-9, 0.2, 5000)
    assert res.converged and 0.0 < res.g[0] < 1.0
    assert abs(res.sk[0] / (res.k[0] ** 2 / 3.0) - 1.0) < 0.05


def test_transport_fits():
    """test_transport.f90:27-71 (subset that concerns the viscosity used here)."""
    for n, m in [(1, 1), (1, 2), (1, 3), (2, 2)]:
        assert abs(transport.sm_knm(n, m, 1.0 - 1e-12) / transport.sm_knm(n, m, 1.0 + 1e-12) - 1) < 5e-3
    assert abs(transport.sm_knm(2, 2, 1.0) - 0.32237) < 1e-4
    assert abs(transport.sm_knm(1, 1, 1e-6) / (-0.25 * math.log(1.4660e-6)) - 1) < 1e-4
    un = O.Units(kind="electron", gamma=1.0, rs=1.86)
    assert abs(un.lambda_tf - 0.6472) < 5e-4
    eta_l, cv, sm = transport.sm_electron_viscosity(un.rs, un.gamma, un.theta, un.lambda_tf)
    assert abs(eta_l - 1.312) < 3e-3 and abs(cv - 0.00467) < 2e-5
    un = O.Units(kind="electron", gamma=0.01, rs=1.0)
    assert abs(un.lambda_tf * math.sqrt(3.0 * 0.01) - 1.0) < 5e-3


def _ion_linear(eta=0.3, xi=0.05, tau=2.0):
    st = hnc.structure_from_hnc(hnc.PairPotential("yukawa", 2.0, 0.1), N0, 4096, 60.0)
    un = O.Units(kind="ion", gamma=2.0, kappa=0.1)
    fun = O.build_functional(un, "ideal,hartree,ry", st.k, st.c)
    return Linear(N0, fun.kernel, 1.0, 4.0 * eta / 3.0 + xi, tau, un.beta), st, un, fun


def test_dispersion_and_chi():
    """test_dispersion.f90:26-72."""
    lin, st, un, fun = _ion_linear()
    for i in range(1, 6):
        q = 0.4 * i
        c = O.interp_linear(st.k, st.c, min(q, st.k[-1]))
        sq = 1.0 / (1.0 - N0 * c)
        assert abs(lin.c2(q) * un.beta * sq - 1.0) < 1e-10
        w, roots = lin.dispersion(q)
        res = w * w - q * q / (3.0 * 2.0 * sq) + 1j * lin.eta_bar * q * q * w / (1.0 - 1j * w * 2.0)
        assert abs(res) < 1e-12
        assert w.real > 0.0 and w.imag < 0.0
        for r in roots:
            assert abs(lin.denominator(q, r) * (1.0 - 1j * r * 2.0)) < 1e-10
    lin.tau = 0.0
    q = 1.0
    nu = q * q * lin.eta_bar
    wns = -0.5j * nu + np.sqrt(complex(q * q * lin.c2(q) - 0.25 * nu * nu))
    assert abs(lin.dispersion(q)[0] - wns) < 1e-12
    assert abs(lin.chi(q, 0.0).real + 1.0 / float(fun.kernel(np.array([q]))[0])) < 1e-10
    assert abs(lin.dsf(q, 0.7) - lin.dsf(q, -0.7)) < 1e-14 and lin.dsf(q, 0.7) > 0.0


def test_analysis_recovers_known_signals():
    t = np.linspace(0.0, 60.0, 20001)
    om, ga = 0.83, -0.041
    y = 1e-4 * np.exp(ga * t) * np.cos(om * t + 0.3)
    o, g, nz = fit_damped_mode(t, y, 5.0)
    assert abs(o / om - 1) < 1e-3 and abs(g / ga - 1) < 1e-2 and nz > 10
    w = 0.7
    a0 = 2e-6 - 1e-6j
    z = a0 * np.exp(-1j * w * t)
    a, sc = fit_driven_response(t, z, w, 30.0)
    assert abs(a - a0) < 1e-15 and sc < 1e-9


def test_structure_cache(tmp_path):
    n0 = 3.0 / (4.0 * np.pi)
    cache = hnc.StructureCache(directory=str(tmp_path))
    pot = hnc.PairPotential("yukawa", 5.0, 1.0)
    a = cache.get(pot, n0, nr=1024, rmax=30.0, tol=1e-8)
    b = cache.get(hnc.PairPotential("yukawa", 5.0, 1.0), n0, nr=1024, rmax=30.0, tol=1e-8)
    assert b is a and cache.hits == 1 and cache.misses == 1
    # a new process-level cache finds the disk entry and returns the same table
    c = hnc.StructureCache(directory=str(tmp_path)).get(pot, n0, nr=1024, rmax=30.0, tol=1e-8)
    np.testing.assert_array_equal(c.c, a.c)
    np.testing.assert_array_equal(c.k, a.k)
    # solve_many: distinct keys solved once each, order kept, equal to direct solves
    reqs = [(hnc.PairPotential("yukawa", g, 1.0), n0, 1024, 30.0, 1e-8, 0.2, 20000, False) for g in (2.0, 5.0, 2.0)]
    out = cache.solve_many(reqs, processes=2)
    direct = hnc.structure_from_hnc(hnc.PairPotential("yukawa", 2.0, 1.0), n0, 1024, 30.0, 1e-8)
    np.testing.assert_allclose(out[0].c, direct.c, rtol=0, atol=1e-12)
    assert out[2] is out[0] and out[1] is cache.get(pot, n0, nr=1024, rmax=30.0, tol=1e-8)
