"""GPU model parity against the CPU oracle on identical inputs (FP64).

For each case: model_rhs at t = 0 (relative max error <= 1e-10 of the field
scale) and the state after a few RK3 steps (relative L2 of n, n - n0, every
momentum component and every stress component <= 1e-11).  Cases cover ions / electrons, Newtonian with and
without viscosity, Maxwell, 1-D / 2-D / 3-D, and the three external
potentials."""
import zlib

import numpy as np
import pytest

from tests import _pairs as P
from oracle import hydft_oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    # name: (params, steps)
    "ion3d_newt_eta0": (dict(nx=32, ny=32, nz=32, lx=2 * np.pi / 0.2, ly=2 * np.pi / 0.2, lz=2 * np.pi / 0.2,
                             gamma=10.0, kappa=1.0, eta=0.0, dt=0.01), 3),
    "ion3d_newt_visc": (dict(nx=16, ny=12, nz=8, lx=20.0, ly=9.0, lz=6.0, gamma=2.0, kappa=0.1,
                             eta=0.3, xi=0.1, dt=0.01), 3),
    "ion2d_maxwell": (dict(nx=32, ny=4, nz=1, lx=20.0, ly=5.0, gamma=2.0, kappa=0.1, eta=0.2,
                           closure="maxwell", tau_model="constant", tau=2.0, dt=0.02), 4),
    "ion3d_maxwell": (dict(nx=16, ny=16, nz=12, lx=12.0, ly=12.0, lz=9.0, gamma=50.0, kappa=2.0, eta=0.2,
                           closure="maxwell", tau_model="constant", tau=0.5, dt=0.005), 3),
    "elec3d": (dict(kind="electron", gamma=1.0, rs=1.86, nx=16, ny=16, nz=16, lx=20.0, ly=20.0, lz=20.0,
                    eta=0.1, dt=0.01), 3),
    "elec1d": (dict(kind="electron", gamma=1.0, rs=1.86, nx=64, ny=1, nz=1, lx=20.0, eta=0.1, dt=0.01), 5),
    # electrons with the Maxwell closure: the relaxation path of the x-line
    # kernel (no functional there) with the gradient term in the model
    "elec3d_maxwell": (dict(kind="electron", gamma=1.0, rs=1.86, nx=16, ny=16, nz=12, lx=20.0, ly=20.0, lz=15.0,
                            eta=0.1, closure="maxwell", tau_model="constant", tau=0.5, dt=0.01), 3),
    "ion1d_driven": (dict(nx=32, lx=20.0, gamma=2.0, kappa=0.1, eta=0.3, vext="driven",
                          vext_amplitude=1e-3, vext_mode=(2, 0, 0), vext_omega=0.4, vext_ramp=0.05,
                          dt=0.02), 5),
    "ion2d_gaussian": (dict(nx=24, ny=16, nz=1, lx=20.0, ly=12.0, gamma=2.0, kappa=0.1, eta=1.0,
                            vext="gaussian", vext_amplitude=1e-3, vext_width=2.0, dt=0.02), 3),
    "ion3d_mode_nodealias": (dict(nx=16, ny=8, nz=8, lx=20.0, ly=10.0, lz=10.0, gamma=2.0, kappa=0.1,
                                  eta=0.5, vext="mode", vext_amplitude=1e-3, vext_mode=(1, 1, 0),
                                  dealias=False, dt=0.02), 3),
    # x lines of 96 points (1.5 points per lane) and a 48-point y axis
    "ion3d_nx96_visc": (dict(nx=96, ny=12, nz=8, lx=30.0, ly=9.0, lz=6.0, gamma=5.0, kappa=1.0,
                             eta=0.3, xi=0.1, dt=0.01), 2),
    "ion3d_maxwell_y48": (dict(nx=16, ny=48, nz=8, lx=12.0, ly=30.0, lz=6.0, gamma=20.0, kappa=1.0, eta=0.2,
                               closure="maxwell", tau_model="constant", tau=0.5, dt=0.005), 2),
    # 128-point x lines (two lines per wave) with the Maxwell stress; the
    # 1-D cases above leave all but one line of a wave empty
    "ion3d_maxwell_nx128": (dict(nx=128, ny=12, nz=4, lx=40.0, ly=9.0, lz=6.0, gamma=50.0, kappa=2.0, eta=0.2,
                                 closure="maxwell", tau_model="constant", tau=0.5, dt=0.005), 2),
    # 256-point x lines (one line per wave, the bench's length) on thin 3-D grids
    "ion3d_nx256_visc": (dict(nx=256, ny=8, nz=4, lx=60.0, ly=6.0, lz=4.0, gamma=5.0, kappa=1.0,
                              eta=0.3, xi=0.1, dt=0.01), 2),
    "elec3d_nx256": (dict(kind="electron", gamma=1.0, rs=1.86, nx=256, ny=8, nz=4, lx=60.0, ly=8.0, lz=6.0,
                          eta=0.1, dt=0.01), 2),
    # driven ions on 3-D grids with 32-point x lines: the fixed x-line forms
    # (ion phase A; phase B with the tabulated wave force, Newtonian and Maxwell)
    "ion3d_driven_nx32": (dict(nx=32, ny=8, nz=8, lx=20.0, ly=10.0, lz=10.0, gamma=2.0, kappa=0.1, eta=0.3,
                               vext="driven", vext_amplitude=1e-3, vext_mode=(1, 1, 1), vext_omega=0.4,
                               vext_ramp=0.05, dt=0.02), 3),
    "ion3d_maxwell_driven_nx64": (dict(nx=64, ny=8, nz=4, lx=20.0, ly=10.0, lz=5.0, gamma=20.0, kappa=1.0,
                                       eta=0.2, closure="maxwell", tau_model="constant", tau=0.5,
                                       vext="driven", vext_amplitude=1e-3, vext_mode=(2, 1, 1), vext_omega=0.7,
                                       vext_ramp=0.05, dt=0.005), 3),
}

# FP32 transforms with the FP64 RK3 state: density error against the FP64
# oracle after the case's steps (relative L2 of n, and of n - n0)
MIXED_TOL_N, MIXED_TOL_DN = 1e-6, 1e-4


def build(name, precision="fp64"):
    from hydft_amd.model import Model
    kw, steps = CASES[name]
    p = P.params(**kw)
    st = P.structure(p.kind, p.gamma, p.kappa, p.rs)
    gm = Model(p, structure=st, precision=precision)
    om = P.oracle_model(p, st)
    evolve = om.stress.evolve
    n, mom, pi = P.smooth_fields(p, seed=zlib.crc32(name.encode()) % 1000, amp_n=0.02, amp_u=0.01, with_pi=evolve)
    gm.set_state(n, mom, pi)
    P.set_oracle_state(om, n, mom, pi)
    return p, gm, om, steps


@pytest.mark.parametrize("name", sorted(CASES))
def test_rhs_matches_oracle(gpu, name):
    p, gm, om, _ = build(name)
    dn, dmom, dpi = gm.rhs(0.3)
    on, omom, opi = om.rhs(0.3)
    scale_n = np.abs(on).max()
    assert np.abs(dn - on).max() <= 1e-10 * scale_n + 1e-18, name
    for d in range(3):
        sc = max(np.abs(omom[d]).max(), 1e-30)
        assert np.abs(dmom[d] - omom[d]).max() <= 1e-10 * max(sc, np.abs(omom[0]).max()), (name, d)
    if om.stress.evolve:
        for c in range(6):
            sc = max(np.abs(opi[c]).max(), 1e-30)
            assert np.abs(dpi[c] - opi[c]).max() <= 1e-9 * sc, (name, c)


@pytest.mark.parametrize("name", sorted(CASES))
def test_steps_match_oracle(gpu, name):
    p, gm, om, steps = build(name)
    om.dt = p.dt
    for _ in range(steps):
        om.advance()
    taken, _ = gm.advance(steps)
    assert taken == steps
    n, mom, pi = gm.get_state()
    n0 = P.N0
    e = {"n": P.rel_l2(n, om.n), "dn": P.rel_l2(n - n0, om.n - n0)}
    for d in range(p.ny > 1 and (3 if p.nz > 1 else 2) or 1):
        e[f"mom{d}"] = P.rel_l2(mom[d], om.mom[d])
    if om.stress.evolve:
        for c in range(6):
            if np.linalg.norm(om.pi[c]) > 0:
                e[f"pi{c}"] = P.rel_l2(pi[c], om.pi[c])
    print(name, " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    for k, v in e.items():
        assert v <= 1e-11, (name, k, v)
    t, dt, step = gm.time
    assert step == steps and abs(t - om.t) <= 1e-14 * max(1.0, om.t)


@pytest.mark.parametrize("name", ["ion3d_newt_visc", "elec3d", "elec1d", "ion2d_gaussian", "ion1d_driven"])
def test_mu_and_budgets_match_oracle(gpu, name):
    p, gm, om, steps = build(name)
    # functional_mu on a non-band-limited positive density
    rng = np.random.default_rng(4)
    n = P.N0 * (1.0 + 0.05 * (rng.random(om.spc.g.shape) - 0.5))
    mu_g = gm.mu(n)
    mu_o = om.fun.mu(om.spc, n, om.spc.fwd(n))
    assert np.abs(mu_g - mu_o).max() <= 1e-11 * np.abs(mu_o).max()
    # diag_budgets after a few steps (the oracle's v_ext is re-evaluated at its last stage time)
    om.dt = p.dt
    for _ in range(steps):
        om.advance()
    gm.advance(steps)
    mass, mom, ekin, fe = gm.budgets()
    omass, omom, oekin, ofe = om.budgets()
    assert abs(mass - omass) <= 1e-13 * abs(omass)
    for d in range(3):
        assert abs(mom[d] - omom[d]) <= 1e-12 * abs(omass) * 1e-2 + 1e-10 * abs(omom[d])
    assert abs(ekin - oekin) <= 1e-9 * abs(oekin)
    assert abs(fe - ofe) <= 1e-11 * abs(ofe)


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_steps_match_oracle(gpu, name):
    p, gm, om, steps = build(name, precision="mixed")
    om.dt = p.dt
    for _ in range(steps):
        om.advance()
    taken, _ = gm.advance(steps)
    assert taken == steps
    n, mom, pi = gm.get_state()
    n0 = P.N0
    assert P.rel_l2(n, om.n) <= MIXED_TOL_N, name
    assert P.rel_l2(n - n0, om.n - n0) <= MIXED_TOL_DN, name
    for d in range(p.ny > 1 and (3 if p.nz > 1 else 2) or 1):
        assert P.rel_l2(mom[d], om.mom[d]) <= MIXED_TOL_DN, (name, d)


def test_mixed_electrons_density_table_and_fallback(gpu):
    """Mixed-precision electrons evaluate the Fermi-Dirac functional from a
    per-model table over |n - n0| <= 0.2 n0 and fall back to the per-point
    evaluation outside it.  A density with a 10 % rms excursion puts points
    on both sides of that bound in every x-line kernel launch; the run must
    still meet the mixed bars against the FP64 oracle."""
    from hydft_amd.model import Model
    kw, _ = CASES["elec3d"]
    p = P.params(**kw)
    st = P.structure(p.kind, p.gamma, p.kappa, p.rs)
    gm = Model(p, structure=st, precision="mixed")
    om = P.oracle_model(p, st)
    n, mom, _ = P.smooth_fields(p, seed=11, amp_n=0.1, amp_u=0.01)
    x = np.abs(n / P.N0 - 1.0)
    assert (x > 0.2).mean() > 0.01 and (x < 0.19).mean() > 0.5  # both paths taken
    gm.set_state(n, mom)
    P.set_oracle_state(om, n, mom)
    om.dt = p.dt
    for _ in range(2):
        om.advance()
    gm.advance(2)
    ng, momg, _ = gm.get_state()
    assert P.rel_l2(ng, om.n) <= MIXED_TOL_N
    assert P.rel_l2(ng - P.N0, om.n - P.N0) <= MIXED_TOL_DN
    for d in range(3):
        assert P.rel_l2(momg[d], om.mom[d]) <= MIXED_TOL_DN, d


def test_profile_classes(gpu):
    """hy_model_profile records an event pair around every launch of a step;
    hy_model_profile_classes(mask) only around the classes in mask, with the
    same launch count and bytes for those classes."""
    import ctypes as C
    from hydft_amd import abi
    p, gm, _, _ = build("ion3d_newt_visc")
    lib = abi.lib().lib
    names = [lib.hy_profile_class_name(c).decode() for c in range(16)]
    xa = names.index("x_phaseA")

    def table():
        la, ms, by = (C.c_int * 16)(), (C.c_double * 16)(), (C.c_double * 16)()
        ncls = lib.hy_model_profile_read(gm.h, 16, la, ms, by)
        return {names[c]: (la[c], ms[c], by[c]) for c in range(ncls) if la[c]}

    abi.call("hy_model_profile", gm.h, 1)
    gm.advance(2)
    full = table()
    assert full["x_phaseA"][0] == 2 * 3 and len(full) > 4 and all(v[1] > 0 for v in full.values())
    abi.call("hy_model_profile_classes", gm.h, 1 << xa)
    gm.advance(2)
    only = table()
    assert list(only) == ["x_phaseA"] and only["x_phaseA"][0] == full["x_phaseA"][0]
    assert only["x_phaseA"][2] == full["x_phaseA"][2] and only["x_phaseA"][1] > 0
    abi.call("hy_model_profile", gm.h, 0)
    gm.advance(1)
    assert table() == {}
