"""GPU parity of the drop-in fft_t / spec_t / stress_t operators.

Known answers of test_fft.f90:20-52 and test_ops.f90:30-71 (same grids, same
tolerances), then seeded random fields against the CPU oracle on several
grid shapes (FP64; tolerance 1e-12 relative to the field scale)."""
import numpy as np
import pytest

from oracle import hydft_oracle as O

pytestmark = pytest.mark.gpu

TWOPI = O.TWOPI


@pytest.fixture(scope="module")
def Spec(gpu):
    from hydft_amd.spectral import Spec
    return Spec


def test_fft_known_answers_1d(Spec):
    """test_fft.f90:20-36."""
    s = Spec(64, 1, 1, TWOPI, TWOPI, TWOPI)
    x = np.arange(64) * (TWOPI / 64)
    u = (1.5 + 0.3 * np.cos(3 * x) + 0.2 * np.sin(5 * x)).reshape(1, 1, 64)
    uh = s.fwd(u)
    v = s.bwd(uh)
    assert np.abs(u - v).max() < 1e-13
    assert abs(uh[0, 0, 0] - 1.5) < 1e-13
    assert abs(uh[0, 0, 3] - 0.15) < 1e-13
    assert abs(uh[0, 0, 5] + 0.1j) < 1e-13
    assert abs(s.mode(u, 3, 0, 0) - 0.15) < 1e-13


def test_fft_golden_known_answers(Spec):
    """The test_fft.f90 known answers as committed data
    (tests/golden/reference_known_answers.json), through the GPU transforms."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))["fft"]
    c = g["one_d"]
    s = Spec(c["n"], 1, 1, c["length"], TWOPI, TWOPI)
    x = np.arange(c["n"]) * (c["length"] / c["n"])
    u = (1.5 + 0.3 * np.cos(3 * x) + 0.2 * np.sin(5 * x)).reshape(1, 1, c["n"])
    uh = s.fwd(u)
    assert np.abs(u - s.bwd(uh)).max() < g["round_trip_tol"]
    for co in c["coefficients"]:
        assert abs(uh[0, 0, co["kx"]] - complex(co["re"], co["im"])) < c["tol"]
    c = g["three_d"]
    s = Spec(*c["shape_xyz"], *c["lengths_xyz"])
    X, Y, Z = O.Grid(*c["shape_xyz"], *c["lengths_xyz"]).coords()
    u = 0.7 + np.cos(2 * X) * np.sin(TWOPI * Y / 4.0) + 0.4 * np.cos(TWOPI * 2.0 * Z)
    assert np.abs(u - s.bwd(s.fwd(u))).max() < g["round_trip_tol"]
    for mo in c["modes"]:
        assert abs(s.mode(u, *mo["m"]) - complex(mo["re"], mo["im"])) < c["tol"]


def test_fft_known_answers_3d(Spec):
    """test_fft.f90:38-52."""
    s = Spec(16, 12, 8, TWOPI, 4.0, 1.0)
    g = O.Grid(16, 12, 8, TWOPI, 4.0, 1.0)
    X, Y, Z = g.coords()
    u = 0.7 + np.cos(2 * X) * np.sin(TWOPI * Y / 4.0) + 0.4 * np.cos(TWOPI * 2.0 * Z)
    v = s.bwd(s.fwd(u))
    assert np.abs(u - v).max() < 1e-13
    assert abs(s.mode(u, 0, 0, 2) - 0.2) < 1e-13
    assert abs(s.mode(u, 2, 1, 0) + 0.25j) < 1e-13


def test_ops_closed_forms(Spec):
    """test_ops.f90:30-71."""
    lx, ly = TWOPI, 4.0
    s = Spec(32, 16, 1, lx, ly, 1.0)
    g = O.Grid(32, 16, 1, lx, ly, 1.0)
    X, Y, _ = g.coords()
    u = np.sin(2 * X) * np.cos(TWOPI * Y / ly)
    du = s.grad(u)
    assert np.abs(du[0] - 2 * np.cos(2 * X) * np.cos(TWOPI * Y / ly)).max() < 1e-12
    assert np.abs(du[1] + (TWOPI / ly) * np.sin(2 * X) * np.sin(TWOPI * Y / ly)).max() < 1e-12
    assert np.abs(du[2]).max() < 1e-14
    ex = -(4.0 + (TWOPI / ly) ** 2) * u
    assert np.abs(s.lap(u) - ex).max() < 1e-11
    v = np.stack([du[0], du[1], np.zeros_like(u)])
    assert np.abs(s.div(v) - ex).max() < 1e-11
    kern = 1.0 / (1.0 + g.k2)
    cu = s.convolve(kern, s.fwd(u))
    assert np.abs(cu - u / (1.0 + 4.0 + (TWOPI / ly) ** 2)).max() < 1e-13
    u2 = np.cos(12 * X) + np.cos(10 * X)
    assert np.abs(s.dealias(u2) - np.cos(10 * X)).max() < 1e-13
    assert abs(s.kmax_dealiased - np.sqrt(100.0 + (TWOPI * 5.0 / ly) ** 2)) < 1e-12
    assert abs(s.mean(3.0 + du[0]) - 3.0) < 1e-13
    np.testing.assert_array_equal(s.mask(), g.mask)


SHAPES = [(32, 32, 32), (16, 12, 8), (48, 24, 4), (64, 1, 1), (32, 16, 1), (96, 8, 12), (8, 256, 4),
          (128, 64, 32)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dealias", [True, False])
def test_ops_vs_oracle(Spec, shape, dealias):
    nx, ny, nz = shape
    L = (20.0, 7.0, 3.0)
    s = Spec(nx, ny, nz, *L, dealias=dealias)
    o = O.Spec(nx, ny, nz, *L, dealias=dealias)
    rng = np.random.default_rng(1000 + nx + 7 * ny + 13 * nz)
    u = rng.standard_normal(o.g.shape)
    tol = 1e-12 * np.abs(u).max() * max(1.0, o.g.kabs.max() ** 2)
    np.testing.assert_array_equal(s.mask(), o.g.mask)
    uh = s.fwd(u)
    np.testing.assert_allclose(uh, o.fwd(u), atol=1e-14 * np.abs(u).max() * 8)
    np.testing.assert_allclose(s.bwd(uh), u, atol=1e-12 * np.abs(u).max())
    for d in range(3):
        np.testing.assert_allclose(s.grad(u)[d], o.grad(u)[d], atol=tol)
    np.testing.assert_allclose(s.lap(u), o.lap(u), atol=tol)
    v = rng.standard_normal((3,) + o.g.shape)
    np.testing.assert_allclose(s.div(v), o.div(list(v)), atol=tol * 3)
    np.testing.assert_allclose(s.dealias(u), o.dealias(u), atol=1e-12 * np.abs(u).max())
    kern = 1.0 / (1.0 + o.g.k2)
    np.testing.assert_allclose(s.convolve(kern, o.fwd(u)), o.convolve(kern, o.fwd(u)), atol=1e-13)
    assert abs(s.mean(u) - o.mean(u)) < 1e-13
    assert abs(s.mode(u, 1, 1, 1) - o.mode(u, 1, 1, 1)) < 1e-14
    assert abs(s.mode(u, 2, -1, -2) - o.mode(u, 2, -1, -2)) < 1e-14
    st = O.Stress("newtonian", 0.7, 0.3, 0.0, 1.0)
    u3 = rng.standard_normal((3,) + o.g.shape)
    pi_o = st.pi0(o, list(u3))
    pi_g = s.stress_pi0(st.eta, st.xi, u3)
    for c in range(6):
        np.testing.assert_allclose(pi_g[c], pi_o[c], atol=tol * 3)
    f_o = st.divergence(o, pi_o)
    f_g = s.stress_div(pi_g)
    for i in range(3):
        np.testing.assert_allclose(f_g[i], f_o[i], atol=tol * 30)
