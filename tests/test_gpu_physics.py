"""The reference's integration tests (conservation, single mode, Maxwell,
driven chi(k,w), ground state) on the GPU model, FP64, with the reference's
parameters and tolerances (test_*.f90 of the reference)."""
import pytest

from tests import _scenarios as S

pytestmark = pytest.mark.gpu


def B(p, st):
    return S.GpuBackend(p, st, "fp64")


@pytest.mark.parametrize("closure,eta,tau", [("newtonian", 2.0, 0.0), ("maxwell", 0.2, 2.0)])
def test_conservation(gpu, closure, eta, tau):
    r = S.conservation(B, closure, eta, tau)
    assert r["mass"] < 1e-12
    assert r["mom"] < 1e-12
    assert r["ekin"] > 0.0
    if closure == "newtonian":
        assert r["energy"] < 1e-8


@pytest.mark.parametrize("kind", ["ion", "electron"])
def test_mode_newtonian(gpu, kind):
    r = S.single_mode(B, kind)
    assert abs(r["om"] / r["w_lin"].real - 1) < 0.01
    assert abs(r["ga"] / r["w_lin"].imag - 1) < 0.03


def test_mode_maxwell(gpu):
    rn = S.single_mode(B, "ion", "newtonian", 0.1, 0.0, tend=120.0, tfit=20.0)
    for tau in (3.0, 0.02):
        r = S.single_mode(B, "ion", "maxwell", 0.1, tau, tend=120.0, tfit=20.0)
        assert abs(r["om"] / r["w_lin"].real - 1) < 0.01
        assert abs(r["ga"] / r["w_lin"].imag - 1) < 0.05
    assert abs(r["om"] / rn["om"] - 1) < 0.01 and abs(r["ga"] / rn["ga"] - 1) < 0.05


@pytest.mark.parametrize("closure,tau", [("newtonian", 0.0), ("maxwell", 1.0)])
def test_driven_mode(gpu, closure, tau):
    r = S.driven(B, closure, tau)
    assert r["err"] < 0.02
    assert r["scatter"] < 0.05


@pytest.mark.parametrize("kind", ["ion", "electron"])
def test_ground_state(gpu, kind):
    r = S.ground_state(B, kind)
    assert r["spread"] < 2e-3
    assert r["err"] < 1e-2
