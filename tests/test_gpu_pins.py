"""GPU checks of the parts that the model and operator tests leave open:
model_compute_dt (model.f90:150-183) against the oracle, spec_grad_hat
(spectral.f90:67-90) per direction, the initial conditions (initial.f90:23-75
and the seeded spectral fields of the benchmark configurations), and the
transforms against rocFFT (torch.fft on the GPU) at 128^3 and 256^3."""
import math

import numpy as np
import pytest

from oracle import hydft_oracle as O
from tests import _pairs as P

pytestmark = pytest.mark.gpu


def _model_pair(precision="fp64", **kw):
    from hydft_amd.model import Model, build_species
    p = P.params(**kw)
    st = P.structure(p.kind, p.gamma, p.kappa, p.rs) if "ry" in p.terms else None
    sp = build_species(p, structure=st)
    return p, sp, Model(p, species=sp, precision=precision), P.oracle_from_species(p, sp)


DT_CASES = {
    # ideal + hartree + ry, eta = 0: only the wave branch
    "ion_wave": dict(nx=32, ny=16, nz=8, lx=20.0, ly=10.0, lz=5.0, gamma=10.0, kappa=1.0),
    # strong viscosity: the viscous branch 2.5/nu_max
    "ion_viscous": dict(nx=32, ny=16, nz=8, lx=20.0, ly=10.0, lz=5.0, gamma=2.0, kappa=0.1, eta=5.0, xi=1.0),
    # Maxwell: eta_l/(m tau) in the wave speed, no viscous branch
    "ion_maxwell": dict(nx=16, ny=16, nz=16, lx=12.0, ly=12.0, lz=12.0, gamma=50.0, kappa=2.0, eta=0.2,
                        closure="maxwell", tau_model="constant", tau=0.5),
    # electrons: tf + gradient + hartree + ry
    "electron": dict(kind="electron", gamma=1.0, rs=1.86, nx=16, ny=16, nz=16, lx=20.0, ly=20.0, lz=20.0,
                     eta=0.1),
}


@pytest.mark.parametrize("name", sorted(DT_CASES))
@pytest.mark.parametrize("flow", [0.0, 0.05, 3.0])
def test_compute_dt_matches_oracle(gpu, name, flow):
    """hy_model_compute_dt equals the oracle's compute_dt on the same state;
    `flow` scales the velocity so that the advective branch dx/u_max wins for
    the largest one."""
    p, sp, m, om = _model_pair(**DT_CASES[name])
    n, mom, pi = P.smooth_fields(p, seed=9, amp_n=0.02, amp_u=1.0, with_pi=om.stress.evolve)
    mom = mom * flow
    m.set_state(n, mom, pi)
    P.set_oracle_state(om, n, mom, pi)
    dt_g = m.compute_dt()
    dt_o = om.compute_dt()
    assert abs(dt_g - dt_o) <= 1e-13 * dt_o, (name, flow, dt_g, dt_o)
    m.close()


def test_compute_dt_branches_are_exercised(gpu):
    """The flows above reach each branch of min(1.7/w_max, 2.5/nu_max, dx/u_max)."""
    p, sp, m, om = _model_pair(**DT_CASES["ion_viscous"])
    om.n = np.full(om.spc.g.shape, om.n0)
    numax = om.eta_l / om.mass * om.spc.g.kmax_dealiased ** 2
    assert abs(om.compute_dt() - om.cfl * 2.5 / numax) <= 1e-15 * om.dt
    m.close()


@pytest.mark.parametrize("shape", [(16, 12, 8), (32, 32, 32), (24, 1, 1), (24, 12, 1)])
def test_grad_hat_each_direction(gpu, shape):
    """hy_spec_grad_hat (spectral.f90:67-90) for every direction, including
    directions beyond ndim (zero), against the oracle."""
    from hydft_amd.spectral import Spec
    nx, ny, nz = shape
    lx, ly, lz = 7.0, 5.0, 3.0
    s = Spec(nx, ny, nz, lx, ly, lz)
    o = O.Spec(nx, ny, nz, lx, ly, lz)
    rng = np.random.default_rng(sum(shape))
    u = rng.random(o.g.shape) - 0.5
    uh = o.fwd(u)
    for d in (1, 2, 3):
        g = s.grad_hat(uh, d)
        ref = o.grad_hat(uh, d - 1)
        if d > o.g.ndim:
            assert not np.any(ref)
        scale = max(np.abs(ref).max(), 1.0)
        assert np.abs(g - ref).max() <= 1e-12 * scale, (shape, d)
    s.close()


def test_initial_mode_and_random_match_oracle(gpu):
    """set_initial_condition 'mode' and 'random' (initial.f90:40-58) equal
    the oracle's ic_mode / ic_random_field on the same deviates."""
    from hydft_amd.host.initial import set_initial_condition
    p = P.params(nx=24, ny=12, nz=8, lx=20.0, ly=9.0, lz=6.0)
    o = O.Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz)
    n0 = P.N0
    a = set_initial_condition(p, n0, kind="mode", amp=1e-3, mode=(2, 1, 1))
    b = O.ic_mode(o, n0, 1e-3, (2, 1, 1))
    assert np.abs(a - b).max() <= 1e-15 * n0
    a = set_initial_condition(p, n0, kind="random", amp=0.02, seed=31)
    b = O.ic_random_field(o, n0, 0.02, np.random.default_rng(31).random(o.g.shape))
    assert np.abs(a - b).max() <= 1e-14 * n0


def _spectrum(p, f):
    return O.Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.dealias).fwd(f)


def _kabs(p):
    g = O.Grid(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.dealias)
    return g.kabs, g.mask


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 24, 16), (64, 64, 1)])
def test_gauss_spectrum_field_properties(gpu, shape):
    """gauss_spectrum_field: real, zero mean, masked, and every kept mode has
    modulus A n0 exp(-q^2/2) (also on the kx = 0 plane)."""
    from hydft_amd.host.initial import gauss_spectrum_field
    nx, ny, nz = shape
    p = P.params(nx=nx, ny=ny, nz=nz, lx=4 * math.pi, ly=4 * math.pi, lz=4 * math.pi)
    n0, amp = P.N0, 1e-2
    d = gauss_spectrum_field(p, n0, amp, 1234) - n0
    c = _spectrum(p, d)
    kabs, mask = _kabs(p)
    want = amp * n0 * np.exp(-0.5 * kabs ** 2) * mask
    want[0, 0, 0] = 0.0
    assert abs(c[0, 0, 0]) <= 1e-15 * n0
    assert np.abs(np.abs(c) - want).max() <= 1e-12 * amp * n0
    # a different seed gives different phases, the same moduli
    c2 = _spectrum(p, gauss_spectrum_field(p, n0, amp, 99) - n0)
    assert np.abs(np.abs(c2) - want).max() <= 1e-12 * amp * n0
    assert np.abs(c2 - c).max() > 1e-3 * amp * n0


def test_shell_rms_field_properties(gpu):
    """shell_rms_field: zero mean, the requested rms, support on
    qlo < |q| < qhi only, one modulus on the whole shell."""
    from hydft_amd.host.initial import shell_rms_field
    p = P.params(nx=48, ny=48, nz=32, lx=2 * math.pi / 0.05, ly=2 * math.pi / 0.05, lz=2 * math.pi / 0.05)
    d = shell_rms_field(p, 5e-3, 0.1, 1.5, 7)
    assert abs(d.mean()) <= 1e-17
    assert abs(math.sqrt(np.mean(d * d)) / 5e-3 - 1.0) <= 1e-13
    c = np.abs(_spectrum(p, d))
    kabs, mask = _kabs(p)
    on = (kabs > 0.1) & (kabs < 1.5) & (mask > 0)
    assert np.abs(c[~on]).max() <= 1e-12 * c[on].max()
    assert c[on].max() / c[on].min() - 1.0 <= 1e-10


@pytest.mark.parametrize("n", [128, 256])
def test_fft_matches_rocfft(gpu, n):
    """hy_fft_r2c / hy_fft_c2r against rocFFT (torch.fft on the GPU) at n^3,
    FP64: forward scaled by 1/N, backward unnormalised, input preserved
    (fft.f90:52-69)."""
    import torch
    from hydft_amd.spectral import Spec
    s = Spec(n, n, n, 10.0, 10.0, 10.0)
    g = torch.Generator(device="cuda").manual_seed(n)
    u = torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g) - 0.5
    u0 = u.clone()
    uh = s.fwd(u)
    ref = torch.fft.rfftn(u) / float(n ** 3)
    assert torch.equal(u, u0)
    err = float((uh - ref).abs().max() / ref.abs().max())
    assert err <= 1e-13, err
    uh0 = uh.clone()
    v = s.bwd(uh)
    assert torch.equal(uh, uh0)
    refv = torch.fft.irfftn(uh, s=(n, n, n)) * float(n ** 3)
    err = float((v - refv).abs().max() / refv.abs().max())
    assert err <= 1e-13, err
    assert float((v - u).abs().max()) <= 1e-13
    torch.cuda.synchronize()
    s.close()
