"""functional_t and its terms on device fields (hy_fun_*), against the
reference's own check and the oracle.

test_kernels.f90:28-69 restated term by term through the device entry
points: every term's linear kernel K_i(k) equals the finite difference of its
mu_i[n] about n0 (relative 2e-4), for ions (ideal, hartree, ry) and electrons
(tf, gradient, hartree, ry and tf, gradient, ry), 64 points over lx = 20,
modes 1..6, amplitude 1e-6.  Then each term's mu and energy against the
oracle's on a 3-D field (relative 1e-12), and the non-positive density error."""
import numpy as np
import pytest

from oracle import hydft_oracle as O
from tests import _pairs as P

pytestmark = pytest.mark.gpu
TWOPI = O.TWOPI


def _setup(kind, terms, shape=(64, 1, 1), lengths=(20.0, 1.0, 1.0)):
    from hydft_amd.functional import Functional
    from hydft_amd.model import build_species
    from hydft_amd.spectral import Spec
    gamma = 1.0 if kind == "electron" else 2.0
    p = P.params(kind=kind, terms=terms, gamma=gamma, kappa=0.1, rs=1.86, nx=shape[0], ny=shape[1], nz=shape[2],
                 lx=lengths[0], ly=lengths[1], lz=lengths[2], hnc_nr=4096, hnc_rmax=60.0,
                 transport="constant", closure="newtonian")
    st = P.structure(kind, gamma, 0.1, 1.86, nr=4096, rmax=60.0) if "ry" in terms else None
    sp = build_species(p, structure=st)
    spec = Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz)
    return p, sp, spec, Functional(spec, sp)


@pytest.mark.parametrize("kind,terms", [("ion", "ideal,hartree,ry"), ("electron", "tf,gradient,hartree,ry"),
                                        ("electron", "tf,gradient,ry")])
def test_term_kernels_equal_finite_differences(gpu, kind, terms):
    """test_kernels.f90:28-69 through hy_fun_add_mu / hy_fun_mu / hy_fun_kernel."""
    p, sp, spec, fun = _setup(kind, terms)
    n0 = sp.units.n0
    eps = 1.0e-6
    x = np.arange(p.nx) * (p.lx / p.nx)
    n0a = np.full(spec.shape, n0)
    n0hat = spec.fwd(n0a)
    zero = np.zeros(spec.shape)
    worst = 0.0
    for m in range(1, 7):
        k = TWOPI * m / p.lx
        n = (n0 * (1.0 + eps * np.cos(k * x))).reshape(spec.shape)
        nhat = spec.fwd(n)
        for it, name in enumerate(fun.names):
            mu = fun.add_mu(it, n, nhat, zero)
            mu0 = fun.add_mu(it, n0a, n0hat, zero)
            kfd = 2.0 * spec.mode(mu - mu0, m, 0, 0).real / (eps * n0)
            kan = fun.kernel(k, it)
            err = abs(kfd - kan) / max(abs(kan), 1.0e-3)
            worst = max(worst, err)
            assert err < 2.0e-4, (kind, terms, m, name, kfd, kan, err)
        mu = fun.mu(n, nhat)
        mu0 = fun.mu(n0a, n0hat)
        kfd = 2.0 * spec.mode(mu - mu0, m, 0, 0).real / (eps * n0)
        kan = fun.kernel(k)
        assert abs(kfd - kan) / abs(kan) < 2.0e-4, (kind, terms, m)
    print(kind, terms, f"worst K(k) vs FD {worst:.2e}")
    fun.close()


@pytest.mark.parametrize("kind,terms", [("ion", "ideal,hartree,ry"), ("electron", "tf,gradient,hartree,ry")])
def test_terms_match_oracle(gpu, kind, terms):
    """add_mu, the functional's mu and each term's energy on a 3-D field against
    the oracle's terms (fe_*.f90 add_mu and energy)."""
    p, sp, spec, fun = _setup(kind, terms, shape=(16, 12, 8), lengths=(20.0, 9.0, 6.0))
    st = sp.structure
    un = O.Units(kind=kind, gamma=p.gamma, kappa=p.kappa, rs=p.rs)
    of = O.build_functional(un, terms, None if st is None else st.k, None if st is None else st.c, p.tfk_gamma)
    ospc = O.Spec(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz)
    of.setup(ospc)
    rng = np.random.default_rng(8)
    n = un.n0 * (1.0 + 0.05 * (rng.random(ospc.g.shape) - 0.5))   # not band-limited: the terms take any field
    nhat = ospc.fwd(n)
    zero = np.zeros(ospc.g.shape)
    for it, t in enumerate(of.terms):
        mg = fun.add_mu(it, n, nhat, zero)
        mo = t.add_mu(ospc, n, nhat, zero)
        assert np.abs(mg - mo).max() <= 1e-12 * max(np.abs(mo).max(), 1e-300), (t.name,)
        eg, eo = fun.energy(n, nhat, it), t.energy(ospc, n, nhat)
        assert abs(eg - eo) <= 1e-12 * abs(eo), (t.name, eg, eo)
        assert abs(fun.kernel(0.7, it) - float(np.atleast_1d(t.kernel(np.array([0.7])))[0])) <= 1e-13 * abs(
            fun.kernel(0.7, it)) + 1e-300
    mu = fun.mu(n, nhat)
    mo = of.mu(ospc, n, nhat)
    assert np.abs(mu - mo).max() <= 1e-12 * np.abs(mo).max()
    assert abs(fun.energy(n, nhat) - of.energy(ospc, n, nhat)) <= 1e-12 * abs(of.energy(ospc, n, nhat))
    fun.close()


def test_nonpositive_density_is_an_error(gpu):
    from hydft_amd import abi
    p, sp, spec, fun = _setup("ion", "ideal,hartree,ry")
    n = np.full(spec.shape, sp.units.n0)
    n[0, 0, 3] = -1e-3
    nhat = spec.fwd(n)
    with pytest.raises(abi.HydftError) as e:
        fun.add_mu("ideal", n, nhat, np.zeros(spec.shape))
    assert e.value.code == abi.HY_ERR_NONPOSITIVE
    # the convolution terms take any density
    fun.add_mu("hartree", n, nhat, np.zeros(spec.shape))
    fun.close()
