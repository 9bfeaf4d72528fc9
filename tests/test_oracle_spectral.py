"""Pins the oracle's grid, FFT and spectral operators (CPU).

Known answers of test_fft.f90:20-52 and test_ops.f90:30-71, a brute-force DFT
on a tiny 3-D grid (independent of numpy's FFT), and the 2/3-rule mask
against the box |m| < (2/3)(N/2) counted by hand."""
import numpy as np

from oracle import hydft_oracle as O

TWOPI = O.TWOPI


def test_fft_known_answers():
    s = O.Spec(64, 1, 1, TWOPI, TWOPI, TWOPI)
    x = s.g.x
    u = (1.5 + 0.3 * np.cos(3 * x) + 0.2 * np.sin(5 * x)).reshape(1, 1, 64)
    uh = s.fwd(u)
    assert np.abs(u - s.bwd(uh)).max() < 1e-13
    assert abs(uh[0, 0, 0] - 1.5) < 1e-13
    assert abs(uh[0, 0, 3] - 0.15) < 1e-13
    assert abs(uh[0, 0, 5] + 0.1j) < 1e-13
    assert abs(s.mode(u, 3, 0, 0) - 0.15) < 1e-13
    s = O.Spec(16, 12, 8, TWOPI, 4.0, 1.0)
    X, Y, Z = s.g.coords()
    u = 0.7 + np.cos(2 * X) * np.sin(TWOPI * Y / 4.0) + 0.4 * np.cos(TWOPI * 2.0 * Z)
    assert np.abs(u - s.bwd(s.fwd(u))).max() < 1e-13
    assert abs(s.mode(u, 0, 0, 2) - 0.2) < 1e-13
    assert abs(s.mode(u, 2, 1, 0) + 0.25j) < 1e-13


def test_fft_matches_brute_force_dft():
    nx, ny, nz = 6, 4, 4
    s = O.Spec(nx, ny, nz, 3.0, 2.0, 5.0)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((nz, ny, nx))
    ref = np.zeros((nz, ny, nx // 2 + 1), complex)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                acc = 0j
                for z in range(nz):
                    for y in range(ny):
                        for x in range(nx):
                            acc += u[z, y, x] * np.exp(-2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
                ref[kz, ky, kx] = acc / (nx * ny * nz)
    np.testing.assert_allclose(s.fwd(u), ref, atol=1e-14)


def test_ops_closed_forms():
    lx, ly = TWOPI, 4.0
    s = O.Spec(32, 16, 1, lx, ly, 1.0)
    X, Y, _ = s.g.coords()
    u = np.sin(2 * X) * np.cos(TWOPI * Y / ly)
    du = s.grad(u)
    assert np.abs(du[0] - 2 * np.cos(2 * X) * np.cos(TWOPI * Y / ly)).max() < 1e-12
    assert np.abs(du[1] + (TWOPI / ly) * np.sin(2 * X) * np.sin(TWOPI * Y / ly)).max() < 1e-12
    assert np.abs(du[2]).max() < 1e-14
    ex = -(4.0 + (TWOPI / ly) ** 2) * u
    assert np.abs(s.lap(u) - ex).max() < 1e-11
    assert np.abs(s.div([du[0], du[1], 0 * u]) - ex).max() < 1e-11
    cu = s.convolve(1.0 / (1.0 + s.g.k2), s.fwd(u))
    assert np.abs(cu - u / (1.0 + 4.0 + (TWOPI / ly) ** 2)).max() < 1e-13
    assert np.abs(s.dealias(np.cos(12 * X) + np.cos(10 * X)) - np.cos(10 * X)).max() < 1e-13
    assert abs(s.g.kmax_dealiased - np.sqrt(100.0 + (TWOPI * 5.0 / ly) ** 2)) < 1e-12
    assert abs(s.mean(3.0 + du[0]) - 3.0) < 1e-13


def test_mask_is_two_thirds_box():
    for n in (8, 12, 16, 32, 48, 64, 96, 128, 256):
        g = O.Grid(n, n, n, 7.0, 7.0, 7.0)
        keep = [m for m in range(n // 2 + 1) if m < (2.0 / 3.0) * (n // 2)]
        # the reference's floating-point test may flip an exact tie; none of these sizes has one
        M = max(keep)
        assert g.mask[0, 0, :].sum() == M + 1
        assert g.mask[:, 0, 0].sum() == 2 * M + 1
        assert g.mask.sum() == (M + 1) * (2 * M + 1) ** 2
    g = O.Grid(16, 8, 1, 1.0, 1.0, 1.0, dealias=False)
    assert g.mask[:, :, -1].sum() == 0 and g.mask[:, 4, :].sum() == 0
    assert g.mask.sum() == 8 * 7


def test_golden_fft_known_answers():
    """The known answers of test_fft.f90 as committed data
    (tests/golden/reference_known_answers.json) against the oracle's r2c."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))["fft"]
    c = g["one_d"]
    s = O.Spec(c["n"], 1, 1, c["length"], TWOPI, TWOPI)
    x = s.g.x
    u = (1.5 + 0.3 * np.cos(3 * x) + 0.2 * np.sin(5 * x)).reshape(1, 1, c["n"])
    uh = s.fwd(u)
    assert np.abs(u - s.bwd(uh)).max() < g["round_trip_tol"]
    for co in c["coefficients"]:
        assert abs(uh[0, 0, co["kx"]] - complex(co["re"], co["im"])) < c["tol"]
    c = g["three_d"]
    nx, ny, nz = c["shape_xyz"]
    lx, ly, lz = c["lengths_xyz"]
    s = O.Spec(nx, ny, nz, lx, ly, lz)
    X, Y, Z = s.g.coords()
    u = 0.7 + np.cos(2 * X) * np.sin(TWOPI * Y / 4.0) + 0.4 * np.cos(TWOPI * 2.0 * Z)
    assert np.abs(u - s.bwd(s.fwd(u))).max() < g["round_trip_tol"]
    for mo in c["modes"]:
        assert abs(s.mode(u, *mo["m"]) - complex(mo["re"], mo["im"])) < c["tol"]
