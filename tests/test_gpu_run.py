"""The time-loop driver (src/hydft.f90:21-55) on the GPU: cadence, tend
clamp, output files, and the final state against the oracle."""
import numpy as np
import pytest

from hydft_amd.host import io
from hydft_amd.host.initial import set_initial_condition

from . import _pairs as P


@pytest.mark.gpu
def test_run_loop_files_and_state(gpu, tmp_path):
    from hydft_amd.model import Model
    from hydft_amd.run import run

    p = P.params(nx=32, lx=20.0, gamma=2.0, kappa=0.1, eta=0.3, dt=0.03, tend=0.5, iprobe=1, idiag=5,
                 iout=10, ic="mode", ic_amplitude=1e-3, ic_mode=(2, 0, 0), probe_mode=(2, 0, 0))
    p.prefix = str(tmp_path / "case")
    st = P.structure(p.kind, p.gamma, p.kappa, p.rs)
    m = Model(p, structure=st, precision="fp64")
    out = run(p, verbose=False, model=m)
    # 16 steps of 0.03, then the last clamped to tend (hydft.f90:44)
    assert out["steps"] == 17
    assert abs(out["t"] - 0.5) <= 1e-14
    probe = io.read_table(out["probe"])
    assert probe.shape == (18, 6)                     # initial row + every step
    diag = io.read_table(out["diag"])
    assert list(diag[:, 0].astype(int)) == [0, 5, 10, 15, 17]
    assert abs(diag[-1, 1] - 0.5) <= 1e-12
    assert abs(diag[-1, 2] - 0.02) <= 1e-12           # the clamped step
    assert [s.rsplit("_", 1)[1] for s in out["fields"]] == ["000000.bin", "000010.bin"]
    # oracle: same initial condition, same step sequence
    om = P.oracle_model(p, st)
    om.n = set_initial_condition(p, P.N0)
    for _ in range(16):
        om.dt = 0.03
        om.advance()
    om.dt = 0.5 - om.t
    om.advance()
    n, mom, _ = m.get_state()
    assert P.rel_l2(n, om.n) <= 1e-11
    assert P.rel_l2(mom[0], om.mom[0]) <= 1e-9
    # last probe row = mode (2,0,0) of the final state, to the table's 10
    # significant digits (es18.9e3)
    a = om.spc.mode(om.n, 2, 0, 0) / P.N0
    assert abs(probe[-1, 1] - a.real) <= 1e-9 * abs(a) and abs(probe[-1, 2] - a.imag) <= 1e-9 * abs(a)
    # the step-10 field file holds the state of step 10
    f = io.read_fields(out["fields"][1])
    assert f["nx"] == 32 and f["ndim"] == 1 and abs(f["t"] - 0.3) <= 1e-12
    assert f["pi"] is None
