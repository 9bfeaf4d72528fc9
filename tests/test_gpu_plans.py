"""The pencil plan autotuner: the column count per workgroup of every y/z
pass shape is measured once and kept (fft.f90:46-47 asks FFTW for measured
plans).  A plan changes how the columns are spread over workgroups, not the
arithmetic of a column: runs give bitwise the same state with the measured
plans as with the untimed defaults, and the batched passes get plans of
their own batch size."""
import ctypes as C

import numpy as np
import pytest

from hydft_amd import abi, serve
from tests import _pairs as P

pytestmark = pytest.mark.gpu
CANDIDATES = {8, 16, 24, 32, 64, 128}


def _plans():
    n = abi.lib().lib.hy_plan_list(0, None)
    buf = (C.c_int * (8 * max(n, 1)))()
    abi.lib().lib.hy_plan_list(n, buf)
    return [tuple(buf[8 * i:8 * i + 8]) for i in range(n)]


def _run(precision):
    from hydft_amd.model import Model
    p = P.params(nx=16, ny=96, nz=48, lx=8.0, ly=48.0, lz=24.0, gamma=10.0, kappa=1.0, eta=0.3, xi=0.1, dt=0.01)
    st = P.structure(p.kind, p.gamma, p.kappa, p.rs)
    m = Model(p, structure=st, precision=precision)
    n, mom, _ = P.smooth_fields(p, seed=5, amp_n=0.02, amp_u=0.01)
    m.set_state(n, mom)
    m.advance(3)
    out = m.get_state()
    m.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_measured_plans_give_the_same_state(gpu, precision):
    try:
        abi.call("hy_plan_reset")
        abi.call("hy_plan_tuning", 0)
        n0, m0, _ = _run(precision)
        assert _plans() == []
        abi.call("hy_plan_tuning", 1)
        n1, m1, _ = _run(precision)
        plans = _plans()
        print(precision, "plans (fp64, L, inverse, A, B, nfields, batch, columns):", plans)
        assert plans and all(p[7] in CANDIDATES for p in plans)
        assert {p[1] for p in plans} >= {96, 48}
        assert np.array_equal(n0, n1) and np.array_equal(m0, m1)
    finally:
        abi.call("hy_plan_tuning", 1)


def test_batched_passes_get_plans_of_their_batch_size(gpu):
    from hydft_amd.model import Batch
    reqs = [r for r in serve.chi_requests(9, seed=330, grid=(24,), steps=20) if r.p.closure == "newtonian"][:3]
    cache = serve._hnc.StructureCache()
    alone = []
    for r in reqs:
        m = serve.make_model(r, "mixed", cache)
        m.advance(20, dt_every=20)
        alone.append(m.get_state()[0])
        m.close()
    abi.call("hy_plan_reset")
    ms = [serve.make_model(r, "mixed", cache) for r in reqs]
    b = Batch(ms, 20, dt_every=20)
    plans = _plans()
    assert any(p[6] == 3 for p in plans), plans
    b.launch(1)
    b.sync()
    for i, m in enumerate(ms):
        assert np.array_equal(m.get_state()[0], alone[i]), i
    b.close()
    for m in ms:
        m.close()
