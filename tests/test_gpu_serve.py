"""Captured steps, batched and served runs on the GPU."""
import numpy as np
import pytest

from hydft_amd import serve

from . import _pairs as P


def _req(rid, **kw):
    r = serve.chi_requests(1, seed=100 + rid, grid=(kw.pop("N", 16),), steps=kw.pop("steps", 60))[0]
    r.rid = rid
    for k, v in kw.items():
        setattr(r.p, k, v)
    return r


@pytest.mark.gpu
def test_graph_replay_equals_advance(gpu):
    """Replays of a captured 20-step sequence take exactly the steps of
    hy_model_advance with the same dt cadence (FP64, bitwise)."""
    req = _req(0, N=16)
    cache = serve._hnc.StructureCache()
    a = serve.make_model(req, "fp64", cache)
    b = serve.make_model(req, "fp64", cache)
    taken, rows_a = a.advance(60, dt_every=20, probe=req.mode)
    assert taken == 60
    b.graph_build(20, dt_every=20, probe=req.mode, ring=60)
    b.graph_launch(3)
    rows_b = b.graph_probe()
    b.check()
    na, ma, _ = a.get_state()
    nb, mb, _ = b.get_state()
    assert np.array_equal(na, nb) and np.array_equal(ma, mb)
    assert np.array_equal(rows_a, rows_b)
    assert a.time == b.time
    # a further eager advance after the replays continues the same cadence
    a.advance(20, dt_every=20)
    b.advance(20, dt_every=20)
    assert np.array_equal(a.get_state()[0], b.get_state()[0])
    a.close(); b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_result_same_alone_and_in_the_mix(gpu, precision):
    """Each request's probe series is the same served alone and served among
    others of different grids and closures (both bitwise: the work of a run
    keeps its own stream order)."""
    reqs = [_req(i, N=(16, 24, 32)[i % 3]) for i in range(6)]
    cache = serve._hnc.StructureCache()
    alone = {}
    for r in reqs[:3]:
        res = serve.Server(precision=precision, cache=cache).serve([r], realtime=False)
        alone[r.rid] = res[0]
    srv = serve.Server(precision=precision, cache=cache, max_slots=6)
    mix = {r.rid: r for r in srv.serve(reqs, realtime=False)}
    assert srv.peak_active >= 4
    for rid, ra in alone.items():
        rb = mix[rid]
        assert np.array_equal(ra.probe, rb.probe), rid
        assert ra.value == rb.value or np.allclose(ra.value["chi"], rb.value["chi"], rtol=0, atol=0)


@pytest.mark.gpu
def test_admission_by_hbm_budget(gpu):
    reqs = [_req(i, N=24, steps=40) for i in range(6)]
    one = serve.estimate_bytes(reqs[0].p, "mixed")
    srv = serve.Server(hbm_budget=2.5 * one, precision="mixed", max_slots=16)
    out = srv.serve(reqs, realtime=False)
    assert len(out) == 6
    assert srv.peak_active <= 2
    assert srv.peak_bytes <= 2.5 * one


@pytest.mark.gpu
def test_served_matches_one_at_a_time(gpu):
    """Served results equal the one-model-at-a-time results of the same
    requests (FP64, bitwise: same kernels, same order per run)."""
    reqs = [_req(i, N=16, steps=40) for i in range(4)]
    cache = serve._hnc.StructureCache()
    a = {r.rid: r for r in serve.one_at_a_time(reqs, precision="fp64", cache=cache)}
    b = {r.rid: r for r in serve.Server(precision="fp64", cache=cache).serve(reqs, realtime=False)}
    for rid in a:
        assert np.array_equal(a[rid].probe, b[rid].probe), rid


@pytest.mark.gpu
def test_dispersion_request_matches_linear_theory(gpu):
    """A served single-mode dispersion request, in the manner of
    test_mode_newtonian.f90 (ions, Gamma = 2, kappa = 0.1, nx = 32, lx = 20,
    mode 2, amplitude 1e-4, cfl 0.1, t = 100, fit from t = 5): frequency within
    1 % and damping within 3 % of the closed form the result carries."""
    from hydft_amd.host.params import Params
    p = Params()
    p.kind, p.gamma, p.kappa = "ion", 2.0, 0.1
    p.nx, p.lx = 32, 20.0
    p.transport, p.eta, p.closure = "constant", 0.1, "newtonian"
    p.hnc_nr, p.hnc_rmax = 4096, 60.0
    p.cfl, p.dt = 0.1, 0.0
    p.ic_amplitude = 1.0e-4
    p.finalize()
    req = serve.Request(0, p, nsteps=1200, kind="dispersion", mode=(2, 0, 0), t_fit=5.0)
    res = serve.Server(precision="fp64", batch=True).serve([req], realtime=False)[0]
    v = res.value
    assert res.probe[-1, 0] > 90.0
    assert abs(v["omega"] / v["omega_linear"] - 1.0) < 0.01, v
    assert abs(v["gamma"] / v["gamma_linear"] - 1.0) < 0.03, v


@pytest.mark.gpu
def test_served_chi_carries_linear_theory(gpu):
    reqs = [_req(i, N=16, steps=40) for i in range(2)]
    out = serve.Server(precision="mixed", batch=True).serve(reqs, realtime=False)
    for r in out:
        assert "chi_linear" in r.value and np.isfinite(r.value["chi_rel_dev"])


@pytest.mark.gpu
def test_advance_keeps_the_graph_ring(gpu):
    """hy_model_advance with more probed steps than a captured graph's ring
    leaves the ring where the graph writes: replays after it record the same
    rows as eager steps of a twin model."""
    req = _req(0, N=16)
    cache = serve._hnc.StructureCache()
    a = serve.make_model(req, "fp64", cache)
    b = serve.make_model(req, "fp64", cache)
    a.graph_build(20, dt_every=20, probe=req.mode, ring=20)
    ta, rows_a = a.advance(40, dt_every=20, probe=req.mode)
    tb, rows_b0 = b.advance(40, dt_every=20, probe=req.mode)
    assert ta == tb == 40 and np.array_equal(rows_a, rows_b0)
    a.graph_launch(1)
    rows_g = a.graph_probe()
    _, rows_b = b.advance(20, dt_every=20, probe=req.mode)
    assert np.array_equal(rows_g, rows_b)
    assert np.array_equal(a.get_state()[0], b.get_state()[0])
    a.close(); b.close()


@pytest.mark.gpu
def test_batch_holds_the_probe_ring(gpu):
    """While a batch holds a model, the model's probe ring cannot grow (the
    batch writes into it); after the batch is destroyed it can."""
    from hydft_amd import abi
    from hydft_amd.model import Batch
    reqs = [_req(i, N=16) for i in range(2)]
    cache = serve._hnc.StructureCache()
    ms = [serve.make_model(r, "fp64", cache) for r in reqs]
    bt = Batch(ms, 20, dt_every=20, probe=(1, 0, 0), ring=20)
    with pytest.raises(abi.HydftError) as e:
        ms[0].graph_build(20, dt_every=20, probe=(1, 0, 0), ring=40)
    assert e.value.code == abi.HY_ERR_INVALID
    ms[0].graph_build(20, dt_every=20, probe=(1, 0, 0), ring=20)  # same ring: allowed
    bt.close()
    ms[0].graph_build(20, dt_every=20, probe=(1, 0, 0), ring=40)
    ms[0].graph_launch(1)
    assert np.all(np.isfinite(ms[0].graph_probe()[:20]))
    for m in ms:
        m.close()


@pytest.mark.gpu
def test_a_failed_request_does_not_stop_the_others(gpu):
    """Requests that fail are reported alone: one whose density turns
    non-positive in its steps (a mode of 0.9 n0 stepped with a fixed dt far
    above the stability limit) with the reference's message (model.f90:91),
    and one whose initial density is not positive (refused at setup); the
    request served beside them gives the result it gives alone."""
    good = _req(0, N=16)
    bad = _req(1, N=16)
    bad.kind, bad.mode = "dispersion", (1, 0, 0)
    bad.p.vext, bad.p.vext_amplitude = "none", 0.0
    bad.p.ic_amplitude, bad.p.dt = 0.9, 50.0
    worse = _req(2, N=16)
    worse.kind, worse.mode = "dispersion", (1, 0, 0)
    worse.p.vext, worse.p.vext_amplitude = "none", 0.0
    worse.p.ic_amplitude = 1.5
    cache = serve._hnc.StructureCache()
    alone = serve.Server(precision="fp64", cache=cache).serve([good], realtime=False)[0]
    srv = serve.Server(precision="fp64", cache=cache, max_slots=4)
    out = {r.rid: r for r in srv.serve([good, bad, worse], realtime=False)}
    assert out[1].error and "non-positive" in out[1].error and out[1].steps == 0
    assert out[2].error and "not positive" in out[2].error
    assert not out[0].error and np.array_equal(out[0].probe, alone.probe)
    assert serve.summary(list(out.values()), 1.0)["failed"] == 2
    one = {r.rid: r for r in serve.one_at_a_time([good, bad, worse], precision="fp64", cache=cache)}
    assert one[1].error and one[2].error and not one[0].error
    assert np.array_equal(one[0].probe, alone.probe)
