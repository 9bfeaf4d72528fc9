"""Batches of same-shape runs: one launch per kernel for all models
(hy_batch_*).  A model's result in a batch is the result it has alone."""
import numpy as np
import pytest

from hydft_amd import abi, serve
from hydft_amd.model import Batch


def _reqs(n, N=16, steps=40, closure=None, seed=300):
    out = []
    for r in serve.chi_requests(3 * n, seed=seed, grid=(N,), steps=steps):
        if closure is None or r.p.closure == closure:
            out.append(r)
    return out[:n]


def _alone(reqs, precision, cache, steps, dt_every):
    res = []
    for r in reqs:
        m = serve.make_model(r, precision, cache)
        taken, rows = m.advance(steps, dt_every=dt_every, probe=r.mode)
        assert taken == steps
        res.append((m.get_state(), rows, m.time))
        m.close()
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("precision,closure", [("fp64", "newtonian"), ("fp64", "maxwell"),
                                               ("mixed", "newtonian"), ("mixed", "maxwell")])
def test_batch_equals_alone(gpu, precision, closure):
    """Every model of a batch ends bitwise where it ends alone: state,
    probe series (CFL dt recomputed every 20 steps) and time."""
    reqs = _reqs(5, closure=closure)
    cache = serve._hnc.StructureCache()
    alone = _alone(reqs, precision, cache, 40, 20)
    ms = [serve.make_model(r, precision, cache) for r in reqs]
    b = Batch(ms, 20, dt_every=20, probe=(1, 0, 0), ring=40)
    b.launch(2)
    b.sync()
    for i, m in enumerate(ms):
        (na, ma, pa), rows, t = alone[i]
        nb, mb, pb = m.get_state()
        assert np.array_equal(na, nb) and np.array_equal(ma, mb), i
        if pa is not None:
            assert np.array_equal(pa, pb), i
        assert np.array_equal(rows, b.probe(i)), i
        assert m.time == t, i
    b.close()
    for m in ms:
        m.close()


@pytest.mark.gpu
def test_batch_models_usable_after(gpu):
    """After the batch is destroyed its models continue with hy_model_advance
    exactly as a model that took all steps alone."""
    reqs = _reqs(3, closure="newtonian", seed=310)
    cache = serve._hnc.StructureCache()
    alone = _alone(reqs, "fp64", cache, 60, 20)
    ms = [serve.make_model(r, "fp64", cache) for r in reqs]
    b = Batch(ms, 20, dt_every=20)
    b.launch(2)
    b.close()
    for i, m in enumerate(ms):
        m.advance(20, dt_every=20)
        assert np.array_equal(m.get_state()[0], alone[i][0][0]), i
        assert m.time == alone[i][2]
        m.close()


@pytest.mark.gpu
def test_batch_rejects_mixed_shapes(gpu):
    cache = serve._hnc.StructureCache()
    a = _reqs(1, N=16, closure="newtonian")[0]
    c = _reqs(1, N=24, closure="newtonian")[0]
    d = _reqs(1, N=16, closure="maxwell")[0]
    for other in (c, d):
        ms = [serve.make_model(a, "fp64", cache), serve.make_model(other, "fp64", cache)]
        with pytest.raises(abi.HydftError) as e:
            Batch(ms, 20, dt_every=20)
        assert e.value.code == abi.HY_ERR_INVALID
        for m in ms:
            m.close()


@pytest.mark.gpu
def test_batch_rejects_a_model_twice(gpu):
    cache = serve._hnc.StructureCache()
    m = serve.make_model(_reqs(1, closure="newtonian")[0], "fp64", cache)
    with pytest.raises(abi.HydftError) as e:
        Batch([m, m], 20, dt_every=20)
    assert e.value.code == abi.HY_ERR_INVALID
    m.close()


@pytest.mark.gpu
def test_batched_server_matches_alone(gpu):
    """The batching server groups same-shape requests; each result equals
    the one-at-a-time result (FP64, bitwise)."""
    reqs = [serve.chi_requests(1, seed=400 + i, grid=((16, 24)[i % 2],), steps=40)[0] for i in range(8)]
    for i, r in enumerate(reqs):
        r.rid = i
    cache = serve._hnc.StructureCache()
    a = {r.rid: r for r in serve.one_at_a_time(reqs, precision="fp64", cache=cache)}
    srv = serve.Server(precision="fp64", cache=cache, batch=True)
    b = {r.rid: r for r in srv.serve(reqs, realtime=False)}
    assert srv.batches_launched < len(reqs)
    for rid in a:
        assert np.array_equal(a[rid].probe, b[rid].probe), rid
        assert a[rid].value["chi"] == b[rid].value["chi"], rid


@pytest.mark.gpu
def test_batches_split_by_kept_box(gpu):
    """48^3 runs whose box lengths put the 2/3-rule boundary mode on
    different sides of the threshold have different kept boxes; the server
    runs one batch per kept box, each result equal to its one-at-a-time run."""
    reqs = [r for r in serve.chi_requests(200, seed=11, grid=(48,), steps=40) if r.p.closure == "newtonian"]
    for r in reqs:
        r.p.terms, r.p.structure = "ideal,hartree", "none"
    boxes = {}
    for r in reqs:
        m = serve.make_model(r, "fp64")
        boxes.setdefault(m.kept(), []).append(r)
        m.close()
        if len(boxes) >= 2 and all(len(v) >= 2 for v in boxes.values()):
            break
    assert len(boxes) >= 2, list(boxes)
    pick = [v[i] for v in list(boxes.values())[:2] for i in range(2)]
    for i, r in enumerate(pick):
        r.rid = i
    srv = serve.Server(precision="fp64", batch=True)
    b = {r.rid: r for r in srv.serve(pick, realtime=False)}
    assert srv.batches_launched == 2
    a = {r.rid: r for r in serve.one_at_a_time(pick, precision="fp64")}
    for rid in a:
        assert np.array_equal(a[rid].probe, b[rid].probe), rid


@pytest.mark.gpu
@pytest.mark.parametrize("precision,closure", [("fp64", "newtonian"), ("mixed", "maxwell")])
def test_batch_runs_of_different_lengths(gpu, precision, closure):
    """Runs of 20, 40 and 60 steps in one batch (hy_batch_set_stops): each ends
    bitwise where it ends alone (state, probe rows, time), while the batch
    takes 60 steps; afterwards a stopped run continues alone as usual."""
    reqs = _reqs(3, closure=closure, seed=320)
    lengths = [20, 60, 40]
    cache = serve._hnc.StructureCache()
    alone = []
    for r, n in zip(reqs, lengths):
        m = serve.make_model(r, precision, cache)
        taken, rows = m.advance(n, dt_every=20, probe=r.mode)
        assert taken == n
        alone.append((m.get_state(), rows, m.time))
        m.close()
    ms = [serve.make_model(r, precision, cache) for r in reqs]
    b = Batch(ms, 20, dt_every=20, probe=(1, 0, 0), ring=60)
    b.set_stops(lengths)
    b.launch(3)
    b.sync()
    for i, (m, n) in enumerate(zip(ms, lengths)):
        (na, ma, pa), rows, t = alone[i]
        nb, mb, pb = m.get_state()
        assert np.array_equal(na, nb) and np.array_equal(ma, mb), i
        if pa is not None:
            assert np.array_equal(pa, pb), i
        assert np.array_equal(rows, b.probe(i)[:n]), i
        assert m.time == t, i
    with pytest.raises(abi.HydftError):
        b.set_stops([10, 20, 20])
    b.close()
    # the 20-step run continues alone: 20 more steps equal a 40-step run alone
    m = ms[0]
    m.advance(20, dt_every=20)
    ref = serve.make_model(reqs[0], precision, cache)
    ref.advance(40, dt_every=20)
    assert np.array_equal(m.get_state()[0], ref.get_state()[0])
    assert m.time == ref.time
    ref.close()
    for m in ms:
        m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_of_uneven_grids(gpu, precision):
    """Batches on 16^3 and 24^3 grids (and of both closures) advanced by one
    graph launch per pass (hy_group_*): every run ends bitwise where it ends
    alone; step limits set on a member batch hold; the batches work on their
    own again once the group is closed."""
    from hydft_amd.model import Group
    cache = serve._hnc.StructureCache()
    sets = [_reqs(3, N=16, closure="newtonian", seed=330), _reqs(2, N=24, closure="maxwell", seed=331),
            _reqs(2, N=24, closure="newtonian", seed=332)]
    lengths = [[40, 40, 20], [40, 40], [40, 40]]
    alone = [[None] * len(s) for s in sets]
    for j, s in enumerate(sets):
        for i, r in enumerate(s):
            m = serve.make_model(r, precision, cache)
            taken, rows = m.advance(lengths[j][i], dt_every=20, probe=r.mode)
            assert taken == lengths[j][i]
            alone[j][i] = (m.get_state(), rows, m.time)
            m.close()
    models = [[serve.make_model(r, precision, cache) for r in s] for s in sets]
    batches = [Batch(ms, 20, dt_every=20, probe=(1, 0, 0), ring=40) for ms in models]
    batches[0].set_stops(lengths[0])
    g = Group(batches)
    g.launch(2)
    g.sync()
    assert g.idle()
    for j, ms in enumerate(models):
        for i, m in enumerate(ms):
            (na, ma, pa), rows, t = alone[j][i]
            nb, mb, pb = m.get_state()
            assert np.array_equal(na, nb) and np.array_equal(ma, mb), (j, i)
            if pa is not None:
                assert np.array_equal(pa, pb), (j, i)
            assert np.array_equal(rows, batches[j].probe(i)[:lengths[j][i]]), (j, i)
            assert m.time == t, (j, i)
    g.close()
    # a member batch launched alone afterwards: 20 more steps equal 60 alone
    batches[2].launch(1)
    batches[2].sync()
    ref = serve.make_model(sets[2][0], precision, cache)
    ref.advance(60, dt_every=20)
    assert np.array_equal(models[2][0].get_state()[0], ref.get_state()[0])
    ref.close()
    for b in batches:
        b.close()
    for ms in models:
        for m in ms:
            m.close()


@pytest.mark.gpu
def test_group_rejects_a_batch_twice(gpu):
    from hydft_amd.model import Group
    cache = serve._hnc.StructureCache()
    ms = [serve.make_model(r, "fp64", cache) for r in _reqs(2, closure="newtonian", seed=333)]
    b = Batch(ms, 20, dt_every=20)
    with pytest.raises(abi.HydftError) as e:
        Group([b, b])
    assert e.value.code == abi.HY_ERR_INVALID
    b.close()
    for m in ms:
        m.close()
