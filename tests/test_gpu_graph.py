"""Graph mode (each step one hipGraph launch) gives the eager path's results,
step after step, for single simulations and for ensembles; re-binning between
replays forces a new capture."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(pic, prec, cells=64, n=40000, seed=11, mcc=True):
    from pic_cuda_amd import configs
    return configs.build(cells, n, prec, mcc=mcc, seed=seed, gas_density=1e22, te=8.0)


def _same_state(pic, a, b, where):
    for f in (pic.FIELD_RHS, pic.FIELD_PHI):
        assert np.array_equal(a["grid"].download(f), b["grid"].download(f)), where
    for s in range(2):
        assert np.array_equal(a["sim"].counters(s), b["sim"].counters(s)), where


def _sorted(x, y, vx, vy, ids):
    o = np.argsort(ids, kind="stable")
    return [v[o] for v in (x, y, vx, vy, ids)]


@pytest.mark.parametrize("prec", [4, 8])
def test_graph_steps_equal_eager_steps(pic, prec):
    a = _case(pic, prec)
    b = _case(pic, prec)
    b["sim"].use_graph(True)
    for k in range(14):
        if k == 7:
            a["sim"].rebin(8, 8)
            b["sim"].rebin(8, 8)
        a["sim"].step(1)
        b["sim"].step(1)
        _same_state(pic, a, b, "step %d" % k)
    rep = b["sim"].stats()
    assert rep.graph_steps >= 12 and rep.graph_captures >= 2, (rep.graph_steps, rep.graph_captures)
    assert a["sim"].stats().graph_steps == 0
    a["sim"].sync()
    b["sim"].sync()
    for s in range(2):
        for u, v in zip(_sorted(*a["particles"].download(s)), _sorted(*b["particles"].download(s))):
            assert np.array_equal(u, v)


def test_graph_ensemble_equals_eager_ensemble(pic):
    sizes = [(64, 30000), (96, 50000), (64, 20000)]
    ea, eb = pic.Ensemble(), pic.Ensemble()
    ca = [_case(pic, 4, cells=c, n=n, seed=20 + i) for i, (c, n) in enumerate(sizes)]
    cb = [_case(pic, 4, cells=c, n=n, seed=20 + i) for i, (c, n) in enumerate(sizes)]
    for c in ca:
        ea.add(c["sim"])
    for c in cb:
        eb.add(c["sim"])
    eb.use_graph(True)
    for k in range(10):
        ea.step(1)
        eb.step(1)
        for x, y in zip(ca, cb):
            _same_state(pic, x, y, "ensemble step %d" % k)
    ea.close()
    eb.close()
