"""Batched ensembles and the job server.

Contract: every member of an ensemble computes bit for bit what it computes
alone (same charge, potential, counters and particle set), whatever the other
members, their sizes and when they joined; the server admits jobs by HBM
footprint and each job's result is the same alone and in the mix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(pic, cells, n, seed, te=4.0, prec=4):
    from pic_cuda_amd import configs
    return configs.build(cells, n, prec, mcc=True, seed=seed, gas_density=1e22, te=te)


def _state(pic, case):
    out = []
    for s in range(2):
        x, y, vx, vy, ids = case["particles"].download(s)
        o = np.lexsort((vy, vx, y, x, ids))
        out.append((x[o], y[o], vx[o], vy[o], ids[o], case["sim"].counters(s)[:11]))
    return out, case["grid"].download(pic.FIELD_PHI)


def _same(a, b):
    (pa, phia), (pb, phib) = a, b
    assert np.array_equal(phia, phib)
    for sa, sb in zip(pa, pb):
        for u, v in zip(sa, sb):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("prec", [4, 8])
def test_ensemble_members_match_alone(pic, prec):
    specs = [(32, 6000, 3, 3.0), (64, 20000, 4, 5.0), (16, 1500, 5, 2.0)]
    alone = []
    for c, n, seed, te in specs:
        k = _case(pic, c, n, seed, te, prec)
        k["sim"].step(5)
        alone.append(_state(pic, k))
    batch = [_case(pic, c, n, seed, te, prec) for c, n, seed, te in specs]
    ens = pic.Ensemble()
    for k in batch:
        ens.add(k["sim"])
    ens.step(5)
    for k, ref in zip(batch, alone):
        _same(_state(pic, k), ref)
    ens.close()


def test_ensemble_with_an_electrode_member_matches_alone(pic):
    """Mixed-mode members without electrodes solve their coarse level exactly
    (sine transform), a member with an electrode keeps the swept cycle; in one
    batch (shared launches, masked kernel variants) each still equals its run
    alone bit for bit."""
    specs = [(32, 6000, 21, 3.0), (64, 20000, 22, 5.0), (64, 15000, 23, 4.0)]

    def electrode(case):
        n = case["grid"].m
        y, x = np.mgrid[0:n, 0:n]
        disc = (x - n / 3) ** 2 + (y - n / 2) ** 2 < (0.12 * n) ** 2
        case["sim"].set_electrodes(disc.astype(np.uint8), np.where(disc, 2.0, 0.0))
    alone = []
    for i, (c, n, seed, te) in enumerate(specs):
        k = _case(pic, c, n, seed, te)
        if i == 1:
            electrode(k)
        k["sim"].step(4)
        alone.append(_state(pic, k))
    batch = [_case(pic, c, n, seed, te) for c, n, seed, te in specs]
    electrode(batch[1])
    ens = pic.Ensemble()
    for k in batch:
        ens.add(k["sim"])
    ens.step(4)
    for k, ref in zip(batch, alone):
        _same(_state(pic, k), ref)
    ens.close()


def test_ensemble_late_join_and_leave(pic):
    a_ref = _case(pic, 32, 5000, 11)
    a_ref["sim"].step(6)
    b_ref = _case(pic, 64, 15000, 12)
    b_ref["sim"].step(3)
    a = _case(pic, 32, 5000, 11)
    b = _case(pic, 64, 15000, 12)
    ens = pic.Ensemble()
    ens.add(a["sim"])
    ens.step(3)
    ens.add(b["sim"])          # joins at its own step 0 while a is at step 3
    ens.step(3)
    _same(_state(pic, a), _state(pic, a_ref))
    _same(_state(pic, b), _state(pic, b_ref))
    ens.remove(a["sim"])
    a["sim"].step(1)           # continues alone after leaving
    a_ref["sim"].step(1)
    _same(_state(pic, a), _state(pic, a_ref))
    ens.close()


def _run_server(pic, jobs, budget=64e9, max_active=64, keep=True):
    import time
    sv = pic.Server(budget, max_active, pic.PREC_FP32, keep_finished=keep)
    ids = [sv.submit(**j) for j in jobs]
    t0 = time.perf_counter()
    for _ in range(100000):
        now = time.perf_counter() - t0
        run, wait = sv.tick(now)
        if run == 0 and wait == 0:
            break
    return sv, ids


def test_server_result_same_alone_and_in_mix(pic):
    target = dict(cells=32, ppc=8.0, steps=12, seed=77, arrival_s=0.0)
    sv1, (j1,) = _run_server(pic, [target])
    others = [dict(cells=64, ppc=6.0, steps=20, seed=5, arrival_s=0.0),
              dict(cells=16, ppc=10.0, steps=3, seed=6, arrival_s=0.0),
              dict(cells=32, ppc=4.0, steps=9, seed=8, arrival_s=0.0)]
    sv2, ids = _run_server(pic, others[:2] + [target] + others[2:])
    j2 = ids[2]
    r1, r2 = sv1.result(j1), sv2.result(j2)
    assert r1.state == r2.state == 2 and r1.steps_done == r2.steps_done == 12
    a = sv1.handles(j1)
    b = sv2.handles(j2)
    assert np.array_equal(a[1].download(pic.FIELD_PHI), b[1].download(pic.FIELD_PHI))
    for s in range(2):
        xa = a[0].download(s)
        xb = b[0].download(s)
        oa = np.lexsort(xa[::-1])
        ob = np.lexsort(xb[::-1])
        for k in range(5):
            assert np.array_equal(xa[k][oa], xb[k][ob])
        assert np.array_equal(a[2].counters(s), b[2].counters(s))
    sv1.close()
    sv2.close()


def test_server_admits_by_footprint(pic):
    # budget fits one job at a time: the second is admitted only after the first retires
    jobs = [dict(cells=32, ppc=8.0, steps=4, seed=1, arrival_s=0.0),
            dict(cells=32, ppc=8.0, steps=4, seed=2, arrival_s=0.0)]
    sv = pic.Server(1.0, 8, pic.PREC_FP32, keep_finished=False)
    with pytest.raises(pic.PicError):
        sv.submit(**jobs[0])                     # larger than the whole budget
    sv.close()
    probe = pic.Server(1e12, 8, pic.PREC_FP32)
    probe.submit(**jobs[0])
    fp = probe.result(0).footprint_bytes
    probe.close()
    sv, ids = _run_server(pic, jobs, budget=1.5 * fp, keep=False)
    r0, r1 = sv.result(ids[0]), sv.result(ids[1])
    assert r0.state == 3 and r1.state == 3
    assert r1.admit_s >= r0.done_s
    assert sv.used_bytes() == 0
    sv.close()
