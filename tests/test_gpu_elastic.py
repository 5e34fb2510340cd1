"""elasticCollisions (mcc.cu:114-150) drop-in: pic_elastic_collisions on a
particle store that belongs to no simulation, one null-collision elastic pass
keyed on (seed, species, id, step), bit-identical to the oracle's pass with the
constants the library builds (pic_xs_build), in FP32 and FP64."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE, ME = 1.602176634e-19, 9.1093837015e-31


def _table():
    # a momentum-transfer-like shape: rises to a peak near 12 eV, then falls
    e = np.concatenate([[0.0], np.geomspace(0.01, 100.0, 120)])
    s = 1e-19 * (0.3 + np.exp(-((np.log(e + 1e-3) - np.log(12.0)) ** 2) / 0.8))
    return e, s


@pytest.mark.parametrize("prec", [4, 8])
def test_elastic_collisions_bitwise_vs_oracle(pic, orc, prec):
    n, nx = 40000, 65
    density, dt, seed, step = 1e21, 3e-9, 7, 11
    rng = np.random.default_rng(23)
    x = rng.uniform(1.0, nx - 2.0, n)
    y = rng.uniform(1.0, nx - 2.0, n)
    en = rng.uniform(0.5, 90.0, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    v = np.sqrt(2 * en * QE / ME)
    vx, vy = v * np.cos(ang), v * np.sin(ang)
    ids = rng.permutation(n).astype(np.uint32)
    parts = pic.Particles2D(n, prec)
    parts.change_species(0, ME, -QE)
    parts.upload(0, x, y, vx, vy, ids)
    e, s = _table()
    pic.elasticCollisions(parts, 0, density, dt, e, s, seed=seed, step=step)

    xs = pic.CrossSections(density, 0.0, points=2048, emax_eV=float(e[-1]))
    xs.add_process(0, pic.ELASTIC, 0.0, e, s)
    m = xs.build(0, [ME], dt)
    dt_ = np.float32 if prec == 4 else np.float64
    sp = orc.Species(prec, x.astype(dt_), y.astype(dt_), vx.astype(dt_), vy.astype(dt_), ids, 0.0, mcc=m)
    orc.push(sp, np.zeros((nx, nx)), nx, nx, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), step, seed, 0, 0, box=False)
    gx, gy, gvx, gvy, gid = parts.download(0)
    o, q = np.argsort(gid), np.argsort(sp.id)
    assert np.array_equal(gid[o], sp.id[q])
    for a, b in ((gx, sp.x), (gy, sp.y), (gvx, sp.vx), (gvy, sp.vy)):
        assert np.array_equal(a[o].astype(dt_), b[q])
    hits = int(sp.counters[orc.C_PROC0])
    assert hits > 500
    # elastic: every speed kept, only the collided particles turned
    sp0 = np.hypot(vx.astype(dt_), vy.astype(dt_)).astype(np.float64)[np.argsort(ids)]
    assert np.allclose(np.hypot(gvx[o], gvy[o]), sp0, rtol=3e-6 if prec == 4 else 1e-13)
    turned = int(np.sum(gvx[o].astype(dt_) != vx.astype(dt_)[np.argsort(ids)]))
    assert turned == hits


def test_elastic_collisions_errors_leave_the_store(pic):
    parts = pic.Particles2D(1000, pic.PREC_FP32)
    parts.change_species(0, ME, -QE)
    rng = np.random.default_rng(3)
    parts.upload(0, rng.uniform(1, 30, 100), rng.uniform(1, 30, 100), rng.normal(size=100) * 1e6,
                 rng.normal(size=100) * 1e6)
    before = parts.download(0)
    e, s = _table()
    with pytest.raises(pic.PicError):
        pic.elasticCollisions(parts, 0, 1e21, 1e-9, e[:1], s[:1])     # table shorter than 2
    with pytest.raises(pic.PicError):
        pic.elasticCollisions(parts, 3, 1e21, 1e-9, e, s)             # no such species
    after = parts.download(0)
    for a, b in zip(before, after):
        assert np.array_equal(np.sort(a), np.sort(b))
    # a later valid call still works on the unchanged store
    pic.elasticCollisions(parts, 0, 1e21, 1e-9, e, s, seed=2, step=1)
    assert parts.count(0) == 100
