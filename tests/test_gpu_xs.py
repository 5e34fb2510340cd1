"""Tabulated cross sections from files through the full step: the csf argon
tables (elastic csf/Ar.txt, ionization csf/Ar+.txt) loaded as LXCat blocks,
every pass bit-identical to the oracle, no majorant violations."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _sorted(x, y, vx, vy, ids):
    o = np.lexsort((vy, vx, y, x, ids))
    return x[o], y[o], vx[o], vy[o], ids[o]


def _lxcat(tmp_path):
    rows = []
    for kind, target, param, name in (("ELASTIC", "Ar", "1.360000e-5", "csf_Ar.txt"),
                                      ("IONIZATION", "Ar -> Ar^+", "1.576000e+1", "csf_Ar+.txt")):
        e, s = np.loadtxt(os.path.join(GOLDEN, name), unpack=True)
        rows += [kind, target, param, "COLUMNS: Energy (eV) | Cross section (m2)", "-----"]
        rows += ["%.6e\t%.6e" % (a, b) for a, b in zip(e, s)]
        rows += ["-----", ""]
    p = tmp_path / "ar.lxcat"
    p.write_text("\n".join(rows))
    return str(p)


@pytest.mark.parametrize("prec", [4, 8])
def test_csf_argon_tables_through_the_step(pic, orc, prec, tmp_path):
    from pic_cuda_amd import configs
    path = _lxcat(tmp_path)

    def procs(xs, e, ion):
        assert xs.load_lxcat(path, e, product_species=ion) == (2, 0)
        xs.add_analytic(ion, pic.CHARGE_EXCHANGE, 0.0, 5e-19)

    case = configs.build(32, 20000, prec, mcc=True, seed=7, gas_density=1e22, te=10.0, processes=procs)
    sim, grid, parts = case["sim"], case["grid"], case["particles"]
    info = sim.mcc_info(0)
    assert info["type"] == [pic.ELASTIC, pic.IONIZATION] and info["threshold"] == [0.0, 15.76]
    sp = []
    dt = np.float32 if prec == 4 else np.float64
    for s in range(2):
        x, y, vx, vy, ids = parts.download(s)
        inf = sim.mcc_info(s)
        kq = pic.ELEMENTARY_CHARGE * (-1 if s == 0 else 1) * case["plasma"].dt / (
            pic.ELECTRON_MASS if s == 0 else 39.948 * pic.AMU)
        sp.append(orc.Species(prec, x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids, kq, mcc=inf,
                              other=inf["product_species"]))
    pl = case["plasma"]
    w = pl.density * pl.length * pl.length / len(sp[0].x)
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * w * qs / (grid.dx * grid.dy)) / pic.EPS0 for q in (-pic.ELEMENTARY_CHARGE, pic.ELEMENTARY_CHARGE)]
    o = orc.OracleSim(sp, grid.m, grid.n, grid.dx, grid.dy, pl.dt, cfac, dict(tol=1e-8, max_cycles=50), seed=7)
    for k in range(6):
        o.phi = grid.download(pic.FIELD_PHI)
        sim.step(1)
        o.step(1, solve=False)
        assert np.array_equal(grid.download(pic.FIELD_RHS), o.rhs), k
        for s in range(2):
            assert np.array_equal(sim.counters(s)[:11], o.sp[s].counters[:11]), (k, s)
    o.phi = grid.download(pic.FIELD_PHI)
    sim.sync()
    o.sync()
    for s in range(2):
        x, y, vx, vy, ids = parts.download(s)
        g = _sorted(x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids)
        r = _sorted(o.sp[s].x, o.sp[s].y, o.sp[s].vx, o.sp[s].vy, o.sp[s].id)
        for u, v in zip(g, r):
            assert np.array_equal(u, v)
    c0 = sim.counters(0)
    assert c0[pic.CNT["process0"]] > 0 and c0[pic.CNT["process1"]] > 0     # elastic and ionization happened
    assert c0[pic.CNT["born_other"]] > 0
    assert c0[pic.CNT["numax_exceeded"]] == 0
