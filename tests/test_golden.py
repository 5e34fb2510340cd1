"""Golden fixtures of a tiny full run (tests/golden/make_golden.py).

CPU: the oracle reproduces the committed fixture bit for bit.
GPU: the HIP path, given the same inputs (seeded initial state, tabulated cross
sections, FP64 multigrid), reproduces the fixture bit for bit in FP32 and FP64.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

KEYS = ("x", "y", "vx", "vy", "id")


def _load(prec):
    return np.load(os.path.join(HERE, "golden", "tiny_step_fp%d.npz" % (8 * prec)))


@pytest.mark.parametrize("prec", [4, 8])
def test_oracle_reproduces_golden(orc, prec):
    inp = make_golden.case_inputs(prec)
    o, rhs = make_golden.run_oracle(inp)
    got = make_golden.pack(o, rhs, inp)
    ref = _load(prec)
    for k in ref.files:
        assert np.array_equal(got[k], ref[k]), k


def test_golden_fixture_sanity():
    ref = _load(4)
    c = ref["s0_counters"]
    c1 = ref["s1_counters"]
    assert c[0] > 0 and c[2] > 0 and c[4] > 0       # absorbed, elastic, ionization
    assert c[6] == c[7] == c[4]                     # one born electron and ion per ionization (none absorbed)
    # particle balance: start - absorbed + born of the species that survived its birth pass
    assert len(ref["s0_id"]) == 300 - c[0] + c[6]
    assert len(ref["s1_id"]) == 300 - c1[0] + c[7]
    assert len(np.unique(ref["s0_id"])) == len(ref["s0_id"])


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bin_w", [(4, 16), (8, 16), (4, 4)])
def test_gpu_reproduces_golden(pic, prec, bin_w):
    inp = make_golden.case_inputs(prec)
    ref = _load(prec)
    nn, n = inp["nn"], inp["n"]
    L = (nn - 1) * inp["dx"]
    grid = pic.Grid2D(nn, nn, L, L)
    parts = pic.Particles2D(4 * n, prec)
    parts.change_species(0, make_golden.ME, -make_golden.QE)
    parts.add_species(make_golden.M_AR, make_golden.QE)
    w = 1e15 * L * L / n
    parts.set_weight(0, w)
    parts.set_weight(1, w)
    for s in range(2):
        parts.init_maxwellian(s, n, nn, nn, inp["vth"][s], inp["seed"], 0, 0)
    walls = pic.Boundaries2D(L, L, nn, nn)
    walls.add_box()
    walls.compile_walls()
    xs = pic.CrossSections(1e22, 0.03, points=inp["mcc"][0]["K"], emax_eV=200.0)
    K = inp["mcc"][0]["K"]
    E = np.arange(K) * (200.0 / (K - 1))
    gas = 1e22
    for s, m in enumerate(inp["mcc"]):
        tab = m["table"].reshape(m["n_proc"], K)
        for q in range(m["n_proc"]):
            xs.add_process(s, m["type"][q], m["threshold"][q], E, tab[q] / gas, m["product_species"])
    mg = pic.mg_params(pic.PRESET_STEP, mode=pic.MG_FP64)
    sim = pic.Simulation(parts, grid, walls, xs, pic.sim_params(dt=inp["dt"], seed=inp["seed"], bin_w=bin_w,
                                                                 bin_h=bin_w, mg=mg))
    for s in range(2):
        info = sim.mcc_info(s)
        assert info["nu_max"] == float(ref["s%d_nu_max" % s])
        assert info["p_null"] == float(ref["s%d_p_null" % s])
        assert np.array_equal(np.array(info["band_nu_max"]), ref["s%d_band_nu_max" % s])
        assert np.array_equal(np.array(info["band_p_null"]), ref["s%d_band_p_null" % s])
    sim.step(8)
    assert np.array_equal(grid.download(pic.FIELD_RHS), ref["rhs"])
    assert np.array_equal(grid.download(pic.FIELD_PHI), ref["phi"])
    sim.sync()
    dt = np.float32 if prec == 4 else np.float64
    for s in range(2):
        x, y, vx, vy, ids = parts.download(s)
        x, y, vx, vy = (a.astype(dt) for a in (x, y, vx, vy))
        o = np.lexsort((vy, vx, y, x, ids))
        for k, a in zip(KEYS, (x, y, vx, vy, ids)):
            assert np.array_equal(a[o], ref["s%d_%s" % (s, k)]), (s, k)
        assert np.array_equal(sim.counters(s)[:11], ref["s%d_counters" % s][:11])
