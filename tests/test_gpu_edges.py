"""Edge cases of the step against the oracle: empty species, particles exactly
on the absorbing box edges, the smallest grids, a dense clump that overflows
its bins into the spill region with re-binning in between, and capacity
errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME = 9.1093837015e-31
M_AR = 39.948 * 1.66053906660e-27
QE = 1.602176634e-19


def _sorted(x, y, vx, vy, ids):
    o = np.lexsort((vy, vx, y, x, ids))
    return x[o], y[o], vx[o], vy[o], ids[o]


def _build(pic, cells, e_state, i_state, prec=4, bin_w=4, mcc=False, dt=1e-10, dx=1e-4, w=1e3, rebin_every=0,
           cap=None):
    """A simulation of uploaded particles (cell units) and the oracle twin."""
    L = cells * dx
    grid = pic.Grid2D(cells + 1, cells + 1, L, L)
    cap = cap or 4 * max(len(e_state[0]), len(i_state[0]), 1) + 8192
    parts = pic.Particles2D(cap, prec)
    parts.change_species(0, ME, -QE)
    ion = parts.add_species(M_AR, QE)
    parts.set_weight(0, w)
    parts.set_weight(ion, w)
    for s, st in enumerate((e_state, i_state)):
        if len(st[0]):
            parts.upload(s, *st)
    bnd = pic.Boundaries2D(L, L, cells + 1, cells + 1)
    bnd.add_box()
    bnd.compile_walls()
    xs = None
    if mcc:
        from pic_cuda_amd import configs
        xs = pic.CrossSections(1e22, 0.03)
        configs.add_argon_processes(xs, 0, ion)
    mg = pic.mg_params(pic.PRESET_STEP, mode=pic.MG_FP64)
    sim = pic.Simulation(parts, grid, bnd, xs, pic.sim_params(dt=dt, bin_w=bin_w, bin_h=bin_w, mcc=1 if mcc else 0,
                                                               mg=mg, rebin_every=rebin_every))
    return dict(grid=grid, particles=parts, walls=bnd, xs=xs, sim=sim, mg=mg, dt=dt, w=w)


def _oracle(pic, orc, c, prec):
    sim, parts, grid = c["sim"], c["particles"], c["grid"]
    dt = np.float32 if prec == 4 else np.float64
    sp = []
    for s, m in enumerate((ME, M_AR)):
        x, y, vx, vy, ids = parts.download(s)
        info = sim.mcc_info(s) if c["xs"] is not None else None
        sp.append(orc.Species(prec, x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids,
                              (-QE if s == 0 else QE) * c["dt"] / m,
                              mcc=info if info and info["n_proc"] else None,
                              other=info["product_species"] if info else -1))
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * c["w"] * qs / (grid.dx * grid.dy)) / pic.EPS0 for q in (-QE, QE)]
    return orc.OracleSim(sp, grid.m, grid.n, grid.dx, grid.dy, c["dt"], cfac, dict(tol=1e-8, max_cycles=50))


def _run_against_oracle(pic, orc, c, prec, passes):
    sim, grid = c["sim"], c["grid"]
    o = _oracle(pic, orc, c, prec)
    for k in range(passes):
        o.phi = grid.download(pic.FIELD_PHI)
        sim.step(1)
        o.step(1, solve=False)
        assert np.array_equal(grid.download(pic.FIELD_RHS), o.rhs), "charge differs at pass %d" % k
        for s in range(2):
            assert np.array_equal(sim.counters(s)[:11], o.sp[s].counters[:11]), (k, s)
    o.phi = grid.download(pic.FIELD_PHI)
    sim.sync()
    o.sync()
    dt = np.float32 if prec == 4 else np.float64
    for s in range(2):
        x, y, vx, vy, ids = c["particles"].download(s)
        g = _sorted(x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids)
        r = _sorted(o.sp[s].x, o.sp[s].y, o.sp[s].vx, o.sp[s].vy, o.sp[s].id)
        assert len(g[0]) == len(r[0])
        for u, v in zip(g, r):
            assert np.array_equal(u, v)
    return o


def _state(x, y, vx, vy):
    n = len(x)
    return (np.asarray(x, float), np.asarray(y, float), np.asarray(vx, float), np.asarray(vy, float),
            np.arange(n, dtype=np.uint32))


def test_empty_simulation_and_empty_species(pic, orc):
    empty = _state([], [], [], [])
    c = _build(pic, 16, empty, empty)
    c["sim"].step(3)
    assert c["sim"].stats().alive == 0
    assert not np.any(c["grid"].download(pic.FIELD_RHS)) and not np.any(c["grid"].download(pic.FIELD_PHI))
    for s in range(2):
        assert len(c["particles"].download(s)[0]) == 0
        assert c["particles"].export_positions(s)[1] == 0
    d = c["sim"].diagnostics()
    assert d["count"] == [0.0, 0.0] and d["kinetic"] == [0.0, 0.0] and d["px"] == [0.0, 0.0]
    assert d["field_energy"] == 0.0 and d["phi_l2"] == 0.0
    # electrons only, ions empty: the oracle agrees pass by pass
    rng = np.random.default_rng(1)
    n = 3000
    e = _state(rng.uniform(0.5, 15.5, n), rng.uniform(0.5, 15.5, n), rng.normal(0, 3e5, n), rng.normal(0, 3e5, n))
    _run_against_oracle(pic, orc, _build(pic, 16, e, empty, prec=8), 8, passes=4)


@pytest.mark.parametrize("prec", [4, 8])
def test_particles_on_the_box_edges(pic, orc, prec):
    """Positions exactly on the edges, corners and one ulp inside, still and
    moving: the box test absorbs x <= 0 or x >= nx - 1 after the drift."""
    cells = 16
    eps = np.finfo(np.float32 if prec == 4 else np.float64).eps * cells
    xs = [0.0, cells, 0.0, cells, 0.0, cells, eps, cells - eps, 8.0, 8.0, 1e-3, cells - 1e-3]
    ys = [8.0, 8.0, 0.0, cells, cells, 0.0, 8.0, 8.0, 0.0, cells, 8.0, 8.0]
    vx = [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -2e4, 2e4]
    vy = [0] * len(xs)
    rng = np.random.default_rng(5)
    n = 2000
    e = _state(np.r_[xs, rng.uniform(0, cells, n)], np.r_[ys, rng.uniform(0, cells, n)],
               np.r_[vx, rng.normal(0, 2e5, n)], np.r_[vy, rng.normal(0, 2e5, n)])
    i = _state(rng.uniform(0, cells, n), rng.uniform(0, cells, n), np.zeros(n), np.zeros(n))
    c = _build(pic, cells, e, i, prec=prec)
    _run_against_oracle(pic, orc, c, prec, passes=3)
    assert c["sim"].counters(0)[pic.CNT["absorbed"]] >= 10


@pytest.mark.parametrize("cells,bin_w", [(2, 2), (4, 2), (8, 4)])
def test_smallest_grids(pic, orc, cells, bin_w):
    rng = np.random.default_rng(cells)
    n = 500
    e = _state(rng.uniform(0, cells, n), rng.uniform(0, cells, n), rng.normal(0, 1e4, n), rng.normal(0, 1e4, n))
    i = _state(rng.uniform(0, cells, n), rng.uniform(0, cells, n), np.zeros(n), np.zeros(n))
    c = _build(pic, cells, e, i, prec=8, bin_w=bin_w)
    _run_against_oracle(pic, orc, c, 8, passes=4)


def test_dense_clump_spills_and_rebins(pic, orc):
    """All electrons start in one cell and fly apart: the bins around the clump
    overflow into the spill region, and re-binning every three steps moves
    them back into bins; every pass still equals the oracle."""
    cells = 32
    rng = np.random.default_rng(9)
    n = 40000
    e = _state(rng.uniform(16.0, 17.0, n), rng.uniform(16.0, 17.0, n), rng.normal(0, 2e5, n),
               rng.normal(0, 2e5, n))
    i = _state(rng.uniform(0, cells, 2000), rng.uniform(0, cells, 2000), np.zeros(2000), np.zeros(2000))
    c = _build(pic, cells, e, i, prec=4, bin_w=4, w=1.0, rebin_every=3)
    _run_against_oracle(pic, orc, c, 4, passes=8)
    assert c["sim"].counters(0)[pic.CNT["spilled"]] > 1000
    assert c["sim"].stats().rebins >= 2
    assert c["sim"].counters(0)[pic.CNT["lost"]] == 0


@pytest.mark.parametrize("cells,bin_w", [(256, 3), (384, 4)])
def test_rebin_with_many_bins(pic, orc, cells, bin_w):
    """Re-binning counts bins in LDS up to 8192 bins and falls back to
    wavefront-aggregated kernels above: 86^2 = 7396 bins of 3 x 3 cells and
    96^2 = 9216 bins of 4 x 4 cells, a clump that spills and is re-binned
    every two steps, every pass equal to the oracle."""
    rng = np.random.default_rng(cells)
    n = 30000
    c0 = cells / 2
    e = _state(rng.uniform(c0, c0 + 2.0, n), rng.uniform(c0, c0 + 2.0, n), rng.normal(0, 2e5, n),
               rng.normal(0, 2e5, n))
    i = _state(rng.uniform(0, cells, 5000), rng.uniform(0, cells, 5000), np.zeros(5000), np.zeros(5000))
    c = _build(pic, cells, e, i, prec=4, bin_w=bin_w, w=1.0, rebin_every=2)
    _run_against_oracle(pic, orc, c, 4, passes=6)
    assert c["sim"].stats().rebins >= 2
    assert c["sim"].counters(0)[pic.CNT["lost"]] == 0


def test_capacity_errors(pic):
    """A store holds max + max/2 + max/8 + 2^20 slots per species; beyond them
    init and upload fail with a capacity error and leave the store as it was."""
    m = 1000
    slots = m + m // 2 + m // 8 + (1 << 20)
    p = pic.Particles2D(m, pic.PREC_FP32)
    with pytest.raises(pic.PicError, match="capacity"):
        p.init_maxwellian(0, slots + 1, 17, 17, 1e5)
    z = np.zeros(slots + 1)
    with pytest.raises(pic.PicError, match="capacity"):
        p.upload(0, z, z, z, z)
    p.init_maxwellian(0, slots - 10, 17, 17, 1e5)
    with pytest.raises(pic.PicError, match="capacity"):
        p.upload(0, np.ones(11), np.ones(11), np.zeros(11), np.zeros(11))
    p.upload(0, np.ones(10), np.ones(10), np.zeros(10), np.zeros(10))
    assert p.count(0) == slots
    p.close()


def test_spill_overflow_is_an_error(pic):
    """2e6 electrons in the first column of 16-cell bins all jump 40 cells in
    one step, into bins that were empty at the last re-binning: their back
    parts hold 20 particles each, and the spill region (the store's capacity
    beyond the bins: N/8 + 2^20 slots plus what the layout left) takes about
    1.3e6, so the rest find no slot.  They are counted as lost and the step
    reports PIC_ERR_CAPACITY instead of going on with less charge."""
    cells, n = 64, 2000000
    rng = np.random.default_rng(12)
    dx, dt = 1e-4, 1e-10
    e = (rng.uniform(1, 15, n), rng.uniform(1, 63, n), np.full(n, 40 * dx / dt), np.zeros(n),
         np.arange(n, dtype=np.uint32))
    i = (rng.uniform(1, 63, 1000), rng.uniform(1, 63, 1000), np.zeros(1000), np.zeros(1000),
         np.arange(1000, dtype=np.uint32))
    c = _build(pic, cells, e, i, bin_w=16, dt=dt, dx=dx, cap=n)
    with pytest.raises(pic.PicError, match="spill region full"):
        c["sim"].step(1)
    lost = int(c["sim"].counters(0)[pic.CNT["lost"]])
    assert lost > 100000
    assert c["sim"].counters(1)[pic.CNT["lost"]] == 0
