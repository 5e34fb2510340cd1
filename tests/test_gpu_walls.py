"""Internal wall segments (detect_intersections, collisions.cu:7-54) with the
per-bin candidate lists: every pass bit-identical to the oracle, which tests
every wall for every particle; identical to the GPU's own full scan; lists
rebuilt on re-binning."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sorted(x, y, vx, vy, ids):
    o = np.lexsort((vy, vx, y, x, ids))
    return x[o], y[o], vx[o], vy[o], ids[o]


def _walls(cells, seed):
    from pic_cuda_amd import configs
    w = configs.random_walls(cells, 40, seed=seed, max_len=20.0)
    # long segments across many bins, one of them on a bin edge
    extra = np.array([[0.5, 3.0, cells - 0.5, cells - 7.0],
                      [16.0, 1.0, 16.0, cells - 1.0],
                      [2.0, cells / 2 + 0.25, cells - 2.0, cells / 2 + 0.25]])
    return np.concatenate([w, extra])


def _case(pic, prec, walls, cells=48, n=30000, bin_w=8, seed=7):
    from pic_cuda_amd import configs
    return configs.build(cells, n, prec, mcc=True, seed=seed, gas_density=1e22, bin_w=bin_w, bin_h=bin_w, te=8.0,
                         walls=walls)


def _oracle_walls(case, walls):
    """The cell-unit walls the library computes: domain units / dx."""
    g = case["grid"]
    w = np.array(walls, dtype=np.float64)
    dom = w * np.array([g.dx, g.dy, g.dx, g.dy])
    return dom / np.array([g.dx, g.dy, g.dx, g.dy])


def _oracle(pic, orc, case, prec, walls, seed=7):
    sim, parts, grid = case["sim"], case["particles"], case["grid"]
    sp = []
    dt = np.float32 if prec == 4 else np.float64
    for s in range(parts.species_num):
        x, y, vx, vy, ids = parts.download(s)
        info = sim.mcc_info(s)
        kq = pic.ELEMENTARY_CHARGE * (-1 if s == 0 else 1) * case["plasma"].dt / (
            pic.ELECTRON_MASS if s == 0 else 39.948 * pic.AMU)
        sp.append(orc.Species(prec, x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids, kq,
                              mcc=info if info["n_proc"] else None, other=info["product_species"]))
    pl = case["plasma"]
    w = pl.density * pl.length * pl.length / len(sp[0].x)
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * w * qs / (grid.dx * grid.dy)) / pic.EPS0 for q in (-pic.ELEMENTARY_CHARGE, pic.ELEMENTARY_CHARGE)]
    mg = case["mg"]
    return orc.OracleSim(sp, grid.m, grid.n, grid.dx, grid.dy, pl.dt, cfac,
                         dict(tol=mg.tol, max_cycles=mg.max_cycles, pre=mg.pre, post=mg.post,
                              coarse_sweeps=mg.coarse_sweeps, omega=mg.omega), seed=seed,
                         walls=_oracle_walls(case, walls))


@pytest.mark.parametrize("prec", [4, 8])
def test_internal_walls_passes_bit_identical(pic, orc, prec):
    walls = _walls(48, seed=5)
    case = _case(pic, prec, walls)
    sim, grid = case["sim"], case["grid"]
    nb, entries, longest = sim.wall_list_stats()
    assert nb == 36 and 0 < longest < len(walls)
    assert entries < nb * len(walls) // 2          # the lists cut the tests per particle
    o = _oracle(pic, orc, case, prec, walls)
    absorbed = 0
    for k in range(5):
        o.phi = grid.download(pic.FIELD_PHI)
        sim.step(1)
        o.step(1, solve=False)
        assert np.array_equal(grid.download(pic.FIELD_RHS), o.rhs), "charge differs at pass %d" % k
        for s in range(2):
            c = sim.counters(s)
            assert np.array_equal(c[:11], o.sp[s].counters[:11]), (k, s, c, o.sp[s].counters)
    o.phi = grid.download(pic.FIELD_PHI)
    sim.sync()
    o.sync()
    dt = np.float32 if prec == 4 else np.float64
    for s in range(2):
        x, y, vx, vy, ids = case["particles"].download(s)
        g = _sorted(x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids)
        r = _sorted(o.sp[s].x, o.sp[s].y, o.sp[s].vx, o.sp[s].vy, o.sp[s].id)
        for u, v in zip(g, r):
            assert np.array_equal(u, v)
        absorbed += int(sim.counters(s)[pic.CNT["absorbed"]])
    assert absorbed > 0


def _absorbed(pic, case, steps):
    if steps:
        case["sim"].step(steps)
    return sum(int(case["sim"].counters(s)[pic.CNT["absorbed"]]) for s in range(2))


def test_lists_equal_full_scan_and_survive_rebinning(pic):
    walls = _walls(48, seed=11)
    a = _case(pic, 4, walls)
    b = _case(pic, 4, walls)
    b["sim"].set_wall_lists(False)
    assert b["sim"].wall_list_stats() == (0, 0, 0)
    c = _case(pic, 4, walls, bin_w=16)
    a["sim"].step(3)
    b["sim"].step(3)
    c["sim"].step(2)
    c["sim"].rebin(4, 4)                 # lists rebuilt for 12 x 12 bins
    assert c["sim"].wall_list_stats()[0] == 144
    c["sim"].step(1)
    for s in range(2):
        assert np.array_equal(a["sim"].counters(s)[:11], b["sim"].counters(s)[:11])
        assert np.array_equal(a["sim"].counters(s)[:11], c["sim"].counters(s)[:11])
    for f in (pic.FIELD_RHS, pic.FIELD_PHI):
        assert np.array_equal(a["grid"].download(f), b["grid"].download(f))
        assert np.array_equal(a["grid"].download(f), c["grid"].download(f))
    # the internal walls absorb more than the box alone
    assert _absorbed(pic, a, 0) > _absorbed(pic, _case(pic, 4, None), 3)


def test_fast_paths_take_every_route(pic, orc):
    """Paths of 0.05 to 40 cells per step over 4 x 4-cell bins with margin 0:
    paths inside their bin (one list), across bin edges (the lists of up to 16
    bins) and longer ones (every wall) -- each pass bit-identical to the
    oracle and to the GPU without lists."""
    cells, dt, N = 64, 1e-10, 20000
    walls = _walls(cells, seed=17)

    def make(lists):
        grid = pic.Grid2D(cells + 1, cells + 1, cells * 1e-4, cells * 1e-4)
        dx = grid.dx
        parts = pic.Particles2D(3 * N, pic.PREC_FP64)
        parts.change_species(0, pic.ELECTRON_MASS, -pic.ELEMENTARY_CHARGE)
        ion = parts.add_species(39.948 * pic.AMU, pic.ELEMENTARY_CHARGE)
        parts.set_weight(0, 1.0)
        parts.set_weight(ion, 1.0)
        rng = np.random.default_rng(21)
        d = np.exp(rng.uniform(np.log(0.05), np.log(40.0), N))
        th = rng.uniform(0, 2 * np.pi, N)
        parts.upload(0, rng.uniform(0.5, cells - 0.5, N), rng.uniform(0.5, cells - 0.5, N),
                     d * np.cos(th) * dx / dt, d * np.sin(th) * dx / dt, ids=np.arange(N))
        parts.upload(ion, rng.uniform(0.5, cells - 0.5, 1000), rng.uniform(0.5, cells - 0.5, 1000),
                     np.zeros(1000), np.zeros(1000), ids=np.arange(1000))
        bnd = pic.Boundaries2D(cells * 1e-4, cells * 1e-4, cells + 1, cells + 1)
        bnd.add_box()
        for (x1, y1, x2, y2) in walls:
            bnd.add_wall(x1 * grid.dx, y1 * grid.dy, x2 * grid.dx, y2 * grid.dy)
        bnd.compile_walls()
        mg = pic.mg_params(pic.PRESET_STEP, mode=pic.MG_FP64)
        sim = pic.Simulation(parts, grid, bnd, None, pic.sim_params(dt=dt, bin_w=4, bin_h=4, mcc=0, mg=mg))
        sim.set_wall_lists(lists, margin=0)
        return dict(grid=grid, particles=parts, walls=bnd, sim=sim, mg=mg)

    a, b = make(True), make(False)
    assert a["sim"].wall_list_stats()[0] == 256
    grid, sim = a["grid"], a["sim"]
    sp = []
    for s_ in range(2):
        x, y, vx, vy, ids = a["particles"].download(s_)
        m = pic.ELECTRON_MASS if s_ == 0 else 39.948 * pic.AMU
        sp.append(orc.Species(8, x, y, vx, vy, ids, pic.ELEMENTARY_CHARGE * (-1 if s_ == 0 else 1) * dt / m))
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * 1.0 * qs / (grid.dx * grid.dy)) / pic.EPS0 for q in (-pic.ELEMENTARY_CHARGE, pic.ELEMENTARY_CHARGE)]
    o = orc.OracleSim(sp, grid.m, grid.n, grid.dx, grid.dy, dt, cfac, dict(tol=1e-8, max_cycles=50),
                      walls=_oracle_walls(a, walls))
    for k in range(3):
        o.phi = grid.download(pic.FIELD_PHI)
        sim.step(1)
        b["sim"].step(1)
        o.step(1, solve=False)
        assert np.array_equal(grid.download(pic.FIELD_RHS), o.rhs), "charge differs at pass %d" % k
        assert np.array_equal(grid.download(pic.FIELD_RHS), b["grid"].download(pic.FIELD_RHS))
        for s_ in range(2):
            assert np.array_equal(sim.counters(s_)[:11], o.sp[s_].counters[:11]), k
            assert np.array_equal(sim.counters(s_)[:11], b["sim"].counters(s_)[:11]), k
    o.phi = grid.download(pic.FIELD_PHI)
    sim.sync()
    o.sync()
    for s_ in range(2):
        g = _sorted(*a["particles"].download(s_))
        r = _sorted(o.sp[s_].x, o.sp[s_].y, o.sp[s_].vx, o.sp[s_].vy, o.sp[s_].id)
        for u, v in zip(g, r):
            assert np.array_equal(u, v)
    # most of the fast electrons hit a wall or the box
    assert sim.counters(0)[pic.CNT["absorbed"]] > N // 10
