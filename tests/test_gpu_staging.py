"""The public device arrays of the reference's Particles2D (r, r_old, v, type,
particles.cuh:7-20), filled the way main.cu:232-238 fills them and handed over
by `last_alive = n` (main.cu:238): the staged entries become the particles of
their species, bit for bit the same values an upload of the same data gives."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _parts(pic, prec, cap):
    from pic_cuda_amd import configs
    p = pic.Particles2D(cap, prec)
    p.change_species(0, configs.ME, -configs.QE)
    p.add_species(configs.M_AR, configs.QE)
    return p


def _grid(pic, cells):
    from pic_cuda_amd import configs
    pl = configs.Plasma(te=2.0, cells=cells)
    return pic.Grid2D(pl.nodes, pl.nodes, pl.length, pl.length), pl


def _sim(pic, p, grid, pl):
    bnd = pic.Boundaries2D(pl.length, pl.length, grid.m, grid.n)
    bnd.add_box()
    bnd.compile_walls()
    sim = pic.Simulation(p, grid, bnd, None, pic.sim_params(dt=pl.dt, seed=3))
    return sim, bnd


def _staged_batch(rng, n, L, vth):
    r = rng.uniform(0.02 * L, 0.98 * L, size=(n, 2))
    v = rng.normal(size=(n, 2)) * vth
    t = rng.integers(0, 2, size=n).astype(np.int32)
    return r, v, t


def _expect(r, v, t, s, dx, dy, np_t, id0):
    sel = np.nonzero(t == s)[0]
    return (np_t(r[sel, 0] / dx), np_t(r[sel, 1] / dy), np_t(v[sel, 0]), np_t(v[sel, 1]),
            (id0 + sel).astype(np.uint32))


def _species(p, s, np_t):
    x, y, vx, vy, ids = p.download(s)
    o = np.argsort(ids)
    return np_t(x[o]), np_t(y[o]), np_t(vx[o]), np_t(vy[o]), ids[o]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_staged_before_and_during_a_simulation(pic, prec):
    prec_v = pic.PREC_FP32 if prec == "fp32" else pic.PREC_FP64
    np_t = np.float32 if prec == "fp32" else np.float64
    grid, pl = _grid(pic, 64)
    dx, dy, L = grid.dx, grid.dy, pl.length
    rng = np.random.default_rng(11)
    p = _parts(pic, prec_v, 60000)
    n1 = 50000
    r, v, t = _staged_batch(rng, n1, L, pl.vte)
    a = p.staging(n1, r_old=True)
    assert a["r"] and a["v"] and a["type"] and a["r_old"]
    # main.cu:232-238: positions into r and r_old, velocities into v, then last_alive
    pic.host_to_device(a["r"], r)
    pic.host_to_device(a["r_old"], r)
    pic.host_to_device(a["v"], v)
    pic.host_to_device(a["type"], t)
    p.set_staged(n1)
    assert p.last_alive == 0              # handed over when a simulation takes them
    sim, bnd = _sim(pic, p, grid, pl)
    for s in range(2):
        got = _species(p, s, np_t)
        exp = _expect(r, v, t, s, dx, dy, np_t, 0)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)
    # the type array is zeroed again: a second fill without types is species 0
    assert not np.any(pic.device_to_host(a["type"], n1, np.int32))
    sim.step(3)
    # a second batch into the running simulation is handed over at once
    n2 = 4000
    r2, v2, _ = _staged_batch(rng, n2, L, pl.vte)
    a2 = p.staging(n2)
    assert a2["r"] == a["r"]              # no reallocation for a smaller batch
    pic.host_to_device(a2["r"], r2)
    pic.host_to_device(a2["v"], v2)
    alive0 = p.count(0)
    p.set_staged(n2)
    assert p.count(0) == alive0 + n2
    x, y, vx, vy, ids = p.download(0)
    new = ids >= n1
    o = np.argsort(ids[new])
    assert np.array_equal(ids[new][o], np.arange(n1, n1 + n2, dtype=np.uint32))
    assert np.array_equal(np_t(x[new][o]), np_t(r2[:, 0] / dx))
    assert np.array_equal(np_t(vy[new][o]), np_t(v2[:, 1]))
    sim.step(2)
    st = sim.stats()
    assert st.alive > 0


def test_staged_type_out_of_range(pic):
    grid, pl = _grid(pic, 32)
    p = _parts(pic, pic.PREC_FP32, 1000)
    a = p.staging(100)
    pic.host_to_device(a["r"], np.full((100, 2), 0.5 * pl.length))
    pic.host_to_device(a["v"], np.zeros((100, 2)))
    t = np.zeros(100, np.int32)
    t[7] = 5
    pic.host_to_device(a["type"], t)
    with pytest.raises(pic.PicError):
        p.commit_staging(100, grid.dx, grid.dy)
    assert p.last_alive == 0
    with pytest.raises(pic.PicError):
        p.set_staged(101)                 # more than the staging area holds
