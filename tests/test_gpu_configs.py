"""The five workloads of the step at their stated sizes, on the GPU.

* plain: 128^2 cells, 1e5 e- + 1e5 Ar+, FP64, no collisions, 2+2 Gauss-Seidel
  multigrid to 1e-8, 100 steps: counts bit-exact against the oracle, potential
  within 1e-10 relative L2.
* full step: 1024^2 cells, 2e7 e- + 2e7 Ar+, FP32 particles, null-collision
  MCC on Ar at 1e21 m^-3: three passes in lockstep with the oracle (OpenMP
  build), charge and counters bit-identical every pass, particle sets
  bit-identical after the owed half step, exact charge sums, nothing lost.
* Poisson: 2049^2, 200 Gaussian blobs + 1e-3 noise: the FP32-in-FP64 solve
  to 1e-10 within 1e-9 relative L2 of the oracle's FP64 solve; the FP64
  solver bit-identical to the oracle at a fixed cycle count.
* ensemble: the 64-case spec (cells 64..512, 1e4..4e6 particles per species,
  Te 1-5 eV, MCC) batched for a few steps: particle balance and nothing lost
  in every member, three sampled members equal to their runs alone.
* serving: the stream's job distribution (cells 64..1024, 100 particles per
  cell, seeds from job ids) with step counts cut short: every job finishes,
  two sampled jobs equal alone and in the mix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME = 9.1093837015e-31


def _oracle(pic, orc, case, prec, seed, omp=True):
    sim, parts, grid = case["sim"], case["particles"], case["grid"]
    dt = np.float32 if prec == 4 else np.float64
    sp = []
    n0 = None
    for s in range(parts.species_num):
        x, y, vx, vy, ids = parts.download(s)
        n0 = n0 or len(x)
        info = sim.mcc_info(s) if case["xs"] is not None else None
        kq = pic.ELEMENTARY_CHARGE * (-1 if s == 0 else 1) * case["plasma"].dt / (
            pic.ELECTRON_MASS if s == 0 else 39.948 * pic.AMU)
        sp.append(orc.Species(prec, x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), ids, kq,
                              mcc=info if info and info["n_proc"] else None,
                              other=info["product_species"] if info else -1))
        del x, y, vx, vy, ids
    pl = case["plasma"]
    w = pl.density * pl.length * pl.length / n0      # the weight configs.build sets
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * w * qs / (grid.dx * grid.dy)) / pic.EPS0 for q in (-pic.ELEMENTARY_CHARGE, pic.ELEMENTARY_CHARGE)]
    mg = case["mg"]
    return orc.OracleSim(sp, grid.m, grid.n, grid.dx, grid.dy, pl.dt, cfac,
                         dict(tol=mg.tol, max_cycles=mg.max_cycles, pre=mg.pre, post=mg.post,
                              coarse_sweeps=mg.coarse_sweeps, omega=mg.omega), seed=seed, omp=omp)


def _key_order(x, ids):
    """order by (id, bits of x as stored): ids are unique apart from born
    particles, whose ids are random draws"""
    xb = np.ascontiguousarray(x).view(np.uint32 if x.dtype == np.float32 else np.uint64).astype(np.uint64)
    if x.dtype == np.float64:
        xb = xb >> np.uint64(32)
    key = (ids.astype(np.uint64) << np.uint64(32)) | xb
    return np.argsort(key, kind="stable")


def _exact_sum(acc):
    """sum of uint64 values as a Python integer (the total exceeds 2^64)"""
    hi = int((acc >> np.uint64(32)).sum(dtype=np.uint64))
    lo = int((acc & np.uint64(0xFFFFFFFF)).sum(dtype=np.uint64))
    return (hi << 32) + lo


def _same_set(gpu, osp, prec):
    dt = np.float32 if prec == 4 else np.float64
    gx, gy, gvx, gvy, gid = gpu
    gx, gy, gvx, gvy = (a.astype(dt) for a in (gx, gy, gvx, gvy))
    assert len(gx) == osp.n
    og, oo = _key_order(gx, gid), _key_order(osp.x, osp.id)
    assert np.array_equal(gid[og], osp.id[oo])
    for a, b in ((gx, osp.x), (gy, osp.y), (gvx, osp.vx), (gvy, osp.vy)):
        assert np.array_equal(a[og], b[oo])


def test_plain_128sq_1e5_fp64_100_steps(pic, orc):
    from pic_cuda_amd import configs
    case = configs.plain_small()
    sim, grid, parts = case["sim"], case["grid"], case["particles"]
    assert parts.count(0) == parts.count(1) == 100000
    o = _oracle(pic, orc, case, 8, seed=1)
    sim.step(100)
    o.step(100)
    phi_g = grid.download(pic.FIELD_PHI)
    err = np.linalg.norm(phi_g - o.phi) / np.linalg.norm(o.phi)
    assert err < 1e-10, err
    sim.sync()
    o.sync()
    for s in range(2):
        c = sim.counters(s)
        assert c[pic.CNT["absorbed"]] == o.sp[s].counters[0] > 0
        assert parts.count(s) == o.sp[s].n
        assert c[pic.CNT["lost"]] == 0
        # the same particles survive, at the same places to FP64 rounding
        gx, gy, gvx, gvy, gid = parts.download(s)
        og, oo = np.argsort(gid), np.argsort(o.sp[s].id)
        assert np.array_equal(gid[og], o.sp[s].id[oo])
        assert np.max(np.abs(gx[og] - o.sp[s].x[oo])) < 1e-9 * 128


def test_full_step_1024sq_4e7_lockstep_passes(pic, orc):
    from pic_cuda_amd import configs
    case = configs.full_step()
    sim, grid, parts = case["sim"], case["grid"], case["particles"]
    n0 = [parts.count(s) for s in range(2)]
    assert n0 == [20000000, 20000000]
    o = _oracle(pic, orc, case, 4, seed=1)
    for k in range(3):
        o.phi = grid.download(pic.FIELD_PHI)
        sim.step(1)
        o.step(1, solve=False)
        assert np.array_equal(grid.download(pic.FIELD_RHS), o.rhs), "charge differs at pass %d" % k
        for s in range(2):
            c = sim.counters(s)
            assert np.array_equal(c[:11], o.sp[s].counters[:11]), (k, s, c, o.sp[s].counters)
            assert c[pic.CNT["lost"]] == 0
            # every survivor of a pass deposits its whole weight: the charge sums exactly
            assert _exact_sum(o.accs[s]) == o.sp[s].n << 40
    o.phi = grid.download(pic.FIELD_PHI)
    sim.sync()
    o.sync()
    c0 = sim.counters(0)
    # electron elastic, excitation, ionization; ion charge exchange (its process 0)
    assert c0[2] > 0 and c0[3] > 0 and c0[4] > 0 and sim.counters(1)[2] > 0
    for s in range(2):
        _same_set(parts.download(s), o.sp[s], 4)
        assert parts.count(s) == o.sp[s].n
    del o


def test_poisson_2049sq_blobs(pic, orc):
    from pic_cuda_amd import configs
    n = 2049
    rhs = configs.poisson_rhs(n)
    h = 1.0 / (n - 1)
    g = pic.Grid2D(n, n, 1.0, 1.0)
    g.upload(pic.FIELD_RHO, rhs)
    rho_p, phi_p, _ = g.device_ptrs()
    mixed = pic.mg_params(pic.PRESET_POISSON)
    assert mixed.tol == 1e-10
    st = pic.Multigrid(n, mixed).solve(phi_p, rho_p, h)
    assert st.rel_residual < 1e-10
    phi_mixed = g.download(pic.FIELD_PHI)
    p64 = pic.mg_params(pic.PRESET_POISSON, mode=pic.MG_FP64)
    phi_o = np.zeros((n, n))
    so = orc.mg_solve(phi_o, rhs, h, pre=p64.pre, post=p64.post, coarse_sweeps=p64.coarse_sweeps, omega=p64.omega,
                      tol=1e-10, max_cycles=p64.max_cycles, omp=True)
    assert so.rel_residual < 1e-10
    err = np.linalg.norm(phi_mixed - phi_o) / np.linalg.norm(phi_o)
    assert err < 1e-9, err
    # the FP64 solver against the oracle at a fixed cycle count: bit for bit
    g.upload(pic.FIELD_PHI, np.zeros((n, n)))
    fixed = pic.mg_params(pic.PRESET_POISSON, mode=pic.MG_FP64, tol=0.0, max_cycles=2)
    st = pic.Multigrid(n, fixed).solve(phi_p, rho_p, h)
    assert st.cycles == 2
    phi_o = np.zeros((n, n))
    orc.mg_solve(phi_o, rhs, h, pre=p64.pre, post=p64.post, coarse_sweeps=p64.coarse_sweeps, omega=p64.omega,
                 tol=0.0, max_cycles=2, omp=True)
    assert np.array_equal(g.download(pic.FIELD_PHI), phi_o)


def _balance(pic, case, n0):
    """live particles = start - absorbed + born (own born electrons, the other
    species' born ions); nothing lost"""
    sim, parts = case["sim"], case["particles"]
    c = [sim.counters(s) for s in range(2)]
    for s in range(2):
        born = c[s][pic.CNT["born_self"]] + c[1 - s][pic.CNT["born_other"]]
        assert parts.count(s) == n0 - c[s][pic.CNT["absorbed"]] + born, (s, c[s])
        assert c[s][pic.CNT["lost"]] == 0


def _state(pic, case):
    out = []
    for s in range(2):
        x, y, vx, vy, ids = case["particles"].download(s)
        o = _key_order(x.astype(np.float32), ids)
        out.append((x[o], y[o], vx[o], vy[o], ids[o], case["sim"].counters(s)[:11]))
    return out, case["grid"].download(pic.FIELD_PHI)


def _same_state(a, b):
    (pa, phia), (pb, phib) = a, b
    assert np.array_equal(phia, phib)
    for sa, sb in zip(pa, pb):
        for u, v in zip(sa, sb):
            assert np.array_equal(u, v)


def test_ensemble_64_cases(pic):
    from pic_cuda_amd import configs
    specs = configs.ensemble_cases(64)
    steps = 3
    built = [configs.build(c, n, pic.PREC_FP32, mcc=True, seed=sd, te=te) for c, n, te, sd in specs]
    ens = pic.Ensemble()
    for k in built:
        ens.add(k["sim"])
    assert ens.size() == 64
    ens.step(steps)
    for k, (c, n, te, sd) in zip(built, specs):
        k["sim"].sync()
        _balance(pic, k, n)
    # sampled members: the smallest, a middle one and the largest grid
    order = sorted(range(64), key=lambda i: (specs[i][0], specs[i][1]))
    for i in (order[0], order[32], order[-1]):
        c, n, te, sd = specs[i]
        alone = configs.build(c, n, pic.PREC_FP32, mcc=True, seed=sd, te=te)
        alone["sim"].step(steps)
        alone["sim"].sync()
        _same_state(_state(pic, built[i]), _state(pic, alone))
        alone["sim"].close()
    ens.close()


def _run_server(pic, jobs, budget=200e9, dt_tick=0.01):
    """ticks on a synthetic clock (dt_tick seconds per tick), so that the
    arrivals interleave with running jobs the same way on every run"""
    sv = pic.Server(budget, 256, pic.PREC_FP32, keep_finished=True)
    ids = [sv.submit(**j) for j in jobs]
    for t in range(1000000):
        run, wait = sv.tick(t * dt_tick)
        if run == 0 and wait == 0:
            break
    return sv, ids


def test_serving_stream_jobs(pic):
    from pic_cuda_amd import configs
    jobs = configs.serve_jobs(16, rate=20.0, seed=9, min_steps=4, max_steps=24)
    sv, ids = _run_server(pic, jobs)
    res = [sv.result(j) for j in ids]
    assert all(r.state == 2 and r.steps_done == j["steps"] for r, j in zip(res, jobs))
    # several jobs ran at once: some admission happened before an earlier job finished
    assert any(res[b].admit_s < res[a].done_s for a in range(len(res)) for b in range(a + 1, len(res)))
    small = sorted(range(len(jobs)), key=lambda k: jobs[k]["cells"])[:2]
    for k in small:
        alone_sv, (ja,) = _run_server(pic, [dict(jobs[k], arrival_s=0.0)])
        a, b = alone_sv.handles(ja), sv.handles(ids[k])
        assert np.array_equal(a[1].download(pic.FIELD_PHI), b[1].download(pic.FIELD_PHI))
        for s in range(2):
            xa, xb = a[0].download(s), b[0].download(s)
            oa = _key_order(xa[0].astype(np.float32), xa[4])
            ob = _key_order(xb[0].astype(np.float32), xb[4])
            for u, v in zip(xa, xb):
                assert np.array_equal(u[oa], v[ob])
            assert np.array_equal(a[2].counters(s), b[2].counters(s))
        alone_sv.close()
    sv.close()
