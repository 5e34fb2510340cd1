#!/usr/bin/env python3
"""Benchmark: the full PIC step at scale on one MI355X.

Workload (one step = one pass of the hot path): 1024^2 cells (1025^2 nodes),
2e7 electrons + 2e7 Ar+ (uniform positions, Maxwellian Te = 2 eV, Ti = 0.03 eV,
n = 1e15 m^-3, dx = 0.5 Debye lengths, dt = 0.2/omega_pe), absorbing box walls
with compaction every step, null-collision MCC (e-Ar elastic, excitation,
ionization; Ar+ charge exchange) on Ar at 1e21 m^-3, FP32 particles, Poisson
solve by FP32 V-cycles (2+2 red-black SOR, factor 1.2) inside FP64 iterative
refinement to ||r||/||rhs|| < 1e-8.
Inputs are generated on the device and stay resident in HBM.

Prints one JSON line: particle-steps/s (sum of live particles over the timed
steps / wall time), the roofline of the dominant kernel (push_kernel: 40
algorithmic bytes per particle, timed with HIP events on the library's stream)
and a CPU baseline (the oracle, OpenMP build and serial build, on a bounded
sample of the same workload).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1024^2+4e7 ptcls; MG-solve ms; ensemble jobs/s, p99 lat"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle-step of the push: x, y, vx, vy (FP32) read and
# written (32 B) and the id (u32) read (4 B). The id is written too (it moves
# with the particle), but that 4 B is traffic the counting rule does not credit.
BYTES_PER_PARTICLE = 36.0


def read_traffic():
    """HBM bytes per push_kernel launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy-kernel calibration), when a summary is committed."""
    path = os.path.join(ROOT, "profiles", "push_kernel_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(gpu_case, omp_steps=4, serial_steps=2):
    """The CPU oracle on the same workload (1024^2 cells, 2e7 + 2e7 particles
    from the same seeded initial state, MCC, FP64 multigrid to 1e-8): one
    untimed first pass, then a few timed full steps, with the OpenMP build on
    all the host threads OpenMP is given and with the serial build on one
    core."""
    from oracle import oracle as orc
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs

    pl = gpu_case["plasma"]
    nn = pl.nodes
    n = gpu_case["n_per_species"]
    sim = gpu_case["sim"]
    mg = gpu_case["mg"]
    w = pl.density * pl.length * pl.length / n
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * w * qs / (pl.dx * pl.dx)) / pic.EPS0 for q in (-pic.ELEMENTARY_CHARGE, pic.ELEMENTARY_CHARGE)]
    infos = [sim.mcc_info(s) for s in range(2)]

    def run(omp, steps):
        sp = []
        for s in range(2):
            vth = pl.vte if s == 0 else pl.vti
            x, y, vx, vy, ids = orc.init_particles(4, n, nn, nn, vth, 1, 0, s, 0)
            kq = pic.ELEMENTARY_CHARGE * (-1 if s == 0 else 1) * pl.dt / (
                pic.ELECTRON_MASS if s == 0 else configs.M_AR)
            sp.append(orc.Species(4, x, y, vx, vy, ids, kq, mcc=infos[s], other=infos[s]["product_species"]))
        o = orc.OracleSim(sp, nn, nn, pl.dx, pl.dx, pl.dt, cfac,
                          dict(tol=mg.tol, max_cycles=mg.max_cycles, pre=mg.pre, post=mg.post,
                               coarse_sweeps=mg.coarse_sweeps, omega=mg.omega), omp=omp)
        o.step(1)                      # the first pass has no kick; full steps follow
        t0 = time.perf_counter()
        psteps = 0
        for _ in range(steps):
            psteps += sum(s.n for s in o.sp)
            o.step(1)
        rate = psteps / (time.perf_counter() - t0)
        del o, sp
        return rate

    threads = orc.lib(True).orc_num_threads()
    v_omp = run(True, omp_steps)
    v_ser = run(False, serial_steps)
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": v_omp, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "oracle (OpenMP, %d threads), %d full steps of the same 1024^2-cell, %.3g e- + %.3g Ar+ "
                      "workload after one untimed first pass (MCC, FP64 multigrid to 1e-8)"
                      % (threads, omp_steps, n, n),
            "serial_value": v_ser, "serial_steps": serial_steps, "nproc": os.cpu_count(), "cpu_model": model}


def poisson(n=2049, cpu=True):
    """Mixed-precision Poisson alone: 2049^2 Dirichlet grid, 200 Gaussian blobs
    + 1e-3 noise, FP32 V-cycles in FP64 refinement vs FP64-only V-cycles, both to
    ||r||/||rhs|| < 1e-10 from a zero start."""
    import numpy as np
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    rhs = configs.poisson_rhs(n)
    h = 1.0 / (n - 1)
    g = pic.Grid2D(n, n, 1.0, 1.0)
    g.upload(pic.FIELD_RHO, rhs)
    rho_p, phi_p, _ = g.device_ptrs()
    out = {}
    sols = {}
    for name, prm in (("mixed", pic.mg_params(pic.PRESET_POISSON)),
                      ("fp64", pic.mg_params(pic.PRESET_POISSON, mode=pic.MG_FP64))):
        mg = pic.Multigrid(n, prm)
        best = None
        for _ in range(3):
            g.upload(pic.FIELD_PHI, np.zeros((n, n)))
            pic.synchronize()
            st = mg.solve(phi_p, rho_p, h)
            best = st if best is None or st.ms < best.ms else best
        sols[name] = g.download(pic.FIELD_PHI)
        cyc = best.cycles
        if name == "mixed":
            # V FP32 V-cycles at 32 B per fine node, K FP64 refinement steps
            # (residual and correction) at 40 B per node, and the first
            # residual of the solve (phi and rhs read, r32 written: 20 B)
            steps = (cyc + prm.inner - 1) // max(1, prm.inner)
            byts = n * n * (32 * cyc + 40 * steps + 20)
        else:
            steps = cyc
            byts = n * n * 64 * cyc
        gbps = byts / (best.ms * 1e-3) / 1e9
        out[name] = {"ms": round(best.ms, 3), "cycles": cyc, "refinement_steps": steps,
                     "rel_residual": best.rel_residual, "launches": best.launches,
                     "algorithmic_GBps": round(gbps, 1), "roofline_frac": round(gbps / HBM_PEAK_GBS, 4)}
    ref = sols["fp64"]
    out["mixed_vs_fp64_rel_l2"] = float(np.linalg.norm(sols["mixed"] - ref) / np.linalg.norm(ref))
    out["speedup_mixed_over_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    if cpu:
        from oracle import oracle as orc
        phi = np.zeros((n, n))
        t0 = time.perf_counter()
        p64 = pic.mg_params(pic.PRESET_POISSON, mode=pic.MG_FP64)
        st = orc.mg_solve(phi, rhs, h, pre=p64.pre, post=p64.post, coarse_sweeps=p64.coarse_sweeps,
                          omega=p64.omega, tol=p64.tol, max_cycles=p64.max_cycles)
        el = time.perf_counter() - t0
        out["cpu_fp64_oracle"] = {"ms": round(el * 1e3, 1), "cycles": st.cycles, "cores": 1}
        out["mixed_vs_cpu_fp64_rel_l2"] = float(np.linalg.norm(sols["mixed"] - phi) / np.linalg.norm(phi))
    return out


def ensemble(n_cases=64, steps=100, sequential=True):
    """Batched ensemble: 64 plasmas, cells log-uniform 64..512 (rounded to GRID_SIZES),
    clamp(100 cells^2, 1e4, 4e6) particles per species, Te ~ U(1,5) eV, MCC on;
    one launch per stage for all cases, against the same cases one after
    another through the single-plasma path."""
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    specs = configs.ensemble_cases(n_cases)
    res = {"cases": n_cases, "steps": steps,
           "particles": int(sum(2 * s[1] for s in specs))}
    if sequential:
        t_seq = 0.0
        for cells, npart, te, seed in specs:
            c = configs.build(cells, npart, pic.PREC_FP32, mcc=True, seed=seed, te=te)
            pic.synchronize()
            t0 = time.perf_counter()
            c["sim"].step(steps)
            pic.synchronize()
            t_seq += time.perf_counter() - t0
            del c
        res["sequential_s"] = round(t_seq, 3)
        res["sequential_jobs_per_s"] = n_cases / t_seq
    built = [configs.build(cells, npart, pic.PREC_FP32, mcc=True, seed=seed, te=te)
             for cells, npart, te, seed in specs]
    ens = pic.Ensemble()
    for c in built:
        ens.add(c["sim"])
    ens.step(1)                      # builds the batch tables; counted below as a step
    pic.synchronize()
    t0 = time.perf_counter()
    ens.step(steps - 1)
    pic.synchronize()
    el = time.perf_counter() - t0
    t_batch = el * steps / (steps - 1)
    res["batched_s"] = round(t_batch, 3)
    res["batched_jobs_per_s"] = n_cases / t_batch
    res["batched_ms_per_step"] = round(1e3 * el / (steps - 1), 3)
    res["batched_particle_steps_per_s"] = res["particles"] * steps / t_batch
    if sequential:
        res["speedup_batched_over_sequential"] = round(t_seq / t_batch, 2)
    # the same ensemble continued with one graph launch per step
    ens.use_graph(True)
    ens.step(1)                      # captures the graph
    pic.synchronize()
    t0 = time.perf_counter()
    ens.step(steps - 1)
    pic.synchronize()
    el = time.perf_counter() - t0
    res["batched_graph_ms_per_step"] = round(1e3 * el / (steps - 1), 3)
    res["batched_graph_jobs_per_s"] = n_cases / (el * steps / (steps - 1))
    if sequential:
        res["speedup_batched_graph_over_sequential"] = round(t_seq / (el * steps / (steps - 1)), 2)
    ens.close()
    return res


def graph(steps=100):
    """Eager launches against one graph launch per step: the plain config
    (128^2 cells, 1e5 + 1e5 particles, FP64, GS 2+2 to 1e-8) where launch
    overhead is a large part of a step, and the full step at scale."""
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    out = {}
    for name, make, n in (("plain_128", lambda: configs.plain_small(), steps),
                          ("full_1024", lambda: configs.full_step(), 40)):
        res = {}
        for mode in ("eager", "graph"):
            c = make()
            sim = c["sim"]
            if mode == "graph":
                sim.use_graph(True)
            sim.step(3)
            pic.synchronize()
            st0 = sim.stats()
            t0 = time.perf_counter()
            sim.step(n)
            pic.synchronize()
            el = time.perf_counter() - t0
            st1 = sim.stats()
            res[mode] = {"ms_per_step": round(1e3 * el / n, 4),
                         "mg_cycles_per_step": (st1.total_mg_cycles - st0.total_mg_cycles) / n,
                         "graph_steps": st1.graph_steps - st0.graph_steps}
            del c
        res["speedup_graph_over_eager"] = round(res["eager"]["ms_per_step"] / res["graph"]["ms_per_step"], 3)
        out[name] = res
    return out


def tune(steps=40):
    """The full step at scale before and after the autotuner (bin tile, SOR
    factor, sweeps per multigrid level)."""
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    c = configs.full_step()
    sim = c["sim"]

    def timed():
        sim.step(3)
        pic.synchronize()
        st0 = sim.stats()
        t0 = time.perf_counter()
        sim.step(steps)
        pic.synchronize()
        el = time.perf_counter() - t0
        st1 = sim.stats()
        return {"ms_per_step": round(1e3 * el / steps, 4),
                "push_ms": round((st1.ms_push - st0.ms_push) / steps, 4),
                "field_ms": round((st1.ms_field - st0.ms_field) / steps, 4),
                "mg_cycles_per_step": (st1.total_mg_cycles - st0.total_mg_cycles) / steps}

    before = timed()
    t0 = time.perf_counter()
    chosen = sim.autotune()
    tune_s = time.perf_counter() - t0
    after = timed()
    return {"before": before, "after": after, "chosen": chosen, "tune_s": round(tune_s, 2),
            "speedup": round(before["ms_per_step"] / after["ms_per_step"], 3)}


def walls(n_walls=256, steps=20):
    """The full step at scale with internal wall segments (random centres and
    directions, lengths 4-64 cells) on top of the absorbing box: per-bin
    candidate lists against every particle testing every wall, and the box
    alone."""
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    segs = configs.random_walls(1024, n_walls, seed=3, min_len=4.0, max_len=64.0)
    out = {"n_walls": n_walls}
    for name, w, lists in (("box_only", None, True), ("lists", segs, True), ("full_scan", segs, False)):
        c = configs.full_step(walls=w)
        sim = c["sim"]
        if not lists:
            sim.set_wall_lists(False)
        sim.step(3)
        pic.synchronize()
        st0 = sim.stats()
        t0 = time.perf_counter()
        sim.step(steps)
        pic.synchronize()
        el = time.perf_counter() - t0
        st1 = sim.stats()
        nb, ne, mx = sim.wall_list_stats()
        out[name] = {"ms_per_step": round(1e3 * el / steps, 4),
                     "push_ms": round((st1.ms_push - st0.ms_push) / steps, 4),
                     "particle_steps_per_s": steps * 0.5 * (st0.alive + st1.alive) / el,
                     "list_entries_per_bin": round(ne / nb, 2) if nb else None, "longest_list": mx}
        del c, sim
    out["push_speedup_lists_over_full_scan"] = round(out["full_scan"]["push_ms"] / out["lists"]["push_ms"], 2)
    return out


def export(reps=20):
    """Position export of the full step at scale (2e7 electrons): float pairs
    into a preallocated device buffer (8 B read + 8 B written per particle),
    and one frame file of both species under /dev/shm."""
    import numpy as np
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    c = configs.full_step()
    sim, parts = c["sim"], c["particles"]
    sim.step(3)
    pic.synchronize()
    dst, n = parts.export_positions(0)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        parts.export_positions(0, dst)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    path = "/dev/shm/pic_bench_frame.bin"
    t0 = time.perf_counter()
    sim.write_frame(path)
    frame_s = time.perf_counter() - t0
    size = os.path.getsize(path)
    os.remove(path)
    return {"particles": n, "export_ms": round(1e3 * t, 4), "export_GBps": round(16.0 * n / t / 1e9, 1),
            "export_frac_hbm": round(16.0 * n / t / 1e9 / HBM_PEAK_GBS, 3),
            "frame_ms": round(1e3 * frame_s, 2), "frame_MB": round(size / 1e6, 1)}


def _serve_stream(jobs, budget):
    """run a job stream through the server on the wall clock; per-job results"""
    import numpy as np
    import pic_cuda_amd as pic
    t_arr = [j["arrival_s"] for j in jobs]
    sv = pic.Server(budget, 256, pic.PREC_FP32)
    ids = [sv.submit(**j) for j in jobs]
    t0 = time.perf_counter()
    ticks = 0
    while True:
        now = time.perf_counter() - t0
        run, wait = sv.tick(now)
        ticks += 1
        if run == 0 and wait == 0:
            break
        if run == 0:
            time.sleep(max(0.0, min(0.05, float(min(t_arr[i] for i in range(len(jobs))
                                                         if sv.result(ids[i]).state == 0)) - now)))
    wall = time.perf_counter() - t0
    results = [sv.result(j) for j in ids]
    step_ms = np.concatenate([sv.step_ms(j) for j in ids])
    comp = np.array([r.done_s - r.arrival_s for r in results])
    wait_s = np.array([r.admit_s - r.arrival_s for r in results])
    psteps = sum(r.particles * r.steps_done for r in results)
    sv.close()
    return {"jobs": len(jobs), "wall_s": round(wall, 3), "jobs_per_s": len(jobs) / wall, "ticks": ticks,
            "particle_steps_per_s": psteps / wall,
            "step_ms_p50": float(np.percentile(step_ms, 50)), "step_ms_p99": float(np.percentile(step_ms, 99)),
            "completion_s_p50": float(np.percentile(comp, 50)), "completion_s_p99": float(np.percentile(comp, 99)),
            "queue_wait_s_p50": float(np.percentile(wait_s, 50)), "queue_wait_s_p99": float(np.percentile(wait_s, 99))}


def serve(n_jobs=24, rate=20.0, budget=200e9, seed=9, max_steps=2000):
    """Serving: Poisson arrivals (rate jobs/s), cells log-uniform 64..1024
    (rounded to GRID_SIZES), 100 particles per cell and species, steps U(200, max),
    seed = job id, admission by an HBM budget, continuous batching.
    stream: the stated arrival rate (20/s, far above what one GPU serves, so the
    queue fills and the run measures saturated throughput); steady: the same
    jobs arriving at 60 % of that throughput (the queue stays short: the
    latencies of a sustained load); sequential: the same jobs one after
    another through the single-plasma path (no batching), the baseline."""
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    jobs = configs.serve_jobs(n_jobs, rate, seed, max_steps=max_steps)
    out = {"jobs": n_jobs, "rate_per_s": rate, "budget_GB": budget / 1e9}
    out["stream"] = _serve_stream(jobs, budget)
    t0 = time.perf_counter()
    for j in jobs:
        c = configs.build(j["cells"], int(round(j["ppc"] * j["cells"] * j["cells"])), pic.PREC_FP32, mcc=True,
                          seed=j["seed"])
        c["sim"].step(j["steps"])
        pic.synchronize()
        c["sim"].close()
        c["particles"].close()
        del c
    seq = time.perf_counter() - t0
    out["sequential"] = {"wall_s": round(seq, 3), "jobs_per_s": n_jobs / seq}
    slow = 0.6 * out["stream"]["jobs_per_s"]
    steady = [dict(j, arrival_s=j["arrival_s"] * rate / slow) for j in jobs]
    out["steady"] = dict(_serve_stream(steady, budget), rate_per_s=round(slow, 3))
    out["speedup_stream_over_sequential"] = round(out["stream"]["jobs_per_s"] / out["sequential"]["jobs_per_s"], 2)
    # flat keys of the saturated stream (the stated workload)
    for k in ("jobs_per_s", "step_ms_p50", "step_ms_p99", "completion_s_p50", "completion_s_p99"):
        out[k] = out["stream"][k]
    return out


def fp64_check(cells, scale, bw, horizons=(1, 5, 20)):
    """Errors of the FP32 step against FP64 on the same workload.
    field_solve: the step's FP32-in-FP64 multigrid potential against an FP64
    multigrid solve of the same charge to 1e-13 (the solver's own error).
    trajectories: the FP32 step (FP32 particles, mixed solve) against the FP64
    step (FP64 particles, FP64 multigrid: the path that equals the FP64 oracle
    bit for bit at test sizes) from the same initial state, at each horizon:
    the potential, energy and momentum diagnostics, live counts and collision
    counts.  Beside it, the FP64 run's own divergence: a second FP64 run from
    the same state with every position moved by one FP32 ulp (the rounding the
    FP32 run starts from).  A few particles cross a wall or a collision
    threshold in one run and not the other, and each moves the potential by
    about 1/sqrt(N) of its norm, so the two potentials drift apart even in
    FP64; the FP32 error is judged against that floor."""
    import numpy as np
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    n = max(1000, int(2e7 * scale))
    a = configs.full_step(scale=scale, cells=cells, bin_w=bw, bin_h=bw)
    state = [a["particles"].download(s) for s in range(2)]
    b = configs.build(cells, n, pic.PREC_FP64, mcc=True, bin_w=bw, bin_h=bw, state=state)
    up = np.float32(np.inf)
    pert = [(np.nextafter(x.astype(np.float32), up).astype(np.float64),
             np.nextafter(y.astype(np.float32), up).astype(np.float64), vx, vy, ids)
            for (x, y, vx, vy, ids) in state]
    c = configs.build(cells, n, pic.PREC_FP64, mcc=True, bin_w=bw, bin_h=bw, state=pert)
    del state, pert

    def rel(x, y):
        return abs(x - y) / abs(y) if y else abs(x - y)

    def collisions(case):
        k = case["sim"].counters(0)
        return int(sum(k[pic.CNT["process%d" % q]] for q in range(4)))

    def l2(p, q):
        return float(np.linalg.norm(p - q) / np.linalg.norm(q))

    per = []
    done = 0
    for hz in horizons:
        for case in (a, b, c):
            case["sim"].step(hz - done)
        done = hz
        pa, pb, pc = (case["grid"].download(pic.FIELD_PHI) for case in (a, b, c))
        per.append({"steps": hz, "phi_rel_l2_fp32_vs_fp64": l2(pa, pb), "phi_rel_l2_fp64_ulp_vs_fp64": l2(pc, pb),
                    "alive": [case["particles"].count(-1) for case in (a, b, c)],
                    "electron_collisions": [collisions(case) for case in (a, b, c)]})
    steps = horizons[-1]
    da, db = a["sim"].diagnostics(), b["sim"].diagnostics()
    pa, pb = a["grid"].download(pic.FIELD_PHI), b["grid"].download(pic.FIELD_PHI)
    # momentum scale of a species: sum of w m |v| ~ sqrt(2 KE M), M = n w m its mass
    pl = configs.Plasma(te=2.0, cells=cells)
    w = pl.density * pl.length * pl.length / n
    masses = (pic.ELECTRON_MASS, 39.948 * pic.AMU)
    # the solver alone: the FP32 run's last charge solved in FP64 to 1e-13,
    # warm-started from the FP32 run's potential
    g = a["grid"]
    ref = pic.Grid2D(g.m, g.n, g.dx * (g.m - 1), g.dy * (g.n - 1))
    ref.upload(pic.FIELD_RHO, g.download(pic.FIELD_RHS))
    ref.upload(pic.FIELD_PHI, pa)
    rho_p, phi_p, _ = ref.device_ptrs()
    mg64 = pic.Multigrid(g.m, pic.mg_params(pic.PRESET_POISSON, mode=pic.MG_FP64, tol=1e-13, max_cycles=100))
    st64 = mg64.solve(phi_p, rho_p, g.dx)
    p64 = ref.download(pic.FIELD_PHI)
    solve = {"phi_rel_l2": l2(pa, p64), "fp64_cycles": st64.cycles, "fp64_rel_residual": st64.rel_residual,
             "step_tol": 1e-8}
    traj = {"steps": steps, "reference": "FP64 particles and FP64 multigrid from the same initial state",
            "phi_rel_l2": l2(pa, pb),
            "kinetic_rel": [rel(da["kinetic"][s], db["kinetic"][s]) for s in range(2)],
            "momentum_diff_rel_to_scale": [
                float(np.hypot(da["px"][s] - db["px"][s], da["py"][s] - db["py"][s]) /
                      np.sqrt(2.0 * db["kinetic"][s] * db["count"][s] * w * masses[s])) for s in range(2)],
            "field_energy_rel": rel(da["field_energy"], db["field_energy"]),
            "alive_fp32": [int(k) for k in da["count"]], "alive_fp64": [int(k) for k in db["count"]],
            "electron_collisions_fp32": collisions(a), "electron_collisions_fp64": collisions(b),
            "horizons": per}
    out = {"field_solve": solve, "trajectories": traj}
    for case in (a, b, c):
        case["sim"].close()
        case["particles"].close()
    return out


def fp64_rate(cells, bw, steps=40):
    """particle-steps/s of the same workload with FP64 particles and the FP64
    multigrid (the reference's precision), for context beside the FP32 line"""
    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    c = configs.build(cells, 20000000, pic.PREC_FP64, mcc=True, bin_w=bw, bin_h=bw)
    sim = c["sim"]
    sim.step(5)
    pic.synchronize()
    st0 = sim.stats()
    t0 = time.perf_counter()
    sim.step(steps)
    pic.synchronize()
    el = time.perf_counter() - t0
    st1 = sim.stats()
    out = {"value": steps * 0.5 * (st0.alive + st1.alive) / el, "ms_per_step": round(1e3 * el / steps, 4),
           "push_ms": round((st1.ms_push - st0.ms_push) / steps, 4),
           "field_ms": round((st1.ms_field - st0.ms_field) / steps, 4), "steps": steps}
    sim.close()
    c["particles"].close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0, help="particle count scale (1.0 = 2e7 per species)")
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--bin", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner (bin tile, SOR factor, cycles per refinement step, sweeps per "
                         "level) after the warm-up steps, before the timed ones")
    ap.add_argument("--workload", default="all",
                    choices=["all", "step", "poisson", "ensemble", "serve", "graph", "tune", "walls", "export"])
    ap.add_argument("--ens-steps", type=int, default=500)
    ap.add_argument("--ens-no-seq", action="store_true",
                    help="ensemble workload: skip the one-after-another runs (for profiling the batched steps)")
    ap.add_argument("--serve-jobs", type=int, default=24)
    a = ap.parse_args()
    if a.workload in ("poisson", "ensemble", "serve", "graph", "tune", "walls", "export"):
        import pic_cuda_amd as pic
        pic.load()
        fn = {"poisson": lambda: poisson(cpu=not a.no_cpu), "ensemble": lambda: ensemble(steps=a.ens_steps, sequential=not a.ens_no_seq),
              "serve": lambda: serve(n_jobs=a.serve_jobs), "graph": lambda: graph(), "tune": lambda: tune(),
              "walls": lambda: walls(), "export": lambda: export()}[a.workload]
        print(json.dumps({"workload": a.workload, **fn()}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; the multi-GPU decomposition is a design only "
              "(domain strips with particle migration and halo exchange)", file=sys.stderr)
        sys.exit(2)

    import pic_cuda_amd as pic
    from pic_cuda_amd import configs
    pic.load()
    t_setup = time.perf_counter()
    case = configs.full_step(scale=a.scale, cells=a.cells, bin_w=a.bin, bin_h=a.bin)
    sim = case["sim"]
    print("bench: setup %.1f s, %d particles" % (time.perf_counter() - t_setup, case["particles"].last_alive),
          file=sys.stderr, flush=True)
    sim.step(a.warmup)
    tuned = None
    if a.tune:
        t_tune = time.perf_counter()
        tuned = sim.autotune()
        tuned["tune_s"] = round(time.perf_counter() - t_tune, 3)
        sim.step(3)   # the tuned layout's first steps settle the cycle count
    pic.synchronize()
    st0 = sim.stats()
    t0 = time.perf_counter()
    sim.step(a.steps)
    pic.synchronize()
    el = time.perf_counter() - t0
    st1 = sim.stats()
    alive_mean = 0.5 * (st0.alive + st1.alive)
    value = a.steps * alive_mean / el
    push_ms = (st1.ms_push - st0.ms_push) / a.steps
    field_ms = (st1.ms_field - st0.ms_field) / a.steps
    achieved = BYTES_PER_PARTICLE * alive_mean / (push_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(), "kernel": "push_kernel",
            "push_ms": round(push_ms, 4)}
    cpu = None
    if not a.no_cpu:
        print("bench: gpu done (%.3g particle-steps/s); timing the CPU baseline" % value, file=sys.stderr,
              flush=True)
        cpu = cpu_baseline(case)
    acc = None
    rate64 = None
    if not a.no_fp64_check:
        print("bench: comparing 20 FP32 steps with the FP64 path", file=sys.stderr, flush=True)
        acc = fp64_check(a.cells, a.scale, a.bin)
        if a.scale == 1.0:
            print("bench: the FP64 rate", file=sys.stderr, flush=True)
            rate64 = fp64_rate(a.cells, a.bin)
    cycles = (st1.total_mg_cycles - st0.total_mg_cycles) / a.steps
    # the whole step by the per-step grid byte rule: 16 B/node charge, 32 B/node
    # per FP32 V-cycle, 40 B/node per FP64 refinement step, 16 B/node E field
    nodes = (a.cells + 1) ** 2
    inner = max(1, case["mg"].inner)
    grid_bytes = nodes * (16 + 32 * cycles + 40 * math.ceil(cycles / inner) + 16)
    step_bytes = BYTES_PER_PARTICLE * alive_mean + grid_bytes
    step_gbps = step_bytes / (el / a.steps) / 1e9
    field_gbps = grid_bytes / (field_ms * 1e-3) / 1e9
    step_roof = {"bytes_per_step": int(step_bytes), "achieved": round(step_gbps, 1), "frac": round(step_gbps / HBM_PEAK_GBS, 4),
                 "field_bytes_per_step": int(grid_bytes), "field_achieved": round(field_gbps, 1),
                 "field_frac": round(field_gbps / HBM_PEAK_GBS, 4), "unit": "GB/s"}
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "full PIC step: %d^2 cells, %.3g e- + %.3g Ar+, null-collision MCC on Ar 1e21 m^-3, "
                               "absorbing box walls, FP32 particles, FP32-in-FP64 multigrid (2+2 red-black SOR 1.2) to 1e-8"
                               % (a.cells, alive_mean / 2, alive_mean / 2),
                   "cells": a.cells, "particles": int(st0.alive), "bin": a.bin},
        "roofline": roof, "step_roofline": step_roof, "cpu_baseline": cpu,
        "field_ms_per_step": round(field_ms, 4), "mg_cycles_per_step": cycles,
        "steps_run_ahead": int(st1.ahead_steps - st0.ahead_steps), "tuned": tuned,
        "alive_start": int(st0.alive), "alive_end": int(st1.alive),
        "fp64_error": acc, "fp64_particles": rate64,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
