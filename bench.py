#!/usr/bin/env python3
"""Benchmark: the full PIC step at scale on one MI355X.

Workload (one step = one pass of the hot path): 1024^2 cells (1025^2 nodes),
2e7 electrons + 2e7 Ar+ (uniform positions, Maxwellian Te = 2 eV, Ti = 0.03 eV,
n = 1e15 m^-3, dx = 0.5 Debye lengths, dt = 0.2/omega_pe), absorbing box walls
with compaction every step, null-collision MCC (e-Ar elastic, excitation,
ionization; Ar+ charge exchange) on Ar at 1e21 m^-3, FP32 particles, Poisson
solve by FP32 V-cycles inside FP64 iterative refinement to ||r||/||rhs|| < 1e-8.
Inputs are generated on the device and stay resident in HBM.

Prints one JSON line: particle-steps/s (sum of live particles over the timed
steps / wall time), the roofline of the dominant kernel (push_kernel: 36
algorithmic bytes per particle, timed with HIP events on the library's stream)
and a CPU baseline (the serial oracle on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1024^2+4e7 ptcls; MG-solve ms; ensemble jobs/s, p99 lat"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 36.0   # x,y,vx,vy FP32 read + write, id read


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


def cpu_baseline(gpu_case, budget_s=12.0):
    """The serial CPU oracle on a 256^2-cell, 1.25e6 + 1.25e6 particle sample of
    the same workload (same particles per cell, physics and solver settings)."""
    import numpy as np
    from oracle import oracle as orc
    from pic_cuda_amd import configs
    import pic_cuda_amd as pic

    cells = 256
    n = 1250000
    pl = configs.Plasma(cells=cells)
    nn = pl.nodes
    sim = gpu_case["sim"]
    sp = []
    for s in range(2):
        info = sim.mcc_info(s)
        vth = pl.vte if s == 0 else pl.vti
        x, y, vx, vy, ids = orc.init_particles(4, n, nn, nn, vth, 1, 0, s, 0)
        kq = pic.ELEMENTARY_CHARGE * (-1 if s == 0 else 1) * pl.dt / (pic.ELECTRON_MASS if s == 0 else configs.M_AR)
        sp.append(orc.Species(4, x, y, vx, vy, ids, kq, mcc=info, other=info["product_species"]))
    w = pl.density * pl.length * pl.length / n
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * w * qs / (pl.dx * pl.dx)) / pic.EPS0 for q in (-pic.ELEMENTARY_CHARGE, pic.ELEMENTARY_CHARGE)]
    o = orc.OracleSim(sp, nn, nn, pl.dx, pl.dx, pl.dt, cfac, dict(tol=1e-8, max_cycles=100))
    o.step(1)                      # first pass has no kick; time full steps after it
    t0 = time.perf_counter()
    steps = 0
    psteps = 0
    while steps < 40 and (time.perf_counter() - t0) < budget_s:
        psteps += sum(s.n for s in o.sp)
        o.step(1)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": psteps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "serial oracle, %d full steps of a 256^2-cell, 1.25e6 e- + 1.25e6 Ar+ sample "
                      "(same ppc, MCC and 1e-8 FP64 multigrid)" % steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0, help="particle count scale (1.0 = 2e7 per species)")
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--bin", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
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
    cycles = (st1.total_mg_cycles - st0.total_mg_cycles) / a.steps
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "full PIC step: %d^2 cells, %.3g e- + %.3g Ar+, null-collision MCC on Ar 1e21 m^-3, "
                               "absorbing box walls, FP32 particles, FP32-in-FP64 multigrid to 1e-8"
                               % (a.cells, alive_mean / 2, alive_mean / 2),
                   "cells": a.cells, "particles": int(st0.alive), "bin": a.bin},
        "roofline": roof, "cpu_baseline": cpu,
        "field_ms_per_step": round(field_ms, 4), "mg_cycles_per_step": cycles,
        "alive_start": int(st0.alive), "alive_end": int(st1.alive),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
