"""The C++ step driver (pic_cuda_amd/pic_run, over pic2d.hpp) runs and agrees
with the same run through the Python mirror."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_driver_matches_python_mirror(pic):
    exe = os.path.join(ROOT, "pic_cuda_amd", "pic_run")
    assert os.path.exists(exe), "build pic_run first (make -C pic_cuda_amd/csrc)"
    out = subprocess.run([exe, "--cells", "64", "--particles", "20000", "--steps", "20", "--mcc", "1", "--seed", "3"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    from pic_cuda_amd import configs
    case = configs.build(64, 20000, pic.PREC_FP32, mcc=True, seed=3)
    case["sim"].step(20)
    case["sim"].sync()
    assert r["alive"] == case["particles"].last_alive
    assert r["absorbed_e"] == int(case["sim"].counters(0)[0])
    assert r["absorbed_i"] == int(case["sim"].counters(1)[0])


def test_cpp_driver_frames_leave_the_run_unchanged(pic, tmp_path):
    """--frames writes the positions every K steps (the updateVBO cadence of
    main.cu); the run's counts equal those of the run without frames."""
    exe = os.path.join(ROOT, "pic_cuda_amd", "pic_run")
    base = [exe, "--cells", "64", "--particles", "20000", "--steps", "20", "--mcc", "1", "--seed", "3"]
    frame = str(tmp_path / "frame.bin")
    runs = [subprocess.run(base + extra, capture_output=True, text=True, timeout=120)
            for extra in ([], ["--frames", frame, "--every", "7"])]
    for out in runs:
        assert out.returncode == 0, out.stderr
    a, b = (json.loads(o.stdout.strip().splitlines()[-1]) for o in runs)
    assert b["frames"] == 3 and a["frames"] == 0
    for k in ("alive", "absorbed_e", "absorbed_i", "collisions_e", "mg_cycles"):
        assert a[k] == b[k], k
    fr = pic.read_frame(frame)
    assert fr["step"] == 20 and len(fr["positions"]) == 2
    assert 0 < sum(len(p) for p in fr["positions"]) <= b["alive"] + 1000


def test_cpp_driver_staged_fill_matches_python_mirror(pic):
    """--staged fills the particles as main.cu:232-238 does (host arrays into
    the public device arrays r, v and type, then `last_alive = n`, through the
    C++ mirror); the same arrays staged through the Python mirror give the same
    run."""
    import numpy as np
    from pic_cuda_amd import configs
    exe = os.path.join(ROOT, "pic_cuda_amd", "pic_run")
    cells, n, steps = 64, 20000, 20
    out = subprocess.run([exe, "--cells", str(cells), "--particles", str(n), "--steps", str(steps), "--mcc", "1",
                          "--seed", "3", "--staged"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    pl = configs.Plasma(cells=cells)
    L = pl.length
    i = np.arange(2 * n, dtype=np.float64)
    vth = np.where(i < n, pl.vte, pl.vti)
    pos = np.empty((2 * n, 2))
    vel = np.empty((2 * n, 2))
    pos[:, 0] = L * (0.02 + 0.96 * np.fmod(i * 0.6180339887498949, 1.0))
    pos[:, 1] = L * (0.02 + 0.96 * np.fmod(i * 0.7548776662466927, 1.0))
    vel[:, 0] = vth * 1.7320508075688772 * (2.0 * np.fmod(i * 0.5698402909980532, 1.0) - 1.0)
    vel[:, 1] = vth * 1.7320508075688772 * (2.0 * np.fmod(i * 0.8191725133961645, 1.0) - 1.0)
    typ = (i >= n).astype(np.int32)
    grid = pic.Grid2D(pl.nodes, pl.nodes, L, L)
    p = pic.Particles2D(2 * n + 4096, pic.PREC_FP32)
    p.change_species(0, configs.ME, -configs.QE)
    p.add_species(configs.M_AR, configs.QE)
    w = pl.density * L * L / n
    p.set_weight(0, w)
    p.set_weight(1, w)
    a = p.staging(2 * n)
    pic.host_to_device(a["r"], pos)
    pic.host_to_device(a["v"], vel)
    pic.host_to_device(a["type"], typ)
    p.set_staged(2 * n)
    bnd = pic.Boundaries2D(L, L, grid.m, grid.n)
    bnd.add_box()
    bnd.compile_walls()
    xs = pic.CrossSections(1e21, 0.03)
    configs.add_argon_processes(xs, 0, 1)
    sim = pic.Simulation(p, grid, bnd, xs, pic.sim_params(dt=pl.dt, seed=3, mcc=1))
    assert p.count(0) == n and p.count(1) == n
    sim.step(steps)
    sim.sync()
    c0, c1 = sim.counters(0), sim.counters(1)
    assert r["alive"] == p.last_alive
    assert r["absorbed_e"] == int(c0[0]) and r["absorbed_i"] == int(c1[0])
    assert r["collisions_e"] == int(c0[2] + c0[3] + c0[4] + c0[5])
    assert r["mg_cycles"] == sim.stats().total_mg_cycles
    assert r["collisions_e"] > 0 and r["absorbed_e"] > 0
