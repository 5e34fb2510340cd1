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
