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
