"""Position export (convertPositionsToFloat, kernels.cu:268-273, as updateVBO
uses it, render.h:384-405) and frame files: the float pairs are the stored
positions rounded to float (cell units) or scaled by the grid spacing first
(domain units); exporting never changes a running simulation."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs_sorted(a):
    a = np.asarray(a, dtype=np.float32).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def _case(pic, prec, seed=3):
    from pic_cuda_amd import configs
    return configs.build(48, 30000, prec, mcc=True, seed=seed, gas_density=1e22, te=8.0)


@pytest.mark.parametrize("prec", [4, 8])
def test_export_matches_download(pic, prec):
    c = _case(pic, prec)
    sim, parts, grid = c["sim"], c["particles"], c["grid"]
    sim.step(4)
    sim.sync()          # download completes the owed half step (its collisions add particles)
    for s in range(2):
        t, n = parts.export_positions(s)
        cell = t.to_numpy(count=2 * n)
        td, nd = parts.export_positions(s, units="domain")
        dom = td.to_numpy(count=2 * nd)
        x, y, *_ = parts.download(s)
        assert n == nd == len(x)
        want = np.stack([x.astype(np.float32), y.astype(np.float32)], axis=1)
        assert np.array_equal(_pairs_sorted(cell), _pairs_sorted(want))
        want_d = np.stack([(x * grid.dx).astype(np.float32), (y * grid.dy).astype(np.float32)], axis=1)
        assert np.array_equal(_pairs_sorted(dom), _pairs_sorted(want_d))
        with pytest.raises(pic.PicError, match="capacity"):
            parts.export_positions(s, pic.DeviceBuffer(8 * (n - 1)))


def test_export_does_not_change_the_run(pic):
    a, b = _case(pic, 4), _case(pic, 4)
    a["sim"].step(3)
    for s in range(2):
        a["particles"].export_positions(s)
        a["particles"].export_positions(s, units="domain")
    a["sim"].step(3)
    b["sim"].step(6)
    for f in (pic.FIELD_RHS, pic.FIELD_PHI):
        assert np.array_equal(a["grid"].download(f), b["grid"].download(f))
    for s in range(2):
        assert np.array_equal(a["sim"].counters(s), b["sim"].counters(s))


def test_frame_file_round_trip(pic, tmp_path):
    c = _case(pic, 4)
    sim, parts = c["sim"], c["particles"]
    sim.step(5)
    path = tmp_path / "frame.bin"
    sim.write_frame(path)
    sim.write_frame(path)                       # replaced in place
    assert not os.path.exists(str(path) + ".tmp")
    fr = pic.read_frame(str(path))
    assert fr["step"] == 5 and len(fr["positions"]) == 2
    L = c["plasma"].length
    assert abs(fr["width"] - L) < 1e-12 * L and abs(fr["height"] - L) < 1e-12 * L
    for s in range(2):
        t, n = parts.export_positions(s, units="domain")
        assert fr["positions"][s].shape == (n, 2)
        assert np.array_equal(_pairs_sorted(fr["positions"][s]), _pairs_sorted(t.to_numpy(count=2 * n)))
        p = fr["positions"][s]
        assert p.min() >= 0.0 and p.max() <= L * (1 + 1e-6)
