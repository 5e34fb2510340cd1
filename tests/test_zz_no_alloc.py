"""pic_particles_export_positions_static, the export that allocates nothing:
called through the C ABI, its float pairs equal those of
pic_particles_export_positions on the same store (as multisets: the order is
unspecified for both), on two calls in a row, the second on a smaller store
(fewer bins and particles than the first left in the shared scratch array)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs_sorted(a):
    a = np.asarray(a, dtype=np.float32).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def _export(pic, fn, parts, species, units, cap):
    buf = pic.DeviceBuffer(8 * max(1, cap))
    n = ctypes.c_int64(-1)
    rc = getattr(pic.load(), fn)(parts._h, int(species), ctypes.c_void_p(buf.ptr), int(cap), int(units),
                                 ctypes.byref(n))
    assert rc == 0, pic.load().pic_last_error()
    out = buf.to_numpy(count=2 * n.value) if n.value > 0 else np.zeros(0, np.float32)
    buf.close()
    return n.value, out


def _case(cells, n, seed, prec):
    from pic_cuda_amd import configs
    return configs.build(cells, n, prec, mcc=True, seed=seed, gas_density=1e22, te=8.0)


@pytest.mark.parametrize("prec", [4, 8])
def test_static_export_equals_export(pic, prec):
    big = _case(96, 60000, 5, prec)
    small = _case(24, 3000, 6, prec)
    for c in (big, small):
        c["sim"].step(3)
    sizes = []
    for c in (big, small):              # the second call: a smaller store
        parts = c["particles"]
        for s in range(2):
            for units in (0, 1):
                cnt = parts.count(s)
                n_ref, ref = _export(pic, "pic_particles_export_positions", parts, s, units, cnt + 16)
                n_st, got = _export(pic, "pic_particles_export_positions_static", parts, s, units, cnt + 16)
                assert n_st == n_ref == cnt > 0
                assert np.array_equal(_pairs_sorted(got), _pairs_sorted(ref))
            sizes.append(cnt)
    assert sizes[2] < sizes[0] and sizes[3] < sizes[1]


def test_static_export_limits(pic):
    c = _case(24, 3000, 7, 4)
    parts = c["particles"]
    cnt = parts.count(0)
    # count query with no buffer, and a buffer too small: the capacity error
    n = ctypes.c_int64(-1)
    assert pic.load().pic_particles_export_positions_static(parts._h, 0, None, 0, 0, ctypes.byref(n)) == 0
    assert n.value == cnt
    buf = pic.DeviceBuffer(8 * (cnt - 1))
    rc = pic.load().pic_particles_export_positions_static(parts._h, 0, ctypes.c_void_p(buf.ptr), cnt - 1, 0,
                                                          ctypes.byref(n))
    assert rc == -4 and b"capacity" in pic.load().pic_last_error()
    buf.close()
    # a store whose bins + spill chunks exceed the static table is refused
    big = pic.Particles2D(4000000, pic.PREC_FP32)
    g = pic.Grid2D(2049, 2049, 1.0, 1.0)     # 512 x 512 bins of 4 x 4 cells
    sim = pic.Simulation(big, g, None, None, pic.sim_params(dt=1e-12, bin_w=4, bin_h=4))
    rc = pic.load().pic_particles_export_positions_static(big._h, 0, None, 0, 0, ctypes.byref(n))
    assert rc == -4 and b"static maximum" in pic.load().pic_last_error()
    sim.close()
