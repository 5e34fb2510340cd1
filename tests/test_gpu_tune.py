"""Sweeps per multigrid level (bit-identical to the oracle in FP64) and the
autotuner: trial pushes leave the simulation unchanged, tuned multigrid
settings still meet the tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mg_gpu(pic, n, rhs, phi0, h, params, pre_l, post_l):
    g = pic.Grid2D(n, n, h * (n - 1), h * (n - 1))
    g.upload(pic.FIELD_PHI, phi0)
    g.upload(pic.FIELD_RHO, rhs)
    rho_p, phi_p, _ = g.device_ptrs()
    mg = pic.Multigrid(n, params)
    mg.set_level_sweeps(pre_l, post_l)
    st = mg.solve(phi_p, rho_p, g.dx)
    return g.download(pic.FIELD_PHI), st


@pytest.mark.parametrize("n", [33, 129, 257])
def test_level_sweeps_fp64_bit_identical(pic, orc, n):
    """Different sweeps on every level, the levels inside the single-workgroup
    coarse cycle included (33: all of them; 257: both kinds)."""
    rng = np.random.default_rng(3 * n)
    rhs = rng.normal(size=(n, n))
    phi0 = np.zeros((n, n))
    phi0[:, -1] = rng.normal(size=n)
    h = 1.0 / (n - 1)
    pre_l = [1, 3, 0, 2, 1, 3, 2, 1]
    post_l = [2, 0, 3, 1, 2, 1, 3, 2]
    params = pic.mg_params(pic.PRESET_STEP, mode=pic.MG_FP64, tol=0.0, max_cycles=3, omega=1.15)
    phi_g, st = _mg_gpu(pic, n, rhs, phi0, h, params, pre_l, post_l)
    phi_o = phi0.copy()
    ost = orc.mg_solve(phi_o, rhs, h, pre=2, post=2, coarse_sweeps=1, omega=1.15, tol=0.0, max_cycles=3,
                       pre_levels=pre_l, post_levels=post_l)
    assert st.cycles == ost.cycles == 3
    assert np.array_equal(phi_g, phi_o)


def _case(pic, seed=13):
    from pic_cuda_amd import configs
    return configs.build(96, 60000, pic.PREC_FP32, mcc=True, seed=seed, gas_density=1e22, te=8.0)


def _sorted(x, y, vx, vy, ids):
    o = np.argsort(ids, kind="stable")
    return [v[o] for v in (x, y, vx, vy, ids)]


def test_bin_tuning_leaves_the_state_unchanged(pic):
    a, b = _case(pic), _case(pic)
    a["sim"].step(3)
    b["sim"].step(3)
    rep = a["sim"].autotune(bins=True, multigrid=False)
    assert rep["trials"] >= 4 and rep["push_ms"] > 0
    a["sim"].step(4)
    b["sim"].step(4)
    for f in (pic.FIELD_RHS, pic.FIELD_PHI):
        assert np.array_equal(a["grid"].download(f), b["grid"].download(f))
    for s in range(2):
        assert np.array_equal(a["sim"].counters(s), b["sim"].counters(s))
    a["sim"].sync()
    b["sim"].sync()
    for s in range(2):
        for u, v in zip(_sorted(*a["particles"].download(s)), _sorted(*b["particles"].download(s))):
            assert np.array_equal(u, v)


def test_multigrid_tuning_keeps_the_tolerance(pic):
    a, b = _case(pic), _case(pic)
    a["sim"].step(3)
    b["sim"].step(3)
    rep = a["sim"].autotune(bins=False, multigrid=True)
    assert rep["trials"] > 10 and rep["solve_cycles"] >= 1
    assert len(rep["pre"]) == rep["levels"] - 1
    assert rep["inner"] in (1, 2, 3)
    # same charge, different solver settings: potentials agree to the tolerance
    a["sim"].step(1)
    b["sim"].step(1)
    assert np.array_equal(a["grid"].download(pic.FIELD_RHS), b["grid"].download(pic.FIELD_RHS))
    pa, pb = a["grid"].download(pic.FIELD_PHI), b["grid"].download(pic.FIELD_PHI)
    assert np.linalg.norm(pa - pb) / np.linalg.norm(pb) < 1e-6
    assert a["sim"].stats().last_mg_rel < 1e-8
