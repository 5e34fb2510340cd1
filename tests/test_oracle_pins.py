"""Pins of the CPU oracle: known-answer vectors, closed forms, invariants and
brute force.  Each check is chosen so that a plausible transcription mistake in
the oracle (a dropped term, a swapped index, a wrong constant) fails it.

These tests run on the CPU only."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- Philox ----
def test_philox_known_answers(orc):
    kat = json.load(open(os.path.join(GOLDEN, "philox4x32_10_kat.json")))
    for v in kat["vectors"]:
        ctr = [int(h, 16) for h in v["ctr"]]
        key = [int(h, 16) for h in v["key"]]
        out = ["%08x" % w for w in orc.philox(ctr, key)]
        assert out == v["out"]


def _philox_py(ctr, key):
    """Philox4x32-10 written from the SC'11 paper's round function."""
    m0, m1, w0, w1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k = list(key)
    for r in range(10):
        if r:
            k = [(k[0] + w0) & 0xFFFFFFFF, (k[1] + w1) & 0xFFFFFFFF]
        p0 = m0 * c[0]
        p1 = m1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & 0xFFFFFFFF, p1 & 0xFFFFFFFF, ((p0 >> 32) ^ c[3] ^ k[1]) & 0xFFFFFFFF,
             p0 & 0xFFFFFFFF]
    return c


def test_philox_random_counters(orc):
    rng = np.random.default_rng(5)
    for _ in range(200):
        ctr = [int(v) for v in rng.integers(0, 2**32, 4)]
        key = [int(v) for v in rng.integers(0, 2**32, 2)]
        assert list(orc.philox(ctr, key)) == _philox_py(ctr, key)


def test_philox2_known_answers(orc):
    """The collision-candidacy generator, Philox2x32-10, against Random123's
    published vectors."""
    kat = json.load(open(os.path.join(GOLDEN, "philox2x32_10_kat.json")))
    for v in kat["vectors"]:
        ctr = [int(h, 16) for h in v["ctr"]]
        out = ["%08x" % w for w in orc.philox2(ctr, int(v["key"], 16))]
        assert out == v["out"]


def test_candidacy_draw_is_uniform_and_keyed(orc):
    """The candidacy uniform of the oracle's MCC (Philox2x32-10 over {id, step},
    key from (seed, case, species)): a chi-square test over 64 cells, no
    correlation between consecutive steps of one particle, and distinct
    streams for neighbouring cases and species."""
    def u(ids, step, key):
        return np.array([orc.philox2([int(i), step], key)[0] >> 8 for i in ids], dtype=np.float64) * 2.0 ** -24

    def key(seed, case, species):
        return int(orc.philox2([seed, case], species)[0])

    ids = np.arange(20000)
    k = key(7, 0, 0)
    a, b = u(ids, 11, k), u(ids, 12, k)
    hist = np.bincount((a * 64).astype(int), minlength=64)
    chi2 = float(((hist - len(a) / 64) ** 2 / (len(a) / 64)).sum())
    assert chi2 < 63 + 5 * math.sqrt(2 * 63)
    assert abs(np.corrcoef(a, b)[0, 1]) < 5 / math.sqrt(len(a))
    keys = {key(7, c, s) for c in range(64) for s in range(4)}
    assert len(keys) == 256
    c = u(ids, 11, key(7, 1, 0))
    assert abs(np.corrcoef(a, c)[0, 1]) < 5 / math.sqrt(len(a))


# ------------------------------------------------------------ det_log -------
def test_det_log_against_libm(orc):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-12, 1.0, 4000), 10.0 ** rng.uniform(-300, 300, 4000),
                         [1.0, 0.5, 2.0, math.e, 0.7071067811865476, 1.4142135623730951, 1e-300, 1e300]])
    for x in xs:
        ref = math.log(x)
        got = orc.det_log(float(x))
        if ref == 0.0:
            assert got == 0.0
        else:
            assert abs(got - ref) <= 2 * math.ulp(ref) + 1e-300, (x, got, ref)


# ------------------------------------------------------------- E field ------
def test_efield_linear_potential_is_uniform(orc):
    nx, ny, dx, dy = 17, 13, 0.25, 0.5
    a, b = 3.0, -1.75
    j, i = np.mgrid[0:ny, 0:nx]
    phi = a * i * dx + b * j * dy
    ex, ey = orc.efield(phi, dx, dy)
    assert np.allclose(ex, -a, rtol=0, atol=1e-12)
    assert np.allclose(ey, -b, rtol=0, atol=1e-12)


def test_efield_quadratic_interior_and_one_sided_edges(orc):
    nx, ny, dx = 11, 9, 0.5
    j, i = np.mgrid[0:ny, 0:nx]
    x = i * dx
    phi = x * x
    ex, ey = orc.efield(phi, dx, dx)
    # central difference of x^2 is exact: -2x
    assert np.allclose(ex[:, 1:-1], -2 * x[:, 1:-1], atol=1e-12)
    # one-sided at the edges (kernels.cu:143-170): -(phi1-phi0)/dx = -dx, -(phi_n-phi_{n-1})/dx
    assert np.allclose(ex[:, 0], -dx, atol=1e-12)
    assert np.allclose(ex[:, -1], -(x[:, -1] ** 2 - x[:, -2] ** 2) / dx, atol=1e-12)
    assert np.allclose(ey, 0.0)


# ------------------------------------------------------------- deposit ------
@pytest.mark.parametrize("prec", [4, 8])
def test_deposit_conserves_charge_exactly(orc, prec):
    rng = np.random.default_rng(2)
    nx = ny = 33
    n = 5000
    x = rng.uniform(0, nx - 1, n)
    y = rng.uniform(0, ny - 1, n)
    acc = orc.deposit(x, y, nx, ny, prec)
    assert int(acc.sum(dtype=np.uint64)) == n * (1 << (2 * orc.QBITS))


def test_deposit_hand_worked_weights(orc):
    nx = ny = 5
    W = 1 << orc.QBITS
    # on a node: everything there
    acc = orc.deposit([2.0], [1.0], nx, ny, 8)
    assert acc[1, 2] == W * W and acc.sum() == W * W
    # cell centre: a quarter on each corner
    acc = orc.deposit([1.5], [2.5], nx, ny, 8)
    q = W * W // 4
    assert acc[2, 1] == q and acc[2, 2] == q and acc[3, 1] == q and acc[3, 2] == q
    # fx = 0.25, fy = 0.75 (exact binary fractions): x weights (3/4,1/4), y (1/4,3/4)
    acc = orc.deposit([0.25], [0.75], nx, ny, 8)
    assert acc[0, 0] == (3 * W // 4) * (W // 4)
    assert acc[0, 1] == (W // 4) * (W // 4)
    assert acc[1, 0] == (3 * W // 4) * (3 * W // 4)
    assert acc[1, 1] == (W // 4) * (3 * W // 4)
    # outside the grid (kernels.cu:53): skipped
    acc = orc.deposit([-0.5, 4.0, 1.0], [1.0, 1.0, 4.2], nx, ny, 8)
    assert acc.sum() == 0


def test_deposit_is_linear(orc):
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 16, 300)
    y = rng.uniform(0, 16, 300)
    a = orc.deposit(x[:100], y[:100], 17, 17, 4)
    b = orc.deposit(x[100:], y[100:], 17, 17, 4)
    c = orc.deposit(x, y, 17, 17, 4)
    assert np.array_equal(a + b, c)


def test_charge_to_rhs(orc):
    acc0 = np.array([1, 2, 3], dtype=np.uint64)
    acc1 = np.array([4, 0, 1 << 40], dtype=np.uint64)
    rhs = orc.charge_to_rhs([acc0, acc1], [-2.0, 0.5])
    assert np.array_equal(rhs, np.array([-2.0 + 2.0, -4.0, -6.0 + 0.5 * 2.0**40]))


# ----------------------------------------------------------- multigrid ------
def _eigen_problem(n):
    h = 1.0 / (n - 1)
    j, i = np.mgrid[0:n, 0:n]
    phi_exact = np.sin(np.pi * i * h) * np.sin(np.pi * j * h)
    lam = -(4.0 / h**2) * (np.sin(np.pi * h / 2) ** 2 + np.sin(np.pi * h / 2) ** 2)
    return h, phi_exact, lam * phi_exact


@pytest.mark.parametrize("n", [17, 65, 129])
def test_mg_discrete_eigenfunction(orc, n):
    h, exact, rhs = _eigen_problem(n)
    phi = np.zeros((n, n))
    st = orc.mg_solve(phi, rhs, h, tol=1e-11, max_cycles=60)
    assert st.rel_residual < 1e-11
    assert np.max(np.abs(phi - exact)) < 1e-9


def _dense_laplacian_solve(rhs, h, bc):
    n = rhs.shape[0]
    m = n - 2
    A = np.zeros((m * m, m * m))
    b = rhs[1:-1, 1:-1].copy().ravel() * h * h
    for jj in range(m):
        for ii in range(m):
            r = jj * m + ii
            A[r, r] = -4.0
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                I, J = ii + di, jj + dj
                if 0 <= I < m and 0 <= J < m:
                    A[r, J * m + I] = 1.0
                else:
                    b[r] -= bc[J + 1, I + 1]
    sol = np.linalg.solve(A, b)
    out = bc.copy()
    out[1:-1, 1:-1] = sol.reshape(m, m)
    return out


def test_mg_matches_dense_direct_solve_with_dirichlet_values(orc):
    n = 17
    rng = np.random.default_rng(4)
    rhs = rng.normal(size=(n, n))
    bc = np.zeros((n, n))
    bc[0, :] = rng.normal(size=n)
    bc[-1, :] = rng.normal(size=n)
    bc[:, 0] = rng.normal(size=n)
    bc[:, -1] = rng.normal(size=n)
    h = 0.1
    ref = _dense_laplacian_solve(rhs, h, bc)
    phi = bc.copy()
    st = orc.mg_solve(phi, rhs, h, tol=1e-13, max_cycles=80)
    assert np.max(np.abs(phi - ref)) < 1e-10 * np.max(np.abs(ref))
    # boundary untouched (Dirichlet)
    assert np.array_equal(phi[0], bc[0]) and np.array_equal(phi[:, -1], bc[:, -1])
    assert st.cycles > 0


def test_mg_ladder_ending_at_4x4(orc):
    # 97 -> 49 -> 25 -> 13 -> 7 -> 4: the 4x4 coarsest level (2x2 unknowns) is
    # swept 32 times; the solve still matches the dense direct solution
    n = 97
    rng = np.random.default_rng(12)
    rhs = rng.normal(size=(n, n))
    rhs[0, :] = rhs[-1, :] = rhs[:, 0] = rhs[:, -1] = 0
    h = 0.05
    phi = np.zeros((n, n))
    st = orc.mg_solve(phi, rhs, h, tol=1e-12, max_cycles=60)
    assert st.rel_residual < 1e-12 and st.cycles < 20
    import ctypes
    res = np.zeros((n, n))
    orc.lib().orc_mg_residual(res.ctypes.data_as(ctypes.c_void_p), phi.ctypes.data_as(ctypes.c_void_p),
                              rhs.ctypes.data_as(ctypes.c_void_p), n, h)
    assert np.linalg.norm(res) < 1e-12 * np.linalg.norm(rhs)


def test_mg_rejects_unsupported_ladder(orc):
    n = 41          # 41 -> 21 -> 11 -> 6: coarsest 6x6 is not supported
    with pytest.raises(RuntimeError):
        orc.mg_solve(np.zeros((n, n)), np.ones((n, n)), 0.1)


def test_mg_vcycle_contraction(orc):
    n = 129
    rng = np.random.default_rng(6)
    rhs = rng.normal(size=(n, n))
    h = 1.0 / (n - 1)
    phi = np.zeros((n, n))
    rels = []
    for k in range(1, 6):
        p = phi.copy()
        st = orc.mg_solve(p, rhs, h, tol=0.0, max_cycles=k)
        rels.append(st.rel_residual)
    for a, b in zip(rels, rels[1:]):
        assert b < 0.2 * a


def test_mg_reference_preset_converges(orc):
    # Multigrid.cu defaults: SOR 1.25, 3 pre / 3 post, 20 coarse sweeps, 5 levels at 129^2
    n = 129
    h, exact, rhs = _eigen_problem(n)
    phi = np.zeros((n, n))
    st = orc.mg_solve(phi, rhs, h, levels=5, pre=3, post=3, coarse_sweeps=20, omega=1.25, tol=1e-9,
                      max_cycles=50)
    assert st.rel_residual < 1e-9


def test_mg_transfer_operators(orc):
    # restriction of a constant interior field is the full-weighting average;
    # prolongation of a bilinear coarse function is exact
    nc = 9
    nf = 2 * (nc - 1) + 1
    J, I = np.mgrid[0:nc, 0:nc]
    coarse = (2.0 * I + 3.0 * J + 1.0).astype(np.float64)
    fine = np.zeros((nf, nf))
    import ctypes
    lib = orc.lib()
    lib.orc_mg_prolong(fine.ctypes.data_as(ctypes.c_void_p), coarse.ctypes.data_as(ctypes.c_void_p), nc)
    jf, i_f = np.mgrid[0:nf, 0:nf]
    assert np.allclose(fine, 2.0 * (i_f / 2.0) + 3.0 * (jf / 2.0) + 1.0)
    f = np.ones((nf, nf))
    c = np.zeros((nc, nc))
    lib.orc_mg_restrict(c.ctypes.data_as(ctypes.c_void_p), f.ctypes.data_as(ctypes.c_void_p), nf)
    assert np.allclose(c[1:-1, 1:-1], 1.0)
    assert np.isclose(c[0, 0], 9.0 / 16.0)      # corner keeps 4+2+2+1 of 16
    assert np.allclose(c[0, 1:-1], 12.0 / 16.0)  # edge keeps 4+2+2+2+1+1 of 16


# --------------------------------------------------------------- walls ------
def _hits_exact(p1, p2, w):
    """Segment intersection in exact rational arithmetic (inclusive)."""
    x1, y1 = map(Fraction, p1)
    x2, y2 = map(Fraction, p2)
    x3, y3, x4, y4 = map(Fraction, w)
    den = (y4 - y3) * (x2 - x1) - (x4 - x3) * (y2 - y1)
    if den == 0:
        return False
    ua = ((x4 - x3) * (y1 - y3) - (y4 - y3) * (x1 - x3)) / den
    ub = ((x2 - x1) * (y1 - y3) - (y2 - y1) * (x1 - x3)) / den
    return 0 <= ua <= 1 and 0 <= ub <= 1


def test_segment_test_brute_force(orc):
    rng = np.random.default_rng(7)
    # coordinates on a dyadic lattice: every intermediate is exact in FP64
    for _ in range(3000):
        p = rng.integers(0, 16, 8) / 4.0
        w = p[4:]
        got = orc.segment_hits(p[0], p[1], p[2], p[3], w, eps=1e-12)
        assert got == _hits_exact(p[:2], p[2:4], w), p


def test_segment_test_special_cases(orc):
    wall = [1.0, 0.0, 1.0, 2.0]
    assert orc.segment_hits(0.5, 1.0, 1.0, 1.0, wall)       # ends on the wall (ua = 1 inclusive)
    assert orc.segment_hits(1.0, 1.0, 1.5, 1.0, wall)       # starts on the wall (ua = 0 inclusive)
    assert orc.segment_hits(0.5, 0.5, 1.5, -0.5, wall)      # through the wall's end point (ub = 0)
    assert not orc.segment_hits(0.5, 3.0, 1.5, 3.0, wall)   # passes beyond the wall
    assert not orc.segment_hits(1.0, 0.5, 1.0, 1.5, wall)   # parallel (collinear) is cut
    assert not orc.segment_hits(0.5, 0.0, 0.5, 2.0, wall)   # parallel


# ----------------------------------------------------------------- push -----
def _species(orc, prec, n, rng, nx, kq=1e-3, v=0.0):
    x = rng.uniform(2, nx - 3, n)
    y = rng.uniform(2, nx - 3, n)
    vx = np.full(n, v)
    vy = np.zeros(n)
    return orc.Species(prec, x, y, vx, vy, np.arange(n, dtype=np.uint32), kq)


@pytest.mark.parametrize("prec", [4, 8])
def test_push_uniform_field_closed_form(orc, prec):
    rng = np.random.default_rng(8)
    nx = 33
    dx = 0.01
    E0 = 250.0
    j, i = np.mgrid[0:nx, 0:nx]
    phi = -E0 * i * dx
    sp = _species(orc, prec, 500, rng, nx, kq=2e-4)
    x0, vx0 = sp.x.copy(), sp.vx.copy()
    dtdx = 1e-3
    orc.push(sp, phi, nx, nx, dx, dx, dtdx, dtdx, dict(kick=True, drift=True), 0, 1, 0, 0)
    order = np.argsort(sp.id)
    tol = 1e-5 if prec == 4 else 1e-12
    dv = 2e-4 * E0
    assert np.allclose(sp.vx[order], vx0 + dv, rtol=tol, atol=tol * dv)
    assert np.allclose(sp.x[order], x0 + (vx0 + dv) * dtdx, rtol=tol, atol=tol)
    assert sp.counters[orc.C_ABSORBED] == 0


def test_push_box_walls_absorb_and_compact_stably(orc):
    nx = 17
    x = np.array([1.0, 15.5, 8.0, 0.3, 10.0])
    vx = np.array([0.0, 1000.0, 0.0, -500.0, 0.0])   # second and fourth leave the box
    sp = orc.Species(8, x, np.full(5, 8.0), vx, np.zeros(5), np.arange(5, dtype=np.uint32), 0.0)
    acc = np.zeros(nx * nx, dtype=np.uint64)
    orc.push(sp, np.zeros((nx, nx)), nx, nx, 1.0, 1.0, 1e-3, 1e-3, dict(drift=True), 0, 1, 0, 0, acc_self=acc)
    assert list(sp.id) == [0, 2, 4]                  # survivors in input order
    assert sp.counters[orc.C_ABSORBED] == 2
    assert int(acc.sum()) == 3 * (1 << 40)


# ------------------------------------------------------------------ MCC -----
def _mcc_const(orc, n_gas, sigma, m, dt, K=64, emax=100.0, types=(0,), thr=(0.0,), extra=None):
    QE = 1.602176634e-19
    de = emax / (K - 1)
    tab = np.full((len(types), K), n_gas * sigma)
    E = np.arange(K) * de
    if extra is not None:
        tab = extra(E)
    efac = 0.5 * m / QE
    vmax = math.sqrt(emax / efac)
    numax = float(np.max(tab.sum(axis=0)) * vmax) * 1.001
    return dict(n_proc=len(types), type=list(types), threshold=list(thr), table=tab.ravel(), K=K,
                inv_de=1.0 / de, nu_max=numax, p_null=-math.expm1(-numax * dt), efac=efac, gas_vth=300.0)


@pytest.mark.parametrize("prec", [4, 8])
def test_mcc_elastic_rate_and_speed(orc, prec):
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    n = 200000
    dt = 2e-9
    m = _mcc_const(orc, 1e21, 1e-19, ME, dt)
    v = math.sqrt(2 * 10.0 * QE / ME)            # 10 eV electrons
    sp = orc.Species(prec, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v), np.zeros(n),
                     np.arange(n, dtype=np.uint32), 0.0, mcc=m)
    orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 3, 9, 0, 0)
    hits = int(sp.counters[orc.C_PROC0])
    p = m["p_null"] * (1e21 * 1e-19 * v) / m["nu_max"]
    mu, sd = n * p, math.sqrt(n * p * (1 - p))
    assert abs(hits - mu) < 5 * sd
    speed = np.hypot(sp.vx, sp.vy)
    assert np.allclose(speed, v, rtol=3e-6 if prec == 4 else 1e-13)
    scattered = sp.vy != 0
    assert scattered.sum() == hits
    ang = np.arctan2(sp.vy[scattered], sp.vx[scattered])
    hist, _ = np.histogram(ang, bins=8, range=(-np.pi, np.pi))
    assert np.all(np.abs(hist - hits / 8) < 6 * np.sqrt(hits / 8))


def test_mcc_excitation_and_ionization_energy_bookkeeping(orc):
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    n = 100000
    dt = 5e-9
    K, emax = 256, 100.0
    E = np.arange(K) * emax / (K - 1)

    def tables(Eg):
        exc = np.where(Eg > 11.5, 1e21 * 4e-20, 0.0)
        ion = np.where(Eg > 15.8, 1e21 * 6e-20, 0.0)
        return np.stack([exc, ion])

    m = _mcc_const(orc, 1e21, 0, ME, dt, K=K, emax=emax, types=(1, 2), thr=(11.5, 15.8), extra=tables)
    e0 = 40.0
    v = math.sqrt(2 * e0 * QE / ME)
    sp = orc.Species(8, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v), np.zeros(n), np.arange(n, dtype=np.uint32),
                     0.0, mcc=m, other=1)
    born = orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 0, 4, 0, 0)
    n_exc, n_ion = int(sp.counters[orc.C_PROC0]), int(sp.counters[orc.C_PROC0 + 1])
    assert n_exc > 100 and n_ion > 100
    assert sp.n == n + n_ion and born is not None and len(born[0]) == n_ion
    energy = m["efac"] * (sp.vx**2 + sp.vy**2)
    prim = energy[:n]
    assert np.sum(np.isclose(prim, e0 - 11.5, rtol=1e-12)) == n_exc
    # primary + secondary energy = E - threshold for every ionization
    ids = sp.id[:n]
    changed = ~np.isclose(prim, e0, rtol=1e-12) & ~np.isclose(prim, e0 - 11.5, rtol=1e-12)
    assert changed.sum() == n_ion
    assert np.isclose(prim[changed].sum() + energy[n:].sum(), n_ion * (e0 - 15.8), rtol=1e-9)
    # born electrons and ions sit at the collision point (no drift in this pass)
    assert np.all(born[0] == 8.0) and np.all(sp.x[n:] == 8.0)
    assert len(set(ids.tolist())) == n


def test_mcc_is_keyed_on_id_and_step(orc):
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    m = _mcc_const(orc, 1e21, 1e-19, ME, 5e-9)
    v = math.sqrt(2 * 5.0 * QE / ME)
    n = 20000

    def run(step, perm):
        sp = orc.Species(4, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v), np.zeros(n),
                         np.arange(n, dtype=np.uint32)[perm], 0.0, mcc=m)
        orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), step, 11, 0, 0)
        o = np.argsort(sp.id)
        return sp.vx[o], sp.vy[o]

    a = run(5, np.arange(n))
    b = run(5, np.random.default_rng(1).permutation(n))
    c = run(6, np.arange(n))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[1], c[1])


# --------------------------------------------------------------- init -------
def test_init_statistics(orc):
    n = 200000
    x, y, vx, vy, ids = orc.init_particles(8, n, 65, 65, 3.0, 1, 0, 0)
    assert x.min() > 0 and x.max() < 64 and y.min() > 0 and y.max() < 64
    assert abs(x.mean() - 32) < 0.2
    assert abs(vx.std() - 3.0) < 0.03 and abs(vy.mean()) < 0.03
    assert np.array_equal(ids, np.arange(n))
    # kurtosis of a Gaussian is 3
    k = np.mean(vx**4) / np.mean(vx**2) ** 2
    assert abs(k - 3.0) < 0.05


def test_init_uniform_statistics(orc):
    """initializeParticles' distributions (main.cu:17-32): uniform box and
    uniform velocity components; moments of U(a, b), independence of the
    components, keyed on the particle id."""
    n = 200000
    x, y, vx, vy, ids = orc.init_uniform(8, n, 12.8, 51.2, 6.4, 25.6, -0.01, 0.01, 3, 0, 0)
    assert x.min() >= 12.8 and x.max() < 51.2 and y.min() >= 6.4 and y.max() < 25.6
    assert vx.min() >= -0.01 and vx.max() < 0.01
    for a, lo, hi in ((x, 12.8, 51.2), (y, 6.4, 25.6), (vx, -0.01, 0.01), (vy, -0.01, 0.01)):
        w = hi - lo
        assert abs(a.mean() - 0.5 * (lo + hi)) < 4 * w / np.sqrt(12 * n)
        assert abs(a.var() / (w * w / 12.0) - 1.0) < 0.02
    for a, b in ((x, y), (vx, vy), (x, vx)):
        assert abs(np.corrcoef(a, b)[0, 1]) < 0.01
    # keyed on id: a later slice with id0 reproduces the same particles
    x2, y2, vx2, vy2, ids2 = orc.init_uniform(8, 1000, 12.8, 51.2, 6.4, 25.6, -0.01, 0.01, 3, 0, 0, id0=5000)
    assert np.array_equal(x2, x[5000:6000]) and np.array_equal(vy2, vy[5000:6000])
    assert np.array_equal(ids2, np.arange(5000, 6000))
    # a separate stream from the Maxwellian init of the same ids
    xm = orc.init_particles(8, 1000, 65, 65, 1.0, 3, 0, 0)[0]
    assert abs(np.corrcoef(xm, x[:1000])[0, 1]) < 0.15


def test_openmp_oracle_pass_equals_serial(orc):
    """The OpenMP build of the oracle (used only for the CPU baseline) gives the
    serial result: same survivors in the same order, same born particles, same
    fixed-point charge and counters."""
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    n, nx = 120000, 33
    dt = 5e-9
    K, emax = 256, 100.0

    def tables(Eg):
        el = np.full_like(Eg, 1e21 * 2e-20)
        ion = np.where(Eg > 15.8, 1e21 * 6e-20, 0.0)
        return np.stack([el, ion])

    m = _mcc_const(orc, 1e21, 0, ME, dt, K=K, emax=emax, types=(0, 2), thr=(0.0, 15.8), extra=tables)
    rng = np.random.default_rng(5)
    x0 = rng.uniform(0.5, nx - 1.5, n)
    y0 = rng.uniform(0.5, nx - 1.5, n)
    v = math.sqrt(2 * 30.0 * QE / ME)
    ang = rng.uniform(0, 2 * np.pi, n)
    phi = rng.normal(size=(nx, nx))
    out = []
    for omp in (False, True):
        sp = orc.Species(4, x0.astype(np.float32), y0.astype(np.float32), (v * np.cos(ang)).astype(np.float32),
                         (v * np.sin(ang)).astype(np.float32), np.arange(n, dtype=np.uint32), -1e3, mcc=m, other=1)
        acc = np.zeros(nx * nx, dtype=np.uint64)
        acc_o = np.zeros(nx * nx, dtype=np.uint64)
        born = orc.push(sp, phi, nx, nx, 1.0, 1.0, 1e-7, 1e-7, dict(kick=True, mcc=True, drift=True), 2, 3, 0, 0,
                        acc_self=acc, acc_other=acc_o, omp=omp)
        out.append((sp, acc, acc_o, born))
    (a, acc_a, acco_a, born_a), (b, acc_b, acco_b, born_b) = out
    assert orc.lib(True).orc_num_threads() >= 1
    assert a.n == b.n and a.n > n // 2
    for u, w in ((a.x, b.x), (a.y, b.y), (a.vx, b.vx), (a.vy, b.vy), (a.id, b.id)):
        assert np.array_equal(u, w)
    assert np.array_equal(acc_a, acc_b) and np.array_equal(acco_a, acco_b)
    assert np.array_equal(a.counters, b.counters) and a.counters[orc.C_PROC0 + 1] > 0
    for u, w in zip(born_a, born_b):
        assert np.array_equal(u, w)


def _bands_const(m, n_gas, sigma, emax, dt):
    """band majorants of a constant cross section: n sigma v at each band's top"""
    QE = 1.602176634e-19
    nu = []
    for b in range(16):
        hi = emax if b == 15 else math.ldexp(emax, b - 14)
        nu.append(n_gas * sigma * math.sqrt(2 * hi * QE / m) * 1.001)
    return nu, [-math.expm1(-v * dt) for v in nu]


def test_mcc_energy_band_edges(orc):
    """Band b holds [emax 2^(b-15), emax 2^(b-14)); band 0 everything below
    emax 2^-14, band 15 everything from emax/2 on.  With only band k made a
    certain candidate (and a majorant no real collision reaches), the
    null-candidate counter says which band a particle fell in."""
    QE = 1.602176634e-19
    m_unit = QE                      # efac = 1/2: E = v^2 / 2 exactly for v = 2^a
    emax = 256.0
    m = _mcc_const(orc, 1.0, 1e-30, m_unit, 1e-9, K=257, emax=emax)
    m["emax"] = emax

    def band_of(vx):
        found = []
        for k in range(16):
            mk = dict(m)
            mk["band_p_null"] = [1.0 if b == k else 0.0 for b in range(16)]
            mk["band_nu_max"] = [1e300] * 16
            sp = orc.Species(8, np.array([8.0]), np.array([8.0]), np.array([vx]), np.array([0.0]),
                             np.array([7], dtype=np.uint32), 0.0, mcc=mk)
            orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 0, 1, 0, 0)
            if sp.counters[orc.C_NULL_CAND]:
                found.append(k)
        assert len(found) == 1
        return found[0]

    for a in range(-4, 6):
        e = 0.5 * 4.0 ** a           # energy of vx = 2^a
        r = e / emax
        want = 0 if r < 2.0 ** -14 else min(15, 15 + int(math.floor(math.log2(r))))
        assert band_of(2.0 ** a) == want
        # just below a band edge: the band underneath (r is a power of two)
        below = np.nextafter(2.0 ** a, 0.0)
        if r <= 2.0 ** -14:
            want_below = 0
        elif r > 0.5:
            want_below = 15
        else:
            want_below = 14 + int(math.log2(r))
        assert band_of(below) == want_below
    assert band_of(0.0) == 0
    assert band_of(float(np.sqrt(2 * 1000.0))) == 15     # above the table end


def test_mcc_banded_rate(orc):
    """10 eV electrons, constant elastic sigma: collisions occur at
    p_null(band) * nu(E) / nu_max(band), the null-collision rate of the
    particle's own band."""
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    n, dt, emax = 200000, 2e-9, 100.0
    m = _mcc_const(orc, 1e21, 1e-19, ME, dt, emax=emax)
    nu_b, p_b = _bands_const(ME, 1e21, 1e-19, emax, dt)
    m.update(emax=emax, band_nu_max=nu_b, band_p_null=p_b)
    e0 = 10.0
    v = math.sqrt(2 * e0 * QE / ME)
    sp = orc.Species(8, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v), np.zeros(n),
                     np.arange(n, dtype=np.uint32), 0.0, mcc=m)
    orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 3, 9, 0, 0)
    b = 15 + int(math.floor(math.log2(e0 / emax)))
    cand = int(sp.counters[orc.C_NULL_CAND])
    assert abs(cand - n * p_b[b]) < 5 * math.sqrt(n * p_b[b])
    hits = int(sp.counters[orc.C_PROC0])
    p = p_b[b] * (1e21 * 1e-19 * v) / nu_b[b]
    assert abs(hits - n * p) < 5 * math.sqrt(n * p * (1 - p))
    # the band's majorant is that of its top energy, 12.5 eV, not of 100 eV
    assert abs(nu_b[b] / m["nu_max"] - math.sqrt(12.5 / 100.0)) < 0.01


def _masked_dense(rhs, mask, phi_fixed, h):
    """direct solve of the 5-point system on the solved nodes, fixed nodes
    (boundary and mask != 0) moved to the right-hand side"""
    n = rhs.shape[0]
    free = np.zeros((n, n), dtype=bool)
    free[1:-1, 1:-1] = mask[1:-1, 1:-1] == 0
    idx = -np.ones((n, n), dtype=int)
    idx[free] = np.arange(free.sum())
    A = np.zeros((free.sum(), free.sum()))
    b = np.zeros(free.sum())
    for j, i in zip(*np.nonzero(free)):
        k = idx[j, i]
        A[k, k] = -4.0 / h**2
        b[k] = rhs[j, i]
        for dj, di in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            jj, ii = j + dj, i + di
            if free[jj, ii]:
                A[k, idx[jj, ii]] = 1.0 / h**2
            else:
                b[k] -= phi_fixed[jj, ii] / h**2
    out = phi_fixed.copy()
    out[free] = np.linalg.solve(A, b)
    return out


def test_mg_masked_all_fluid_equals_unmasked(orc):
    n = 65
    rng = np.random.default_rng(41)
    rhs = rng.normal(size=(n, n))
    a, b = np.zeros((n, n)), np.zeros((n, n))
    orc.mg_solve(a, rhs, 1.0 / (n - 1), tol=0.0, max_cycles=3, omega=1.2)
    orc.mg_solve(b, rhs, 1.0 / (n - 1), tol=0.0, max_cycles=3, omega=1.2, mask=np.zeros((n, n), np.uint8))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [33, 49])
def test_mg_embedded_electrode_matches_direct_solve(orc, n):
    """A Dirichlet disc at +1 V and a solid block kept at -0.5 V inside a
    grounded box with a charge blob: the masked solve reaches the direct solve
    of the masked 5-point system; fixed nodes keep their values."""
    h = 1.0 / (n - 1)
    y, x = np.mgrid[0:n, 0:n]
    mask = np.zeros((n, n), np.uint8)
    disc = (x - 0.35 * n) ** 2 + (y - 0.5 * n) ** 2 < (0.12 * n) ** 2
    mask[disc] = 1
    block = (x > 0.6 * n) & (x < 0.75 * n) & (y > 0.3 * n) & (y < 0.45 * n)
    mask[block] = 2
    phi_bc = np.where(disc, 1.0, 0.0)
    rhs = -50.0 * np.exp(-((x - 0.6 * n) ** 2 + (y - 0.7 * n) ** 2) / (0.05 * n) ** 2)
    phi = np.zeros((n, n))
    phi[block] = -0.5
    st = orc.mg_solve(phi, rhs, h, tol=1e-12, max_cycles=200, omega=1.0, mask=mask, phi_bc=phi_bc)
    assert st.rel_residual < 1e-12
    assert np.all(phi[disc] == 1.0) and np.all(phi[block] == -0.5)
    fixed = np.where(disc, 1.0, np.where(block, -0.5, 0.0))
    ref = _masked_dense(rhs, mask, fixed, h)
    assert np.max(np.abs(phi - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


# ------------------------------------------------- gather weights, selection --
@pytest.mark.parametrize("prec", [4, 8])
def test_push_gather_quadratic_potential_closed_form(orc, prec):
    """CIC gather (kernels.cu:106-127) against a closed form that sees every
    weight and node.  phi = a X^2 + b Y^2 + c X Y (X = i dx, Y = j dy) has
    central differences equal to its gradient at every interior node, so the
    node field is E = -(2a X + c Y, 2b Y + c X): linear in x and y, which
    bilinear interpolation reproduces exactly anywhere inside a cell of
    interior nodes.  Swapped weights or nodes (w10 <-> w01, a node one row
    off), a missing term or a transposed grid give a different E wherever
    fx != fy.  dx != dy and nx != ny catch swapped spacings and axes."""
    nx, ny = 33, 25
    dx, dy = 0.013, 0.007
    a, b, c = 3.0e3, -1.7e3, 2.3e3
    j, i = np.mgrid[0:ny, 0:nx]
    X, Y = i * dx, j * dy
    phi = a * X * X + b * Y * Y + c * X * Y
    rng = np.random.default_rng(81)
    n = 4000
    # cells whose four nodes are interior: ix in [1, nx-3], iy in [1, ny-3]
    x = rng.uniform(1.0, nx - 2.0, n)
    y = rng.uniform(1.0, ny - 2.0, n)
    vx0 = rng.normal(size=n)
    vy0 = rng.normal(size=n)
    kq = 1e-4
    sp = orc.Species(prec, x, y, vx0, vy0, np.arange(n, dtype=np.uint32), kq)
    xs, ys, vxs, vys = sp.x.astype(np.float64), sp.y.astype(np.float64), sp.vx.copy(), sp.vy.copy()
    orc.push(sp, phi, nx, ny, dx, dy, 0.0, 0.0, dict(kick=True), 0, 1, 0, 0, box=False)
    o = np.argsort(sp.id)
    Ex = -(2 * a * xs * dx + c * ys * dy)
    Ey = -(2 * b * ys * dy + c * xs * dx)
    scale = kq * max(np.abs(Ex).max(), np.abs(Ey).max())
    tol = 2e-6 if prec == 4 else 1e-11
    assert np.max(np.abs(sp.vx[o] - (vxs + kq * Ex))) < tol * max(scale, 1.0)
    assert np.max(np.abs(sp.vy[o] - (vys + kq * Ey))) < tol * max(scale, 1.0)
    # the field varies by far more than the tolerance across a cell
    assert kq * 2 * abs(a) * dx > 100 * tol * max(scale, 1.0)


def _mcc_tables(K, emax, procs, below=0.0):
    """procs: list of (type, threshold eV, n*sigma) with constant sigma above
    threshold and `below` times that under it"""
    return lambda Eg: np.stack([np.where(Eg >= thr, ns, below * ns) for _, thr, ns in procs])


def test_mcc_process_selection_frequencies(orc):
    """Null-collision selection among several processes (the build's
    replacement of collision_kernel, mcc.cu:56-85, for a set of channels):
    a candidate with R = u1 nu_max takes the first process whose cumulative
    rate exceeds R, so process k is chosen with probability
    p_null * nu_k(E) / nu_max, and a process above the particle's energy
    never.  Checked within binomial bounds for electrons above both inelastic
    thresholds and between them.  The tables hold the same sigma below the
    thresholds too, so only the threshold test keeps a process out.  A
    non-cumulative test (R < nu_k), reusing the candidacy uniform, or
    ignoring thresholds each fails here."""
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    K, emax, dt = 256, 100.0, 2e-9
    procs = [(0, 0.0, 1e21 * 1e-19), (1, 11.5, 1e21 * 3e-20), (2, 15.8, 1e21 * 5e-20)]
    m = _mcc_const(orc, 1e21, 0, ME, dt, K=K, emax=emax, types=[p[0] for p in procs],
                   thr=[p[1] for p in procs], extra=_mcc_tables(K, emax, procs, below=1.0))
    for e0, active in ((40.0, (0, 1, 2)), (13.0, (0, 1))):
        n = 300000
        v = math.sqrt(2 * e0 * QE / ME)
        sp = orc.Species(8, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v), np.zeros(n),
                         np.arange(n, dtype=np.uint32), 0.0, mcc=m, other=1)
        orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 4, 21, 0, 0)
        cand = int(sp.counters[orc.C_NULL_CAND])
        assert abs(cand - n * m["p_null"]) < 5 * math.sqrt(n * m["p_null"])
        for k, (_, thr, ns) in enumerate(procs):
            hits = int(sp.counters[orc.C_PROC0 + k])
            if k not in active:
                assert hits == 0
                continue
            p = m["p_null"] * ns * v / m["nu_max"]
            assert abs(hits - n * p) < 5 * math.sqrt(n * p * (1 - p)), (e0, k, hits, n * p)
            assert hits > 1000


def _gauss_moments_ok(vx, vy, sigma):
    """mean 0, variance sigma^2, kurtosis 3 per component and uncorrelated
    components, each within 5 standard errors"""
    n = len(vx)
    for a in (vx, vy):
        assert abs(a.mean()) < 5 * sigma / math.sqrt(n)
        var = a.var()
        assert abs(var / sigma**2 - 1) < 5 * math.sqrt(2.0 / n)
        kurt = np.mean((a - a.mean()) ** 4) / var**2
        assert abs(kurt - 3.0) < 5 * math.sqrt(24.0 / n)
    assert abs(np.corrcoef(vx, vy)[0, 1]) < 5 / math.sqrt(n)


def test_mcc_charge_exchange_takes_gas_velocity(orc):
    """Ar+ charge exchange: the ion takes the velocity of a gas atom, a 2D
    Maxwellian of thermal speed gas_vth per component (Gaussian with mean 0,
    variance gas_vth^2, kurtosis 3, independent components), whatever its
    velocity before."""
    QE, MAR = 1.602176634e-19, 6.6335209e-26
    K, emax, dt = 64, 100.0, 1e-7
    m = _mcc_const(orc, 1e21, 5e-19, MAR, dt, K=K, emax=emax, types=(3,), thr=(0.0,))
    assert m["gas_vth"] == 300.0
    n = 200000
    v0 = 20000.0
    sp = orc.Species(8, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v0), np.zeros(n),
                     np.arange(n, dtype=np.uint32), 0.0, mcc=m)
    orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 2, 17, 0, 1)
    hit = sp.vx != v0
    hits = int(sp.counters[orc.C_PROC0])
    assert hit.sum() == hits and hits > 50000
    _gauss_moments_ok(sp.vx[hit], sp.vy[hit], 300.0)
    assert np.all(sp.vy[~hit] == 0.0)


def test_mcc_ionization_products(orc):
    """Ionization (the north-star channel beyond mcc.cu:88-112): of the energy
    left after the threshold, E - thr, the primary keeps a uniform fraction
    and the secondary electron the rest; both leave isotropically, and the
    born ion takes a gas-atom velocity (Gaussian, gas_vth per component)."""
    QE, ME = 1.602176634e-19, 9.1093837015e-31
    K, emax, dt, thr = 256, 100.0, 3e-9, 15.8
    procs = [(2, thr, 1e21 * 6e-20)]
    m = _mcc_const(orc, 1e21, 0, ME, dt, K=K, emax=emax, types=(2,), thr=(thr,),
                   extra=_mcc_tables(K, emax, procs))
    n = 150000
    e0 = 90.0
    v = math.sqrt(2 * e0 * QE / ME)
    sp = orc.Species(8, np.full(n, 8.0), np.full(n, 8.0), np.full(n, v), np.zeros(n),
                     np.arange(n, dtype=np.uint32), 0.0, mcc=m, other=1)
    born = orc.push(sp, np.zeros((17, 17)), 17, 17, 1.0, 1.0, 0.0, 0.0, dict(mcc=True), 1, 5, 0, 0)
    n_ion = int(sp.counters[orc.C_PROC0])
    assert n_ion > 40000 and len(born[0]) == n_ion and sp.n == n + n_ion
    er = e0 - thr
    prim = m["efac"] * (sp.vx[:n] ** 2 + sp.vy[:n] ** 2)
    ionized = ~np.isclose(prim, e0, rtol=1e-12)
    assert ionized.sum() == n_ion
    frac = np.sort(prim[ionized] / er)
    # Kolmogorov-Smirnov distance to U(0, 1); 1.95/sqrt(n) is the 0.1 % level
    cdf = np.arange(1, n_ion + 1) / n_ion
    d = max(np.max(cdf - frac), np.max(frac - (cdf - 1.0 / n_ion)))
    assert d < 1.95 / math.sqrt(n_ion)
    # the secondary takes the rest, in the order the primaries were ionized
    sec = m["efac"] * (sp.vx[n:] ** 2 + sp.vy[n:] ** 2)
    assert np.allclose(prim[ionized] + sec, er, rtol=1e-12, atol=1e-12 * er)
    # directions of primary and secondary: uniform angles, independent of each other
    for vx_, vy_ in ((sp.vx[:n][ionized], sp.vy[:n][ionized]), (sp.vx[n:], sp.vy[n:])):
        ang = np.arctan2(vy_, vx_)
        hist, _ = np.histogram(ang, bins=12, range=(-np.pi, np.pi))
        assert np.all(np.abs(hist - n_ion / 12) < 5 * np.sqrt(n_ion / 12))
    a1 = np.arctan2(sp.vy[:n][ionized], sp.vx[:n][ionized])
    a2 = np.arctan2(sp.vy[n:], sp.vx[n:])
    assert abs(np.corrcoef(np.cos(a1), np.cos(a2))[0, 1]) < 5 / math.sqrt(n_ion)
    # born ions: gas-atom velocities
    _gauss_moments_ok(born[2], born[3], 300.0)
    assert np.all(born[0] == 8.0) and np.all(born[1] == 8.0)
