"""Workload configurations of the 2D electrostatic PIC step (SI units).

``plain_small``  128^2 cells, 1e5 e- + 1e5 Ar+, no collisions, FP64, 100 steps.
``full_step``    1024^2 cells, 2e7 e- + 2e7 Ar+, null-collision MCC on an Ar
                 background of 1e21 m^-3, FP32 particles, 1000 steps.
``poisson``      2049^2 Dirichlet grid with 200 Gaussian charge blobs + noise.
``ensemble_cases``  64 plasmas, cells log-uniform 64..512, 100 particles per
                 cell and species (clamped to 1e4..4e6), Te ~ U(1, 5) eV, MCC on.
``serve_jobs``   a Poisson-arrival stream of jobs, cells log-uniform 64..1024,
                 100 particles per cell and species, steps U(200, 2000).

Plasma: n = 1e15 m^-3, Te = 2 eV, Ti = 0.03 eV, dx = 0.5 Debye lengths,
dt = 0.2 / omega_pe, grounded absorbing walls on the four box edges.
Synthetic cross sections: elastic e-Ar 1e-19 m^2 (constant); excitation
11.5 eV and ionization 15.8 eV with the Lotz-like shape
sigma = peak * e * ln(u)/u (u = E/threshold, maximum `peak` at u = e); Ar+
charge exchange 5e-19 m^2 (constant).  Neutral gas at 0.03 eV.
"""
import math

import numpy as np

from . import (AMU, ELEMENTARY_CHARGE as QE, ELECTRON_MASS as ME, EPS0, Boundaries2D, CrossSections, Grid2D,
               Particles2D, Simulation, mg_params, sim_params, PREC_FP32, PREC_FP64, ELASTIC, EXCITATION,
               IONIZATION, CHARGE_EXCHANGE, MG_FP64, MG_MIXED, PRESET_STEP, PRESET_STEP_SOR)

M_AR = 39.948 * AMU


class Plasma:
    """Derived constants of a uniform electron/Ar+ plasma."""

    def __init__(self, density=1e15, te=2.0, ti=0.03, cells=128, dx_debye=0.5, dt_wpe=0.2):
        self.density, self.te, self.ti, self.cells = density, te, ti, cells
        self.debye = math.sqrt(EPS0 * te / (density * QE))
        self.dx = dx_debye * self.debye
        self.length = cells * self.dx
        self.wpe = math.sqrt(density * QE * QE / (EPS0 * ME))
        self.dt = dt_wpe / self.wpe
        self.vte = math.sqrt(te * QE / ME)
        self.vti = math.sqrt(ti * QE / M_AR)
        self.nodes = cells + 1


def add_argon_processes(xs, electron=0, ion=1):
    """The synthetic e-Ar / Ar+-Ar process set."""
    xs.add_analytic(electron, ELASTIC, 0.0, 1e-19)
    xs.add_analytic(electron, EXCITATION, 11.5, 2e-21)
    xs.add_analytic(electron, IONIZATION, 15.8, 3e-20, product_species=ion)
    xs.add_analytic(ion, CHARGE_EXCHANGE, 0.0, 5e-19)


def build(cells, n_per_species, precision, mcc, seed=1, case_id=0, mg=None, bin_w=16, bin_h=16, halo=2,
          gas_density=1e21, gas_t=0.03, te=2.0, capacity=None, walls=None, processes=None, state=None,
          rebin_every=0):
    """Create particles, grid, walls, cross sections and the step driver.
    `walls`: internal wall segments (x1, y1, x2, y2) in cell units, added to
    the absorbing box. `processes(xs, electron, ion)` fills the cross-section
    set (default: the synthetic e-Ar / Ar+-Ar set). `state`: per species
    (x, y, vx, vy, ids) in cell units, uploaded instead of the Maxwellian
    initialization (the macro-particle weight still follows n_per_species)."""
    pl = Plasma(te=te, cells=cells)
    nn = pl.nodes
    grid = Grid2D(nn, nn, pl.length, pl.length)
    cap = capacity or int(n_per_species * 1.02) + 4096
    parts = Particles2D(cap, precision)
    parts.change_species(0, ME, -QE)
    ion = parts.add_species(M_AR, QE)
    w = pl.density * pl.length * pl.length / n_per_species
    parts.set_weight(0, w)
    parts.set_weight(ion, w)
    if state is None:
        parts.init_maxwellian(0, n_per_species, nn, nn, pl.vte, seed, case_id, 0)
        parts.init_maxwellian(ion, n_per_species, nn, nn, pl.vti, seed, case_id, 0)
    else:
        for s, arrays in enumerate(state):
            parts.upload(s, *arrays)
    bnd = Boundaries2D(pl.length, pl.length, nn, nn)
    bnd.add_box()
    for (x1, y1, x2, y2) in (walls if walls is not None else ()):
        bnd.add_wall(x1 * grid.dx, y1 * grid.dy, x2 * grid.dx, y2 * grid.dy)
    bnd.compile_walls()
    xs = None
    if mcc:
        xs = CrossSections(gas_density, gas_t)
        (processes or add_argon_processes)(xs, 0, ion)
    if mg is None:
        mg = mg_params(PRESET_STEP_SOR, mode=MG_MIXED if precision == PREC_FP32 else MG_FP64)
    sim = Simulation(parts, grid, bnd, xs, sim_params(dt=pl.dt, seed=seed, case_id=case_id, bin_w=bin_w,
                                                        bin_h=bin_h, halo=halo, mcc=1 if mcc else 0, mg=mg,
                                                        rebin_every=rebin_every))
    return dict(plasma=pl, grid=grid, particles=parts, walls=bnd, xs=xs, sim=sim, mg=mg, n_per_species=n_per_species)


def plain_small(scale=1.0, **kw):
    """128^2 cells, 1e5 + 1e5 particles, FP64, no collisions, multigrid 2+2
    red-black Gauss-Seidel to 1e-8."""
    n = max(1000, int(1e5 * scale))
    kw.setdefault("mg", mg_params(PRESET_STEP, mode=MG_FP64))
    return build(128, n, PREC_FP64, mcc=False, **kw)


def full_step(scale=1.0, **kw):
    n = max(1000, int(2e7 * scale))
    cells = kw.pop("cells", 1024)
    return build(cells, n, PREC_FP32, mcc=True, **kw)


def random_walls(cells, n, seed=1, min_len=1.0, max_len=24.0):
    """`n` internal wall segments in cell units: centres uniform over the
    interior, directions uniform, lengths uniform in [min_len, max_len]."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(1, cells - 1, n)
    cy = rng.uniform(1, cells - 1, n)
    th = rng.uniform(0, np.pi, n)
    ln = rng.uniform(min_len, max_len, n)
    hx, hy = 0.5 * ln * np.cos(th), 0.5 * ln * np.sin(th)
    return np.stack([cx - hx, cy - hy, cx + hx, cy + hy], axis=1)


def poisson_rhs(n=2049, blobs=200, seed=1, noise=1e-3):
    """Gaussian charge blobs (centres uniform, widths U(5,50) cells, amplitudes
    N(0,1)) plus uniform noise in [-noise, noise]; zero on the boundary."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, n - 1, blobs)
    cy = rng.uniform(0, n - 1, blobs)
    w = rng.uniform(5, 50, blobs)
    a = rng.normal(0, 1, blobs)
    idx = np.arange(n, dtype=np.float64)
    rhs = np.zeros((n, n))
    for k in range(blobs):
        lo_x, hi_x = int(max(0, cx[k] - 6 * w[k])), int(min(n, cx[k] + 6 * w[k] + 1))
        lo_y, hi_y = int(max(0, cy[k] - 6 * w[k])), int(min(n, cy[k] + 6 * w[k] + 1))
        gx = np.exp(-0.5 * ((idx[lo_x:hi_x] - cx[k]) / w[k]) ** 2)
        gy = np.exp(-0.5 * ((idx[lo_y:hi_y] - cy[k]) / w[k]) ** 2)
        rhs[lo_y:hi_y, lo_x:hi_x] += a[k] * np.outer(gy, gx)
    rhs += rng.uniform(-noise, noise, (n, n))
    rhs[0, :] = rhs[-1, :] = rhs[:, 0] = rhs[:, -1] = 0.0
    return rhs


# cells per side with a multigrid ladder down to 3x3 or 4x4 nodes (32*2^k or
# 32*3*2^k): log-uniform draws are rounded to the nearest in log space
GRID_SIZES = (64, 96, 128, 192, 256, 384, 512, 768, 1024)


def round_cells(c, lo, hi):
    sizes = [s for s in GRID_SIZES if lo <= s <= hi]
    return min(sizes, key=lambda s: abs(np.log(s) - np.log(c)))


def ensemble_cases(n_cases=64, seed=4):
    """(cells, particles per species, Te eV, seed) of the batched ensemble:
    cells log-uniform in [64, 512] (rounded to GRID_SIZES), clamp(100 cells^2,
    1e4, 4e6) particles per species, Te ~ U(1, 5) eV, seed 1000 + case."""
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n_cases):
        cells = round_cells(np.exp(rng.uniform(np.log(64), np.log(512))), 64, 512)
        npart = int(min(4e6, max(1e4, 100 * cells * cells)))
        te = float(rng.uniform(1.0, 5.0))
        cases.append((cells, npart, te, 1000 + k))
    return cases


def serve_jobs(n_jobs=24, rate=20.0, seed=9, min_steps=200, max_steps=2000):
    """Job descriptions of the serving stream: Poisson arrivals at `rate` per
    second, cells log-uniform in [64, 1024] (rounded to GRID_SIZES), 100
    particles per cell and species, steps U(min_steps, max_steps), seed = job
    index.  Dicts of pic_cuda_amd.Server.submit arguments."""
    rng = np.random.default_rng(seed)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    jobs = []
    for k in range(n_jobs):
        cells = round_cells(np.exp(rng.uniform(np.log(64), np.log(1024))), 64, 1024)
        steps = int(rng.integers(min_steps, max_steps + 1))
        jobs.append(dict(cells=cells, ppc=100.0, steps=steps, seed=k, arrival_s=float(t_arr[k])))
    return jobs
