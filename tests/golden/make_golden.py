"""Generate tests/golden/tiny_step_*.npz: a tiny full PIC run of the oracle.

Case: 17x17 nodes, 300 e- + 300 Ar+ from the seeded initial state, Te = 8 eV,
box walls, null-collision MCC with tables and energy-band majorants computed
here in numpy from the analytic shapes (elastic 1e-19, excitation 11.5 eV /
2e-21, ionization 15.8 eV / 3e-20 Lotz-like, Ar+ charge exchange 5e-19, gas
1e22 m^-3), FP64 multigrid 2+2 to 1e-8, 8 steps and the final half kick.
Stored: the MCC inputs, the per-species final particles sorted by (id, x, y,
vx, vy), the counters, the last charge rhs and phi.

The fixture pins the oracle against later edits (tests/test_golden.py) and
feeds the GPU test that runs the same inputs through the HIP path.

    python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

QE, ME, AMU, EPS0 = 1.602176634e-19, 9.1093837015e-31, 1.66053906660e-27, 8.8541878128e-12
M_AR = 39.948 * AMU


def case_inputs(prec):
    n_cells, n = 16, 300
    te, dens = 8.0, 1e15
    debye = math.sqrt(EPS0 * te / (dens * QE))
    dx = 0.5 * debye
    L = n_cells * dx
    dt = 0.2 / math.sqrt(dens * QE * QE / (EPS0 * ME))
    nn = n_cells + 1
    K, emax, gas, tgas = 512, 200.0, 1e22, 0.03
    de = emax / (K - 1)
    E = np.arange(K) * de

    def lotz(peak, thr):
        u = E / thr
        with np.errstate(divide="ignore", invalid="ignore"):
            s = np.where(u > 1.0, peak * math.e * np.log(u) / u, 0.0)
        return s

    def mcc(m, tabs, types, thr, partner_mass, product):
        # majorant of the interpolated collision frequency per energy band:
        # band 0 below emax*2^-14, band b in [emax*2^(b-15), emax*2^(b-14)),
        # band 15 up to emax; the supremum over the band, x1.001
        tab = np.stack(tabs) * gas
        efac = 0.5 * m / QE

        def nu(e):
            x = e / de
            j = int(math.floor(x))
            t = x - j
            if j >= K - 1:
                j, t = K - 2, 1.0
            v = math.sqrt(2 * e * QE / m)
            return sum((tab[q, j] + t * (tab[q, j + 1] - tab[q, j])) * v for q in range(len(types)) if e >= thr[q])

        def band_max(lo, hi):
            # supremum over [lo, hi]: cut at table nodes and thresholds; on a
            # piece sigma_tot = A + B E, so nu = (A + B E) sqrt(E) peaks inside
            # only at E = -A / (3B) when A > 0 > B
            cut = [lo, hi] + [j * de for j in range(int(math.ceil(lo / de)), K) if lo < j * de < hi]
            cut += [t for t in thr if lo < t < hi]
            cut.sort()
            mx = max(nu(e) for e in cut)
            for a, b in zip(cut[:-1], cut[1:]):
                mid = 0.5 * (a + b)
                if not b > a or mid / de >= K - 1:
                    continue
                j = int(math.floor(mid / de))
                A = B = 0.0
                for q in range(len(types)):
                    if mid >= thr[q]:
                        A += tab[q, j] - j * (tab[q, j + 1] - tab[q, j])
                        B += (tab[q, j + 1] - tab[q, j]) / de
                if A > 0.0 and B < 0.0 and a < -A / (3.0 * B) < b:
                    mx = max(mx, nu(-A / (3.0 * B)))
            return mx

        band_nu = []
        for b in range(16):
            lo = 0.0 if b == 0 else math.ldexp(emax, b - 15)
            hi = emax if b == 15 else math.ldexp(emax, b - 14)
            band_nu.append(band_max(lo, hi) * 1.001)
        numax = max(band_nu)
        return dict(n_proc=len(types), type=list(types), threshold=list(thr), table=tab.ravel(), K=K,
                    inv_de=1.0 / de, nu_max=numax, p_null=-math.expm1(-numax * dt), efac=efac,
                    gas_vth=math.sqrt(tgas * QE / partner_mass), product_species=product, emax=emax,
                    band_nu_max=band_nu, band_p_null=[-math.expm1(-v * dt) for v in band_nu])

    m_e = mcc(ME, [np.full(K, 1e-19), lotz(2e-21, 11.5), lotz(3e-20, 15.8)], (0, 1, 2), (0.0, 11.5, 15.8), M_AR, 1)
    m_i = mcc(M_AR, [np.full(K, 5e-19)], (3,), (0.0,), M_AR, -1)
    w = dens * L * L / n
    qs = 1.0 / float(1 << 40)
    cfac = [-(q * w * qs / (dx * dx)) / EPS0 for q in (-QE, QE)]
    kq = [(-QE) * dt / ME, QE * dt / M_AR]
    vth = [math.sqrt(te * QE / ME), math.sqrt(0.03 * QE / M_AR)]
    return dict(prec=prec, nn=nn, n=n, dx=dx, dt=dt, mcc=[m_e, m_i], cfac=cfac, kq=kq, vth=vth, seed=21)


def run_oracle(inp, steps=8):
    from oracle import oracle as orc
    sp = []
    for s in range(2):
        x, y, vx, vy, ids = orc.init_particles(inp["prec"], inp["n"], inp["nn"], inp["nn"], inp["vth"][s],
                                               inp["seed"], 0, s, 0)
        sp.append(orc.Species(inp["prec"], x, y, vx, vy, ids, inp["kq"][s], mcc=inp["mcc"][s],
                              other=inp["mcc"][s]["product_species"]))
    o = orc.OracleSim(sp, inp["nn"], inp["nn"], inp["dx"], inp["dx"], inp["dt"], inp["cfac"],
                      dict(tol=1e-8, max_cycles=100), seed=inp["seed"])
    o.step(steps)
    rhs = o.rhs.copy()
    o.sync()
    return o, rhs


def pack(o, rhs, inp):
    out = {}
    for s, sp in enumerate(o.sp):
        order = np.lexsort((sp.vy, sp.vx, sp.y, sp.x, sp.id))
        for k in ("x", "y", "vx", "vy", "id"):
            out["s%d_%s" % (s, k)] = getattr(sp, k)[order]
        out["s%d_counters" % s] = sp.counters.copy()
        out["s%d_table" % s] = inp["mcc"][s]["table"]
        for k in ("inv_de", "nu_max", "p_null", "efac", "gas_vth"):
            out["s%d_%s" % (s, k)] = np.float64(inp["mcc"][s][k])
        for k in ("band_nu_max", "band_p_null"):
            out["s%d_%s" % (s, k)] = np.asarray(inp["mcc"][s][k], dtype=np.float64)
    out["rhs"] = rhs
    out["phi"] = o.phi
    return out


def main():
    for prec in (4, 8):
        inp = case_inputs(prec)
        o, rhs = run_oracle(inp)
        path = os.path.join(HERE, "tiny_step_fp%d.npz" % (8 * prec))
        np.savez_compressed(path, **pack(o, rhs, inp))
        print("wrote", path, [sp.n for sp in o.sp], [int(c) for c in o.sp[0].counters[:9]])


if __name__ == "__main__":
    main()
