"""ctypes bindings of the serial CPU oracle (oracle/pic_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the CPU baseline and never
as the product path.  The product (pic_cuda_amd) never imports this module.

``OracleSim`` restates the step driver updateSimulationCUDA (main.cu:37-140) on
flat per-species arrays: the same pass structure as the GPU driver (the kick and
collisions of step n-1 fused in front of the drift of step n, a kick-only pass
to finish), so counters and fixed-point charge can be compared bit for bit.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_OMP = None

ELASTIC, EXCITATION, IONIZATION, CHARGE_EXCHANGE = 0, 1, 2, 3
QBITS = 20
C_ABSORBED, C_NULL_CAND, C_PROC0 = 0, 1, 2
C_BORN_SELF, C_BORN_OTHER, C_BORN_ABSORBED = 6, 7, 8
C_NUMAX_EXCEEDED, C_REJECT_FALLBACK = 9, 10


class PushParams(ctypes.Structure):
    _fields_ = [
        ("prec", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("kq", ctypes.c_double),
        ("dtdx", ctypes.c_double), ("dtdy", ctypes.c_double),
        ("do_kick", ctypes.c_int), ("do_mcc", ctypes.c_int), ("do_drift", ctypes.c_int),
        ("box_walls", ctypes.c_int), ("n_walls", ctypes.c_int),
        ("walls", ctypes.POINTER(ctypes.c_double)), ("wall_eps", ctypes.c_double),
        ("seed", ctypes.c_uint32), ("step", ctypes.c_uint32), ("case_id", ctypes.c_uint32),
        ("species", ctypes.c_uint32),
        ("n_proc", ctypes.c_int), ("proc_type", ctypes.c_int * 4), ("proc_thr", ctypes.c_double * 4),
        ("tab", ctypes.POINTER(ctypes.c_double)), ("K", ctypes.c_int),
        ("inv_de", ctypes.c_double), ("nu_max", ctypes.c_double), ("p_null", ctypes.c_double),
        ("efac", ctypes.c_double), ("gas_vth", ctypes.c_double),
        ("ion_species", ctypes.c_uint32), ("ion_efac_unused", ctypes.c_double),
        ("inv_emax", ctypes.c_double), ("band_nu_max", ctypes.c_double * 16), ("band_p_null", ctypes.c_double * 16),
    ]


class MGParams(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("pre", ctypes.c_int), ("post", ctypes.c_int),
                ("coarse_sweeps", ctypes.c_int), ("omega", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_cycles", ctypes.c_int), ("per_level", ctypes.c_int), ("pre_level", ctypes.c_int * 16),
                ("post_level", ctypes.c_int * 16)]


class MGStats(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("rel_residual", ctypes.c_double), ("rhs_norm", ctypes.c_double)]


def build():
    """Compile the oracle libraries (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(omp=False):
    global _LIB, _LIB_OMP
    name = "libpic_oracle_omp.so" if omp else "libpic_oracle.so"
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    d, i, u32, i64 = ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_int64
    vp = ctypes.c_void_p
    lib.orc_philox4x32_10.argtypes = [vp, vp, vp]
    lib.orc_philox2x32_10.argtypes = [vp, u32, vp]
    lib.orc_det_log.restype = d
    lib.orc_det_log.argtypes = [d]
    lib.orc_efield.argtypes = [vp, i, i, d, d, vp, vp]
    lib.orc_deposit.argtypes = [i, vp, vp, i64, i, i, vp]
    lib.orc_charge_to_rhs.argtypes = [vp, vp, i, i64, vp]
    lib.orc_mg_solve.restype = i
    lib.orc_mg_solve.argtypes = [vp, vp, i, i, d, ctypes.POINTER(MGParams), ctypes.POINTER(MGStats)]
    lib.orc_mg_solve_masked.restype = i
    lib.orc_mg_solve_masked.argtypes = [vp, vp, i, i, d, ctypes.POINTER(MGParams), vp, vp, ctypes.POINTER(MGStats)]
    lib.orc_mg_smooth.argtypes = [vp, vp, i, d, d, i]
    lib.orc_mg_residual.argtypes = [vp, vp, vp, i, d]
    lib.orc_mg_restrict.argtypes = [vp, vp, i]
    lib.orc_mg_prolong.argtypes = [vp, vp, i]
    lib.orc_push.restype = i64
    lib.orc_push.argtypes = [ctypes.POINTER(PushParams), vp, vp, vp, vp, vp, vp, i64, i64, vp,
                             vp, vp, vp, vp, vp, i64, ctypes.POINTER(i64), vp, vp]
    lib.orc_segment_hits.restype = i
    lib.orc_segment_hits.argtypes = [d, d, d, d, vp, i, d]
    lib.orc_init_particles.argtypes = [i, i64, i, i, d, u32, u32, u32, u32, vp, vp, vp, vp, vp]
    lib.orc_init_uniform.argtypes = [i, i64, d, d, d, d, d, d, u32, u32, u32, u32, vp, vp, vp, vp, vp]
    lib.orc_num_threads.restype = i
    if omp:
        _LIB_OMP = lib
    else:
        _LIB = lib
    return lib


def lib(omp=False):
    if omp:
        return _LIB_OMP or _load(True)
    return _LIB or _load(False)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------------------------------------------------------------------
def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def philox2(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    o = np.zeros(2, dtype=np.uint32)
    lib().orc_philox2x32_10(_p(c), int(key) & 0xFFFFFFFF, _p(o))
    return o


def det_log(x):
    return lib().orc_det_log(float(x))


def efield(phi, dx, dy):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    ny, nx = phi.shape
    ex = np.zeros_like(phi)
    ey = np.zeros_like(phi)
    lib().orc_efield(_p(phi), nx, ny, dx, dy, _p(ex), _p(ey))
    return ex, ey


def deposit(x, y, nx, ny, prec):
    dt = np.float32 if prec == 4 else np.float64
    x = np.ascontiguousarray(x, dtype=dt)
    y = np.ascontiguousarray(y, dtype=dt)
    acc = np.zeros(nx * ny, dtype=np.uint64)
    lib().orc_deposit(prec, _p(x), _p(y), len(x), nx, ny, _p(acc))
    return acc.reshape(ny, nx)


def charge_to_rhs(accs, cfacs):
    accs = [np.ascontiguousarray(a, dtype=np.uint64).ravel() for a in accs]
    ptrs = (ctypes.c_void_p * len(accs))(*[a.ctypes.data for a in accs])
    cf = np.asarray(cfacs, dtype=np.float64)
    rhs = np.zeros(accs[0].size, dtype=np.float64)
    lib().orc_charge_to_rhs(ctypes.cast(ptrs, ctypes.c_void_p), _p(cf), len(accs), rhs.size, _p(rhs))
    return rhs


def mg_solve(phi, rhs, h, levels=0, pre=2, post=2, coarse_sweeps=1, omega=1.0, tol=1e-8, max_cycles=100, omp=False,
             pre_levels=None, post_levels=None, mask=None, phi_bc=None):
    """solve_multigrid restated (Multigrid.cu:573-642); phi updated in place.
    omp: the OpenMP build (colour sweeps split over threads; same values).
    pre_levels/post_levels: sweeps of each level above the coarsest (finest
    first) instead of pre/post.  mask/phi_bc: embedded electrodes (uint8 mask,
    0 solved, 1 Dirichlet at phi_bc, 2 solid)."""
    assert phi.dtype == np.float64 and phi.flags.c_contiguous
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    n = phi.shape[0]
    p = MGParams(levels, pre, post, coarse_sweeps, omega, tol, max_cycles)
    if pre_levels is not None or post_levels is not None:
        p.per_level = 1
        for l in range(16):
            p.pre_level[l] = pre_levels[l] if pre_levels is not None and l < len(pre_levels) else pre
            p.post_level[l] = post_levels[l] if post_levels is not None and l < len(post_levels) else post
    st = MGStats()
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        phi_bc = np.ascontiguousarray(np.zeros_like(phi) if phi_bc is None else phi_bc, dtype=np.float64)
        rc = lib(omp).orc_mg_solve_masked(_p(phi), _p(rhs), n, n, h, ctypes.byref(p), _p(mask), _p(phi_bc),
                                          ctypes.byref(st))
    else:
        rc = lib(omp).orc_mg_solve(_p(phi), _p(rhs), n, n, h, ctypes.byref(p), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("orc_mg_solve failed: %d" % rc)
    return st


def init_particles(prec, n, nx, ny, vth, seed, case_id, species, id0=0):
    dt = np.float32 if prec == 4 else np.float64
    x, y, vx, vy = (np.zeros(n, dtype=dt) for _ in range(4))
    ids = np.zeros(n, dtype=np.uint32)
    lib().orc_init_particles(prec, n, nx, ny, vth, seed, case_id, species, id0, _p(x), _p(y), _p(vx), _p(vy),
                             _p(ids))
    return x, y, vx, vy, ids


def init_uniform(prec, n, x0, x1, y0, y1, v0, v1, seed, case_id, species, id0=0):
    """initializeParticles (main.cu:17-32) made reproducible: uniform box
    positions (cell units) and uniform velocities."""
    dt = np.float32 if prec == 4 else np.float64
    x, y, vx, vy = (np.zeros(n, dtype=dt) for _ in range(4))
    ids = np.zeros(n, dtype=np.uint32)
    lib().orc_init_uniform(prec, n, x0, x1, y0, y1, v0, v1, seed, case_id, species, id0, _p(x), _p(y), _p(vx),
                           _p(vy), _p(ids))
    return x, y, vx, vy, ids


def segment_hits(x1, y1, x2, y2, walls, eps=1e-10):
    w = np.ascontiguousarray(walls, dtype=np.float64).ravel()
    return bool(lib().orc_segment_hits(x1, y1, x2, y2, _p(w), len(w) // 4, eps))


# ---------------------------------------------------------------------------
class Species:
    """Flat particle arrays of one species plus its constants."""

    def __init__(self, prec, x, y, vx, vy, ids, kq, mcc=None, other=-1):
        self.prec = prec
        self.dt = np.float32 if prec == 4 else np.float64
        self.x = np.ascontiguousarray(x, dtype=self.dt)
        self.y = np.ascontiguousarray(y, dtype=self.dt)
        self.vx = np.ascontiguousarray(vx, dtype=self.dt)
        self.vy = np.ascontiguousarray(vy, dtype=self.dt)
        self.id = np.ascontiguousarray(ids, dtype=np.uint32)
        self.kq = kq
        self.mcc = mcc            # dict of the null-collision constants or None
        self.other = other
        self.counters = np.zeros(16, dtype=np.uint64)
        self._bufs = None         # capacity arrays that x..id are views of after a pass
        self._born = None

    @property
    def n(self):
        return len(self.x)

    def _buffers(self, need):
        """Arrays of capacity >= need holding the particles in front; x..id
        become views of them (no copy when they already are)."""
        n = self.n
        b = self._bufs
        fresh = b is None or len(b[0]) < need or not all(
            v.base is buf and v.ctypes.data == buf.ctypes.data for v, buf in zip(self._views(), b))
        if fresh:
            cap = max(need, int(need * 1.25))
            b = [np.empty(cap, dtype=self.dt) for _ in range(4)] + [np.empty(cap, dtype=np.uint32)]
            for v, buf in zip(self._views(), b):
                buf[:n] = v
            self._bufs = b
        return b

    def _views(self):
        return (self.x, self.y, self.vx, self.vy, self.id)

    def _set_len(self, k):
        b = self._bufs
        self.x, self.y, self.vx, self.vy, self.id = (buf[:k] for buf in b)

    def append(self, x, y, vx, vy, ids):
        n, m = self.n, len(x)
        b = self._buffers(n + m)
        for buf, v in zip(b, (x, y, vx, vy, ids)):
            buf[n:n + m] = v
        self._set_len(n + m)


def push(sp, phi, nx, ny, dx, dy, dtdx, dtdy, flags, step, seed, case_id, species_idx, box=True, walls=None,
         wall_eps=1e-10, acc_self=None, acc_other=None, omp=False):
    """One fused pass over a species (kick/MCC/drift/walls/deposit/compaction).

    flags: dict(kick, mcc, drift).  Returns (born-ion arrays or None)."""
    n = sp.n
    born_cap = n + 16
    cap = n + born_cap
    arrs = sp._buffers(cap)
    cap = len(arrs[0])
    idb = arrs[4]
    if sp._born is None or len(sp._born[0]) < born_cap:
        bc = max(born_cap, int(born_cap * 1.25))
        sp._born = [np.empty(bc, dtype=sp.dt) for _ in range(4)] + [np.empty(bc, dtype=np.uint32)]
    xo, yo, vxo, vyo, ido = sp._born
    born_cap = len(xo)
    if acc_self is None:
        acc_self = np.zeros(nx * ny, dtype=np.uint64)
    if acc_other is None:
        acc_other = np.zeros(nx * ny, dtype=np.uint64)
    p = PushParams()
    p.prec = sp.prec
    p.nx, p.ny, p.dx, p.dy, p.kq, p.dtdx, p.dtdy = nx, ny, dx, dy, sp.kq, dtdx, dtdy
    p.do_kick, p.do_drift = int(flags.get("kick", 0)), int(flags.get("drift", 0))
    p.do_mcc = int(flags.get("mcc", 0) and sp.mcc is not None and sp.mcc["n_proc"] > 0)
    p.box_walls = int(box)
    wall_arr = None
    if walls is not None and len(walls):
        wall_arr = np.ascontiguousarray(walls, dtype=np.float64).ravel()
        p.n_walls = len(wall_arr) // 4
        p.walls = wall_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    p.wall_eps = wall_eps
    p.seed, p.step, p.case_id, p.species = seed, step & 0xFFFFFFFF, case_id, species_idx
    tab = None
    if sp.mcc is not None and sp.mcc["n_proc"] > 0:
        m = sp.mcc
        p.n_proc = m["n_proc"]
        for q in range(m["n_proc"]):
            p.proc_type[q] = m["type"][q]
            p.proc_thr[q] = m["threshold"][q]
        tab = np.ascontiguousarray(m["table"], dtype=np.float64)
        p.tab = tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        p.K, p.inv_de, p.nu_max, p.p_null = m["K"], m["inv_de"], m["nu_max"], m["p_null"]
        p.efac, p.gas_vth = m["efac"], m["gas_vth"]
        # energy-banded majorant; a dict without bands has one majorant for all
        p.inv_emax = 1.0 / m["emax"] if m.get("emax") else 1.0
        for b in range(16):
            p.band_nu_max[b] = m["band_nu_max"][b] if "band_nu_max" in m else m["nu_max"]
            p.band_p_null[b] = m["band_p_null"][b] if "band_p_null" in m else m["p_null"]
    n_other = ctypes.c_int64(0)
    k = lib(omp).orc_push(ctypes.byref(p), _p(phi), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), _p(idb),
                          n, cap, _p(acc_self), _p(xo), _p(yo), _p(vxo), _p(vyo), _p(ido), born_cap,
                          ctypes.byref(n_other), _p(acc_other), _p(sp.counters))
    sp._set_len(k)
    no = n_other.value
    if no:
        return (xo[:no].copy(), yo[:no].copy(), vxo[:no].copy(), vyo[:no].copy(), ido[:no].copy())
    return None


class OracleSim:
    """Serial restatement of the GPU step driver (pic_sim_step / pic_sim_sync)."""

    def __init__(self, species, nx, ny, dx, dy, dt, cfac, mg, seed=1, case_id=0, box=True, walls=None,
                 wall_eps=1e-10, phi=None, omp=False):
        self.sp = species
        self.nx, self.ny, self.dx, self.dy, self.dt = nx, ny, dx, dy, dt
        self.cfac = cfac
        self.mg = dict(mg)
        self.seed, self.case_id = seed, case_id
        self.box, self.walls, self.wall_eps = box, walls, wall_eps
        self.phi = np.zeros((ny, nx)) if phi is None else np.array(phi, dtype=np.float64)
        self.rhs = np.zeros((ny, nx))
        self.step_no = 0
        self.pending = False
        self.omp = omp
        self.last_mg = None
        self.accs = None

    def _pass(self, kick, mcc, drift, deposit=True):
        nn = self.nx * self.ny
        accs = [np.zeros(nn, dtype=np.uint64) for _ in self.sp]
        born = {}
        for s, sp in enumerate(self.sp):
            other_acc = accs[sp.other] if (sp.other >= 0 and deposit) else np.zeros(nn, dtype=np.uint64)
            b = push(sp, self.phi, self.nx, self.ny, self.dx, self.dy, self.dt / self.dx, self.dt / self.dy,
                     dict(kick=kick, mcc=mcc, drift=drift), self.step_no - 1, self.seed, self.case_id, s,
                     box=self.box, walls=self.walls, wall_eps=self.wall_eps,
                     acc_self=accs[s] if deposit else np.zeros(nn, dtype=np.uint64), acc_other=other_acc,
                     omp=self.omp)
            if b is not None:
                born.setdefault(sp.other, []).append(b)
        for s, lst in born.items():
            sp = self.sp[s]
            for (x, y, vx, vy, ids) in lst:
                sp.append(x, y, vx, vy, ids)
        return accs

    def step(self, n=1, solve=True):
        for _ in range(n):
            kick = self.pending
            accs = self._pass(kick=kick, mcc=kick, drift=True)
            self.accs = accs
            self.step_no += 1
            self.rhs = charge_to_rhs(accs, self.cfac).reshape(self.ny, self.nx)
            if solve:
                self.last_mg = mg_solve(self.phi, self.rhs, self.dx, omp=self.omp, **self.mg)
            self.pending = True

    def sync(self):
        if self.pending:
            self._pass(kick=True, mcc=True, drift=False, deposit=False)
            self.pending = False
