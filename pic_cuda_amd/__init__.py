"""pic_cuda_amd -- MI355X (gfx950) 2D electrostatic PIC time step.

Python mirror of the reference's host interfaces, over the C ABI of
``libpic_amd.so`` (declared in include/pic_amd.h):

=======================  =====================================================
this module              reference (dimad285/PIC_CUDA)
=======================  =====================================================
``Particles2D``          particles.cuh:3-39 / particles.cu:112-261
``Grid2D``               grid.cuh:7-24 / grid.cu:10-53
``Boundaries2D``         boundaries.h:7-23 / boundaries.cu:3-25
``solve_multigrid``      Multigrid.hpp:3-11 / Multigrid.cu:573-642
``Multigrid``            the solver context (Multigrid.cu:30-187 pool)
``read_cross_section``   mcc.cu:17-41
``CrossSections``        cross_sections, mcc.hpp:6-16
``elasticCollisions``    mcc.cu:114-150
``Simulation.step``      updateSimulationCUDA, main.cu:37-140
=======================  =====================================================

Positions cross the ABI in cell units; the Python classes accept domain units
(metres) when ``units="domain"`` and convert with the grid spacing.

The library is required: importing works without it, but every call raises
``RuntimeError`` when ``libpic_amd.so`` is missing.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpic_amd.so")
_lib = None

# constants mirrored from include/pic_amd.h
PREC_FP32, PREC_FP64 = 4, 8
ELASTIC, EXCITATION, IONIZATION, CHARGE_EXCHANGE = 0, 1, 2, 3
FIELD_RHO, FIELD_PHI, FIELD_EX, FIELD_EY, FIELD_RHS = 0, 1, 2, 3, 4
MG_FP64, MG_MIXED = 0, 1
PRESET_REFERENCE, PRESET_STEP, PRESET_POISSON, PRESET_STEP_SOR = 0, 1, 2, 3
TUNE_BINS, TUNE_MG = 1, 2
CNT = dict(absorbed=0, null_candidates=1, process0=2, process1=3, process2=4, process3=5, born_self=6,
           born_other=7, born_absorbed=8, numax_exceeded=9, reject_fallback=10, spilled=11, lost=12,
           global_deposit=13)

ELEMENTARY_CHARGE = 1.602176634e-19
ELECTRON_MASS = 9.1093837015e-31
AMU = 1.66053906660e-27
EPS0 = 8.8541878128e-12


class MGParams(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("pre", ctypes.c_int), ("post", ctypes.c_int),
                ("coarse_sweeps", ctypes.c_int), ("omega", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_cycles", ctypes.c_int), ("mode", ctypes.c_int), ("inner", ctypes.c_int),
                ("coarse_direct", ctypes.c_int)]


class MGStats(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("rel_residual", ctypes.c_double), ("rhs_norm", ctypes.c_double),
                ("ms", ctypes.c_double), ("launches", ctypes.c_int)]


class SimParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("eps0", ctypes.c_double), ("seed", ctypes.c_uint32),
                ("case_id", ctypes.c_uint32), ("bin_w", ctypes.c_int), ("bin_h", ctypes.c_int),
                ("halo", ctypes.c_int), ("mcc", ctypes.c_int), ("rebin_every", ctypes.c_int),
                ("wall_eps", ctypes.c_double), ("mg", MGParams)]


class SimReport(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("alive", ctypes.c_int64), ("last_mg_cycles", ctypes.c_int),
                ("last_mg_rel", ctypes.c_double), ("total_mg_cycles", ctypes.c_int64), ("rebins", ctypes.c_int),
                ("ms_push", ctypes.c_double), ("ms_field", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("graph_steps", ctypes.c_int64), ("graph_captures", ctypes.c_int64), ("ms_graph", ctypes.c_double),
                ("ahead_steps", ctypes.c_int64), ("ahead_held", ctypes.c_int64)]


class Diagnostics(ctypes.Structure):
    _fields_ = [("n_species", ctypes.c_int), ("count", ctypes.c_double * 4), ("kinetic", ctypes.c_double * 4),
                ("px", ctypes.c_double * 4), ("py", ctypes.c_double * 4), ("field_energy", ctypes.c_double),
                ("phi_l2", ctypes.c_double)]


class TuneReport(ctypes.Structure):
    _fields_ = [("bin_w", ctypes.c_int), ("bin_h", ctypes.c_int), ("push_ms", ctypes.c_double),
                ("omega", ctypes.c_double), ("levels", ctypes.c_int), ("pre", ctypes.c_int * 16),
                ("post", ctypes.c_int * 16), ("solve_ms", ctypes.c_double), ("solve_cycles", ctypes.c_int),
                ("trials", ctypes.c_int), ("inner", ctypes.c_int)]


class MccInfo(ctypes.Structure):
    _fields_ = [("n_proc", ctypes.c_int), ("K", ctypes.c_int), ("type", ctypes.c_int * 4),
                ("threshold", ctypes.c_double * 4), ("inv_de", ctypes.c_double), ("nu_max", ctypes.c_double),
                ("p_null", ctypes.c_double), ("efac", ctypes.c_double), ("gas_vth", ctypes.c_double),
                ("product_species", ctypes.c_int), ("emax", ctypes.c_double), ("band_nu_max", ctypes.c_double * 16),
                ("band_p_null", ctypes.c_double * 16)]


class JobDesc(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int), ("ppc", ctypes.c_double), ("steps", ctypes.c_int), ("seed", ctypes.c_uint32),
                ("te_eV", ctypes.c_double), ("density", ctypes.c_double), ("gas_density", ctypes.c_double),
                ("arrival_s", ctypes.c_double), ("mcc", ctypes.c_int)]


class XsProcess(ctypes.Structure):
    _fields_ = [("species", ctypes.c_int), ("type", ctypes.c_int), ("product_species", ctypes.c_int),
                ("n_points", ctypes.c_int), ("threshold_eV", ctypes.c_double), ("peak_m2", ctypes.c_double)]


class JobResult(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("steps_done", ctypes.c_int), ("step_ms_mean", ctypes.c_double),
                ("step_ms_max", ctypes.c_double), ("particles", ctypes.c_int64), ("alive_end", ctypes.c_int64),
                ("footprint_bytes", ctypes.c_double)]


# every symbol declared in include/pic_amd.h, with its ctypes signature
_vp, _i, _d, _u32, _i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint32, ctypes.c_int64
_P = ctypes.POINTER
SIGNATURES = {
    "pic_last_error": (ctypes.c_char_p, []),
    "pic_version": (_i, [_P(_i), _P(_i)]),
    "pic_device_synchronize": (_i, []),
    "pic_device_alloc": (_i, [_i64, _P(_vp)]),
    "pic_device_free": (_i, [_vp]),
    "pic_device_to_host": (_i, [_vp, _vp, _i64]),
    "pic_host_to_device": (_i, [_vp, _vp, _i64]),
    "pic_particles_staging": (_i, [_vp, _i64, _P(_vp), _P(_vp), _P(_vp), _P(_vp)]),
    "pic_particles_set_staged": (_i, [_vp, _i64]),
    "pic_particles_commit_staging": (_i, [_vp, _i64, _d, _d]),
    "pic_particles_create": (_i, [_i64, _i, _P(_vp)]),
    "pic_particles_destroy": (_i, [_vp]),
    "pic_particles_add_species": (_i, [_vp, _d, _d, _P(_i)]),
    "pic_particles_change_species": (_i, [_vp, _i, _d, _d]),
    "pic_particles_set_weight": (_i, [_vp, _i, _d]),
    "pic_particles_num_species": (_i, [_vp, _P(_i)]),
    "pic_particles_last_alive": (_i, [_vp, _i, _P(_i64)]),
    "pic_particles_upload": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp]),
    "pic_particles_download": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp, _P(_i64)]),
    "pic_particles_init_maxwellian": (_i, [_vp, _i, _i64, _i, _i, _d, _u32, _u32, _u32]),
    "pic_particles_init_uniform": (_i, [_vp, _i, _i64, _d, _d, _d, _d, _d, _d, _u32, _u32, _u32]),
    "pic_particles_remove_collided": (_i, [_vp]),
    "pic_grid_create": (_i, [_i, _i, _d, _d, _P(_vp)]),
    "pic_grid_destroy": (_i, [_vp]),
    "pic_grid_info": (_i, [_vp, _P(_i), _P(_i), _P(_d), _P(_d)]),
    "pic_grid_device_ptrs": (_i, [_vp, _P(_vp), _P(_vp), _P(_vp)]),
    "pic_grid_download": (_i, [_vp, _i, _vp]),
    "pic_grid_upload": (_i, [_vp, _i, _vp]),
    "pic_grid_compute_E": (_i, [_vp]),
    "pic_walls_create": (_i, [_d, _d, _i, _i, _P(_vp)]),
    "pic_walls_add": (_i, [_vp, _d, _d, _d, _d]),
    "pic_walls_add_box": (_i, [_vp]),
    "pic_walls_compile": (_i, [_vp]),
    "pic_walls_count": (_i, [_vp, _P(_i)]),
    "pic_walls_destroy": (_i, [_vp]),
    "pic_mg_default_params": (_i, [_i, _P(MGParams)]),
    "pic_mg_create": (_i, [_i, _P(MGParams), _P(_vp)]),
    "pic_mg_solve": (_i, [_vp, _vp, _vp, _d, _P(MGStats)]),
    "pic_mg_destroy": (_i, [_vp]),
    "pic_solve_multigrid": (_i, [_vp, _vp, _i, _i, _i, _d, _d]),
    "pic_read_cross_section": (_i, [ctypes.c_char_p, _i, _vp, _vp, _P(_i)]),
    "pic_xs_create": (_i, [_d, _d, _P(_vp)]),
    "pic_xs_add_process": (_i, [_vp, _i, _i, _d, _i, _vp, _vp, _i]),
    "pic_xs_add_analytic": (_i, [_vp, _i, _i, _d, _d, _i]),
    "pic_xs_set_table": (_i, [_vp, _i, _d]),
    "pic_xs_destroy": (_i, [_vp]),
    "pic_read_table": (_i, [ctypes.c_char_p, _i, _vp, _vp, _P(_i)]),
    "pic_xs_load_lxcat": (_i, [_vp, ctypes.c_char_p, _i, _i, _P(_i), _P(_i)]),
    "pic_xs_process_count": (_i, [_vp, _P(_i)]),
    "pic_xs_build": (_i, [_vp, _i, _i, _vp, _d, _P(MccInfo), _i, _vp]),
    "pic_xs_get_process": (_i, [_vp, _i, _P(XsProcess), _i, _vp, _vp]),
    "pic_elastic_collisions": (_i, [_vp, _i, _d, _d, _i, _vp, _vp, _u32, _u32]),
    "pic_sim_default_params": (_i, [_P(SimParams)]),
    "pic_sim_create": (_i, [_vp, _vp, _vp, _vp, _P(SimParams), _P(_vp)]),
    "pic_sim_step": (_i, [_vp, _i]),
    "pic_sim_sync": (_i, [_vp]),
    "pic_sim_counters": (_i, [_vp, _i, _vp]),
    "pic_sim_stats": (_i, [_vp, _P(SimReport)]),
    "pic_sim_diagnostics": (_i, [_vp, _P(Diagnostics)]),
    "pic_sim_mcc_info": (_i, [_vp, _i, _P(MccInfo), _vp, _i]),
    "pic_sim_rebin": (_i, [_vp, _i, _i]),
    "pic_sim_use_graph": (_i, [_vp, _i]),
    "pic_sim_set_wall_lists": (_i, [_vp, _i, _i]),
    "pic_sim_write_frame": (_i, [_vp, ctypes.c_char_p]),
    "pic_particles_export_positions": (_i, [_vp, _i, _vp, _i64, _i, _P(_i64)]),
    "pic_particles_export_positions_static": (_i, [_vp, _i, _vp, _i64, _i, _P(_i64)]),
    "pic_sim_wall_list_stats": (_i, [_vp, _P(_i), _P(ctypes.c_longlong), _P(_i)]),
    "pic_sim_autotune": (_i, [_vp, _i, ctypes.POINTER(TuneReport)]),
    "pic_sim_set_level_sweeps": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pic_mg_set_level_sweeps": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pic_mg_set_electrodes": (_i, [_vp, _vp, _vp]),
    "pic_sim_set_electrodes": (_i, [_vp, _vp, _vp]),
    "pic_sim_destroy": (_i, [_vp]),
    "pic_ensemble_create": (_i, [_P(_vp)]),
    "pic_ensemble_add": (_i, [_vp, _vp]),
    "pic_ensemble_remove": (_i, [_vp, _vp]),
    "pic_ensemble_size": (_i, [_vp, _P(_i)]),
    "pic_ensemble_step": (_i, [_vp, _i]),
    "pic_ensemble_use_graph": (_i, [_vp, _i]),
    "pic_ensemble_destroy": (_i, [_vp]),
    "pic_server_create": (_i, [_d, _i, _i, _P(SimParams), _i, _P(_vp)]),
    "pic_server_submit": (_i, [_vp, _P(JobDesc), _P(_i)]),
    "pic_server_tick": (_i, [_vp, _d, _P(_i), _P(_i)]),
    "pic_server_job_result": (_i, [_vp, _i, _P(JobResult)]),
    "pic_server_job_step_ms": (_i, [_vp, _i, _i, _vp, _P(_i)]),
    "pic_server_job_handles": (_i, [_vp, _i, _P(_vp), _P(_vp), _P(_vp)]),
    "pic_server_release": (_i, [_vp, _i]),
    "pic_server_used_bytes": (_i, [_vp, _P(_d)]),
    "pic_server_destroy": (_i, [_vp]),
}


def load(path=None):
    """Load libpic_amd.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libpic_amd.so not found at %s: run __graft_entry__.build() "
                           "(make -C pic_cuda_amd/csrc)" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class PicError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        msg = load().pic_last_error()
        raise PicError("pic error %d: %s" % (rc, msg.decode() if msg else ""))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _devptr(t):
    """Device pointer of a torch tensor, an int address, or None."""
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    return ctypes.c_void_p(t.data_ptr())


def synchronize():
    _check(load().pic_device_synchronize())


def host_to_device(dev_ptr, array):
    """Copy a contiguous numpy array to a device address (after the library's work)."""
    a = np.ascontiguousarray(array)
    _check(load().pic_host_to_device(ctypes.c_void_p(dev_ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes))


def device_to_host(dev_ptr, count, dtype):
    """Copy `count` elements of `dtype` from a device address into a new numpy array."""
    out = np.empty(int(count), dtype=dtype)
    _check(load().pic_device_to_host(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dev_ptr), out.nbytes))
    return out


class DeviceBuffer:
    """Device memory from the library (no torch needed): `ptr` for the entry
    points that take device pointers, to_numpy() to read it back."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self._p = ctypes.c_void_p()
        _check(load().pic_device_alloc(self.nbytes, ctypes.byref(self._p)))

    @property
    def ptr(self):
        return self._p.value

    def data_ptr(self):
        return self._p.value

    def to_numpy(self, dtype=np.float32, count=None):
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        _check(load().pic_device_to_host(out.ctypes.data_as(ctypes.c_void_p), self._p, n * dt.itemsize))
        return out

    def close(self):
        if self._p.value and _lib is not None:
            _lib.pic_device_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
class Particles2D:
    """Particles2D(max_particles) -- particles.cu:112-150.

    Species 0 exists with m = 1, q = 0.1 like the reference constructor.
    ``max_particles`` is the capacity of each species."""

    _owned = True

    @classmethod
    def _view(cls, handle, precision, grid=None):
        """Non-owning wrapper of a handle held elsewhere (e.g. by a job server)."""
        o = cls.__new__(cls)
        o._h = ctypes.c_void_p(handle)
        o._owned = False
        o.precision = precision
        o.max_particles = 0
        o.grid = grid
        return o

    def __init__(self, max_particles, precision=PREC_FP32):
        self._h = ctypes.c_void_p()
        _check(load().pic_particles_create(int(max_particles), int(precision), ctypes.byref(self._h)))
        self.precision = precision
        self.max_particles = int(max_particles)
        self.grid = None

    def close(self):
        if self._h:
            _check(load().pic_particles_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._owned and self._h and _lib is not None:
                _lib.pic_particles_destroy(self._h)
        except Exception:
            pass

    def staging(self, n_max, r_old=False):
        """The reference's public device arrays r, r_old, v, type
        (particles.cuh:7-20) as a staging area of n_max entries: a dict of
        device addresses (r, v: (x, y) double pairs in domain units; type:
        int32 species index, zeroed).  Fill them with `host_to_device`."""
        r, ro, v, t = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(load().pic_particles_staging(self._h, int(n_max), ctypes.byref(r),
                                            ctypes.byref(ro) if r_old else None, ctypes.byref(v), ctypes.byref(t)))
        return {"r": r.value, "r_old": ro.value if r_old else None, "v": v.value, "type": t.value}

    def set_staged(self, n):
        """particles.last_alive = n (main.cu:238): the first n staged entries
        become particles (when a simulation takes them, or at once in one)."""
        _check(load().pic_particles_set_staged(self._h, int(n)))

    def commit_staging(self, n, dx, dy):
        _check(load().pic_particles_commit_staging(self._h, int(n), float(dx), float(dy)))

    def add_species(self, m, q):
        """add_species(double m, double q) -- argument order (m, q), particles.cu:174."""
        idx = ctypes.c_int()
        _check(load().pic_particles_add_species(self._h, float(m), float(q), ctypes.byref(idx)))
        return idx.value

    def change_species(self, species, m, q):
        _check(load().pic_particles_change_species(self._h, int(species), float(m), float(q)))

    def set_weight(self, species, w):
        _check(load().pic_particles_set_weight(self._h, int(species), float(w)))

    @property
    def species_num(self):
        n = ctypes.c_int()
        _check(load().pic_particles_num_species(self._h, ctypes.byref(n)))
        return n.value

    def count(self, species=-1):
        n = ctypes.c_int64()
        _check(load().pic_particles_last_alive(self._h, int(species), ctypes.byref(n)))
        return n.value

    @property
    def last_alive(self):
        return self.count(-1)

    def _scale(self, units):
        if units == "cell":
            return 1.0, 1.0
        if self.grid is None:
            raise ValueError("domain units need the grid: set particles.grid")
        return self.grid.dx, self.grid.dy

    def upload(self, species, x, y, vx, vy, ids=None, units="cell"):
        sx, sy = self._scale(units)
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64) / sx)
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float64) / sy)
        vx = np.ascontiguousarray(vx, dtype=np.float64)
        vy = np.ascontiguousarray(vy, dtype=np.float64)
        idp = None
        if ids is not None:
            ids = np.ascontiguousarray(ids, dtype=np.uint32)
            idp = _ptr(ids)
        _check(load().pic_particles_upload(self._h, int(species), len(x), _ptr(x), _ptr(y), _ptr(vx), _ptr(vy),
                                           idp))

    def download(self, species, units="cell"):
        n = ctypes.c_int64()
        _check(load().pic_particles_download(self._h, int(species), 0, None, None, None, None, None,
                                             ctypes.byref(n)))
        cap = n.value + 1024
        x, y, vx, vy = (np.zeros(cap) for _ in range(4))
        ids = np.zeros(cap, dtype=np.uint32)
        _check(load().pic_particles_download(self._h, int(species), cap, _ptr(x), _ptr(y), _ptr(vx), _ptr(vy),
                                             _ptr(ids), ctypes.byref(n)))
        k = n.value
        sx, sy = self._scale(units)
        return x[:k] * sx, y[:k] * sy, vx[:k], vy[:k], ids[:k]

    def export_positions(self, species, dst=None, units="cell"):
        """Live positions as float32 (x, y) pairs on the device
        (convertPositionsToFloat, kernels.cu:268-273): into `dst` (a
        DeviceBuffer) or, with dst None, a new DeviceBuffer of the right size.
        Returns (buffer, count); the order is unspecified."""
        u = {"cell": 0, "domain": 1}[units]
        n = _i64()
        if dst is None:
            _check(load().pic_particles_export_positions(self._h, int(species), None, 0, u, ctypes.byref(n)))
            dst = DeviceBuffer(8 * max(1, n.value))
        _check(load().pic_particles_export_positions(self._h, int(species), _devptr(dst.ptr), dst.nbytes // 8, u,
                                                     ctypes.byref(n)))
        return dst, n.value

    def init_uniform(self, species, n, x_range, y_range, v_range, seed=1, case_id=0, id0=0):
        """initializeParticles (main.cu:17-32), reproducible: positions uniform
        in x_range x y_range (cell units), velocity components uniform in v_range."""
        _check(load().pic_particles_init_uniform(self._h, int(species), int(n), float(x_range[0]), float(x_range[1]),
                                                 float(y_range[0]), float(y_range[1]), float(v_range[0]),
                                                 float(v_range[1]), int(seed), int(case_id), int(id0)))

    def init_maxwellian(self, species, n, nx, ny, vth, seed=1, case_id=0, id0=0):
        _check(load().pic_particles_init_maxwellian(self._h, int(species), int(n), int(nx), int(ny), float(vth),
                                                    int(seed), int(case_id), int(id0)))

    def removeCollidedParticles(self):
        _check(load().pic_particles_remove_collided(self._h))


class Grid2D:
    """Grid2D(M, N, X, Y) -- grid.cu:10-20 (dx = X/(M-1)); fields start at zero."""

    _owned = True

    @classmethod
    def _view(cls, handle):
        o = cls.__new__(cls)
        o._h = ctypes.c_void_p(handle)
        o._owned = False
        o._info()
        return o

    def __init__(self, M, N, X, Y):
        self._h = ctypes.c_void_p()
        _check(load().pic_grid_create(int(M), int(N), float(X), float(Y), ctypes.byref(self._h)))
        self._info()

    def _info(self):
        m, n, dx, dy = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _check(load().pic_grid_info(self._h, ctypes.byref(m), ctypes.byref(n), ctypes.byref(dx), ctypes.byref(dy)))
        self.m, self.n, self.dx, self.dy = m.value, n.value, dx.value, dy.value
        self.x, self.y = (self.m - 1) * self.dx, (self.n - 1) * self.dy

    def __del__(self):
        try:
            if self._owned and self._h and _lib is not None:
                _lib.pic_grid_destroy(self._h)
        except Exception:
            pass

    def device_ptrs(self):
        r, p, e = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(load().pic_grid_device_ptrs(self._h, ctypes.byref(r), ctypes.byref(p), ctypes.byref(e)))
        return r.value, p.value, e.value

    def download(self, field):
        out = np.zeros((self.n, self.m))
        _check(load().pic_grid_download(self._h, int(field), _ptr(out)))
        return out

    def upload(self, field, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.n, self.m)
        _check(load().pic_grid_upload(self._h, int(field), _ptr(a)))

    @property
    def rho(self):
        return self.download(FIELD_RHO)

    @property
    def phi(self):
        return self.download(FIELD_PHI)

    def computeElectricField(self):
        """computeElectricField (kernels.cu:134-191) into the E member."""
        _check(load().pic_grid_compute_E(self._h))
        return self.download(FIELD_EX), self.download(FIELD_EY)


class Boundaries2D:
    """Boundaries2D(width, height, nx, ny) -- boundaries.h:7-23."""

    def __init__(self, width, height, nx, ny):
        self._h = ctypes.c_void_p()
        _check(load().pic_walls_create(float(width), float(height), int(nx), int(ny), ctypes.byref(self._h)))

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.pic_walls_destroy(self._h)
        except Exception:
            pass

    def add_wall(self, x1, y1, x2, y2):
        _check(load().pic_walls_add(self._h, float(x1), float(y1), float(x2), float(y2)))

    def add_box(self):
        _check(load().pic_walls_add_box(self._h))

    def compile_walls(self):
        _check(load().pic_walls_compile(self._h))

    @property
    def wall_count(self):
        n = ctypes.c_int()
        _check(load().pic_walls_count(self._h, ctypes.byref(n)))
        return n.value


def mg_params(preset=PRESET_STEP, **over):
    p = MGParams()
    _check(load().pic_mg_default_params(int(preset), ctypes.byref(p)))
    for k, v in over.items():
        setattr(p, k, v)
    return p


class Multigrid:
    """Multigrid solver context for an n x n grid (replaces the global pool)."""

    def __init__(self, n, params=None, **over):
        self._h = ctypes.c_void_p()
        self.params = params if params is not None else mg_params(PRESET_STEP, **over)
        _check(load().pic_mg_create(int(n), ctypes.byref(self.params), ctypes.byref(self._h)))
        self.n = n

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.pic_mg_destroy(self._h)
        except Exception:
            pass

    def solve(self, phi_dev, rhs_dev, h):
        st = MGStats()
        _check(load().pic_mg_solve(self._h, _devptr(phi_dev), _devptr(rhs_dev), float(h), ctypes.byref(st)))
        return st

    def set_level_sweeps(self, pre, post):
        """Sweeps of each level, finest first."""
        n = len(pre)
        a = (ctypes.c_int * max(1, n))(*pre)
        b = (ctypes.c_int * max(1, n))(*post)
        _check(load().pic_mg_set_level_sweeps(self._h, n, a, b))

    def set_electrodes(self, mask, phi_bc=None):
        """Embedded electrodes: mask (n x n) 0 solved, 1 Dirichlet at phi_bc,
        2 solid; None removes them."""
        _check(load().pic_mg_set_electrodes(self._h, *_electrode_args(mask, phi_bc, self.n)))


def _electrode_args(mask, phi_bc, n):
    if mask is None:
        return (None, None)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    if m.size != n * n:
        raise PicError("electrode mask must have n*n entries")
    b = None if phi_bc is None else np.ascontiguousarray(phi_bc, dtype=np.float64)
    if b is not None and b.size != n * n:
        raise PicError("phi_bc must have n*n entries")
    # the arrays are copied by the library before the call returns
    return (m.ctypes.data_as(ctypes.c_void_p), None if b is None else b.ctypes.data_as(ctypes.c_void_p))


def solve_multigrid(phi_d, rho_d, nx, ny, levels, h, omega=1.25):
    """solve_multigrid(phi, rho, nx, ny, levels, h, omega) -- Multigrid.hpp:3-11 (device tensors)."""
    _check(load().pic_solve_multigrid(_devptr(phi_d), _devptr(rho_d), int(nx), int(ny), int(levels), float(h),
                                      float(omega)))


def read_cross_section(path):
    """read_cross_section(filename) -- mcc.cu:17-41; returns (energies, sigmas)."""
    n = ctypes.c_int()
    _check(load().pic_read_cross_section(path.encode(), 0, None, None, ctypes.byref(n)))
    e = np.zeros(n.value)
    s = np.zeros(n.value)
    _check(load().pic_read_cross_section(path.encode(), n.value, _ptr(e), _ptr(s), ctypes.byref(n)))
    return e, s


def read_table(path):
    """Two-column table like read_cross_section, numbers also as m*10^{e};
    returns (energies, sigmas)."""
    n = ctypes.c_int()
    _check(load().pic_read_table(path.encode(), 0, None, None, ctypes.byref(n)))
    e = np.zeros(n.value)
    s = np.zeros(n.value)
    _check(load().pic_read_table(path.encode(), n.value, _ptr(e), _ptr(s), ctypes.byref(n)))
    return e, s


def read_frame(path):
    """A frame file of Simulation.write_frame: dict(step, width, height,
    positions=[float32 array (n, 2) per species])."""
    raw = open(path, "rb").read()
    if raw[:4] != b"PICF":
        raise ValueError("not a frame file: %s" % path)
    version, step, ns, units = np.frombuffer(raw, np.uint32, 1, 4)[0], np.frombuffer(raw, np.uint64, 1, 8)[0], \
        np.frombuffer(raw, np.uint32, 1, 16)[0], np.frombuffer(raw, np.uint32, 1, 20)[0]
    if version != 1 or units != 1:
        raise ValueError("unsupported frame version or units")
    width, height = np.frombuffer(raw, np.float64, 2, 24)
    off, pos = 40, []
    for _ in range(int(ns)):
        c = int(np.frombuffer(raw, np.uint64, 1, off)[0])
        pos.append(np.frombuffer(raw, np.float32, 2 * c, off + 8).reshape(c, 2).copy())
        off += 8 + 8 * c
    return dict(step=int(step), width=float(width), height=float(height), positions=pos)


def _mcc_dict(info, tab):
    return dict(n_proc=info.n_proc, K=info.K, type=list(info.type)[:info.n_proc],
                threshold=list(info.threshold)[:info.n_proc], inv_de=info.inv_de, nu_max=info.nu_max,
                p_null=info.p_null, efac=info.efac, gas_vth=info.gas_vth,
                product_species=info.product_species, table=tab[:info.n_proc * info.K], emax=info.emax,
                band_nu_max=list(info.band_nu_max), band_p_null=list(info.band_p_null))


class CrossSections:
    """Background gas and collision processes (cross_sections, mcc.hpp:6-16)."""

    def __init__(self, gas_density, gas_temperature_eV, points=2048, emax_eV=200.0):
        self._h = ctypes.c_void_p()
        _check(load().pic_xs_create(float(gas_density), float(gas_temperature_eV), ctypes.byref(self._h)))
        _check(load().pic_xs_set_table(self._h, int(points), float(emax_eV)))

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.pic_xs_destroy(self._h)
        except Exception:
            pass

    def add_process(self, species, ptype, threshold_eV, energies, sigmas, product_species=-1):
        e = np.ascontiguousarray(energies, dtype=np.float64)
        s = np.ascontiguousarray(sigmas, dtype=np.float64)
        _check(load().pic_xs_add_process(self._h, int(species), int(ptype), float(threshold_eV), len(e), _ptr(e),
                                         _ptr(s), int(product_species)))

    def add_analytic(self, species, ptype, threshold_eV, peak_m2, product_species=-1):
        _check(load().pic_xs_add_analytic(self._h, int(species), int(ptype), float(threshold_eV), float(peak_m2),
                                          int(product_species)))

    def load_lxcat(self, path, species, product_species=-1):
        """Every block of an LXCat file as a process of `species`; returns
        (added, skipped attachment blocks)."""
        a, k = _i(), _i()
        _check(load().pic_xs_load_lxcat(self._h, path.encode(), int(species), int(product_species), ctypes.byref(a),
                                        ctypes.byref(k)))
        return a.value, k.value

    def build(self, species, masses, dt):
        """The null-collision constants and sampled table a simulation builds
        for `species` (masses of every species in kg); host only."""
        m = np.ascontiguousarray(masses, dtype=np.float64)
        info = MccInfo()
        _check(load().pic_xs_build(self._h, int(species), len(m), _ptr(m), float(dt), ctypes.byref(info), 0, None))
        tab = np.zeros(max(1, info.n_proc * info.K))
        _check(load().pic_xs_build(self._h, int(species), len(m), _ptr(m), float(dt), ctypes.byref(info), tab.size,
                                   _ptr(tab)))
        return _mcc_dict(info, tab)

    def processes(self):
        """The processes held, in order: dicts with species, type,
        product_species, threshold_eV, peak_m2, energies, sigmas."""
        n = _i()
        _check(load().pic_xs_process_count(self._h, ctypes.byref(n)))
        out = []
        for k in range(n.value):
            info = XsProcess()
            _check(load().pic_xs_get_process(self._h, k, ctypes.byref(info), 0, None, None))
            e = np.zeros(info.n_points)
            s = np.zeros(info.n_points)
            _check(load().pic_xs_get_process(self._h, k, ctypes.byref(info), info.n_points, _ptr(e), _ptr(s)))
            out.append(dict(species=info.species, type=info.type, product_species=info.product_species,
                            threshold_eV=info.threshold_eV, peak_m2=info.peak_m2, energies=e, sigmas=s))
        return out


def elasticCollisions(particles, species, density, dt, energies, sigmas, seed=1, step=0):
    """elasticCollisions(particles, density, dt, xs) -- mcc.cu:114-150 (null-collision, Philox)."""
    e = np.ascontiguousarray(energies, dtype=np.float64)
    s = np.ascontiguousarray(sigmas, dtype=np.float64)
    _check(load().pic_elastic_collisions(particles._h, int(species), float(density), float(dt), len(e), _ptr(e),
                                         _ptr(s), int(seed), int(step)))


def sim_params(**over):
    p = SimParams()
    _check(load().pic_sim_default_params(ctypes.byref(p)))
    for k, v in over.items():
        if k == "mg":
            p.mg = v
        else:
            setattr(p, k, v)
    return p


class Simulation:
    """The step driver (updateSimulationCUDA, main.cu:37-140) bound to its objects."""

    _owned = True

    @classmethod
    def _view(cls, handle, particles, grid):
        o = cls.__new__(cls)
        o._h = ctypes.c_void_p(handle)
        o._owned = False
        o.particles, o.grid, o.walls, o.xs, o.params = particles, grid, None, None, None
        return o

    def __init__(self, particles, grid, walls=None, xs=None, params=None, **over):
        self.params = params if params is not None else sim_params(**over)
        self._h = ctypes.c_void_p()
        self.particles, self.grid, self.walls, self.xs = particles, grid, walls, xs
        particles.grid = grid
        _check(load().pic_sim_create(particles._h, grid._h, walls._h if walls else None, xs._h if xs else None,
                                     ctypes.byref(self.params), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            _check(load().pic_sim_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._owned and self._h and _lib is not None:
                _lib.pic_sim_destroy(self._h)
        except Exception:
            pass

    def step(self, n=1):
        _check(load().pic_sim_step(self._h, int(n)))

    def sync(self):
        _check(load().pic_sim_sync(self._h))

    def counters(self, species):
        out = np.zeros(16, dtype=np.uint64)
        _check(load().pic_sim_counters(self._h, int(species), _ptr(out)))
        return out

    def stats(self):
        r = SimReport()
        _check(load().pic_sim_stats(self._h, ctypes.byref(r)))
        return r

    def diagnostics(self):
        """Energy and momentum of the synchronized state: per species count,
        kinetic energy and momentum (lists), field energy and |phi|_2."""
        d = Diagnostics()
        _check(load().pic_sim_diagnostics(self._h, ctypes.byref(d)))
        ns = d.n_species
        return dict(count=list(d.count[:ns]), kinetic=list(d.kinetic[:ns]), px=list(d.px[:ns]),
                    py=list(d.py[:ns]), field_energy=d.field_energy, phi_l2=d.phi_l2)

    def mcc_info(self, species):
        info = MccInfo()
        _check(load().pic_sim_mcc_info(self._h, int(species), ctypes.byref(info), None, 0))
        tab = np.zeros(max(1, info.n_proc * info.K))
        _check(load().pic_sim_mcc_info(self._h, int(species), ctypes.byref(info), _ptr(tab), tab.size))
        return _mcc_dict(info, tab)

    def rebin(self, bin_w=0, bin_h=0):
        _check(load().pic_sim_rebin(self._h, int(bin_w), int(bin_h)))

    def use_graph(self, on=True):
        """Run each step after the first as one graph launch."""
        _check(load().pic_sim_use_graph(self._h, 1 if on else 0))

    def write_frame(self, path):
        """Every species' positions (domain units) into a frame file, replaced
        atomically; read it back with read_frame."""
        _check(load().pic_sim_write_frame(self._h, str(path).encode()))

    def set_wall_lists(self, on=True, margin=2):
        """Per-bin candidate walls for the wall test (on by default): each bin
        lists the walls within margin + 1 cells of it."""
        _check(load().pic_sim_set_wall_lists(self._h, 1 if on else 0, int(margin)))

    def wall_list_stats(self):
        """(bins, list entries, longest list) of the candidate walls."""
        nb, ne, mx = _i(), ctypes.c_longlong(), _i()
        _check(load().pic_sim_wall_list_stats(self._h, ctypes.byref(nb), ctypes.byref(ne), ctypes.byref(mx)))
        return nb.value, ne.value, mx.value

    def autotune(self, bins=True, multigrid=True):
        """Time candidate bin tiles and multigrid settings on the device and
        keep the fastest; returns what was chosen."""
        r = TuneReport()
        what = (TUNE_BINS if bins else 0) | (TUNE_MG if multigrid else 0)
        _check(load().pic_sim_autotune(self._h, what, ctypes.byref(r)))
        nl = max(0, r.levels - 1)
        return {"bin": (r.bin_w, r.bin_h), "push_ms": r.push_ms, "omega": r.omega, "levels": r.levels,
                "pre": list(r.pre)[:nl], "post": list(r.post)[:nl], "solve_ms": r.solve_ms,
                "solve_cycles": r.solve_cycles, "trials": r.trials, "inner": r.inner}

    def set_electrodes(self, mask, phi_bc=None):
        """Embedded electrodes in the field solve (see Multigrid.set_electrodes)."""
        self._electrodes = _electrode_args(mask, phi_bc, self.grid.m)
        _check(load().pic_sim_set_electrodes(self._h, *self._electrodes))

    def set_level_sweeps(self, pre, post):
        """Sweeps of each multigrid level, finest first."""
        n = len(pre)
        a = (ctypes.c_int * max(1, n))(*pre)
        b = (ctypes.c_int * max(1, n))(*post)
        _check(load().pic_sim_set_level_sweeps(self._h, n, a, b))


def updateSimulation(sim, steps):
    """updateSimulationCUDA(particles, grid, boundaries, step, dt, steps, timing)."""
    sim.step(steps)
    return sim.stats()


class Ensemble:
    """Independent plasmas stepped together, one launch per stage for all."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(load().pic_ensemble_create(ctypes.byref(self._h)))
        self.members = []

    def add(self, sim):
        _check(load().pic_ensemble_add(self._h, sim._h))
        self.members.append(sim)

    def remove(self, sim):
        _check(load().pic_ensemble_remove(self._h, sim._h))
        self.members.remove(sim)

    def step(self, n=1):
        _check(load().pic_ensemble_step(self._h, int(n)))

    def size(self):
        n = _i()
        _check(load().pic_ensemble_size(self._h, ctypes.byref(n)))
        return n.value

    def use_graph(self, on=True):
        _check(load().pic_ensemble_use_graph(self._h, 1 if on else 0))

    def close(self):
        if self._h:
            _check(load().pic_ensemble_destroy(self._h))
            self._h = ctypes.c_void_p()
            self.members = []

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.pic_ensemble_destroy(self._h)
        except Exception:
            pass


class Server:
    """Continuous batching of PIC jobs, admitted by HBM footprint."""

    def __init__(self, hbm_budget_bytes, max_active=256, precision=PREC_FP32, params=None, keep_finished=False,
                 **over):
        self.params = params if params is not None else sim_params(**over)
        self.precision = precision
        self._h = ctypes.c_void_p()
        _check(load().pic_server_create(float(hbm_budget_bytes), int(max_active), int(precision),
                                        ctypes.byref(self.params), int(bool(keep_finished)), ctypes.byref(self._h)))

    def submit(self, cells, ppc, steps, seed, arrival_s=0.0, mcc=True, te_eV=2.0, density=1e15, gas_density=1e21):
        d = JobDesc(int(cells), float(ppc), int(steps), int(seed), float(te_eV), float(density), float(gas_density),
                    float(arrival_s), int(bool(mcc)))
        jid = ctypes.c_int()
        _check(load().pic_server_submit(self._h, ctypes.byref(d), ctypes.byref(jid)))
        return jid.value

    def tick(self, now_s):
        a, w = ctypes.c_int(), ctypes.c_int()
        _check(load().pic_server_tick(self._h, float(now_s), ctypes.byref(a), ctypes.byref(w)))
        return a.value, w.value

    def result(self, jid):
        r = JobResult()
        _check(load().pic_server_job_result(self._h, int(jid), ctypes.byref(r)))
        return r

    def step_ms(self, jid):
        n = ctypes.c_int()
        _check(load().pic_server_job_step_ms(self._h, int(jid), 0, None, ctypes.byref(n)))
        out = np.zeros(max(1, n.value), dtype=np.float32)
        _check(load().pic_server_job_step_ms(self._h, int(jid), out.size, _ptr(out), ctypes.byref(n)))
        return out[:n.value]

    def handles(self, jid):
        """(Particles2D, Grid2D, Simulation) views of a running or finished job."""
        p, g, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(load().pic_server_job_handles(self._h, int(jid), ctypes.byref(p), ctypes.byref(g), ctypes.byref(s)))
        grid = Grid2D._view(g.value)
        parts = Particles2D._view(p.value, self.precision, grid)
        return parts, grid, Simulation._view(s.value, parts, grid)

    def release(self, jid):
        _check(load().pic_server_release(self._h, int(jid)))

    def used_bytes(self):
        u = ctypes.c_double()
        _check(load().pic_server_used_bytes(self._h, ctypes.byref(u)))
        return u.value

    def close(self):
        if self._h:
            _check(load().pic_server_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.pic_server_destroy(self._h)
        except Exception:
            pass
