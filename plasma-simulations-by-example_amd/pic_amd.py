"""Python binding of libpicamd (ctypes over the C ABI in include/picamd.h).

Mirrors the reference class API (ch3/ver2/World.h, Species.h, PotentialSolver.h):
``World``, ``Species``, ``PotentialSolver``.  Every call goes to the HIP library; there is
no CPU fallback -- without the built library or a GPU the constructors raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import subprocess
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpicamd.so")

PIC_OK, PIC_NOT_CONVERGED = 0, 1
MIXED, FP64 = 0, 1
REFLECT, ABSORB = 0, 1
GS, PCG, QN, SOR_LINEAR, PCG_SYM, MG = 0, 1, 2, 3, 4, 5
PHI, RHO, NODE_VOL, EF, OBJECT_ID = 0, 1, 2, 3, 4
T_PUSH, T_DEPOSIT, T_SORT, T_SOLVE, T_EF, T_RHO, T_INJECT = range(7)
T_NAMES = ["push", "deposit", "sort", "solve", "ef", "rho", "inject", "push_kernel"]

# every symbol include/picamd.h declares
EXPORTS = [
    "pic_abi_version", "pic_last_error", "pic_device_count",
    "pic_world_create", "pic_world_destroy", "pic_world_set_extents", "pic_world_add_sphere",
    "pic_world_add_inlet", "pic_world_get_field", "pic_world_set_field",
    "pic_world_compute_charge_density", "pic_world_get_pe", "pic_world_sync", "pic_world_stream",
    "pic_world_set_timing", "pic_world_get_timers",
    "pic_species_create", "pic_species_destroy", "pic_species_add_particles",
    "pic_species_set_particles", "pic_species_get_particles", "pic_species_get_np",
    "pic_species_advance", "pic_species_compute_number_density", "pic_species_get_den",
    "pic_species_get_den_ave", "pic_species_update_averages", "pic_species_diag", "pic_species_diag_noalloc",
    "pic_species_sort", "pic_species_inject_beam", "pic_species_load_box_device",
    "pic_solver_create", "pic_solver_destroy", "pic_solver_set_reference_values",
    "pic_solver_solve", "pic_solver_get_info", "pic_solver_solve_qn", "pic_solver_compute_ef",
    "pic_adjoint_create", "pic_adjoint_destroy", "pic_adjoint_forward", "pic_adjoint_backward",
    "pic_adjoint_stats",
    "pic_ensemble_create", "pic_ensemble_destroy", "pic_ensemble_set_particles",
    "pic_ensemble_load_box", "pic_ensemble_get_particles", "pic_ensemble_sort",
    "pic_ensemble_set_tuning", "pic_ensemble_solve_initial", "pic_ensemble_step",
    "pic_ensemble_get_field", "pic_ensemble_solver_status", "pic_ensemble_diag",
    "pic_ensemble_stats",
    "pic_tuning_load", "pic_tuning_lookup", "pic_species_set_launch", "pic_solver_set_mg",
    "pic_species_set_frame_velocity",
]


class Box(C.Structure):
    """pic_box: nodes per axis and extents of one grounded box."""
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("x0", C.c_double * 3), ("xm", C.c_double * 3)]


class AdjointConfig(C.Structure):
    """pic_adjoint_config (include/picamd.h)."""
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("x0", C.c_double * 3), ("xm", C.c_double * 3),
                ("sphere_c", C.c_double * 3), ("sphere_r", C.c_double),
                ("dt", C.c_double), ("steps", C.c_int), ("checkpoint_every", C.c_int),
                ("mass", C.c_double), ("charge", C.c_double), ("n0", C.c_double),
                ("T_ion", C.c_double), ("n_init", C.c_long), ("n_inject", C.c_long),
                ("soft_width", C.c_double), ("cg_tol", C.c_double), ("nr_tol", C.c_double),
                ("adj_rel_tol", C.c_double), ("seed", C.c_uint64)]

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_u32p = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class PicError(RuntimeError):
    pass


_lib = None
_exiting = False


def _at_exit():
    # objects still alive at interpreter exit are not released through the library: the
    # HIP runtime may already be tearing down, and the process frees the device anyway
    global _exiting
    _exiting = True


atexit.register(_at_exit)


def build():
    """Compiles libpicamd.so (hipcc, gfx950) in the package directory."""
    subprocess.run(["make", "-s", "-C", HERE, "lib"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PicError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        sig = {
            "pic_abi_version": (C.c_int, []),
            "pic_last_error": (C.c_char_p, []),
            "pic_device_count": (C.c_int, [C.POINTER(C.c_int)]),
            "pic_world_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
            "pic_world_destroy": (C.c_int, [_vp]),
            "pic_world_set_extents": (C.c_int, [_vp, _dp, _dp]),
            "pic_world_add_sphere": (C.c_int, [_vp, _dp, C.c_double, C.c_double]),
            "pic_world_add_inlet": (C.c_int, [_vp]),
            "pic_world_get_field": (C.c_int, [_vp, C.c_int, _dp]),
            "pic_world_set_field": (C.c_int, [_vp, C.c_int, _dp]),
            "pic_world_compute_charge_density": (C.c_int, [_vp, C.POINTER(_vp), C.c_int]),
            "pic_world_get_pe": (C.c_int, [_vp, _dp]),
            "pic_world_sync": (C.c_int, [_vp]),
            "pic_world_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
            "pic_world_set_timing": (C.c_int, [_vp, C.c_int]),
            "pic_world_get_timers": (C.c_int, [_vp, _dp, _lp, C.c_int]),
            "pic_species_create": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                             C.c_long, C.c_int, C.POINTER(_vp)]),
            "pic_species_destroy": (C.c_int, [_vp]),
            "pic_species_add_particles": (C.c_int, [_vp, C.c_long, _dp, _dp, _dp, C.c_double,
                                                    _lp]),
            "pic_species_set_particles": (C.c_int, [_vp, C.c_long, _dp, _dp, _dp, _u32p]),
            "pic_species_get_particles": (C.c_int, [_vp, _dp, _dp, _dp, _u32p]),
            "pic_species_get_np": (C.c_int, [_vp, _lp]),
            "pic_species_advance": (C.c_int, [_vp, C.c_int, C.c_double]),
            "pic_species_compute_number_density": (C.c_int, [_vp]),
            "pic_species_get_den": (C.c_int, [_vp, _dp]),
            "pic_species_get_den_ave": (C.c_int, [_vp, _dp]),
            "pic_species_update_averages": (C.c_int, [_vp]),
            "pic_species_diag": (C.c_int, [_vp, _dp]),
            "pic_species_diag_noalloc": (C.c_int, [_vp, _dp]),
            "pic_species_sort": (C.c_int, [_vp]),
            "pic_species_inject_beam": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double,
                                                  C.c_double, C.c_uint64, C.c_uint64, _lp]),
            "pic_species_load_box_device": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_long,
                                                      C.c_double, C.c_double, C.c_uint64]),
            "pic_solver_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.POINTER(_vp)]),
            "pic_solver_destroy": (C.c_int, [_vp]),
            "pic_solver_set_reference_values": (C.c_int, [_vp, C.c_double, C.c_double,
                                                          C.c_double]),
            "pic_solver_solve": (C.c_int, [_vp, _dp, C.POINTER(C.c_int)]),
            "pic_solver_get_info": (C.c_int, [_vp, C.POINTER(C.c_int)]),
            "pic_solver_solve_qn": (C.c_int, [_vp]),
            "pic_solver_compute_ef": (C.c_int, [_vp]),
            "pic_adjoint_create": (C.c_int, [C.POINTER(AdjointConfig), C.POINTER(_vp)]),
            "pic_adjoint_destroy": (C.c_int, [_vp]),
            "pic_adjoint_forward": (C.c_int, [_vp, _dp, _dp, _vp, _vp]),
            "pic_adjoint_backward": (C.c_int, [_vp, C.c_double, _vp, _vp, _dp]),
            "pic_adjoint_stats": (C.c_int, [_vp, _dp]),
            "pic_ensemble_create": (C.c_int, [C.c_int, C.POINTER(Box), C.c_int, _dp, _dp, _lp,
                                              C.c_int, C.c_int, C.POINTER(_vp)]),
            "pic_ensemble_destroy": (C.c_int, [_vp]),
            "pic_ensemble_set_particles": (C.c_int, [_vp, C.c_int, C.c_int, C.c_long, _dp, _dp,
                                                     _dp]),
            "pic_ensemble_load_box": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                                C.c_uint64]),
            "pic_ensemble_get_particles": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp,
                                                     _u32p]),
            "pic_ensemble_sort": (C.c_int, [_vp]),
            "pic_ensemble_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                  C.c_int]),
            "pic_ensemble_solve_initial": (C.c_int, [_vp, C.c_int, C.c_double]),
            "pic_ensemble_step": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_double]),
            "pic_ensemble_get_field": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
            "pic_ensemble_solver_status": (C.c_int, [_vp, C.POINTER(C.c_int), _dp]),
            "pic_ensemble_diag": (C.c_int, [_vp, _dp]),
            "pic_ensemble_stats": (C.c_int, [_vp, _dp]),
            "pic_tuning_load": (C.c_int, [C.c_char_p]),
            "pic_tuning_lookup": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_char_p,
                                            C.c_double, _dp]),
            "pic_species_set_launch": (C.c_int, [_vp, C.c_long, C.c_int]),
            "pic_species_set_frame_velocity": (C.c_int, [_vp, _dp]),
            "pic_solver_set_mg": (C.c_int, [_vp, C.c_int, C.c_int, C.c_long]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, allow_nc=False):
    if rc == PIC_OK or (allow_nc and rc == PIC_NOT_CONVERGED):
        return rc
    msg = lib().pic_last_error().decode(errors="replace")
    raise PicError(f"libpicamd error {rc}: {msg}")


def _d(a):
    return a.ctypes.data_as(_dp)


def _v3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def device_count():
    n = C.c_int(0)
    _check(lib().pic_device_count(C.byref(n)))
    return n.value


class World:
    """World(ni, nj, nk) with node fields on the GPU (ch3/ver2/World.h:38-147)."""

    def __init__(self, ni, nj, nk):
        self.h = _vp()
        # species and solvers of this world: closed before it, whatever order the garbage
        # collector finalizes them in
        self._children = weakref.WeakSet()
        _check(lib().pic_world_create(ni, nj, nk, C.byref(self.h)))
        self.ni, self.nj, self.nk = ni, nj, nk
        self.nn = ni * nj * nk
        self.x0 = self.xm = self.dh = None

    def close(self):
        if self.h and not _exiting:
            for c in list(self._children):
                c.close()
            lib().pic_world_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_extents(self, x0, xm):
        x0 = _v3(x0)
        xm = _v3(xm)
        _check(lib().pic_world_set_extents(self.h, _d(x0), _d(xm)))
        self.x0, self.xm = x0, xm
        self.dh = (xm - x0) / (np.array([self.ni, self.nj, self.nk]) - 1)

    def add_sphere(self, c, r, phi_s):
        _check(lib().pic_world_add_sphere(self.h, _d(_v3(c)), r, phi_s))

    def add_inlet(self):
        _check(lib().pic_world_add_inlet(self.h))

    def field(self, which):
        out = np.zeros(self.nn * (3 if which == EF else 1))
        _check(lib().pic_world_get_field(self.h, which, _d(out)))
        return out

    def set_field(self, which, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.nn * (3 if which == EF else 1)
        _check(lib().pic_world_set_field(self.h, which, _d(v)))

    def compute_charge_density(self, species):
        arr = (_vp * len(species))(*[s.h for s in species])
        _check(lib().pic_world_compute_charge_density(self.h, arr, len(species)))

    def get_pe(self):
        pe = C.c_double(0)
        _check(lib().pic_world_get_pe(self.h, C.byref(pe)))
        return pe.value

    def sync(self):
        _check(lib().pic_world_sync(self.h))

    def stream(self):
        s = _vp()
        _check(lib().pic_world_stream(self.h, C.byref(s)))
        return s.value

    def set_timing(self, on=True):
        _check(lib().pic_world_set_timing(self.h, 1 if on else 0))

    def timers(self, reset=False):
        ms = np.zeros(len(T_NAMES))
        calls = np.zeros(len(T_NAMES), dtype=np.int64)
        _check(lib().pic_world_get_timers(self.h, _d(ms), calls.ctypes.data_as(_lp),
                                          1 if reset else 0))
        return {name: (ms[i], int(calls[i])) for i, name in enumerate(T_NAMES)}


class Species:
    """Species(name, mass, charge, mpw0, world) (ch3/ver2/Species.h:22-27)."""

    def __init__(self, world: World, mass, charge, mpw0=0.0, precision=MIXED, capacity=1 << 16,
                 track_ids=False, name=""):
        self.world = world
        self.name = name
        self.mass, self.charge, self.mpw0 = mass, charge, mpw0
        self.precision = precision
        self.track_ids = track_ids
        self.h = _vp()
        _check(lib().pic_species_create(world.h, mass, charge, mpw0, precision, int(capacity),
                                        1 if track_ids else 0, C.byref(self.h)))
        world._children.add(self)

    def close(self):
        if self.h and not _exiting and self.world.h:
            lib().pic_species_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_particles(self, pos, vel, mpw, dt):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        mpw = np.ascontiguousarray(mpw, dtype=np.float64).reshape(-1)
        n = C.c_long(0)
        _check(lib().pic_species_add_particles(self.h, len(mpw), _d(pos), _d(vel), _d(mpw), dt,
                                               C.byref(n)))
        return n.value

    def set_particles(self, pos, vel, mpw, ids=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        mpw = np.ascontiguousarray(mpw, dtype=np.float64).reshape(-1)
        idp = None
        if ids is not None:
            ids = np.ascontiguousarray(ids, dtype=np.uint32)
            idp = ids.ctypes.data_as(_u32p)
        _check(lib().pic_species_set_particles(self.h, len(mpw), _d(pos), _d(vel), _d(mpw), idp))

    def np(self):
        n = C.c_long(0)
        _check(lib().pic_species_get_np(self.h, C.byref(n)))
        return n.value

    def particles(self):
        """(pos (n,3), vel (n,3), mpw (n,), ids or None) in device order."""
        n = self.np()
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        mpw = np.zeros(n)
        ids = np.zeros(n, dtype=np.uint32) if self.track_ids else None
        if n:
            _check(lib().pic_species_get_particles(
                self.h, _d(pos), _d(vel), _d(mpw),
                ids.ctypes.data_as(_u32p) if ids is not None else None))
        return pos, vel, mpw, ids

    def advance(self, bc, dt):
        _check(lib().pic_species_advance(self.h, bc, dt))

    def compute_number_density(self):
        _check(lib().pic_species_compute_number_density(self.h))

    def den(self):
        out = np.zeros(self.world.nn)
        _check(lib().pic_species_get_den(self.h, _d(out)))
        return out

    def den_ave(self):
        out = np.zeros(self.world.nn)
        _check(lib().pic_species_get_den_ave(self.h, _d(out)))
        return out

    def update_averages(self):
        _check(lib().pic_species_update_averages(self.h))

    def diag(self):
        """[real_count, px, py, pz, KE]"""
        out = np.zeros(5)
        _check(lib().pic_species_diag(self.h, _d(out)))
        return out

    def diag_noalloc(self):
        """diag() through the entry point that allocates nothing (static device scratch)."""
        out = np.zeros(5)
        _check(lib().pic_species_diag_noalloc(self.h, _d(out)))
        return out

    def sort(self):
        _check(lib().pic_species_sort(self.h))

    def set_launch(self, push_chunk, xcd_stripe):
        """Push launch shape: particles per workgroup, map entries per XCD stripe."""
        _check(lib().pic_species_set_launch(self.h, int(push_chunk), int(xcd_stripe)))

    def set_frame_velocity(self, u):
        """PIC_MIXED velocities are stored as v - u in FP32 (a beam: u = its drift)."""
        _check(lib().pic_species_set_frame_velocity(self.h, _d(_v3(u))))

    def inject_beam(self, v_drift, den, T, dt, seed, counter):
        n = C.c_long(0)
        _check(lib().pic_species_inject_beam(self.h, v_drift, den, T, dt, seed, counter,
                                             C.byref(n)))
        return n.value

    def load_box_device(self, x1, x2, num_den, num_mp, v_drift=0.0, T=0.0, seed=1):
        _check(lib().pic_species_load_box_device(self.h, _d(_v3(x1)), _d(_v3(x2)), num_den,
                                                 int(num_mp), v_drift, T, seed))


class PotentialSolver:
    """PotentialSolver(world, type, max_it, tol) (ch3/ver2/PotentialSolver.h:46-50)."""

    def __init__(self, world: World, stype=GS, max_it=10000, tol=1e-4):
        self.world = world
        self.h = _vp()
        _check(lib().pic_solver_create(world.h, stype, max_it, tol, C.byref(self.h)))
        world._children.add(self)

    def close(self):
        if self.h and not _exiting and self.world.h:
            lib().pic_solver_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reference_values(self, phi0, Te0, n0):
        _check(lib().pic_solver_set_reference_values(self.h, phi0, Te0, n0))

    def solve(self):
        """Returns (converged, residual, iterations)."""
        r = C.c_double(0)
        it = C.c_int(0)
        rc = _check(lib().pic_solver_solve(self.h, C.byref(r), C.byref(it)), allow_nc=True)
        return rc == PIC_OK, r.value, it.value

    def info(self):
        """What the last solve did: {"newton", "iters", "fallbacks", "converged"}."""
        a = (C.c_int * 4)()
        _check(lib().pic_solver_get_info(self.h, a))
        return {"newton": a[0], "iters": a[1], "fallbacks": a[2], "converged": bool(a[3])}

    def set_mg(self, sweeps=0, gal_from=0, fuse_max=0):
        """Multigrid shape (0: the tuning table's, else the default); rebuilt by the next solve."""
        _check(lib().pic_solver_set_mg(self.h, int(sweeps), int(gal_from), int(fuse_max)))

    def solve_qn(self):
        _check(lib().pic_solver_solve_qn(self.h))

    def compute_ef(self):
        _check(lib().pic_solver_compute_ef(self.h))


class DeviceArray:
    """n doubles of device memory from the HIP runtime libpicamd uses (hipMalloc through
    ctypes), for callers of the entry points that take device pointers without torch."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            lib()  # the library's runtime is loaded first; bind that same file
            # (torch, when imported, maps its own copy from its wheel: not that one)
            path = "libamdhip64.so.7"
            with open("/proc/self/maps") as f:
                for line in f:
                    if "libamdhip64.so" in line and "/" in line and "site-packages" not in line:
                        path = line[line.index("/"):].strip()
                        break
            h = C.CDLL(path)
            h.hipMalloc.argtypes = [C.POINTER(_vp), C.c_size_t]
            h.hipFree.argtypes = [_vp]
            h.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, n):
        self.n = int(n)
        self.ptr = _vp()
        rc = self.hip().hipMalloc(C.byref(self.ptr), 8 * max(self.n, 1))
        if rc != 0:
            raise PicError(f"hipMalloc of {self.n} doubles failed ({rc})")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(a.size)
        cls.hip().hipDeviceSynchronize()
        if cls.hip().hipMemcpy(d.ptr, a.ctypes.data_as(_vp), 8 * a.size, 1) != 0:  # host to device
            raise PicError("hipMemcpy to the device failed")
        return d

    def get(self):
        out = np.empty(self.n)
        self.hip().hipDeviceSynchronize()
        if self.hip().hipMemcpy(out.ctypes.data_as(_vp), self.ptr, 8 * self.n, 2) != 0:  # device to host
            raise PicError("hipMemcpy to the host failed")
        return out

    @property
    def address(self):
        return self.ptr.value or 0

    def close(self):
        if self.ptr and not _exiting:
            self.hip().hipFree(self.ptr)
        self.ptr = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SphereAdjointRun:
    """The differentiable flow past a sphere (pic_adjoint_*): K steps of the ch3/ver2 loop in
    FP64 with soft absorption, and their hand-written adjoint over checkpoints.

    forward(theta, den_ptr, phi_ptr) runs the steps at theta = (drift [m/s], sphere
    potential [V], Te [eV]) and returns the ion current to the sphere; den_ptr / phi_ptr
    (device addresses of nn doubles, or 0) receive the density and potential after the last
    step.  backward(dcurrent, dden_ptr, dphi_ptr) returns dL/dtheta at the last forward's
    theta.  The torch.autograd.Function in adjoint.py wraps these two calls."""

    def __init__(self, **cfg):
        c = AdjointConfig()
        for k, v in cfg.items():
            if k in ("x0", "xm", "sphere_c"):
                getattr(c, k)[:] = [float(t) for t in v]
            else:
                setattr(c, k, v)
        self.cfg = c
        self.nn = c.ni * c.nj * c.nk
        self.h = _vp()
        _check(lib().pic_adjoint_create(C.byref(c), C.byref(self.h)))
        self.theta = None

    def close(self):
        if self.h and not _exiting:
            lib().pic_adjoint_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, theta, den_ptr=0, phi_ptr=0):
        th = np.ascontiguousarray(theta, dtype=np.float64).reshape(3)
        cur = C.c_double(0)
        _check(lib().pic_adjoint_forward(self.h, _d(th), C.byref(cur), _vp(den_ptr or None),
                                         _vp(phi_ptr or None)))
        self.theta = th.copy()
        return cur.value

    def backward(self, dcurrent, dden_ptr=0, dphi_ptr=0):
        g = np.zeros(3)
        _check(lib().pic_adjoint_backward(self.h, float(dcurrent), _vp(dden_ptr or None),
                                          _vp(dphi_ptr or None), _d(g)))
        return g

    def stats(self):
        """[forward ms, backward ms, Newton solves of the forward, adjoint CG iterations]"""
        out = np.zeros(4)
        _check(lib().pic_adjoint_stats(self.h, _d(out)))
        return out

    def wake_mask(self, z0=0.22, z1=0.30, r=0.05):
        """Nodes behind the sphere: the cylinder of radius r about its axis, z in [z0, z1]
        (flat order U = k*ni*nj + j*ni + i)."""
        c = self.cfg
        dh = [(c.xm[a] - c.x0[a]) / (n - 1) for a, n in enumerate((c.ni, c.nj, c.nk))]
        k, j, i = np.meshgrid(np.arange(c.nk), np.arange(c.nj), np.arange(c.ni), indexing="ij")
        x = c.x0[0] + dh[0] * i - c.sphere_c[0]
        y = c.x0[1] + dh[1] * j - c.sphere_c[1]
        z = c.x0[2] + dh[2] * k
        return ((x * x + y * y <= r * r) & (z >= z0) & (z <= z1)).ravel()


class Ensemble:
    """Many independent ch2 grounded boxes stepped together (pic_ensemble_*): one launch per
    phase over a box-descriptor table, each step one hipGraph replay.

    boxes: list of (nodes (ni, nj, nk), x0, xm); species: list of (mass, charge);
    counts[b][s]: particles of species s in box b."""

    PHI, RHO, EF = 0, 1, 2

    def __init__(self, boxes, species, counts, precision=MIXED, track_ids=False):
        nb, ns = len(boxes), len(species)
        arr = (Box * nb)()
        self.nodes = []
        for b, (n, x0, xm) in enumerate(boxes):
            arr[b].ni, arr[b].nj, arr[b].nk = (int(t) for t in n)
            arr[b].x0[:] = [float(t) for t in x0]
            arr[b].xm[:] = [float(t) for t in xm]
            self.nodes.append(tuple(int(t) for t in n))
        mass = np.ascontiguousarray([m for m, _ in species], dtype=np.float64)
        charge = np.ascontiguousarray([q for _, q in species], dtype=np.float64)
        cnt = np.ascontiguousarray(np.asarray(counts, dtype=np.int64).reshape(nb * ns))
        self.nbox, self.nsp = nb, ns
        self.counts = cnt.reshape(nb, ns)
        self.h = _vp()
        _check(lib().pic_ensemble_create(nb, arr, ns, _d(mass), _d(charge),
                                         cnt.ctypes.data_as(_lp), precision, int(track_ids),
                                         C.byref(self.h)))

    def close(self):
        if self.h and not _exiting:
            lib().pic_ensemble_destroy(self.h)
        self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nn(self, b):
        ni, nj, nk = self.nodes[b]
        return ni * nj * nk

    def set_particles(self, b, s, pos, vel, mpw):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mpw = np.ascontiguousarray(mpw, dtype=np.float64)
        _check(lib().pic_ensemble_set_particles(self.h, b, s, len(mpw), _d(pos), _d(vel), _d(mpw)))

    def load_box(self, b, s, x1, x2, num_den, seed=1):
        _check(lib().pic_ensemble_load_box(self.h, b, s, _d(_v3(x1)), _d(_v3(x2)), num_den, seed))

    def particles(self, b, s, ids=False):
        n = int(self.counts[b, s])
        pos, vel, mpw = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        idv = np.zeros(n, dtype=np.uint32) if ids else None
        _check(lib().pic_ensemble_get_particles(self.h, b, s, _d(pos), _d(vel), _d(mpw),
                                                idv.ctypes.data_as(_u32p) if ids else None))
        return (pos, vel, mpw, idv) if ids else (pos, vel, mpw)

    def sort(self):
        _check(lib().pic_ensemble_sort(self.h))

    def set_tuning(self, chunk_particles=4096, window_nodes=6144, sor_threads=1024, sort_every=20,
                   use_graph=True):
        _check(lib().pic_ensemble_set_tuning(self.h, chunk_particles, window_nodes, sor_threads,
                                             sort_every, int(use_graph)))

    def solve_initial(self, max_it=10000, tol=1e-4):
        _check(lib().pic_ensemble_solve_initial(self.h, max_it, tol))

    def step(self, nsteps, dt, max_it=10000, tol=1e-4):
        _check(lib().pic_ensemble_step(self.h, nsteps, dt, max_it, tol))

    def field(self, b, which):
        n = self.nn(b) * (3 if which == self.EF else 1)
        out = np.zeros(n)
        _check(lib().pic_ensemble_get_field(self.h, b, which, _d(out)))
        return out

    def den(self, b, s):
        out = np.zeros(self.nn(b))
        _check(lib().pic_ensemble_get_field(self.h, b, 10 + s, _d(out)))
        return out

    def solver_status(self):
        it = (C.c_int * self.nbox)()
        res = np.zeros(self.nbox)
        _check(lib().pic_ensemble_solver_status(self.h, it, _d(res)))
        return np.array(it[:]), res

    def diag(self):
        out = np.zeros(5 * self.nbox * self.nsp)
        _check(lib().pic_ensemble_diag(self.h, _d(out)))
        return out.reshape(self.nbox, self.nsp, 5)

    def stats(self):
        """{"ms_per_step", "graph_kernels", "particles", "nodes", "push_workgroups"}"""
        out = np.zeros(5)
        _check(lib().pic_ensemble_stats(self.h, _d(out)))
        return {"ms_per_step": out[0], "graph_kernels": int(out[1]), "particles": int(out[2]),
                "nodes": int(out[3]), "push_workgroups": int(out[4])}


def tuning_load(path=None):
    """(Re)loads the tuning table (None: PICAMD_TUNING, else tuning/table.txt beside the
    library)."""
    _check(lib().pic_tuning_load(path.encode() if path else None))


def tuning_lookup(kind, nodes, particles, key, default):
    """The tuning table's value of key for kind nearest to (nodes, particles)."""
    v = C.c_double(0)
    _check(lib().pic_tuning_lookup(kind.encode(), float(nodes), float(particles), key.encode(),
                                   float(default), C.byref(v)))
    return v.value


def tuned_ensemble_params(ens):
    """The launch shapes the library took from the tuning table for this ensemble."""
    st = ens.stats()
    n, p = st["nodes"], max(st["particles"], 1)
    return dict(chunk_particles=int(tuning_lookup("ensemble", n, p, "chunk", 4096)),
                window_nodes=int(tuning_lookup("ensemble", n, p, "window", 6144)),
                sor_threads=int(tuning_lookup("ensemble", n, p, "sor_threads", 1024)),
                sort_every=int(tuning_lookup("ensemble", n, p, "sort_every", 20)))
