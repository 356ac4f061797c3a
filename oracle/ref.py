"""ctypes wrapper around the reference's own classes compiled into oracle/_ref.

TEST INFRASTRUCTURE ONLY (pins the oracle; bench.py's cpu_baseline times it).
``oracle/Makefile`` (``make ref``) compiles the reference's unchanged World / Species /
PotentialSolver / Source sources together with oracle/ref_driver.cpp; the reference
tree is needed only for that build, the resulting .so files are self-contained.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)

CHAPTERS = {"ch2": "libref_ch2.so", "ch3": "libref_ch3.so", "mt": "libref_mt.so"}
_libs = {}

GS, PCG, QN = 0, 1, 2


def available(ch="ch2"):
    return os.path.exists(os.path.join(REF_DIR, CHAPTERS[ch]))


def build(reference_root="/root/reference"):
    if not os.path.isdir(reference_root):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={reference_root}"], check=True)
    return True


def _load(ch):
    if ch in _libs:
        return _libs[ch]
    path = os.path.join(REF_DIR, CHAPTERS[ch])
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
    L = C.CDLL(path, mode=os.RTLD_LOCAL)
    sig = {
        "ref_chapter": (C.c_int, []),
        "ref_reseed": (None, [C.c_uint]),
        "ref_rnd": (C.c_double, []),
        "ref_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int]),
        "ref_destroy": (None, [C.c_void_p]),
        "ref_add_species": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double,
                                      C.c_double]),
        "ref_add_particles": (None, [C.c_void_p, C.c_int, C.c_long, _dp, _dp, _dp]),
        "ref_set_particles": (None, [C.c_void_p, C.c_int, C.c_long, _dp, _dp, _dp]),
        "ref_get_np": (C.c_long, [C.c_void_p, C.c_int]),
        "ref_get_particles": (None, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
        "ref_advance": (None, [C.c_void_p, C.c_int]),
        "ref_compute_number_density": (None, [C.c_void_p, C.c_int]),
        "ref_compute_charge_density": (None, [C.c_void_p]),
        "ref_advance_time": (C.c_int, [C.c_void_p]),
        "ref_get_ts": (C.c_int, [C.c_void_p]),
        "ref_get_field": (C.c_int, [C.c_void_p, C.c_int, _dp]),
        "ref_set_field": (C.c_int, [C.c_void_p, C.c_int, _dp]),
        "ref_make_solver": (None, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
        "ref_set_reference_values": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
        "ref_solve": (C.c_int, [C.c_void_p]),
        "ref_compute_ef": (None, [C.c_void_p]),
        "ref_get_pe": (C.c_double, [C.c_void_p]),
        "ref_species_diag": (None, [C.c_void_p, C.c_int, _dp]),
        "ref_time_particle_steps": (C.c_double, [C.c_void_p, C.c_int, _lp]),
        "ref_time_field_steps": (C.c_double, [C.c_void_p, C.c_int]),
    }
    if ch == "ch2":
        sig["ref_load_box"] = (None, [C.c_void_p, C.c_int, _dp, _dp, C.c_double, C.c_int])
        sig["ref_load_box_qs"] = (None, [C.c_void_p, C.c_int, _dp, _dp, C.c_double, _ip])
    if ch != "ch2":
        sig.update({
            "ref_add_sphere": (None, [C.c_void_p, _dp, C.c_double, C.c_double]),
            "ref_add_inlet": (None, [C.c_void_p]),
            "ref_add_cold_source": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
            "ref_sample_sources": (None, [C.c_void_p]),
            "ref_steady_state": (C.c_int, [C.c_void_p]),
            "ref_update_averages": (None, [C.c_void_p, C.c_int]),
        })
    if ch == "ch3":
        sig["ref_output_fields"] = (None, [C.c_void_p])
        sig["ref_output_diag"] = (None, [C.c_void_p])
    if ch == "mt":
        sig["ref_set_num_threads"] = (None, [C.c_void_p, C.c_int])
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _libs[ch] = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _v3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


PHI, RHO, NODE_VOL, EF, OBJECT_ID = 0, 1, 2, 3, 4


class RefSim:
    """One reference World plus its species, solver and sources."""

    def __init__(self, ch, ni, nj, nk, x0, xm, dt=1e-10, num_ts=0):
        self.L = _load(ch)
        self.ch = ch
        self.ni, self.nj, self.nk = ni, nj, nk
        self.nn = ni * nj * nk
        self.h = self.L.ref_create(ni, nj, nk, _d(_v3(x0)), _d(_v3(xm)), dt, num_ts)
        self.nsp = 0

    def close(self):
        if self.h:
            self.L.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reseed(self, seed):
        self.L.ref_reseed(seed)

    def add_species(self, name, mass, charge, mpw0=0.0):
        self.nsp += 1
        return self.L.ref_add_species(self.h, name.encode(), mass, charge, mpw0)

    def load_box(self, sp, x1, x2, num_den, num_mp):
        self.L.ref_load_box(self.h, sp, _d(_v3(x1)), _d(_v3(x2)), num_den, num_mp)

    def load_box_qs(self, sp, x1, x2, num_den, num_mp):
        n = np.ascontiguousarray(np.asarray(num_mp, dtype=np.int32))
        self.L.ref_load_box_qs(self.h, sp, _d(_v3(x1)), _d(_v3(x2)), num_den,
                               n.ctypes.data_as(_ip))

    def add_particles(self, sp, pos, vel, mpw):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mpw = np.ascontiguousarray(mpw, dtype=np.float64)
        self.L.ref_add_particles(self.h, sp, len(mpw), _d(pos), _d(vel), _d(mpw))

    def set_particles(self, sp, pos, vel, mpw):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mpw = np.ascontiguousarray(mpw, dtype=np.float64)
        self.L.ref_set_particles(self.h, sp, len(mpw), _d(pos), _d(vel), _d(mpw))

    def np_(self, sp):
        return self.L.ref_get_np(self.h, sp)

    def particles(self, sp):
        n = self.np_(sp)
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        mpw = np.zeros(n)
        self.L.ref_get_particles(self.h, sp, _d(pos), _d(vel), _d(mpw))
        return pos, vel, mpw

    def advance(self, sp):
        self.L.ref_advance(self.h, sp)

    def compute_number_density(self, sp):
        self.L.ref_compute_number_density(self.h, sp)

    def compute_charge_density(self):
        self.L.ref_compute_charge_density(self.h)

    def advance_time(self):
        return bool(self.L.ref_advance_time(self.h))

    def field(self, which):
        out = np.zeros(self.nn * (3 if which == EF else 1))
        if self.L.ref_get_field(self.h, which, _d(out)) != 0:
            raise ValueError(which)
        return out

    def den(self, sp):
        return self.field(100 + sp)

    def den_ave(self, sp):
        return self.field(200 + sp)

    def set_field(self, which, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if self.L.ref_set_field(self.h, which, _d(v)) != 0:
            raise ValueError(which)

    def make_solver(self, stype=GS, max_it=10000, tol=1e-4):
        self.L.ref_make_solver(self.h, stype, max_it, tol)

    def set_reference_values(self, phi0, Te0, n0):
        self.L.ref_set_reference_values(self.h, phi0, Te0, n0)

    def solve(self):
        return bool(self.L.ref_solve(self.h))

    def compute_ef(self):
        self.L.ref_compute_ef(self.h)

    def pe(self):
        return self.L.ref_get_pe(self.h)

    def species_diag(self, sp):
        """[np, real_count, px, py, pz, KE]"""
        out = np.zeros(6)
        self.L.ref_species_diag(self.h, sp, _d(out))
        return out

    def add_sphere(self, c, r, phi):
        self.L.ref_add_sphere(self.h, _d(_v3(c)), r, phi)

    def add_inlet(self):
        self.L.ref_add_inlet(self.h)

    def add_cold_source(self, sp, v_drift, den):
        return self.L.ref_add_cold_source(self.h, sp, v_drift, den)

    def sample_sources(self):
        self.L.ref_sample_sources(self.h)

    def steady_state(self):
        return bool(self.L.ref_steady_state(self.h))

    def update_averages(self, sp):
        self.L.ref_update_averages(self.h, sp)

    def output_fields(self):
        """Output::fields (ch3/ver2/Output.cpp:12-81) into results/ of the working directory."""
        self.L.ref_output_fields(self.h)

    def output_diag(self):
        """Output::diagOutput (ch3/ver2/Output.cpp:97-132): runtime_diags.csv of the working
        directory, opened by the first call of the process."""
        self.L.ref_output_diag(self.h)

    def set_num_threads(self, n):
        self.L.ref_set_num_threads(self.h, n)

    def time_particle_steps(self, nsteps):
        ps = C.c_long(0)
        t = self.L.ref_time_particle_steps(self.h, nsteps, C.byref(ps))
        return t, ps.value

    def time_field_steps(self, nsteps):
        return self.L.ref_time_field_steps(self.h, nsteps)
