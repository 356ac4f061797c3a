"""FP64 CPU oracle for the ES-PIC inner loop (ctypes over oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module, and only as the checker.
The product library (libpicamd.so) never routes through it.

Every function restates one reference function with the reference's own loop order
and floating-point operation order (see oracle/pic_oracle.c for the file:line each
follows).  How each function is pinned (tests/test_oracle_*.py):

* bit-exact against the reference itself, compiled from its own sources into
  ``oracle/_ref`` (oracle/Makefile, ``make ref``) -- every function below except
  ``warm_beam_sample`` (ch4 is not built: its Species/Collisions pull in the DSMC
  stack that is out of scope);
* against committed goldens under tests/golden/ produced by that same build;
* against mathematics that a plausible mistake would violate: CIC exactness on affine
  fields (gather), partition of unity / sum(den*V) = sum(mpw) (scatter), the
  uniform-E closed form and |v| invariance under reflection (push), closed-form node
  volumes, a discrete manufactured solution (linear SOR), the QN closed form, exact
  gradients of quadratics (computeEF), the uniform-density Boltzmann solution on a
  pure-Neumann box (nonlinear GS, NR-PCG), the MT19937 10000th-output known answer
  (RNG), and Maxwellian moments (warm_beam_sample).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EPS_0 = 8.85418782e-12
QE = 1.602176565e-19
AMU = 1.660538921e-27
ME = 9.10938215e-31
K = 1.380648e-23

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


class Grid(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("x0", C.c_double * 3), ("xm", C.c_double * 3),
                ("dh", C.c_double * 3), ("xc", C.c_double * 3)]


class MT19937(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


class DsphConfig(C.Structure):
    """orc_dsph_config (field order of the GPU's pic_adjoint_config)."""
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("x0", C.c_double * 3), ("xm", C.c_double * 3),
                ("sphere_c", C.c_double * 3), ("sphere_r", C.c_double),
                ("dt", C.c_double), ("steps", C.c_int), ("checkpoint_every", C.c_int),
                ("mass", C.c_double), ("charge", C.c_double), ("n0", C.c_double),
                ("T_ion", C.c_double), ("n_init", C.c_long), ("n_inject", C.c_long),
                ("soft_width", C.c_double), ("cg_tol", C.c_double), ("nr_tol", C.c_double),
                ("adj_rel_tol", C.c_double), ("seed", C.c_uint64)]


def dsph_config(**kw):
    c = DsphConfig()
    for k, v in kw.items():
        if k in ("x0", "xm", "sphere_c"):
            getattr(c, k)[:] = [float(t) for t in v]
        else:
            setattr(c, k, v)
    return c


_lib = None


def build():
    """Compiles oracle/liboracle.so with gcc (no -march: no FMA contraction)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_grid_init": (None, [C.POINTER(Grid), C.c_int, C.c_int, C.c_int, _dp, _dp]),
            "orc_node_volumes": (None, [C.POINTER(Grid), _dp]),
            "orc_gather3": (None, [C.POINTER(Grid), _dp, _dp, _dp]),
            "orc_gather1": (C.c_double, [C.POINTER(Grid), _dp, _dp]),
            "orc_scatter": (None, [C.POINTER(Grid), _dp, _dp, C.c_double]),
            "orc_push_reflect": (None, [C.POINTER(Grid), C.c_long, _dp, _dp, _dp, C.c_double,
                                        C.c_double, C.c_double]),
            "orc_push_absorb": (C.c_long, [C.POINTER(Grid), C.c_long, _dp, _dp, _dp, _lp, _dp,
                                           C.c_double, C.c_double, C.c_double, _dp, C.c_double]),
            "orc_number_density": (None, [C.POINTER(Grid), C.c_long, _dp, _dp, _dp, _dp]),
            "orc_add_particle": (C.c_int, [C.POINTER(Grid), _dp, C.c_double, C.c_double,
                                           C.c_double, _dp, _dp, C.c_double, _lp, _dp, _dp, _dp]),
            "orc_load_box": (None, [C.POINTER(Grid), C.POINTER(MT19937), _dp, _dp, C.c_double,
                                    C.c_int, _dp, C.c_double, C.c_double, C.c_double, _lp, _dp,
                                    _dp, _dp]),
            "orc_load_box_qs": (None, [C.POINTER(Grid), _dp, _dp, C.c_double, _ip, _dp,
                                       C.c_double, C.c_double, C.c_double, _lp, _dp, _dp, _dp]),
            "orc_cold_beam_sample": (C.c_int, [C.POINTER(Grid), C.POINTER(MT19937), C.c_double,
                                               C.c_double, C.c_double, _dp, C.c_double,
                                               C.c_double, C.c_double, C.c_long, _lp, _dp, _dp,
                                               _dp]),
            "orc_warm_beam_sample": (C.c_int, [C.POINTER(Grid), C.POINTER(MT19937), C.c_double,
                                               C.c_double, C.c_double, C.c_double, _dp,
                                               C.c_double, C.c_double, C.c_double, C.c_long, _lp,
                                               _dp, _dp, _dp]),
            "orc_solve_sor_linear": (C.c_int, [C.POINTER(Grid), _dp, _dp, C.c_int, C.c_double,
                                               _dp, _ip]),
            "orc_solve_gs_boltzmann": (C.c_int, [C.POINTER(Grid), _dp, _dp, _ip, C.c_double,
                                                 C.c_double, C.c_double, C.c_int, C.c_double,
                                                 _dp, _ip]),
            "orc_solve_qn": (None, [C.POINTER(Grid), _dp, _dp, _ip, C.c_double, C.c_double,
                                    C.c_double]),
            "orc_solve_nrpcg": (C.c_int, [C.POINTER(Grid), _dp, _dp, _ip, C.c_double, C.c_double,
                                          C.c_double, C.c_int, C.c_double, _dp]),
            "orc_compute_ef": (None, [C.POINTER(Grid), _dp, _dp]),
            "orc_add_sphere": (None, [C.POINTER(Grid), _ip, _dp, _dp, C.c_double, C.c_double]),
            "orc_add_inlet": (None, [C.POINTER(Grid), _ip, _dp]),
            "orc_get_pe": (C.c_double, [C.POINTER(Grid), _dp, _dp]),
            "orc_species_diag": (None, [C.c_long, _dp, _dp, C.c_double, _dp]),
            "orc_mt_seed": (None, [C.POINTER(MT19937), C.c_uint32]),
            "orc_mt_next": (C.c_uint32, [C.POINTER(MT19937)]),
            "orc_mt_uniform": (C.c_double, [C.POINTER(MT19937)]),
            "orc_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32,
                                         C.POINTER(C.c_uint32)]),
            "orc_solve_nr_sym": (C.c_int, [C.POINTER(Grid), _dp, _dp, _ip, C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_int, _dp]),
            "orc_dsph_run": (C.c_int, [C.POINTER(DsphConfig), _dp, _dp, _dp, _dp, _dp]),
            "orc_update_average": (None, [C.c_long, _dp, _dp, _ip]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _vec(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


class OGrid:
    """World geometry: node counts, extents, spacing (World::setExtents)."""

    def __init__(self, ni, nj, nk, x0, xm):
        self.g = Grid()
        x0 = _vec(x0)
        xm = _vec(xm)
        lib().orc_grid_init(C.byref(self.g), ni, nj, nk, _d(x0), _d(xm))
        self.ni, self.nj, self.nk = ni, nj, nk
        self.nn = ni * nj * nk
        self.x0 = np.array(self.g.x0[:])
        self.xm = np.array(self.g.xm[:])
        self.dh = np.array(self.g.dh[:])
        self.xc = np.array(self.g.xc[:])

    @property
    def ref(self):
        return C.byref(self.g)

    def node_volumes(self):
        nv = np.zeros(self.nn)
        lib().orc_node_volumes(self.ref, _d(nv))
        return nv

    def zeros(self, comps=1):
        return np.zeros(self.nn * comps)


class Rng:
    """std::mt19937 + uniform_real_distribution<double>(0,1) as the reference's Rnd."""

    def __init__(self, seed):
        self.s = MT19937()
        lib().orc_mt_seed(C.byref(self.s), seed)

    def next_u32(self):
        return lib().orc_mt_next(C.byref(self.s))

    def uniform(self):
        return lib().orc_mt_uniform(C.byref(self.s))


class Particles:
    """AoS-in-arrays particle storage (pos, vel: (cap,3); mpw: (cap,))."""

    def __init__(self, cap):
        self.pos = np.zeros((cap, 3))
        self.vel = np.zeros((cap, 3))
        self.mpw = np.zeros(cap)
        self.np = C.c_long(0)
        self.cap = cap

    @property
    def n(self):
        return self.np.value

    def view(self):
        n = self.n
        return self.pos[:n].copy(), self.vel[:n].copy(), self.mpw[:n].copy()

    def set(self, pos, vel, mpw):
        n = len(mpw)
        self.pos[:n] = pos
        self.vel[:n] = vel
        self.mpw[:n] = mpw
        self.np.value = n


def gather3(g: OGrid, ef3, lc):
    out = np.zeros(3)
    lib().orc_gather3(g.ref, _d(ef3), _d(_vec(lc)), _d(out))
    return out


def gather1(g: OGrid, f, lc):
    return lib().orc_gather1(g.ref, _d(f), _d(_vec(lc)))


def scatter(g: OGrid, f, lc, v):
    lib().orc_scatter(g.ref, _d(f), _d(_vec(lc)), v)


def push_reflect(g: OGrid, P: Particles, ef3, dt, charge, mass):
    lib().orc_push_reflect(g.ref, P.n, _d(P.pos), _d(P.vel), _d(ef3), dt, charge, mass)


def push_absorb(g: OGrid, P: Particles, ef3, dt, charge, mass, sphere_c, sphere_r, ids=None):
    c = _vec(sphere_c)
    idp = ids.ctypes.data_as(_lp) if ids is not None else None
    n = lib().orc_push_absorb(g.ref, P.n, _d(P.pos), _d(P.vel), _d(P.mpw), idp, _d(ef3), dt,
                              charge, mass, _d(c), sphere_r * sphere_r)
    P.np.value = n
    return n


def number_density(g: OGrid, pos, mpw, nv):
    den = np.zeros(g.nn)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mpw = np.ascontiguousarray(mpw, dtype=np.float64)
    lib().orc_number_density(g.ref, len(mpw), _d(pos), _d(mpw), _d(nv), _d(den))
    return den


def charge_density(charges, dens):
    """World::computeChargeDensity: rho = sum over charged species of den*q."""
    rho = np.zeros_like(dens[0])
    for q, d in zip(charges, dens):
        if q == 0:
            continue
        rho = rho + d * q
    return rho


def add_particles(g: OGrid, P: Particles, ef3, dt, charge, mass, pos, vel, mpw):
    for p in range(len(mpw)):
        x = _vec(pos[p])
        v = _vec(vel[p])
        lib().orc_add_particle(g.ref, _d(ef3), dt, charge, mass, _d(x), _d(v), float(mpw[p]),
                               C.byref(P.np), _d(P.pos), _d(P.vel), _d(P.mpw))


def load_box(g: OGrid, P: Particles, rng: Rng, x1, x2, num_den, num_mp, ef3, dt, charge, mass):
    lib().orc_load_box(g.ref, C.byref(rng.s), _d(_vec(x1)), _d(_vec(x2)), num_den, num_mp,
                       _d(ef3), dt, charge, mass, C.byref(P.np), _d(P.pos), _d(P.vel), _d(P.mpw))


def load_box_qs(g: OGrid, P: Particles, x1, x2, num_den, num_mp, ef3, dt, charge, mass):
    nmp = np.ascontiguousarray(np.asarray(num_mp, dtype=np.int32))
    lib().orc_load_box_qs(g.ref, _d(_vec(x1)), _d(_vec(x2)), num_den, _i(nmp), _d(ef3), dt,
                          charge, mass, C.byref(P.np), _d(P.pos), _d(P.vel), _d(P.mpw))


def cold_beam_sample(g: OGrid, P: Particles, rng: Rng, v_drift, den, mpw0, ef3, dt, charge,
                     mass):
    r = lib().orc_cold_beam_sample(g.ref, C.byref(rng.s), v_drift, den, mpw0, _d(ef3), dt,
                                   charge, mass, P.cap, C.byref(P.np), _d(P.pos), _d(P.vel),
                                   _d(P.mpw))
    if r < 0:
        raise RuntimeError("particle capacity exceeded")
    return r


def cold_beam_draws(g: OGrid, rng: Rng, v_drift, den, mpw0, dt):
    """The random draws of ColdBeamSource::sample (ch3/ver2/Source.cpp:4-27) without its
    addParticle calls: the count (int)(n v A dt / mpw0 + rnd()), then x and y of each
    particle on the z = x0 face.  Returns the (n, 3) positions; the caller adds them with
    velocity (0, 0, v_drift) and weight mpw0."""
    Lx = g.dh[0] * (g.ni - 1)
    Ly = g.dh[1] * (g.nj - 1)
    num_real = den * v_drift * (Lx * Ly) * dt
    n = int(num_real / mpw0 + rng.uniform())
    pos = np.empty((n, 3))
    for i in range(n):
        pos[i, 0] = g.x0[0] + rng.uniform() * Lx
        pos[i, 1] = g.x0[1] + rng.uniform() * Ly
        pos[i, 2] = g.x0[2]
    return pos


def warm_beam_sample(g: OGrid, P: Particles, rng: Rng, v_drift, den, T, mpw0, ef3, dt, charge,
                     mass):
    r = lib().orc_warm_beam_sample(g.ref, C.byref(rng.s), v_drift, den, T, mpw0, _d(ef3), dt,
                                   charge, mass, P.cap, C.byref(P.np), _d(P.pos), _d(P.vel),
                                   _d(P.mpw))
    if r < 0:
        raise RuntimeError("particle capacity exceeded")
    return r


def solve_sor_linear(g: OGrid, phi, rho, max_it, tol):
    """Returns (converged, l2, iterations); phi updated in place."""
    l2 = C.c_double(0)
    its = C.c_int(0)
    ok = lib().orc_solve_sor_linear(g.ref, _d(phi), _d(rho), max_it, tol, C.byref(l2),
                                    C.byref(its))
    return bool(ok), l2.value, its.value


def solve_gs_boltzmann(g: OGrid, phi, rho, object_id, phi0, Te0, n0, max_it, tol):
    l2 = C.c_double(0)
    its = C.c_int(0)
    oid = np.ascontiguousarray(object_id, dtype=np.int32)
    ok = lib().orc_solve_gs_boltzmann(g.ref, _d(phi), _d(rho), _i(oid), phi0, Te0, n0, max_it,
                                      tol, C.byref(l2), C.byref(its))
    return bool(ok), l2.value, its.value


def solve_qn(g: OGrid, phi, rho, object_id, phi0, Te0, n0):
    oid = np.ascontiguousarray(object_id, dtype=np.int32)
    lib().orc_solve_qn(g.ref, _d(phi), _d(rho), _i(oid), phi0, Te0, n0)


def solve_nrpcg(g: OGrid, phi, rho, object_id, phi0, Te0, n0, max_it, tol):
    """Returns (converged, stats[nr_its, pcg_its_total, gs_fallbacks, norm])."""
    oid = np.ascontiguousarray(object_id, dtype=np.int32)
    stats = np.zeros(4)
    ok = lib().orc_solve_nrpcg(g.ref, _d(phi), _d(rho), _i(oid), phi0, Te0, n0, max_it, tol,
                               _d(stats))
    return bool(ok), stats


def compute_ef(g: OGrid, phi):
    ef = np.zeros(3 * g.nn)
    lib().orc_compute_ef(g.ref, _d(phi), _d(ef))
    return ef


def add_sphere(g: OGrid, object_id, phi, c, r, phi_s):
    lib().orc_add_sphere(g.ref, _i(object_id), _d(phi), _d(_vec(c)), r, phi_s)


def add_inlet(g: OGrid, object_id, phi):
    lib().orc_add_inlet(g.ref, _i(object_id), _d(phi))


def get_pe(g: OGrid, ef3, nv):
    return lib().orc_get_pe(g.ref, _d(ef3), _d(nv))


def species_diag(vel, mpw, mass):
    """Returns [real_count, px, py, pz, KE]."""
    out = np.zeros(5)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    mpw = np.ascontiguousarray(mpw, dtype=np.float64)
    lib().orc_species_diag(len(mpw), _d(vel), _d(mpw), mass, _d(out))
    return out


def update_average(ave, inst, samples):
    """Field::updateAverage (ch3/ver2/Field.h:214-220) in place; returns the new sample count."""
    n = C.c_int(samples)
    lib().orc_update_average(len(ave), _d(ave), _d(np.ascontiguousarray(inst, dtype=np.float64)),
                             C.byref(n))
    return n.value


class SteadyState:
    """World::steadyState (ch3/ver2/World.cpp:129-158).  Each call forms the totals of the
    species' real counts, of the magnitudes of their z momenta and of PE plus their kinetic
    energies, in that order of summation; the state turns steady when all three differ from
    the previous call's by less than 1e-3 relative, and stays steady (the totals are then no
    longer formed).  species_diags holds one [real_count, px, py, pz, KE] row per species."""

    TOL = 1e-3

    def __init__(self):
        self.steady = False
        self.last = (0.0, 0.0, 0.0)

    def __call__(self, species_diags, pe):
        if self.steady:
            return True
        mass, mom, en = 0.0, 0.0, pe
        for d in species_diags:
            mass += d[0]
            mom += abs(d[3])
            en += d[4]
        lm, lp, le = self.last
        if abs((mass - lm) / mass) < self.TOL and abs((mom - lp) / mom) < self.TOL and \
                abs((en - le) / en) < self.TOL:
            self.steady = True
        self.last = (mass, mom, en)
        return self.steady


def philox4x32_10(ctr, key):
    """Philox4x32-10 of a 4-word counter under a 2-word key."""
    c = (C.c_uint32 * 4)(*[int(t) & 0xFFFFFFFF for t in ctr])
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF, out)
    return list(out)


def solve_nr_sym(g: OGrid, phi, rho, object_id, phi0, Te0, n0, cg_tol, nr_tol, max_cg=20000):
    """Newton on the Neumann-reduced system with Jacobi-CG; phi updated in place.
    Returns (converged, [newton its, cg its, 0, last |y|rms])."""
    st = np.zeros(4)
    oid = np.ascontiguousarray(object_id, dtype=np.int32)
    ok = lib().orc_solve_nr_sym(g.ref, _d(phi), _d(np.ascontiguousarray(rho)), _i(oid),
                                phi0, Te0, n0, cg_tol, nr_tol, max_cg, _d(st))
    return bool(ok), st


def dsph_run(cfg: DsphConfig, theta):
    """The differentiable sphere run at theta = (drift, sphere potential, Te).  Returns
    (current, den and phi after the last step, [solves, cg its, live weight, sphere weight])."""
    nn = cfg.ni * cfg.nj * cfg.nk
    den, phi = np.zeros(nn), np.zeros(nn)
    cur = C.c_double(0)
    st = np.zeros(4)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    ok = lib().orc_dsph_run(C.byref(cfg), _d(th), C.byref(cur), _d(den), _d(phi), _d(st))
    if not ok:
        raise RuntimeError("oracle differentiable run: a Newton solve did not converge")
    return cur.value, den, phi, st
