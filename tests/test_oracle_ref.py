"""Live cross-checks of the oracle against the reference built from its sources (CPU).

Skipped where oracle/_ref is not built.  Complements the committed goldens with shapes the
goldens do not cover (anisotropic grids, the threaded ch9/MT build used as a CPU baseline).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R

QE, AMU, ME = O.QE, O.AMU, O.ME
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "plasma-simulations-by-example_amd")

needs_ref = pytest.mark.skipif(not (R.available("ch2") and R.available("ch3") and R.available("mt")),
                               reason="oracle/_ref not built (make -C oracle ref)")


@needs_ref
def test_anisotropic_box_push_deposit_sor_bit_exact():
    n, x0, xm, dt = (17, 13, 23), (-0.05, 0.0, -0.2), (0.11, 0.06, 0.3), 5e-10
    sim = R.RefSim("ch2", *n, x0, xm, dt, 5)
    e = sim.add_species("e-", ME, -QE)
    rng = np.random.default_rng(3)
    g = O.OGrid(*n, x0, xm)
    pos = g.x0 + rng.random((3000, 3)) * (g.xm - g.x0)
    vel = rng.uniform(-3e6, 3e6, (3000, 3))
    mpw = rng.uniform(1, 2, 3000) * 1e3
    sim.set_particles(e, pos, vel, mpw)
    P = O.Particles(3000)
    P.set(pos, vel, mpw)
    rho_seed = rng.normal(0, 1, g.nn) * 1e-9
    sim.set_field(R.RHO, rho_seed)
    sim.make_solver(0, 5000, 1e-3)
    sim.solve()
    sim.compute_ef()
    phi = g.zeros()
    O.solve_sor_linear(g, phi, rho_seed, 5000, 1e-3)
    ef = O.compute_ef(g, phi)
    nv = g.node_volumes()
    for _ in range(5):
        sim.advance(e)
        sim.compute_number_density(e)
        sim.compute_charge_density()
        sim.solve()
        sim.compute_ef()
        O.push_reflect(g, P, ef, dt, -QE, ME)
        den = O.number_density(g, P.pos[:P.n], P.mpw[:P.n], nv)
        rho = O.charge_density([-QE], [den])
        O.solve_sor_linear(g, phi, rho, 5000, 1e-3)
        ef = O.compute_ef(g, phi)
    assert np.array_equal(P.view()[0], sim.particles(e)[0])
    assert np.array_equal(phi, sim.field(R.PHI))
    assert np.array_equal(ef, sim.field(R.EF))


@needs_ref
def test_threaded_reference_matches_serial():
    """ch9/MT (the threaded CPU baseline) pushes like the serial code; its deposit sums
    per-thread buffers, so densities agree to rounding."""
    n, x0, xm, dt = (21, 21, 41), (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4), 1e-7
    rng = np.random.default_rng(4)
    g = O.OGrid(*n, x0, xm)
    npart = 40000
    pos = g.x0 + rng.random((npart, 3)) * (g.xm - g.x0)
    vel = np.zeros((npart, 3))
    vel[:, 2] = 7000.0
    mpw = np.full(npart, 50.0)
    out = {}
    for ch in ("ch3", "mt"):
        sim = R.RefSim(ch, *n, x0, xm, dt, 10)
        sim.add_sphere((0, 0, 0.15), 0.05, -100.0)
        sim.add_inlet()
        sp = sim.add_species("O+", 16 * AMU, QE, 50.0)
        if ch == "mt":
            sim.set_num_threads(4)
        phi, _ = np.zeros(g.nn), None
        sim.set_field(R.EF, np.tile([10.0, -5.0, 20.0], g.nn))
        sim.set_particles(sp, pos, vel, mpw)
        for _ in range(3):
            sim.advance(sp)
            sim.compute_number_density(sp)
        out[ch] = (sim.np_(sp), sim.den(sp))
        sim.close()
    assert out["ch3"][0] == out["mt"][0]
    np.testing.assert_allclose(out["mt"][1], out["ch3"][1], rtol=1e-12, atol=1e-3)


REF_ROOT = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(REF_ROOT), reason="the reference tree is not present")
@pytest.mark.parametrize("main", ["ch2/Main.cpp", "ch3/ver2/Main.cpp"])
def test_unchanged_reference_main_builds_against_dropin(main, tmp_path):
    """The reference's Main.cpp, unchanged, compiles against the drop-in classes and links
    with libpicamd_host (the recipe of oracle/Makefile's `mains`): every header it includes
    by a quoted name resolves to the drop-in's, none to the reference's own beside it."""
    host = os.path.join(PKG, "host")
    if not os.path.exists(os.path.join(PKG, "libpicamd_host.so")):
        pytest.skip("libpicamd_host.so not built")
    src = os.path.join(REF_ROOT, main)
    flags = ["-std=c++17", "-I", os.path.join(ROOT, "include"), "-iquote", host, "-x", "c++"]
    deps = subprocess.run(["g++", *flags, "-M", "-"], stdin=open(src), cwd=tmp_path,
                          capture_output=True, text=True, check=True).stdout
    local = [d for d in deps.replace("\\\n", " ").split()[1:] if not d.startswith("/usr/")]
    names = {os.path.basename(d) for d in local}
    assert {"World.h", "Species.h", "PotentialSolver.h", "Output.h", "Field.h"} <= names
    assert all(d.startswith((host, os.path.join(ROOT, "include"))) for d in local), local
    exe = tmp_path / "main"
    r = subprocess.run(["g++", "-O1", *flags, "-o", str(exe), "-", "-L", PKG, "-lpicamd_host",
                        "-lpicamd", f"-Wl,-rpath,{PKG}"], stdin=open(src), cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert exe.exists()
