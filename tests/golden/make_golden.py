"""Writes the golden vectors under tests/golden/ by running the reference itself.

The reference's C++ (World / Species / PotentialSolver / Source of ch2 and ch3/ver2) is
compiled from its own sources into oracle/_ref by oracle/Makefile (``make -C oracle ref``);
this script drives those libraries through oracle/ref.py and stores inputs and outputs
only.  Rerun with:  python tests/golden/make_golden.py

Files
  rng_mt19937.npz      reference Rnd stream after reseeding (seeds 1 and 5489); first
                       particles of Species::loadParticlesBox (ch2)
  box_ch2.npz          the ch2 grounded box (21^3, 80k O+ / 10k e-, seed 1,
                       dt 2e-10, SOR w=1.4 tol 1e-2): initial solve + 10 steps; fields, a
                       particle subset, per-step diagnostics
  qs_ch2.npz           Species::loadParticlesBoxQS on an 11^3 particle lattice
  sphere_ch3.npz       ch3/ver2 sphere (13x13x25), cold beam (seed 7), 12 steps for each
                       solver type GS / PCG / QN
  hist_box.npz         the grounded box as box_ch2 but SOR tol 1e-8, 2000 steps: per-step diagnostics
                       (getRealCount, getMomentum, getKE of both species) and getPE
  hist_ch3.npz         ch3/ver2/Main.cpp's sphere run (21x21x41, -100 V, cold beam 7 km/s,
                       seed 7) with the QN solver, dt 2e-7, mpw0 200, 400 steps: per-step
                       particle count, diagnostics, getPE and World::steadyState (averages
                       from the step it holds); den_ave and phi at the end
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref as R  # noqa: E402

QE = 1.602176565e-19
AMU = 1.660538921e-27
ME = 9.10938215e-31

BOX = dict(n=(21, 21, 21), x0=(-0.1, -0.1, -0.1), xm=(0.1, 0.1, 0.1), dt=2e-10)
SPH = dict(n=(13, 13, 25), x0=(-0.1, -0.1, 0.0), xm=(0.1, 0.1, 0.4), dt=1e-7)


def gen_rng():
    sim = R.RefSim("ch2", 3, 3, 3, (0, 0, 0), (1, 1, 1))
    out = {}
    for seed in (1, 5489):
        sim.reseed(seed)
        out[f"u_seed{seed}"] = np.array([sim.L.ref_rnd() for _ in range(64)])
    sim.close()
    sim = R.RefSim("ch2", *BOX["n"], BOX["x0"], BOX["xm"], BOX["dt"], 10)
    sp = sim.add_species("O+", 16 * AMU, QE)
    sim.reseed(1)
    sim.load_box(sp, BOX["x0"], BOX["xm"], 1e11, 200)
    pos, vel, mpw = sim.particles(sp)
    out.update(load_pos=pos, load_mpw=mpw)
    sim.close()
    np.savez_compressed(os.path.join(HERE, "rng_mt19937.npz"), **out)


def gen_box():
    sim = R.RefSim("ch2", *BOX["n"], BOX["x0"], BOX["xm"], BOX["dt"], 10)
    i = sim.add_species("O+", 16 * AMU, QE)
    e = sim.add_species("e-", ME, -QE)
    sim.reseed(1)
    sim.load_box(i, BOX["x0"], BOX["xm"], 1e11, 80000)
    sim.load_box(e, (0.0, 0.0, 0.0), BOX["xm"], 1e11, 10000)
    sim.make_solver(0, 10000, 1e-2)
    sim.solve()
    sim.compute_ef()
    diag, pe = [], []
    while sim.advance_time():
        for s in (i, e):
            sim.advance(s)
            sim.compute_number_density(s)
        sim.compute_charge_density()
        sim.solve()
        sim.compute_ef()
        diag.append(np.stack([sim.species_diag(i), sim.species_diag(e)]))
        pe.append(sim.pe())
    pi, vi, _ = sim.particles(i)
    pe_, ve, _ = sim.particles(e)
    np.savez_compressed(
        os.path.join(HERE, "box_ch2.npz"),
        steps=len(pe), phi=sim.field(R.PHI), rho=sim.field(R.RHO), ef=sim.field(R.EF),
        den_i=sim.den(i), den_e=sim.den(e), ion_idx=np.arange(0, 80000, 97),
        ion_pos=pi[::97], ion_vel=vi[::97], ele_idx=np.arange(0, 10000, 13), ele_pos=pe_[::13],
        ele_vel=ve[::13], diag=np.array(diag), pe=np.array(pe))
    sim.close()


def gen_qs():
    sim = R.RefSim("ch2", 21, 21, 21, (-0.1, -0.1, 0.0), (0.1, 0.1, 0.2), 2e-10, 0)
    sp = sim.add_species("O+", 16 * AMU, QE)
    sim.load_box_qs(sp, (-0.1, -0.1, 0.0), (0.1, 0.1, 0.2), 1e11, (11, 11, 11))
    pos, vel, mpw = sim.particles(sp)
    np.savez_compressed(os.path.join(HERE, "qs_ch2.npz"), pos=pos, mpw=mpw)
    sim.close()


def gen_sphere():
    out = {}
    for stype, name in ((R.GS, "gs"), (R.PCG, "pcg"), (R.QN, "qn")):
        sim = R.RefSim("ch3", *SPH["n"], SPH["x0"], SPH["xm"], SPH["dt"], 400)
        sim.add_sphere((0, 0, 0.15), 0.05, -100.0)
        sim.add_inlet()
        ion = sim.add_species("O+", 16 * AMU, QE, 50.0)
        sim.add_cold_source(ion, 7000.0, 1e10)
        sim.make_solver(stype, 20000 if stype == R.GS else 10000, 1e-4)
        if name == "gs":
            out["phi_ctor"] = sim.field(R.PHI)  # buildMatrix -> solveQN with default refs
            out["object_id"] = sim.field(R.OBJECT_ID)
        sim.set_reference_values(0.0, 1.5, 1e10)
        sim.reseed(7)
        sim.solve()
        sim.compute_ef()
        out[f"{name}_phi0"] = sim.field(R.PHI)
        nps, diags = [], []
        for _ in range(12):
            sim.advance_time()
            sim.sample_sources()
            sim.advance(ion)
            sim.compute_number_density(ion)
            sim.compute_charge_density()
            sim.solve()
            sim.compute_ef()
            nps.append(sim.np_(ion))
            diags.append(sim.species_diag(ion))
        pos, vel, mpw = sim.particles(ion)
        out[f"{name}_np"] = np.array(nps)
        out[f"{name}_diag"] = np.array(diags)
        out[f"{name}_phi"] = sim.field(R.PHI)
        out[f"{name}_ef"] = sim.field(R.EF)
        out[f"{name}_den"] = sim.den(ion)
        out[f"{name}_rho"] = sim.field(R.RHO)
        out[f"{name}_pos"] = pos[::41]
        out[f"{name}_vel"] = vel[::41]
        sim.close()
    np.savez_compressed(os.path.join(HERE, "sphere_ch3.npz"), **out)


def gen_hist_box(steps=2000, tol=1e-8):
    sim = R.RefSim("ch2", *BOX["n"], BOX["x0"], BOX["xm"], BOX["dt"], steps)
    i = sim.add_species("O+", 16 * AMU, QE)
    e = sim.add_species("e-", ME, -QE)
    sim.reseed(1)
    sim.load_box(i, BOX["x0"], BOX["xm"], 1e11, 80000)
    sim.load_box(e, (0.0, 0.0, 0.0), BOX["xm"], 1e11, 10000)
    sim.make_solver(0, 10000, tol)
    sim.solve()
    sim.compute_ef()
    diag, pe = [], []
    while sim.advance_time():
        for s in (i, e):
            sim.advance(s)
            sim.compute_number_density(s)
        sim.compute_charge_density()
        sim.solve()
        sim.compute_ef()
        diag.append(np.stack([sim.species_diag(i), sim.species_diag(e)]))
        pe.append(sim.pe())
    np.savez_compressed(os.path.join(HERE, "hist_box.npz"), diag=np.array(diag), pe=np.array(pe),
                        tol=tol, steps=steps)
    sim.close()


def gen_hist_ch3(steps=400, dt=2e-7, mpw0=200.0):
    # ch3/ver2/Main.cpp:18-41 (grid, extents, sphere, inlet, beam, QN) at twice its dt and four
    # times its macro-particle weight, so that the flow crosses the domain and reaches
    # World::steadyState within the 400 steps
    n, x0, xm = (21, 21, 41), (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    sim = R.RefSim("ch3", *n, x0, xm, dt, steps)
    sim.add_sphere((0, 0, 0.15), 0.05, -100.0)
    sim.add_inlet()
    ion = sim.add_species("O+", 16 * AMU, QE, mpw0)
    sim.add_cold_source(ion, 7000.0, 1e10)
    sim.make_solver(R.QN, 20000, 1e-4)
    sim.set_reference_values(0.0, 1.5, 1e10)
    sim.reseed(7)
    sim.solve()
    sim.compute_ef()
    nps, diags, pes, steady = [], [], [], []
    while sim.advance_time():
        sim.sample_sources()
        sim.advance(ion)
        sim.compute_number_density(ion)
        sim.compute_charge_density()
        sim.solve()
        sim.compute_ef()
        st = sim.steady_state()
        if st:
            sim.update_averages(ion)
        nps.append(sim.np_(ion))
        diags.append(sim.species_diag(ion))
        pes.append(sim.pe())
        steady.append(st)
    np.savez_compressed(os.path.join(HERE, "hist_ch3.npz"), np=np.array(nps), diag=np.array(diags),
                        pe=np.array(pes), steady=np.array(steady, dtype=np.int8),
                        den_ave=sim.den_ave(ion), phi=sim.field(R.PHI), steps=steps, dt=dt,
                        mpw0=mpw0)
    sim.close()


if __name__ == "__main__":
    if not (R.available("ch2") and R.available("ch3")):
        R.build()
    gen_rng()
    gen_qs()
    gen_box()
    gen_sphere()
    gen_hist_box()
    gen_hist_ch3()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
