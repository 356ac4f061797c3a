"""The FP64 oracle reproduces the reference's own outputs bit for bit (CPU).

The vectors in tests/golden/ were written by tests/golden/make_golden.py from the
reference's C++ compiled from its sources (oracle/_ref).  Here the oracle's restatement
replays the same inputs and must match exactly: same loop order, same operation order.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
QE, AMU, ME = O.QE, O.AMU, O.ME


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_rng_stream_matches_reference():
    G = load("rng_mt19937.npz")
    for seed in (1, 5489):
        r = O.Rng(seed)
        u = np.array([r.uniform() for _ in range(64)])
        assert np.array_equal(u, G[f"u_seed{seed}"])


def test_load_box_matches_reference():
    G = load("rng_mt19937.npz")
    g = O.OGrid(21, 21, 21, (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1))
    P = O.Particles(200)
    O.load_box(g, P, O.Rng(1), g.x0, g.xm, 1e11, 200, g.zeros(3), 2e-10, QE, 16 * AMU)
    pos, vel, mpw = P.view()
    assert np.array_equal(pos, G["load_pos"])
    assert np.array_equal(mpw, G["load_mpw"])


def test_load_box_qs_matches_reference():
    G = load("qs_ch2.npz")
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.2)
    g = O.OGrid(21, 21, 21, x0, xm)
    P = O.Particles(11 ** 3)
    O.load_box_qs(g, P, x0, xm, 1e11, (11, 11, 11), g.zeros(3), 2e-10, QE, 16 * AMU)
    pos, vel, mpw = P.view()
    assert np.array_equal(pos, G["pos"])
    assert np.array_equal(mpw, G["mpw"])


def test_grounded_box_ten_steps_bit_exact():
    """ch2/Main.cpp loop on the grounded box (21^3, 80k O+ / 10k e-): load, solve, then advance/deposit/rho/SOR/EF."""
    G = load("box_ch2.npz")
    x0, xm, dt = (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), 2e-10
    g = O.OGrid(21, 21, 21, x0, xm)
    nv = g.node_volumes()
    rng = O.Rng(1)
    ef = g.zeros(3)
    Pi, Pe = O.Particles(80000), O.Particles(10000)
    O.load_box(g, Pi, rng, x0, xm, 1e11, 80000, ef, dt, QE, 16 * AMU)
    O.load_box(g, Pe, rng, (0, 0, 0), xm, 1e11, 10000, ef, dt, -QE, ME)
    phi = g.zeros()
    O.solve_sor_linear(g, phi, g.zeros(), 10000, 1e-2)
    ef = O.compute_ef(g, phi)
    diag, pe = [], []
    for _ in range(int(G["steps"])):
        O.push_reflect(g, Pi, ef, dt, QE, 16 * AMU)
        di = O.number_density(g, Pi.pos[:Pi.n], Pi.mpw[:Pi.n], nv)
        O.push_reflect(g, Pe, ef, dt, -QE, ME)
        de = O.number_density(g, Pe.pos[:Pe.n], Pe.mpw[:Pe.n], nv)
        rho = O.charge_density([QE, -QE], [di, de])
        O.solve_sor_linear(g, phi, rho, 10000, 1e-2)
        ef = O.compute_ef(g, phi)
        d = []
        for P, m in ((Pi, 16 * AMU), (Pe, ME)):
            s = O.species_diag(P.vel[:P.n], P.mpw[:P.n], m)
            d.append(np.concatenate([[P.n], s]))
        diag.append(np.stack(d))
        pe.append(O.get_pe(g, ef, nv))
    assert np.array_equal(phi, G["phi"])
    assert np.array_equal(rho, G["rho"])
    assert np.array_equal(ef, G["ef"])
    assert np.array_equal(di, G["den_i"])
    assert np.array_equal(de, G["den_e"])
    assert np.array_equal(Pi.pos[G["ion_idx"]], G["ion_pos"])
    assert np.array_equal(Pi.vel[G["ion_idx"]], G["ion_vel"])
    assert np.array_equal(Pe.pos[G["ele_idx"]], G["ele_pos"])
    assert np.array_equal(Pe.vel[G["ele_idx"]], G["ele_vel"])
    assert np.array_equal(np.array(diag), G["diag"])
    assert np.array_equal(np.array(pe), G["pe"])


@pytest.mark.parametrize("name", ["gs", "pcg", "qn"])
def test_sphere_twelve_steps_bit_exact(name):
    """ch3/ver2/Main.cpp loop on a 13x13x25 sphere world with cold-beam injection."""
    G = load("sphere_ch3.npz")
    x0, xm, dt = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4), 1e-7
    g = O.OGrid(13, 13, 25, x0, xm)
    nv = g.node_volumes()
    oid = np.zeros(g.nn, dtype=np.int32)
    phi = g.zeros()
    O.add_sphere(g, oid, phi, (0, 0, 0.15), 0.05, -100.0)
    O.add_inlet(g, oid, phi)
    assert np.array_equal(oid.astype(float), G["object_id"])
    rho = g.zeros()
    O.solve_qn(g, phi, rho, oid, 0, 1.5, 1e12)  # PotentialSolver ctor: buildMatrix -> solveQN
    assert np.array_equal(phi, G["phi_ctor"])

    def solve():
        if name == "gs":
            O.solve_gs_boltzmann(g, phi, rho, oid, 0, 1.5, 1e10, 20000, 1e-4)
        elif name == "pcg":
            O.solve_nrpcg(g, phi, rho, oid, 0, 1.5, 1e10, 10000, 1e-4)
        else:
            O.solve_qn(g, phi, rho, oid, 0, 1.5, 1e10)

    solve()
    assert np.array_equal(phi, G[f"{name}_phi0"])
    ef = O.compute_ef(g, phi)
    rng = O.Rng(7)
    P = O.Particles(200000)
    nps, diags = [], []
    for _ in range(12):
        O.cold_beam_sample(g, P, rng, 7000.0, 1e10, 50.0, ef, dt, QE, 16 * AMU)
        O.push_absorb(g, P, ef, dt, QE, 16 * AMU, (0, 0, 0.15), 0.05)
        den = O.number_density(g, P.pos[:P.n], P.mpw[:P.n], nv)
        rho = O.charge_density([QE], [den])
        solve()
        ef = O.compute_ef(g, phi)
        nps.append(P.n)
        diags.append(np.concatenate([[P.n], O.species_diag(P.vel[:P.n], P.mpw[:P.n], 16 * AMU)]))
    assert np.array_equal(np.array(nps), G[f"{name}_np"])
    assert np.array_equal(np.array(diags), G[f"{name}_diag"])
    assert np.array_equal(phi, G[f"{name}_phi"])
    assert np.array_equal(ef, G[f"{name}_ef"])
    assert np.array_equal(den, G[f"{name}_den"])
    assert np.array_equal(rho, G[f"{name}_rho"])
    assert np.array_equal(P.pos[:P.n][::41], G[f"{name}_pos"])
    assert np.array_equal(P.vel[:P.n][::41], G[f"{name}_vel"])


def sphere_history_world():
    """ch3/ver2/Main.cpp:18-45: the 21x21x41 sphere world, the PotentialSolver constructor's
    QN solve with its default n0, then setReferenceValues(0, 1.5, 1e10) and solve()."""
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    g = O.OGrid(21, 21, 41, x0, xm)
    oid = np.zeros(g.nn, dtype=np.int32)
    phi = g.zeros()
    O.add_sphere(g, oid, phi, (0, 0, 0.15), 0.05, -100.0)
    O.add_inlet(g, oid, phi)
    O.solve_qn(g, phi, g.zeros(), oid, 0, 1.5, 1e12)
    O.solve_qn(g, phi, g.zeros(), oid, 0, 1.5, 1e10)
    return g, oid, phi


def test_cold_beam_draws_match_sample():
    """The host-side draws of ColdBeamSource::sample (ch3/ver2/Source.cpp:4-27) followed by
    addParticle give the oracle's cold_beam_sample, draw for draw, over several calls."""
    g, oid, phi = sphere_history_world()
    ef = O.compute_ef(g, phi)
    dt, mpw0, m = 2e-7, 200.0, 16 * AMU
    ra, rb = O.Rng(7), O.Rng(7)
    Pa, Pb = O.Particles(20000), O.Particles(20000)
    for _ in range(5):
        O.cold_beam_sample(g, Pa, ra, 7000.0, 1e10, mpw0, ef, dt, QE, m)
        pos = O.cold_beam_draws(g, rb, 7000.0, 1e10, mpw0, dt)
        n = len(pos)
        O.add_particles(g, Pb, ef, dt, QE, m, pos, np.tile([0.0, 0.0, 7000.0], (n, 1)),
                        np.full(n, mpw0))
    assert Pa.n == Pb.n > 10000
    for a, b in zip(Pa.view(), Pb.view()):
        assert np.array_equal(a, b)
    assert ra.uniform() == rb.uniform()  # both streams consumed the same number of draws


def test_sphere_history_400_steps_bit_exact():
    """ch3/ver2/Main.cpp's main loop (:48-80) for 400 steps at dt 2e-7 and mpw0 200 with the
    QN solver: per-step particle count, getRealCount / getMomentum / getKE, getPE,
    World::steadyState and, from the step it holds, Species::updateAverages; den_ave and phi
    at the end.  Every value bit-identical to the compiled reference's."""
    G = load("hist_ch3.npz")
    steps, dt, mpw0 = int(G["steps"]), float(G["dt"]), float(G["mpw0"])
    g, oid, phi = sphere_history_world()
    nv = g.node_volumes()
    ef = O.compute_ef(g, phi)
    rng = O.Rng(7)
    P = O.Particles(600000)
    steady = O.SteadyState()
    den_ave, samples = g.zeros(), 0
    nps, diags, pes, flags = [], [], [], []
    for _ in range(len(G["np"])):
        O.cold_beam_sample(g, P, rng, 7000.0, 1e10, mpw0, ef, dt, QE, 16 * AMU)
        O.push_absorb(g, P, ef, dt, QE, 16 * AMU, (0, 0, 0.15), 0.05)
        den = O.number_density(g, P.pos[:P.n], P.mpw[:P.n], nv)
        rho = O.charge_density([QE], [den])
        O.solve_qn(g, phi, rho, oid, 0, 1.5, 1e10)
        ef = O.compute_ef(g, phi)
        d = O.species_diag(P.vel[:P.n], P.mpw[:P.n], 16 * AMU)
        pe = O.get_pe(g, ef, nv)
        st = steady([d], pe)
        if st:
            samples = O.update_average(den_ave, den, samples)
        nps.append(P.n)
        diags.append(np.concatenate([[P.n], d]))
        pes.append(pe)
        flags.append(st)
    assert len(nps) == steps + 1
    assert np.array_equal(np.array(nps), G["np"])
    assert np.array_equal(np.array(diags), G["diag"])
    assert np.array_equal(np.array(pes), G["pe"])
    assert np.array_equal(np.array(flags, dtype=np.int8), G["steady"])
    assert 0 < samples < steps  # the averages started inside the run
    assert np.array_equal(den_ave, G["den_ave"])
    assert np.array_equal(phi, G["phi"])


def test_grounded_box_history_2000_steps_bit_exact():
    """ch2/Main.cpp's loop on the grounded box with the SOR tolerance at 1e-8, 2000 steps: the
    per-step diagnostics of both species and getPE bit-identical to the compiled
    reference's."""
    G = load("hist_box.npz")
    tol = float(G["tol"])
    x0, xm, dt = (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), 2e-10
    g = O.OGrid(21, 21, 21, x0, xm)
    nv = g.node_volumes()
    rng = O.Rng(1)
    ef = g.zeros(3)
    Pi, Pe = O.Particles(80000), O.Particles(10000)
    O.load_box(g, Pi, rng, x0, xm, 1e11, 80000, ef, dt, QE, 16 * AMU)
    O.load_box(g, Pe, rng, (0, 0, 0), xm, 1e11, 10000, ef, dt, -QE, ME)
    phi = g.zeros()
    O.solve_sor_linear(g, phi, g.zeros(), 10000, tol)
    ef = O.compute_ef(g, phi)
    diag, pe = [], []
    for _ in range(len(G["pe"])):
        dens = []
        for P, q, m in ((Pi, QE, 16 * AMU), (Pe, -QE, ME)):
            O.push_reflect(g, P, ef, dt, q, m)
            dens.append(O.number_density(g, P.pos[:P.n], P.mpw[:P.n], nv))
        rho = O.charge_density([QE, -QE], dens)
        O.solve_sor_linear(g, phi, rho, 10000, tol)
        ef = O.compute_ef(g, phi)
        diag.append(np.stack([np.concatenate([[P.n], O.species_diag(P.vel[:P.n], P.mpw[:P.n], m)])
                              for P, m in ((Pi, 16 * AMU), (Pe, ME))]))
        pe.append(O.get_pe(g, ef, nv))
    assert np.array_equal(np.array(diag), G["diag"])
    assert np.array_equal(np.array(pe), G["pe"])
