"""Batched ensemble of grounded boxes (pic_ensemble_*) against the oracle, box by box.

Each box is a ch2 world (ch2/Main.cpp:14-75): O+ ions over the box and electrons over its
upper octant, specular walls (ch2/Species.cpp:32-36), linear SOR w = 1.4
(ch2/PotentialSolver.cpp:11-67), computeEF (:70-110).  Every GPU step is one replay of the
captured graph; between replays the host reads the particles and fields back (the program's
own eager work), and each replay is checked against the oracle run on the same inputs:
  - particles: FP64 bit-identical to orc_push_reflect with the E the GPU pushed with (by id:
    the counting sort reorders them); mixed within FP32 rounding;
  - density and rho against Species::computeNumberDensity / World::computeChargeDensity of
    the oracle on the GPU's particles;
  - E bit-identical to computeEF of the GPU's phi; phi passes the reference's stopping test.
Box sizes span every SOR path: boxes whose phi fits in LDS (9^3 to 25x19x29) and boxes whose
phi does not (33^3); by default all of them then run in the blocked launches, and with
PIC_ENS_SOR_BLOCKS=0 the first kind is solved in LDS and the second by half-sweep launches.
"""
import numpy as np
import pytest

import pic_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SPECIES = [(16 * O.AMU, O.QE), (O.ME, -O.QE)]
# (nodes, box width): cubes and anisotropic boxes, 9^3 (phi in LDS) to 33^3 (phi in HBM)
BOXES = [((9, 9, 9), 0.05), ((13, 11, 17), 0.1), ((17, 17, 17), 0.2), ((21, 21, 21), 0.2),
         ((33, 33, 33), 0.15), ((25, 19, 29), 0.12)]
DT = 2e-10
TOL = 1e-6


def build(precision, seed=3, ids=True, graph=True):
    rng = np.random.default_rng(seed)
    boxes, counts, loads = [], [], []
    for b, (n, L) in enumerate(BOXES):
        x0, xm = np.full(3, -L / 2), np.full(3, L / 2)
        boxes.append((n, x0, xm))
        dens = 10 ** rng.uniform(10, 12)
        ni, ne = int(rng.integers(3000, 12000)), int(rng.integers(1000, 6000))
        counts.append([ni, ne])
        pi = x0 + rng.random((ni, 3)) * (xm - x0)
        pe = rng.random((ne, 3)) * (L / 2)  # the octant [0, L/2]^3, as ch2/Main.cpp
        vi = rng.normal(0, 2e3, (ni, 3))
        ve = rng.normal(0, 1e6, (ne, 3))
        wi = np.full(ni, dens * L ** 3 / ni)
        we = np.full(ne, dens * (L / 2) ** 3 / ne)
        loads.append([(pi, vi, wi), (pe, ve, we)])
    ens = P.Ensemble(boxes, SPECIES, counts, precision, track_ids=ids)
    for b in range(len(boxes)):
        for s in range(2):
            ens.set_particles(b, s, *loads[b][s])
    # small chunks and frequent sorts: several workgroups per range, the sort on every path
    ens.set_tuning(chunk_particles=1024, window_nodes=4096, sor_threads=1024, sort_every=2,
                   use_graph=graph)
    grids = [O.OGrid(*n, x0, xm) for (n, x0, xm) in boxes]
    return ens, grids, loads


def by_id(parts):
    pos, vel, mpw, ids = parts
    o = np.argsort(ids)
    return pos[o], vel[o], mpw[o]


@pytest.mark.parametrize("sor", ["blocks", "half_sweeps"])
@pytest.mark.parametrize("prec", [P.FP64, P.MIXED])
def test_ensemble_steps_match_oracle_box_by_box(prec, sor, monkeypatch):
    """sor: every box in the blocked SOR launches (the default when some box's phi does not
    fit in LDS), or the boxes that fit solved in LDS by one workgroup each and the others by
    one launch per half-sweep (PIC_ENS_SOR_BLOCKS=0)."""
    if sor == "half_sweeps":
        monkeypatch.setenv("PIC_ENS_SOR_BLOCKS", "0")
    ens, grids, loads = build(prec)
    ens.solve_initial(20000, TOL)
    it, res = ens.solver_status()
    assert (it > 0).all() and (res < TOL).all(), (it, res)
    orc = [[[a.copy() for a in loads[b][s]] for s in range(2)] for b in range(len(grids))]
    for step in range(4):
        efs = [ens.field(b, P.Ensemble.EF) for b in range(len(grids))]
        for b, g in enumerate(grids):
            for s, (m, q) in enumerate(SPECIES):
                Po = O.Particles(len(orc[b][s][2]))
                Po.set(*orc[b][s])
                O.push_reflect(g, Po, efs[b], DT, q, m)
                orc[b][s] = [a.copy() for a in Po.view()]
        ens.step(1, DT, 20000, TOL)
        it, res = ens.solver_status()
        assert (it > 0).all(), it  # every box's SOR converged
        for b, g in enumerate(grids):
            nv = g.node_volumes()
            dens = []
            for s in range(2):
                gp, gv, gw = by_id(ens.particles(b, s, ids=True))
                po, vo, wo = orc[b][s]
                if prec == P.FP64:
                    assert np.array_equal(gp, po) and np.array_equal(gv, vo), (step, b, s)
                else:
                    # cell + FP32 offset: a few FP32 ulps of a cell per step
                    np.testing.assert_allclose(gp, po, rtol=0, atol=2e-6 * g.dh.max())
                    np.testing.assert_allclose(gv, vo, rtol=0, atol=2e-6 * np.abs(vo).max())
                ref = O.number_density(g, gp, gw, nv)
                tol = 1e-12 if prec == P.FP64 else 1e-5
                den = ens.den(b, s)
                assert np.linalg.norm(den - ref) <= tol * np.linalg.norm(ref), (step, b, s)
                dens.append(den)
            rho = ens.field(b, P.Ensemble.RHO)
            rho_ref = O.charge_density([q for _, q in SPECIES], dens)
            assert np.linalg.norm(rho - rho_ref) <= 1e-14 * np.linalg.norm(rho_ref)
            phi = ens.field(b, P.Ensemble.PHI)
            # the reference's stopping test holds at the GPU's phi (ch2/PotentialSolver.cpp:48-63)
            assert O.solve_sor_linear(g, phi.copy(), rho, 1, TOL)[0], (step, b)
            ef = ens.field(b, P.Ensemble.EF)
            ef_ref = O.compute_ef(g, phi)
            if prec == P.FP64:
                assert np.array_equal(ef, ef_ref), (step, b)
            else:
                np.testing.assert_allclose(ef, ef_ref, rtol=1e-6, atol=1e-6 * np.abs(ef_ref).max())


def test_graph_replays_match_eager_steps():
    """The captured step replayed, with host reads in between, against the same launches
    issued eagerly, on identical inputs (FP64: the two differ only by the order of the
    deposit's atomic additions)."""
    a, grids, _ = build(P.FP64, seed=5, graph=True)
    b, _, _ = build(P.FP64, seed=5, graph=False)
    for e in (a, b):
        e.solve_initial(20000, TOL)
    for step in range(3):
        a.step(1, DT, 20000, TOL)
        b.step(1, DT, 20000, TOL)
        for k in range(len(grids)):
            pa = a.field(k, P.Ensemble.PHI)
            pb = b.field(k, P.Ensemble.PHI)
            assert np.linalg.norm(pa - pb) <= 1e-9 * np.linalg.norm(pb), (step, k)
            for s in range(2):
                xa, va, _ = by_id(a.particles(k, s, ids=True))
                xb, vb, _ = by_id(b.particles(k, s, ids=True))
                np.testing.assert_allclose(xa, xb, rtol=0, atol=1e-12 * grids[k].dh.max())
                np.testing.assert_allclose(va, vb, rtol=1e-9, atol=1e-9 * np.abs(vb).max())
    st = a.stats()
    assert st["graph_kernels"] >= 4 and st["particles"] == sum(int(c) for c in a.counts.ravel())


def test_blocked_sor_matches_half_sweep_launches(monkeypatch):
    """The blocked SOR (the default: every box in blocks of 8^3 nodes, one launch per two
    iterations from an LDS window into phi's second buffer) against PIC_ENS_SOR_BLOCKS=0 (the
    boxes whose phi fits in LDS solved there by one workgroup each, the 33^3 box by one launch
    per half-sweep), on identical inputs: the same iteration count for every box at every
    solve, and phi alike to the deposit's atomic-order differences.  Most boxes end in partial
    blocks (interiors of 7 to 31 nodes)."""
    runs = []
    for mode in ("1", "0"):
        monkeypatch.setenv("PIC_ENS_SOR_BLOCKS", mode)
        e, grids, _ = build(P.FP64, seed=7)
        e.solve_initial(20000, TOL)
        runs.append(e)
    a, b = runs
    assert np.array_equal(a.solver_status()[0], b.solver_status()[0])
    for step in range(3):
        a.step(1, DT, 20000, TOL)
        b.step(1, DT, 20000, TOL)
        ia, ra = a.solver_status()
        ib, rb = b.solver_status()
        assert np.array_equal(ia, ib), (step, ia, ib)
        assert (ia > 0).all(), ia  # every box converged
        for k in range(len(grids)):
            pa = a.field(k, P.Ensemble.PHI)
            pb = b.field(k, P.Ensemble.PHI)
            assert np.linalg.norm(pa - pb) <= 1e-9 * np.linalg.norm(pb), (step, k)


def test_device_load_and_diagnostics():
    """loadParticlesBox on the device (ch2/Species.cpp:74-97): every particle inside its
    range's box, weights num_den V / n, zero velocity; the per-range diagnostics sums match
    the downloaded particles."""
    boxes = [((17, 17, 17), (-0.1,) * 3, (0.1,) * 3), ((9, 13, 11), (0, 0, 0), (0.05, 0.08, 0.06))]
    ens = P.Ensemble(boxes, SPECIES, [[50000, 20000], [7000, 3000]], P.MIXED)
    for b, (_, x0, xm) in enumerate(boxes):
        ens.load_box(b, 0, x0, xm, 1e11, seed=9)
        mid = tuple(0.5 * (np.array(x0) + np.array(xm)))
        ens.load_box(b, 1, mid, xm, 1e11, seed=9)
    d = ens.diag()
    for b, (_, x0, xm) in enumerate(boxes):
        for s in range(2):
            pos, vel, w = ens.particles(b, s)
            lo = np.array(x0) if s == 0 else 0.5 * (np.array(x0) + np.array(xm))
            assert (pos >= lo - 1e-9).all() and (pos < np.array(xm)).all()
            V = np.prod(np.array(xm) - lo)
            np.testing.assert_allclose(w, 1e11 * V / len(w), rtol=1e-6)
            assert np.all(vel == 0)
            np.testing.assert_allclose(d[b, s, 0], w.sum(), rtol=1e-12)
            # uniform: the mean position sits at the box centre (5 sigma)
            sig = (np.array(xm) - lo) / np.sqrt(12 * len(w))
            assert np.all(np.abs(pos.mean(0) - 0.5 * (lo + np.array(xm))) < 5 * sig)
