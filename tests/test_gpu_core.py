"""GPU parity of the hot-path kernels against the FP64 oracle (through the C ABI).

Tolerances: FP64 species push with the oracle's arithmetic (FMA contraction off), so their
positions and velocities are bit-identical to it for a given E; FP64 node kernels agree to
rounding (computeEF and node volumes bit-exact; deposits to 1e-13 relative, since atomic
summation order differs).  PIC_MIXED species store the position as cell + FP32 offset and push in
FP32: positions within 2e-4 cells (RMS 1e-5 cells), velocities within 2e-5 of the velocity
scale, densities within 1e-5 relative L2 over a 10-step horizon -- far inside the 1e-4
relative field-error budget.  Where the RMS bound comes from: an FP32 velocity carries a
relative rounding of about 6e-8 per step, and at up to 3 cells per step (3e7 m/s, 1e-9 s,
1 cm cells) that moves the position by about 2e-7 cells per step; over 10 steps the errors
add to an RMS of about 2e-6 cells, which is what the runs measure (2.1e-6 at most), and the
bound leaves a factor of 5.
"""
import numpy as np
import pytest

import pic_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C1_X0 = (-0.1, -0.1, -0.1)
C1_XM = (0.1, 0.1, 0.1)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_world(n=(21, 21, 21), x0=C1_X0, xm=C1_XM):
    w = P.World(*n)
    w.set_extents(x0, xm)
    g = O.OGrid(*n, x0, xm)
    return w, g


def smooth_field(g, seed, amp):
    """E from a smooth random potential (so it is a plausible field), FP64 interleaved."""
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(np.arange(g.ni), np.arange(g.nj), np.arange(g.nk), indexing="ij")
    phi = np.zeros((g.ni, g.nj, g.nk))
    for _ in range(4):
        a, b, c = rng.uniform(0.5, 3, 3)
        ph = rng.uniform(0, 6.28, 3)
        phi += np.sin(a * i / g.ni * 6.28 + ph[0]) * np.cos(b * j / g.nj * 6.28 + ph[1]) * \
            np.sin(c * k / g.nk * 6.28 + ph[2])
    phi = np.ascontiguousarray(phi.transpose(2, 1, 0)).reshape(-1) * amp
    return phi, O.compute_ef(g, phi)


def random_particles(g, n, seed, vmax, box=None):
    rng = np.random.default_rng(seed)
    lo = g.x0 if box is None else np.asarray(box[0])
    hi = g.xm if box is None else np.asarray(box[1])
    pos = lo + rng.random((n, 3)) * (hi - lo)
    vel = rng.uniform(-vmax, vmax, (n, 3))
    mpw = rng.uniform(0.5, 1.5, n) * 1e4
    return pos, vel, mpw


def test_node_volumes_sphere_inlet_exact():
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    w, g = make_world((21, 21, 41), x0, xm)
    w.add_sphere((0, 0, 0.15), 0.05, -100.0)
    w.add_inlet()
    oid = np.zeros(g.nn, dtype=np.int32)
    phi = g.zeros()
    O.add_sphere(g, oid, phi, (0, 0, 0.15), 0.05, -100.0)
    O.add_inlet(g, oid, phi)
    assert np.array_equal(w.field(P.NODE_VOL), g.node_volumes())
    assert np.array_equal(w.field(P.OBJECT_ID), oid.astype(np.float64))
    assert np.array_equal(w.field(P.PHI), phi)


def test_compute_ef_bit_exact_and_pe():
    w, g = make_world((23, 19, 17))
    phi, _ = smooth_field(g, 3, 50.0)
    w.set_field(P.PHI, phi)
    s = P.PotentialSolver(w, P.SOR_LINEAR, 10, 1e-2)
    s.compute_ef()
    ef = w.field(P.EF)
    ref = O.compute_ef(g, phi)
    assert np.array_equal(ef, ref)
    pe = w.get_pe()
    assert abs(pe - O.get_pe(g, ref, g.node_volumes())) <= 1e-13 * abs(pe)


def test_solve_qn_matches_closed_form():
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    w, g = make_world((13, 13, 25), x0, xm)
    w.add_sphere((0, 0, 0.15), 0.05, -100.0)
    w.add_inlet()
    rng = np.random.default_rng(5)
    rho = rng.uniform(0, 3e10, g.nn) * O.QE
    rho[::7] = 0  # exercises the 1e-6 floor
    w.set_field(P.RHO, rho)
    s = P.PotentialSolver(w, P.QN, 10, 1e-4)
    s.set_reference_values(0.0, 1.5, 1e10)
    assert s.solve()[0]
    oid = np.zeros(g.nn, dtype=np.int32)
    phi = g.zeros()
    O.add_sphere(g, oid, phi, (0, 0, 0.15), 0.05, -100.0)
    O.add_inlet(g, oid, phi)
    O.solve_qn(g, phi, rho, oid, 0.0, 1.5, 1e10)
    np.testing.assert_allclose(w.field(P.PHI), phi, rtol=1e-14, atol=1e-13)


def _push_case(prec, bc, steps, dt, vmax, sphere=None, n=20000, grid=(21, 21, 21), x0=C1_X0,
               xm=C1_XM, seed=11, amp=200.0, charge=-O.QE, mass=O.ME, box=None, drift=None,
               frame=None, frame_after=False):
    w, g = make_world(grid, x0, xm)
    phi, ef = smooth_field(g, seed, amp)
    if sphere is not None:
        w.add_sphere(*sphere)
    w.set_field(P.EF, ef)
    sp = P.Species(w, mass, charge, 1e4, prec, capacity=n, track_ids=True)
    pos, vel, mpw = random_particles(g, n, seed + 1, vmax, box)
    if drift is not None:
        vel = vel + np.asarray(drift)
    if sphere is not None:  # start outside the sphere
        keep = np.sum((pos - np.asarray(sphere[0])) ** 2, axis=1) > sphere[1] ** 2 * 1.02
        pos, vel, mpw = pos[keep], vel[keep], mpw[keep]
    ids = np.arange(len(mpw), dtype=np.uint32)
    if frame is not None and not frame_after:
        sp.set_frame_velocity(frame)
    sp.set_particles(pos, vel, mpw, ids)
    if frame is not None and frame_after:
        sp.set_frame_velocity(frame)
    # the device stores what it was given: read back the starting state (mixed rounds it)
    p0, v0, w0, i0 = sp.particles()
    order = np.argsort(i0)
    Pr = O.Particles(len(w0))
    Pr.set(p0[order], v0[order], w0[order])
    rid = i0[order].astype(np.int64)
    nv = g.node_volumes()
    for _ in range(steps):
        sp.advance(bc, dt)
        sp.compute_number_density()
        if bc == P.REFLECT:
            O.push_reflect(g, Pr, ef, dt, charge, mass)
        else:
            O.push_absorb(g, Pr, ef, dt, charge, mass, sphere[0] if sphere else (0, 0, -9),
                          sphere[1] if sphere else 0.0, ids=rid)
    gp, gv, gw, gi = sp.particles()
    o = np.argsort(gi)
    gp, gv, gw, gi = gp[o], gv[o], gw[o], gi[o]
    rp, rv, rw = Pr.view()
    ro = np.argsort(rid[:Pr.n])
    rp, rv, rw, ri = rp[ro], rv[ro], rw[ro], rid[:Pr.n][ro]
    den_ref = O.number_density(g, Pr.pos[:Pr.n], Pr.mpw[:Pr.n], nv)
    return dict(g=g, gp=gp, gv=gv, gw=gw, gi=gi, rp=rp, rv=rv, rw=rw, ri=ri, den=sp.den(),
                den_ref=den_ref, diag=sp.diag(), sp=sp, w=w)


@pytest.mark.parametrize("vmax,dt", [(2e5, 1e-9), (3e7, 1e-9)])
def test_push_reflect_fp64_matches_oracle(vmax, dt):
    r = _push_case(P.FP64, P.REFLECT, 10, dt, vmax)
    assert np.array_equal(r["gi"], r["ri"])
    assert np.array_equal(r["gp"], r["rp"]) and np.array_equal(r["gv"], r["rv"])
    assert rel_l2(r["den"], r["den_ref"]) < 1e-13
    d = O.species_diag(r["rv"], r["rw"], O.ME)
    np.testing.assert_allclose(r["diag"], d, rtol=1e-12, atol=1e-40)


@pytest.mark.parametrize("vmax,dt", [(2e5, 1e-9), (3e7, 1e-9)])
def test_push_reflect_mixed_matches_oracle(vmax, dt):
    r = _push_case(P.MIXED, P.REFLECT, 10, dt, vmax)
    g = r["g"]
    assert np.array_equal(r["gi"], r["ri"])
    dcell = np.abs(r["gp"] - r["rp"]) / g.dh
    assert dcell.max() < 2e-4, dcell.max()
    assert np.sqrt(np.mean(dcell ** 2)) < 1e-5
    # FP32 velocities: errors are relative to the velocity scale, not to each component
    vscale = np.abs(r["rv"]).max()
    assert np.abs(r["gv"] - r["rv"]).max() < 2e-5 * vscale
    assert rel_l2(r["gv"], r["rv"]) < 2e-6
    assert rel_l2(r["den"], r["den_ref"]) < 1e-5
    d = O.species_diag(r["rv"], r["rw"], O.ME)
    mom_scale = O.ME * np.sum(r["rw"][:, None] * np.abs(r["rv"]))
    np.testing.assert_allclose(r["diag"][[0, 4]], d[[0, 4]], rtol=1e-5)
    np.testing.assert_allclose(r["diag"][1:4], d[1:4], rtol=0, atol=1e-5 * mom_scale)


# 17^3 nodes = 2x2x2 tiles of 8^3 cells; 200,000 particles in the first tile, where one push
# workgroup takes 32,768 of them: seven workgroups share the tile (stayer slots by global
# atomics, later chunks' windows added to the far-deposit array), and movers cross into
# the neighbouring tiles
SPLIT = dict(n=200_000, grid=(17, 17, 17), box=((-0.1, -0.1, -0.1), (-0.004, -0.004, -0.004)))


def test_push_split_tile_fp64_matches_oracle():
    r = _push_case(P.FP64, P.REFLECT, 6, 1e-9, 3e6, **SPLIT)
    assert len(r["gi"]) == SPLIT["n"] and np.array_equal(r["gi"], r["ri"])
    assert np.array_equal(r["gp"], r["rp"]) and np.array_equal(r["gv"], r["rv"])
    assert rel_l2(r["den"], r["den_ref"]) < 1e-13
    # the particles left the first tile's cells
    assert np.any(r["rp"] > -0.0)


def test_push_split_tile_mixed_matches_oracle():
    r = _push_case(P.MIXED, P.REFLECT, 6, 1e-9, 3e6, **SPLIT)
    g = r["g"]
    assert len(r["gi"]) == SPLIT["n"] and np.array_equal(r["gi"], r["ri"])
    dcell = np.abs(r["gp"] - r["rp"]) / g.dh
    assert dcell.max() < 2e-4, dcell.max()
    assert rel_l2(r["den"], r["den_ref"]) < 1e-5


def test_push_absorb_sphere_fp64_matches_oracle():
    sphere = ((0.0, 0.0, 0.0), 0.05, -100.0)
    r = _push_case(P.FP64, P.ABSORB, 10, 2e-9, 3e6, sphere=sphere)
    assert np.array_equal(r["gi"], r["ri"])
    assert np.array_equal(r["gp"], r["rp"]) and np.array_equal(r["gv"], r["rv"])
    assert rel_l2(r["den"], r["den_ref"]) < 1e-13


def test_push_absorb_sphere_mixed_matches_oracle():
    sphere = ((0.0, 0.0, 0.0), 0.05, -100.0)
    r = _push_case(P.MIXED, P.ABSORB, 10, 2e-9, 3e6, sphere=sphere)
    # a kill decision may flip only for a particle within FP32 rounding of a wall/sphere
    common = np.intersect1d(r["gi"], r["ri"])
    assert len(r["ri"]) - len(common) <= 2 and len(r["gi"]) - len(common) <= 2
    gm = np.isin(r["gi"], common)
    rm = np.isin(r["ri"], common)
    dcell = np.abs(r["gp"][gm] - r["rp"][rm]) / r["g"].dh
    assert dcell.max() < 2e-4, dcell.max()
    assert np.sqrt(np.mean(dcell ** 2)) < 1e-5
    assert rel_l2(r["den"], r["den_ref"]) < 1e-4


@pytest.mark.parametrize("bc", [P.REFLECT, P.ABSORB])
def test_push_mixed_frame_velocity_matches_oracle(bc):
    """A drifting mixed species with its frame velocity at the drift
    (pic_species_set_frame_velocity): every interface still sees v, the push still matches
    the oracle, and the FP32 velocity error falls from the scale of the drift (3e6 m/s) to
    that of the deviations from it (2e5 m/s; a weak field keeps them there over the steps)."""
    drift = (0.0, 0.0, 3e6)
    sphere = ((0.0, 0.0, 0.0), 0.05, -100.0) if bc == P.ABSORB else None
    kw = dict(sphere=sphere, drift=drift, amp=0.5)
    plain = _push_case(P.MIXED, bc, 10, 1e-9, 2e5, **kw)
    framed = _push_case(P.MIXED, bc, 10, 1e-9, 2e5, frame=drift, **kw)
    errs = []
    for r in (plain, framed):
        common = np.intersect1d(r["gi"], r["ri"])
        assert len(r["ri"]) - len(common) <= 2 and len(r["gi"]) - len(common) <= 2
        gm, rm = np.isin(r["gi"], common), np.isin(r["ri"], common)
        dcell = np.abs(r["gp"][gm] - r["rp"][rm]) / r["g"].dh
        assert dcell.max() < 2e-4, dcell.max()
        assert rel_l2(r["den"], r["den_ref"]) < 1e-4
        d = O.species_diag(r["rv"], r["rw"], O.ME)
        np.testing.assert_allclose(r["diag"][[0, 3, 4]], d[[0, 3, 4]], rtol=1e-6)
        # the drift's component: x and y hold only the deviations in either form
        errs.append(float(np.median(np.abs(r["gv"][gm, 2] - r["rv"][rm, 2]))))
    assert errs[1] < errs[0] / 4, errs


def test_frame_velocity_change_keeps_particles():
    """Setting the frame on a species that holds particles re-expresses them: read-back
    velocities and diagnostics unchanged to FP32 rounding of the stored value."""
    drift = (1e5, -2e5, 3e6)
    a = _push_case(P.MIXED, P.REFLECT, 0, 1e-9, 2e5, drift=drift, frame=drift)
    b = _push_case(P.MIXED, P.REFLECT, 0, 1e-9, 2e5, drift=drift, frame=drift, frame_after=True)
    assert np.array_equal(a["gi"], b["gi"])
    np.testing.assert_allclose(b["gv"], a["gv"], rtol=0, atol=0.5)  # ulp(3e6) in FP32: 0.25
    np.testing.assert_allclose(b["diag"], a["diag"], rtol=1e-6)
    b["sp"].set_frame_velocity((0.0, 0.0, 0.0))
    v0 = b["sp"].particles()[1][np.argsort(b["sp"].particles()[3])]
    np.testing.assert_allclose(v0, b["gv"], rtol=0, atol=0.5)


# The wide-rim push (a three-cell window rim for the tiles whose particles left the two-cell
# one in the last push; off by default): forced for every tile, and chosen per tile from the
# far counts (PIC_PUSH_WIDE=all / auto); without it the default two-cell rim.  At
# 3e7 m/s and 1e-9 s a particle moves up to 3 cells a step, so many land outside the two-cell
# rim, and some outside the three-cell one too.  Those far deposits are deferred: the push
# marks such a particle's slot, and k_far_arrivals deposits it through its destination tile
# (the density checks) and clears the mark; the weights read back are checked as well.
@pytest.mark.parametrize("mode", ["all", "auto", "off"])
def test_push_wide_rim_fp64_matches_oracle(mode, monkeypatch):
    monkeypatch.setenv("PIC_PUSH_WIDE", mode)
    r = _push_case(P.FP64, P.REFLECT, 10, 1e-9, 3e7)
    assert np.array_equal(r["gi"], r["ri"])
    assert np.array_equal(r["gp"], r["rp"]) and np.array_equal(r["gv"], r["rv"])
    assert np.array_equal(r["gw"], r["rw"])
    assert rel_l2(r["den"], r["den_ref"]) < 1e-13
    d = O.species_diag(r["rv"], r["rw"], O.ME)
    np.testing.assert_allclose(r["diag"], d, rtol=1e-12, atol=1e-40)


@pytest.mark.parametrize("mode", ["all", "auto", "off"])
def test_push_wide_rim_mixed_matches_oracle(mode, monkeypatch):
    monkeypatch.setenv("PIC_PUSH_WIDE", mode)
    r = _push_case(P.MIXED, P.REFLECT, 10, 1e-9, 3e7)
    assert np.array_equal(r["gi"], r["ri"])
    assert np.array_equal(r["gw"], r["rw"].astype(np.float32).astype(np.float64))  # FP32 weights
    dcell = np.abs(r["gp"] - r["rp"]) / r["g"].dh
    assert dcell.max() < 2e-4, dcell.max()
    assert rel_l2(r["den"], r["den_ref"]) < 1e-5


def test_push_wide_rim_split_tile_matches_oracle(monkeypatch):
    # seven workgroups share the first tile; with the wide rim every one of them adds its
    # window beyond the slot to the far-deposit array
    monkeypatch.setenv("PIC_PUSH_WIDE", "all")
    r = _push_case(P.FP64, P.REFLECT, 6, 1e-9, 3e6, **SPLIT)
    assert len(r["gi"]) == SPLIT["n"] and np.array_equal(r["gi"], r["ri"])
    assert np.array_equal(r["gp"], r["rp"]) and np.array_equal(r["gv"], r["rv"])
    assert np.array_equal(r["gw"], r["rw"])
    assert rel_l2(r["den"], r["den_ref"]) < 1e-13


def test_sor_linear_converges_to_oracle_criterion():
    w, g = make_world()
    rng = np.random.default_rng(2)
    rho = rng.normal(0, 1, g.nn) * 1e11 * O.QE
    w.set_field(P.RHO, rho)
    w.set_field(P.PHI, g.zeros())
    s = P.PotentialSolver(w, P.SOR_LINEAR, 10000, 1e-2)
    ok, res, its = s.solve()
    assert ok and res < 1e-2
    phi = w.field(P.PHI)
    # the reference's own stopping test, evaluated in FP64 by the oracle on the GPU result
    chk = phi.copy()
    ok2, l2, its2 = O.solve_sor_linear(g, chk, rho, 1, 1e-2)
    assert ok2, l2
    # tight solve on both sides: same fixed point
    w.set_field(P.PHI, g.zeros())
    s2 = P.PotentialSolver(w, P.SOR_LINEAR, 100000, 1e-9)
    assert s2.solve()[0]
    ref = g.zeros()
    assert O.solve_sor_linear(g, ref, rho, 100000, 1e-9)[0]
    assert rel_l2(w.field(P.PHI), ref) < 1e-9


def _sphere_world(n=(13, 13, 25)):
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    w, g = make_world(n, x0, xm)
    w.add_sphere((0, 0, 0.15), 0.05, -100.0)
    w.add_inlet()
    oid = np.zeros(g.nn, dtype=np.int32)
    phi = g.zeros()
    O.add_sphere(g, oid, phi, (0, 0, 0.15), 0.05, -100.0)
    O.add_inlet(g, oid, phi)
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.5, 1.5, g.nn) * 1e10 * O.QE
    return w, g, oid, phi, rho


def test_gs_boltzmann_same_fixed_point():
    w, g, oid, phi, rho = _sphere_world()
    O.solve_qn(g, phi, g.zeros(), oid, 0, 1.5, 1e12)  # buildMatrix's solveQN guess
    w.set_field(P.PHI, phi)
    w.set_field(P.RHO, rho)
    s = P.PotentialSolver(w, P.GS, 200000, 1e-9)
    s.set_reference_values(0, 1.5, 1e10)
    ok, res, its = s.solve()
    assert ok, (res, its)
    ref = phi.copy()
    assert O.solve_gs_boltzmann(g, ref, rho, oid, 0, 1.5, 1e10, 200000, 1e-9)[0]
    assert rel_l2(w.field(P.PHI), ref) < 1e-9


def test_nrpcg_matches_oracle():
    w, g, oid, phi, rho = _sphere_world()
    O.solve_qn(g, phi, g.zeros(), oid, 0, 1.5, 1e12)
    w.set_field(P.PHI, phi)
    w.set_field(P.RHO, rho)
    s = P.PotentialSolver(w, P.PCG, 10000, 1e-4)
    s.set_reference_values(0, 1.5, 1e10)
    ok, res, its = s.solve()
    ref = phi.copy()
    ok_r, st = O.solve_nrpcg(g, ref, rho, oid, 0, 1.5, 1e10, 10000, 1e-4)
    assert ok == ok_r
    # same algorithm; only the dot-product summation order differs
    assert abs(its - st[1]) <= 2, (its, st)
    assert rel_l2(w.field(P.PHI), ref) < 1e-8


def test_nrpcg_sym_same_solution_as_reference_nrpcg():
    """PIC_SOLVER_PCG_SYM (Neumann nodes eliminated) converges to the solution the
    reference's NR-PCG converges to; both stop on RMS(y) < 1e-3 V."""
    w, g, oid, phi, rho = _sphere_world()
    O.solve_qn(g, phi, g.zeros(), oid, 0, 1.5, 1e12)
    w.set_field(P.PHI, phi)
    w.set_field(P.RHO, rho)
    s = P.PotentialSolver(w, P.PCG_SYM, 10000, 1e-8)
    s.set_reference_values(0, 1.5, 1e10)
    ok, res, its = s.solve()
    assert ok and res < 1e-3
    ref = phi.copy()
    ok_r, st = O.solve_nrpcg(g, ref, rho, oid, 0, 1.5, 1e10, 10000, 1e-8)
    assert ok_r
    np.testing.assert_allclose(w.field(P.PHI), ref, rtol=0, atol=2e-3)


def test_mg_same_solution_as_reference_nrpcg():
    """PIC_SOLVER_MG (multigrid-preconditioned CG inside the same Newton iteration) reaches
    the reference NR-PCG's solution, and the PCG_SYM solution of the same reduced system."""
    w, g, oid, phi, rho = _sphere_world()
    O.solve_qn(g, phi, g.zeros(), oid, 0, 1.5, 1e12)
    w.set_field(P.PHI, phi)
    w.set_field(P.RHO, rho)
    s = P.PotentialSolver(w, P.MG, 10000, 1e-8)
    s.set_reference_values(0, 1.5, 1e10)
    ok, res, its = s.solve()
    assert ok and res < 1e-3
    mg = w.field(P.PHI)
    ref = phi.copy()
    ok_r, st = O.solve_nrpcg(g, ref, rho, oid, 0, 1.5, 1e10, 10000, 1e-8)
    assert ok_r
    np.testing.assert_allclose(mg, ref, rtol=0, atol=2e-3)
    w.set_field(P.PHI, phi)
    s2 = P.PotentialSolver(w, P.PCG_SYM, 10000, 1e-8)
    s2.set_reference_values(0, 1.5, 1e10)
    ok2, _, its2 = s2.solve()
    assert ok2
    assert rel_l2(mg, w.field(P.PHI)) < 1e-8
    assert its < its2 / 3, (its, its2)


@pytest.mark.parametrize("n", [(21, 21, 21), (33, 33, 65), (65, 65, 129)])
def test_mg_iterations_grid_independent(n):
    """CG iterations with the V-cycle stay nearly constant as the grid is refined (Jacobi
    PCG's grow with the grid width), on grids that coarsen fully (33x33x65 -> 3x3x5) and on
    one whose coarsening stops early (21^3 -> 11^3 -> 6^3); same solution as PCG_SYM."""
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    rng = np.random.default_rng(5)
    nn = n[0] * n[1] * n[2]
    rho = rng.uniform(0.5, 1.5, nn) * 1e10 * O.QE
    out = {}
    for typ in (P.MG, P.PCG_SYM):
        w, g = make_world(n, x0, xm)
        w.add_sphere((0, 0, 0.15), 0.05, -100.0)
        w.add_inlet()
        w.set_field(P.RHO, rho)
        s = P.PotentialSolver(w, typ, 20000, 1e-6)
        s.set_reference_values(0, 1.5, 1e10)
        ok, res, its = s.solve()
        assert ok, (typ, res, its)
        out[typ] = (w.field(P.PHI), its)
    assert rel_l2(out[P.MG][0], out[P.PCG_SYM][0]) < 1e-8
    # four Newton steps from phi = 0 at every size here
    assert out[P.MG][1] <= 4 * 30, out[P.MG][1]
    assert out[P.MG][1] < out[P.PCG_SYM][1], (out[P.MG][1], out[P.PCG_SYM][1])


def test_mg_galerkin_operators_follow_the_geometry():
    """The multigrid's Galerkin coarse operators are built once per geometry.  A solver that
    solved before a sphere was added must rebuild them, and one whose sphere is re-added at a
    new potential may keep them: either way its solve is bit-identical, in result and
    iteration count, to that of a new solver on a world with the same geometry and state.  (A
    stale operator would still converge, to the same solution within the tolerance, but in
    other iterations and bits.)"""
    x0, xm = (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4)
    n = (33, 33, 65)
    rho = np.random.default_rng(7).uniform(0.5, 1.5, n[0] * n[1] * n[2]) * 1e10 * O.QE
    w, g = make_world(n, x0, xm)
    w.add_inlet()
    w.set_field(P.RHO, rho)
    s = P.PotentialSolver(w, P.MG, 20000, 1e-6)
    s.set_reference_values(0, 1.5, 1e10)
    spheres = []

    def solve_both():
        phi0 = w.field(P.PHI)
        ok, res, its = s.solve()
        assert ok, (res, its)
        wf, _ = make_world(n, x0, xm)
        for sph in spheres:
            wf.add_sphere(*sph)
        wf.add_inlet()
        wf.set_field(P.RHO, rho)
        wf.set_field(P.PHI, phi0)
        sf = P.PotentialSolver(wf, P.MG, 20000, 1e-6)
        sf.set_reference_values(0, 1.5, 1e10)
        okf, resf, itsf = sf.solve()
        assert okf and itsf == its, (its, itsf)
        assert np.array_equal(w.field(P.PHI), wf.field(P.PHI))

    solve_both()
    for phi_s in (-50.0, -100.0):  # a new sphere, then the same sphere at a new potential
        spheres[:] = [((0, 0, 0.15), 0.05, phi_s)]
        w.add_sphere(*spheres[0])
        solve_both()
