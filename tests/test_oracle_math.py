"""Pins the FP64 oracle against mathematics (CPU).

Each check is chosen so that a plausible slip -- a dropped term, a wrong sign, a swapped
index, a transposed axis -- fails it, independently of the reference build.
"""
import numpy as np
import pytest

from oracle import oracle as O

QE, AMU, ME, EPS_0 = O.QE, O.AMU, O.ME, O.EPS_0


def grid(n=(9, 11, 13), x0=(-0.1, -0.2, 0.0), xm=(0.3, 0.2, 0.5)):
    return O.OGrid(*n, x0, xm)


def node_xyz(g):
    k, j, i = np.meshgrid(np.arange(g.nk), np.arange(g.nj), np.arange(g.ni), indexing="ij")
    x = g.x0[0] + i.reshape(-1) * g.dh[0]
    y = g.x0[1] + j.reshape(-1) * g.dh[1]
    z = g.x0[2] + k.reshape(-1) * g.dh[2]
    return x, y, z


def test_mt19937_known_answer():
    r = O.Rng(5489)
    for _ in range(9999):
        r.next_u32()
    assert r.next_u32() == 4123659995  # the value the C++ standard fixes for mt19937


def test_philox4x32_10_known_answers():
    """The oracle's Philox4x32-10 (the counter-based generator of the device injection and
    of the differentiable run) against the published known-answer vectors of the Random123
    library (Salmon et al., SC11): zero counter and key, all-ones, and the pi-digit case."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        assert O.philox4x32_10(ctr, key) == list(want), (ctr, key)


def test_node_volumes_closed_form():
    g = grid()
    nv = g.node_volumes()
    assert np.isclose(nv.sum(), np.prod(g.xm - g.x0), rtol=1e-13)
    full = np.prod(g.dh)
    assert nv.max() == full and nv.min() == full / 8


def test_gather_exact_on_affine_field():
    """CIC reproduces an affine field exactly, per component (catches swapped corners,
    wrong weights, a transposed index)."""
    g = grid()
    x, y, z = node_xyz(g)
    A = np.array([[1.5, -2.0, 0.7, 3.0], [0.2, 0.9, -1.1, -4.0], [-0.3, 0.4, 2.2, 1.0]])
    ef = np.stack([A[c, 0] * x + A[c, 1] * y + A[c, 2] * z + A[c, 3] for c in range(3)], axis=1)
    ef = np.ascontiguousarray(ef).reshape(-1)
    rng = np.random.default_rng(0)
    for _ in range(200):
        lc = rng.random(3) * (np.array([g.ni, g.nj, g.nk]) - 1 - 1e-9)
        X = g.x0 + lc * g.dh
        e = O.gather3(g, ef, lc)
        np.testing.assert_allclose(e, A[:, :3] @ X + A[:, 3], rtol=1e-12, atol=1e-12)


def test_scatter_partition_of_unity_and_mass():
    g = grid()
    rng = np.random.default_rng(1)
    f = g.zeros()
    total = 0.0
    for _ in range(100):
        lc = rng.random(3) * (np.array([g.ni, g.nj, g.nk]) - 1 - 1e-9)
        v = rng.random()
        O.scatter(g, f, lc, v)
        total += v
    assert np.isclose(f.sum(), total, rtol=1e-13)
    # number density: sum(den * node_vol) == sum(mpw)
    n = 5000
    pos = g.x0 + rng.random((n, 3)) * (g.xm - g.x0)
    mpw = rng.random(n) * 1e4
    nv = g.node_volumes()
    den = O.number_density(g, pos, mpw, nv)
    assert np.isclose(np.sum(den * nv), mpw.sum(), rtol=1e-12)
    # the first moment is preserved too (CIC conserves the centre of charge)
    x, y, z = node_xyz(g)
    assert np.isclose(np.sum(den * nv * x), np.sum(mpw * pos[:, 0]), rtol=1e-11)


def test_push_uniform_field_closed_form():
    """Leapfrog in a uniform field: v_n = v0 + n a dt, x_n = x0 + n v0 dt + a dt^2 n(n+1)/2."""
    g = grid((11, 11, 11), (0, 0, 0), (1, 1, 1))
    E = np.array([3.0, -2.0, 1.5])
    ef = np.tile(E, g.nn)
    P = O.Particles(1)
    x0 = np.array([[0.5, 0.5, 0.5]])
    v0 = np.array([[1e-3, -2e-3, 0.5e-3]])
    P.set(x0, v0, np.array([1.0]))
    q, m, dt = 2.0, 1e3, 1e-2
    a = q / m * E
    n = 20
    for _ in range(n):
        O.push_reflect(g, P, ef, dt, q, m)
    np.testing.assert_allclose(P.vel[0], v0[0] + n * a * dt, rtol=1e-12)
    np.testing.assert_allclose(P.pos[0], x0[0] + n * v0[0] * dt + a * dt * dt * n * (n + 1) / 2,
                               rtol=1e-12)


def test_reflection_conserves_speed_and_bounds():
    g = grid((11, 11, 11), (0, 0, 0), (1, 1, 1))
    ef = g.zeros(3)
    rng = np.random.default_rng(2)
    n = 2000
    P = O.Particles(n)
    P.set(rng.random((n, 3)), rng.uniform(-30, 30, (n, 3)), np.ones(n))
    s0 = np.abs(P.vel[:n]).copy()
    for _ in range(25):
        O.push_reflect(g, P, ef, 0.01, 1.0, 1.0)
        assert np.all(P.pos[:n] >= 0) and np.all(P.pos[:n] <= 1)
    np.testing.assert_array_equal(np.abs(P.vel[:n]), s0)


def test_absorb_kills_exactly_the_outside_and_sphere():
    g = grid((11, 11, 11), (0, 0, 0), (1, 1, 1))
    ef = g.zeros(3)
    rng = np.random.default_rng(3)
    n = 4000
    P = O.Particles(n)
    pos = rng.random((n, 3))
    vel = rng.uniform(-20, 20, (n, 3))
    P.set(pos, vel, np.ones(n))
    ids = np.arange(n)
    c, r = np.array([0.5, 0.5, 0.5]), 0.2
    O.push_absorb(g, P, ef, 0.01, 1.0, 1.0, c, r, ids=ids)
    new = pos + vel * 0.01
    alive = np.all((new >= 0) & (new < 1), axis=1) & (np.sum((new - c) ** 2, axis=1) > r * r)
    assert P.n == alive.sum()
    assert set(ids[:P.n]) == set(np.nonzero(alive)[0])


def test_linear_sor_manufactured_solution():
    """rho = -eps0 * lap_h(phi*) with phi* = 0 on the walls -> SOR returns phi*."""
    g = grid((12, 10, 9))
    rng = np.random.default_rng(4)
    phi_star = rng.normal(0, 1, (g.nk, g.nj, g.ni))
    phi_star[0], phi_star[-1] = 0, 0
    phi_star[:, 0], phi_star[:, -1] = 0, 0
    phi_star[:, :, 0], phi_star[:, :, -1] = 0, 0
    idx2 = 1 / g.dh ** 2
    lap = np.zeros_like(phi_star)
    p = phi_star
    lap[1:-1, 1:-1, 1:-1] = (idx2[0] * (p[1:-1, 1:-1, 2:] - 2 * p[1:-1, 1:-1, 1:-1] + p[1:-1, 1:-1, :-2])
                             + idx2[1] * (p[1:-1, 2:, 1:-1] - 2 * p[1:-1, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1])
                             + idx2[2] * (p[2:, 1:-1, 1:-1] - 2 * p[1:-1, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1]))
    rho = (-EPS_0 * lap).reshape(-1)
    phi = g.zeros()
    ok, l2, its = O.solve_sor_linear(g, phi, rho, 100000, 1e-9)
    assert ok
    np.testing.assert_allclose(phi, phi_star.reshape(-1), atol=1e-9)


def test_qn_closed_form():
    g = grid()
    rng = np.random.default_rng(5)
    rho = rng.uniform(0, 2, g.nn) * 1e10 * QE
    rho[::5] = 0
    oid = np.zeros(g.nn, dtype=np.int32)
    oid[::7] = 1
    phi = np.full(g.nn, -7.0)
    O.solve_qn(g, phi, rho, oid, 0.3, 1.5, 1e10)
    expect = 0.3 + 1.5 * np.log(np.maximum(rho / (1e10 * QE), 1e-6))
    np.testing.assert_allclose(phi[oid == 0], expect[oid == 0], rtol=1e-14)
    assert np.all(phi[oid > 0] == -7.0)


def test_ef_exact_on_quadratics():
    """Central and second-order one-sided differences are exact for quadratics."""
    g = grid()
    x, y, z = node_xyz(g)
    phi = 2 * x * x - 3 * y * y + 0.5 * z * z + 4 * x - y + 2 * z + 1
    ef = O.compute_ef(g, phi).reshape(-1, 3)
    np.testing.assert_allclose(ef[:, 0], -(4 * x + 4), atol=1e-10)
    np.testing.assert_allclose(ef[:, 1], -(-6 * y - 1), atol=1e-10)
    np.testing.assert_allclose(ef[:, 2], -(z + 2), atol=1e-10)


@pytest.mark.parametrize("solver", ["gs", "nrpcg"])
def test_boltzmann_uniform_density_solution(solver):
    """Pure-Neumann box, uniform ion density n_i: phi = phi0 + Te ln(n_i / n0)."""
    g = grid((9, 9, 13), (0, 0, 0), (0.08, 0.08, 0.12))
    oid = np.zeros(g.nn, dtype=np.int32)
    ni, n0, Te, phi0 = 2.5e10, 1e10, 1.5, 0.2
    rho = np.full(g.nn, QE * ni)
    phi = np.zeros(g.nn)
    if solver == "gs":
        ok, l2, its = O.solve_gs_boltzmann(g, phi, rho, oid, phi0, Te, n0, 200000, 1e-8)
    else:
        ok, st = O.solve_nrpcg(g, phi, rho, oid, phi0, Te, n0, 10000, 1e-8)
    assert ok
    np.testing.assert_allclose(phi, phi0 + Te * np.log(ni / n0), atol=2e-6)


def test_cold_beam_count_and_placement():
    g = O.OGrid(11, 11, 21, (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4))
    rng = O.Rng(3)
    P = O.Particles(400000)
    ef = g.zeros(3)
    dt, mpw0, den, v = 1e-7, 50.0, 1e10, 7000.0
    counts = [O.cold_beam_sample(g, P, rng, v, den, mpw0, ef, dt, QE, 16 * AMU) for _ in range(40)]
    expect = den * v * 0.04 * dt / mpw0  # n v A dt / mpw0 = 5600 exactly representable
    assert abs(np.mean(counts) - expect) < 1.0
    pos, vel, _ = P.view()
    assert np.all(pos[:, 2] == 0.0)
    assert np.all(vel[:, 2] == v) and np.all(vel[:, :2] == 0)


def test_warm_beam_moments():
    """WarmBeamSource: drift + isotropic speed from sums of uniforms (ch4/Species.cpp)."""
    g = O.OGrid(11, 11, 21, (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4))
    rng = O.Rng(4)
    P = O.Particles(300000)
    T, m, v = 1000.0, 16 * AMU, 7000.0
    for _ in range(40):
        O.warm_beam_sample(g, P, rng, v, 1e10, T, 50.0, g.zeros(3), 1e-7, QE, m)
    _, vel, _ = P.view()
    rel = vel - np.array([0, 0, v])
    assert np.all(np.abs(rel.mean(axis=0)) < 10.0)          # isotropic: zero mean
    vth = np.sqrt(2 * O.K * T / m)
    speed = np.linalg.norm(rel, axis=1)
    # E|v1|^2 per component of the sum of 3 uniforms minus 1.5 is 1/4 -> speed^2 mean:
    # (9/6) * 3 * vth^2 / 4 = 9/8 vth^2
    assert abs(np.mean(speed ** 2) / (9 / 8 * vth ** 2) - 1) < 0.02
