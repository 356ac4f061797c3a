"""pic_species_diag_noalloc: the diagnostics reduction with its partial sums in the library's
static device array, called through the C ABI, against pic_species_diag on the same
particles (Species::getRealCount / getMomentum / getKE, ch3/ver2/Species.cpp:84-108).

After a host write of the particles both entry points run the same full reduction (the same
per-block partials summed in the same fixed order), so the outputs must be bit-identical.
"""
import numpy as np
import pytest

import pic_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def loaded_species(n_nodes, n, prec, seed):
    x0, xm = np.array([-0.1, -0.1, 0.0]), np.array([0.1, 0.1, 0.4])
    w = P.World(*n_nodes)
    w.set_extents(x0, xm)
    rng = np.random.default_rng(seed)
    pos = x0 + rng.random((n, 3)) * (xm - x0) * 0.999
    vel = rng.normal(0.0, 700.0, (n, 3)) + [0.0, 0.0, 7000.0]
    mpw = rng.uniform(10.0, 50.0, n)
    sp = P.Species(w, 16 * O.AMU, O.QE, 20.0, prec, capacity=n + 1024)
    sp.set_particles(pos, vel, mpw)
    ref = O.species_diag(vel, mpw, 16 * O.AMU)
    return w, sp, ref


@pytest.mark.parametrize("prec", [P.FP64, P.MIXED])
def test_diag_noalloc_matches_diag_on_two_calls(prec):
    # first call: 33x33x65 nodes, 300k particles; second: a smaller world and input
    wa, a, ref_a = loaded_species((33, 33, 65), 300_000, prec, 1)
    got = a.diag_noalloc()
    assert np.array_equal(got, a.diag()), (got, a.diag())
    tol = 1e-12 if prec == P.FP64 else 1e-6
    assert np.allclose(got, ref_a, rtol=tol, atol=0), (got, ref_a)

    wb, b, ref_b = loaded_species((17, 17, 17), 5_000, prec, 2)
    got_b = b.diag_noalloc()
    assert np.array_equal(got_b, b.diag()), (got_b, b.diag())
    assert np.allclose(got_b, ref_b, rtol=tol, atol=0)

    # the array keeps the last call's contents: a call on the larger species again must not
    # see any of them
    again = a.diag_noalloc()
    assert np.array_equal(again, got)


def test_diag_noalloc_after_push_matches_full_reduction():
    """After advance() pic_species_diag reads the push's own partials (another summation
    order); the static-scratch reduction over the pushed particles agrees to rounding."""
    w, sp, _ = loaded_species((33, 33, 65), 100_000, P.FP64, 3)
    sp.advance(P.ABSORB, 1e-7)
    fused = sp.diag()
    full = sp.diag_noalloc()
    assert np.allclose(full, fused, rtol=1e-12, atol=0), (full, fused)
    pos, vel, mpw, _ = sp.particles()
    ref = O.species_diag(vel, mpw, 16 * O.AMU)
    assert np.allclose(full, ref, rtol=1e-12, atol=0)


def test_diag_noalloc_rejects_grids_beyond_its_scratch():
    # 256x256x1024 cells = 131072 tiles of 8^3 cells: more than the 131008 the array holds
    w = P.World(257, 257, 1025)
    w.set_extents((-0.1, -0.1, 0.0), (0.1, 0.1, 0.8))
    sp = P.Species(w, 16 * O.AMU, O.QE, 20.0, P.MIXED, capacity=1024)
    with pytest.raises(P.PicError, match="static scratch"):
        sp.diag_noalloc()
    assert np.all(np.isfinite(sp.diag()))
