"""Workloads shared by bench.py and the autotuner: the sphere case of the 257x257x513 runs
and the batched ensemble of grounded boxes."""
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Case:
    x0 = (-0.1, -0.1, 0.0)
    xm = (0.1, 0.1, 0.4)
    sphere_c = (0.0, 0.0, 0.15)
    sphere_r = 0.05
    phi_s = -100.0
    n0 = 1e10
    v_drift = 7000.0
    Te = 1.5
    dt = 1e-7


def initial_potential(P, w, solver_tol=1e-2):
    """Initial phi for a domain preloaded with a uniform plasma.  Undamped Newton from
    phi = 0 around a -100 V sphere diverges (the exp(phi/Te) term overshoots); the
    reference never meets that state because its domain fills from vacuum.  Continuation
    in the sphere potential (-25, -50, -75, -100 V) with a loose-tolerance Newton solve reaches
    the sheath solution; the timed steps then use the reference solver settings."""
    s = P.PotentialSolver(w, P.MG, 5000, solver_tol)
    s.set_reference_values(0.0, Case.Te, Case.n0)
    for k in range(1, 5):
        phi_s = Case.phi_s * k / 4
        w.add_sphere(Case.sphere_c, Case.sphere_r, phi_s)
        t0 = time.time()
        ok, res, its = s.solve()
        log(f"[bench]   continuation phi_s={phi_s:.0f} V: converged {ok}, {its} CG its, "
            f"{time.time() - t0:.1f}s")
    s.close()


def ensemble_boxes(nbox, seed=1, scale=1.0):
    """The batched workload: nbox grounded boxes (ch2/Main.cpp) drawn from a seed.  Nodes per
    axis log-uniform in [17, 65] (17^3 to 65^3 nodes), particles per species log-uniform in
    [1e5, 4e6], widths uniform in [0.05, 0.2] m, densities log-uniform in [1e10, 1e12] m^-3.
    O+ ions fill the box, electrons its upper octant [0, L/2]^3, all at rest."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nbox):
        n = int(round(np.exp(rng.uniform(np.log(17), np.log(65)))))
        L = rng.uniform(0.05, 0.2)
        dens = 10 ** rng.uniform(10, 12)
        counts = [int(np.exp(rng.uniform(np.log(1e5), np.log(4e6))) * scale) for _ in range(2)]
        boxes.append(dict(n=n, L=L, den=dens, counts=counts))
    return boxes


