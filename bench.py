#!/usr/bin/env python3
"""ES-PIC inner-loop benchmark on one MI355X.

Workload (default): flow past a sphere on a 257x257x513 grid over [-0.1,0.1]^2 x [0,0.4] m,
sphere r=0.05 m at (0,0,0.15) at -100 V, inlet k=0 at 0 V; 1e9 O+ macroparticles loaded
uniformly with a 7 km/s drift along z plus a 1000 K isotropic thermal spread (Philox,
seed 1), a warm beam injected at z=0 every step (n0=1e10 m^-3); Boltzmann electrons
(n0=1e10, Te=1.5 eV, phi0=0) solved by Newton + CG (tol 1e-4, NR tol 1e-3) on the
Neumann-reduced system with a multigrid V-cycle as the preconditioner, dt=1e-7 s.

One step = inject -> advance (fused push + CIC gather + CIC deposit + rebin) ->
computeNumberDensity -> computeChargeDensity -> solve -> computeEF, the call sequence of
ch3/ver2/Main.cpp:50-70.  The headline value is particle-steps/s of push+gather+deposit
(HIP events around those launches on the world's stream), mixed precision; the same state
is then run with FP64 particles.  Poisson ms/step is reported beside it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "plasma-simulations-by-example_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

from workloads import Case, ensemble_boxes, initial_potential  # noqa: E402

QE = 1.602176565e-19
AMU = 1.660538921e-27
K_B = 1.380648e-23
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
# Algorithmic bytes of one push+gather+deposit launch: a particle must be read (cell, 3
# offsets, 3 velocities, weight: 32 B) and its new cell, offsets and velocities written
# (28 B; the weight never changes), and E read once per node (FP32, 12 B).  The kernel
# itself writes 32 B per particle, weight included, since every particle moves to a new
# slot each step.
BYTES_PER_PARTICLE_STEP = 60
BYTES_PER_NODE_E = 12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=float, default=1e9)
    ap.add_argument("--grid", type=str, default="257,257,513")
    ap.add_argument("--no-fp64", action="store_true", help="skip the all-FP64 comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference timing")
    ap.add_argument("--cpu-particles", type=float, default=1e7)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--T-ion", type=float, default=1000.0)
    ap.add_argument("--horizon", type=int, default=None,
                    help="steps in all (untimed beyond warmup+steps) before the mixed-vs-FP64 "
                         "field and history comparison (default 100; 0 with --no-fp64)")
    ap.add_argument("--push-only", action="store_true",
                    help="profiling mode: the same particle workload, pushes only (no solve), "
                         "then one Field_::updateAverage pass (a plain streaming kernel of known "
                         "bytes) that calibrates the HBM byte counters")
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed",
                    help="particle form of --push-only")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiable sphere case (forward + adjoint gradient)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the differentiable sphere case and print its JSON")
    ap.add_argument("--adjoint-grid", type=str, default="81,81,161")
    ap.add_argument("--adjoint-particles", type=float, default=1e7)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--no-adjoint-fd", action="store_true",
                    help="skip the finite-difference check of the adjoint gradient")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched ensemble of grounded boxes")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched ensemble and print its JSON")
    ap.add_argument("--ensemble-boxes", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-scale", type=float, default=1.0,
                    help="multiplies every box's particle counts (1: the 1e5-4e6 range)")
    ap.add_argument("--no-worlds", action="store_true",
                    help="skip timing the same boxes as separate worlds")
    ap.add_argument("--ensemble-eager", action="store_true",
                    help="launch the ensemble step's kernels one by one instead of replaying its "
                         "captured graphs (the same kernels; for profilers that cannot follow "
                         "graph launches)")
    return ap.parse_args()


def push_only(P, args, grid):
    """Loads the workload's particles and runs `steps` advance+computeNumberDensity steps in
    the field of the sheath-free potential (phi = 0 with the sphere at -100 V), printing
    the push kernel's average duration.  Used under rocprofv3 --pmc."""
    ni, nj, nk = grid
    w = P.World(ni, nj, nk)
    w.set_extents(Case.x0, Case.xm)
    w.add_sphere(Case.sphere_c, Case.sphere_r, Case.phi_s)
    w.add_inlet()
    vol = np.prod(np.subtract(Case.xm, Case.x0))
    prec = P.MIXED if args.precision == "mixed" else P.FP64
    ions = P.Species(w, 16 * AMU, QE, Case.n0 * vol / args.particles, prec,
                     capacity=int(args.particles * 1.02))
    ions.set_frame_velocity((0.0, 0.0, Case.v_drift))  # as the beam source sets it
    ions.load_box_device(Case.x0, Case.xm, Case.n0, int(args.particles), Case.v_drift, args.T_ion,
                         seed=1)
    s = P.PotentialSolver(w, P.QN, 1, 1e-4)
    s.compute_ef()
    w.set_timing(True)
    for i in range(args.warmup + args.steps):
        if i == args.warmup:
            w.timers(reset=True)
        t_before = w.timers()["push"][0]
        ions.advance(P.ABSORB, Case.dt)
        ions.compute_number_density()
        w.sync()
        log(f"[bench] push-only {args.precision} step {i}: push {w.timers()['push'][0] - t_before:.3f} ms")
    tm = w.timers(reset=True)
    n = ions.np()
    # calibration of the HBM byte counters: Field_::updateAverage (k_avg) is a plain
    # streaming kernel over the nodes that reads den and den_ave (16 B/node) and writes
    # den_ave (8 B/node).  Torch kernels are avoided here: launching one under a counter
    # collection pass crashed the profiler's dispatch interception.
    ions.update_averages()
    w.sync()
    print(json.dumps({"push_only": True, "precision": args.precision, "particles": n, "push_avg_ms": tm["push"][0] / tm["push"][1],
                      "push_kernel_avg_ms": tm["push_kernel"][0] / tm["push_kernel"][1],
                      "deposit_avg_ms": tm["deposit"][0] / tm["deposit"][1],
                      "calibration": {"kernel": "k_avg", "read_bytes": 16 * ni * nj * nk,
                                      "write_bytes": 8 * ni * nj * nk}}), flush=True)


_COPY_PROBE = """
import torch
nbytes, reps = 1 << 31, 10
a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
b = torch.empty_like(a)
for _ in range(2):
    torch.mul(a, 1.0, out=b)
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record()
for _ in range(reps):
    torch.mul(a, 1.0, out=b)
e1.record()
torch.cuda.synchronize()
print(2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9)
"""


def copy_bandwidth():
    """Achievable HBM bandwidth on this box: a plain elementwise copy kernel (torch) over
    2 GiB, read + write counted, timed with events; reported beside the roofline peak.  It
    runs in a child process: torch ships its own HIP runtime, and two runtimes in one
    process do not share the device reliably."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, "-c", _COPY_PROBE], capture_output=True, text=True, timeout=300)
        return float(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except (subprocess.TimeoutExpired, ValueError, IndexError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(P, grid, nparticles, precision, steps, warmup, T_ion, want_fields=False, phi=None, horizon=0):
    """Builds the sphere case and times `steps` steps after `warmup` untimed ones, then runs
    on untimed to `horizon` steps in all (when larger), recording the diagnostics history of
    every step.  Returns a dict of results."""
    ni, nj, nk = grid
    w = P.World(ni, nj, nk)
    w.set_extents(Case.x0, Case.xm)
    w.add_sphere(Case.sphere_c, Case.sphere_r, Case.phi_s)
    w.add_inlet()
    vol = np.prod(np.subtract(Case.xm, Case.x0))
    mpw0 = Case.n0 * vol / nparticles
    ions = P.Species(w, 16 * AMU, QE, mpw0, precision, capacity=int(nparticles * 1.02), name="O+")
    ions.set_frame_velocity((0.0, 0.0, Case.v_drift))  # as the beam source sets it
    t0 = time.time()
    ions.load_box_device(Case.x0, Case.xm, Case.n0, int(nparticles), Case.v_drift, T_ion, seed=1)
    w.sync()
    t_load = time.time() - t0
    ions.compute_number_density()
    w.compute_charge_density([ions])
    os.environ["PIC_VERBOSE"] = "1"  # one line per Newton iteration while setting up
    if phi is None:
        initial_potential(P, w)
    else:
        w.set_field(P.PHI, phi)
    phi_init = w.field(P.PHI)
    # Newton + CG with the reference's tolerances, on the Neumann-reduced system (the
    # reference's one-sided Neumann rows leave a non-symmetric Jacobian on which CG stalls
    # at this size), preconditioned by a multigrid V-cycle (PIC_SOLVER_MG; the Jacobi
    # preconditioner of the reference, PIC_SOLVER_PCG_SYM, reaches the same solution in
    # ~100x more CG iterations)
    solver = P.PotentialSolver(w, P.MG, 10000, 1e-4)
    solver.set_reference_values(0.0, Case.Te, Case.n0)
    solver.compute_ef()

    counter = [0]

    def step():
        ions.inject_beam(Case.v_drift, Case.n0, T_ion, Case.dt, seed=1, counter=counter[0])
        counter[0] += 1
        ions.advance(P.ABSORB, Case.dt)
        ions.compute_number_density()
        w.compute_charge_density([ions])
        ok, res, its = solver.solve()
        solver.compute_ef()
        return ok, its

    log(f"[bench] {'mixed' if precision == P.MIXED else 'fp64'}: loaded {nparticles:.3g} "
        f"particles in {t_load:.1f}s; warmup {warmup}")
    for i in range(warmup):
        ok, its = step()
        log(f"[bench]   warmup {i}: cg its {its} converged {ok} np {ions.np():.4g}")
    os.environ["PIC_VERBOSE"] = "0"
    w.set_timing(True)
    w.timers(reset=True)
    np_sum = 0
    pcg = []
    w.sync()
    t0 = time.perf_counter()
    push_prev = 0.0
    for i in range(steps):
        n_before = ions.np()
        ok, its = step()
        np_sum += n_before
        pcg.append(its)
        push_now = w.timers()["push"][0]
        log(f"[bench]   step {i}: cg its {its} converged {ok}, push {push_now - push_prev:.3f} ms")
        push_prev = push_now
    w.sync()
    wall = time.perf_counter() - t0
    tm = w.timers(reset=True)
    w.set_timing(False)
    push_ms, push_calls = tm["push"]
    dep_ms, _ = tm["deposit"]
    out = dict(
        np_final=ions.np(), particle_steps=np_sum, wall_s=wall, ms_per_step=1e3 * wall / steps,
        push_ms=push_ms, deposit_ms=dep_ms, push_calls=push_calls,
        inject_ms=tm["inject"][0], solve_ms=tm["solve"][0], ef_ms=tm["ef"][0], rho_ms=tm["rho"][0],
        pcg_iters=pcg, load_s=t_load, nn=ni * nj * nk,
    )
    out["particle_steps_per_s"] = np_sum / ((push_ms + dep_ms) * 1e-3)
    out["push_avg_ms"] = push_ms / max(push_calls, 1)
    # the fused push kernel alone (the push phase less its segment plan)
    out["push_kernel_avg_ms"] = tm["push_kernel"][0] / max(tm["push_kernel"][1], 1)
    out["phi_init"] = phi_init
    # untimed continuation to the comparison horizon; diagnostics of every step
    hist = []
    late_push = []  # push time of each continuation step (the wake fills in over ~100 steps)
    late_kernel = []  # the push kernel's HIP-event time of each continuation step
    late_np = []  # particles pushed at each continuation step
    w.set_timing(True)
    w.timers(reset=True)
    for i in range(warmup + steps, max(horizon, warmup + steps)):
        late_np.append(ions.np())
        ok, its = step()
        if not ok:
            log(f"[bench]   step {i}: Newton did not converge")
        tl = w.timers(reset=True)
        late_push.append(tl["push"][0])
        late_kernel.append(tl["push_kernel"][0] / max(tl["push_kernel"][1], 1))
        d = ions.diag()  # getRealCount, getMomentum (3), getKE
        hist.append([float(ions.np()), d[0], d[3], d[4], w.get_pe()])
        if (i + 1) % 20 == 0:
            log(f"[bench]   {'mixed' if precision == P.MIXED else 'fp64'} step {i + 1}: np {ions.np():.4g}, "
                f"KE {d[4]:.6e}, PE {hist[-1][4]:.6e}, push {np.mean(late_push[-20:]):.2f} ms (mean of last 20)")
    w.set_timing(False)
    out["history"] = np.array(hist)
    out["late_push_ms"] = late_push
    out["late_kernel_ms"] = late_kernel
    out["late_np"] = late_np
    out["steps_total"] = max(horizon, warmup + steps)
    if want_fields:
        out["phi"] = w.field(P.PHI)
        out["ef"] = w.field(P.EF)
        out["den"] = ions.den()
    ions.close()
    solver.close()
    w.close()
    return out


def adjoint_case(P, args):
    """The differentiable sphere case: K steps forward in FP64 with soft absorption, then the
    hand-written adjoint over checkpoints every M steps, for the loss
    L = I/I0 + mean_wake(den^2)/D0 (ion current to the sphere plus the wake density L2,
    normalised at theta0).  Optionally checked against central finite differences of the
    forward run (eps = 1e-4 relative)."""
    ni, nj, nk = (int(x) for x in args.adjoint_grid.split(","))
    n_init = int(args.adjoint_particles)
    vol = np.prod(np.subtract(Case.xm, Case.x0))
    area = (Case.xm[0] - Case.x0[0]) * (Case.xm[1] - Case.x0[1])
    # the beam's weight per macro-particle equals the fill's: J = v A dt n_init / V
    n_inject = int(round(Case.v_drift * area * Case.dt * n_init / vol))
    cfg = dict(ni=ni, nj=nj, nk=nk, x0=Case.x0, xm=Case.xm, sphere_c=Case.sphere_c,
               sphere_r=Case.sphere_r, dt=Case.dt, steps=args.adjoint_steps,
               checkpoint_every=args.adjoint_every, mass=16 * AMU, charge=QE, n0=Case.n0,
               T_ion=args.T_ion, n_init=n_init, n_inject=n_inject, soft_width=1.0,
               cg_tol=1e-6, nr_tol=1e-8, adj_rel_tol=1e-8, seed=1)
    run = P.SphereAdjointRun(**cfg)
    mask = run.wake_mask()
    nn = ni * nj * nk
    den = P.DeviceArray(nn)
    theta = np.array([Case.v_drift, Case.phi_s, Case.Te])
    cur = run.forward(theta, den.address)
    d = den.get()
    fwd_ms = run.stats()[0]
    solves = run.stats()[2]
    I0, D0 = cur, float(np.mean(d[mask] ** 2))
    dden = P.DeviceArray.from_host(np.where(mask, 2 * d / (mask.sum() * D0), 0.0))
    grad = run.backward(1.0 / I0, dden.address)
    st = run.stats()
    log(f"[bench] adjoint {ni}x{nj}x{nk}, {n_init:.3g} ions + {n_inject}/step, K={args.adjoint_steps}: "
        f"forward {fwd_ms:.0f} ms, backward {st[1]:.0f} ms, dL/dln(theta) {grad * np.abs(theta)}")
    slots = n_init + n_inject * (args.adjoint_steps + 1) / 2.0  # mean slots per step
    out = {
        "workload": f"differentiable sphere {ni}x{nj}x{nk}, {n_init:.3g} O+ (FP64) + {n_inject} "
                    f"injected per step, K={args.adjoint_steps}, checkpoints every {args.adjoint_every}",
        "theta": theta.tolist(), "loss": "I/I0 + mean_wake(den^2)/D0 at theta0",
        "current_A": cur, "forward_ms": fwd_ms, "backward_ms": st[1],
        "backward_over_forward": st[1] / fwd_ms,
        "newton_solves_forward": solves, "adjoint_cg_iterations": st[3],
        "forward_particle_steps_per_s": slots * args.adjoint_steps / (fwd_ms * 1e-3),
        "grad": grad.tolist(), "dL_dln_theta": (grad * np.abs(theta)).tolist(),
    }
    if not args.no_adjoint_fd:
        def L(t):
            c_ = run.forward(t, den.address)
            dd = den.get()
            return c_ / I0 + float(np.mean(dd[mask] ** 2)) / D0
        s_adj = grad * np.abs(theta)
        out["fd_check"] = {}
        for eps in (1e-4, 1e-5):
            fd = np.zeros(3)
            for a in range(3):
                h = eps * abs(theta[a])
                tp, tm = theta.copy(), theta.copy()
                tp[a] += h
                tm[a] -= h
                fd[a] = (L(tp) - L(tm)) / (2 * h)
            s_fd = fd * np.abs(theta)
            out["fd_check"][f"eps_rel_{eps:g}"] = {
                "fd_dL_dln_theta": s_fd.tolist(),
                "rel_err": (np.abs(s_adj - s_fd) / np.abs(s_fd)).tolist(),
                "err_over_max": (np.abs(s_adj - s_fd) / np.abs(s_fd).max()).tolist()}
            log(f"[bench] adjoint fd check eps {eps:g}: {out['fd_check'][f'eps_rel_{eps:g}']}")
    dden.close()
    den.close()
    run.close()
    return out


ENS_DT = 2e-10  # ch2/Main.cpp:19
ENS_TOL = 1e-4  # ch2/Main.cpp:38


def ensemble_case(P, args):
    """The 64-box ensemble: every box's ch2 step (push + gather + deposit of both species,
    density, rho, SOR, computeEF) in one hipGraph replay, against the same boxes run as
    separate worlds through the per-world entry points."""
    spec = [(16 * AMU, QE), (9.10938215e-31, -QE)]
    bx = ensemble_boxes(args.ensemble_boxes, 1, args.ensemble_scale)
    boxes = [((b["n"],) * 3, (-b["L"] / 2,) * 3, (b["L"] / 2,) * 3) for b in bx]
    counts = [b["counts"] for b in bx]
    ens = P.Ensemble(boxes, spec, counts, P.MIXED)
    for i, b in enumerate(bx):
        L = b["L"]
        ens.load_box(i, 0, (-L / 2,) * 3, (L / 2,) * 3, b["den"], seed=1)
        ens.load_box(i, 1, (0.0,) * 3, (L / 2,) * 3, b["den"], seed=1)
    tuned = P.tuned_ensemble_params(ens)  # applied by the library at creation
    if args.ensemble_eager:
        ens.set_tuning(chunk_particles=tuned["chunk_particles"], window_nodes=tuned["window_nodes"],
                       sor_threads=tuned["sor_threads"], sort_every=tuned["sort_every"], use_graph=False)
    t0 = time.time()
    ens.solve_initial(10000, ENS_TOL)
    log(f"[bench] ensemble: {len(bx)} boxes, {ens.stats()['particles']:.3g} particles, "
        f"{ens.stats()['nodes']:.3g} nodes; initial solve {time.time() - t0:.1f}s; tuning {tuned}")
    ens.step(args.warmup, ENS_DT, 10000, ENS_TOL)
    ens.step(args.ensemble_steps, ENS_DT, 10000, ENS_TOL)
    st = ens.stats()
    it, _ = ens.solver_status()
    nparts = st["particles"]
    out = {
        "workload": f"{len(bx)} grounded boxes (seed 1), 17^3-65^3 nodes, 1e5-4e6 particles per "
                    f"species x {args.ensemble_scale:g}, mixed precision, SOR tol {ENS_TOL}",
        "particles": nparts, "nodes": st["nodes"], "ms_per_step": st["ms_per_step"],
        "particle_steps_per_s": nparts / (st["ms_per_step"] * 1e-3),
        "graph_kernels_per_step": st["graph_kernels"], "launches_per_step": 1,
        "push_workgroups": st["push_workgroups"], "tuning": tuned,
        "sor_iterations_last_step": {"max": int(np.max(np.abs(it))), "mean": float(np.mean(np.abs(it))),
                                     "all_converged": bool((it > 0).all())},
    }
    log(f"[bench] ensemble: {st['ms_per_step']:.2f} ms/step, {out['particle_steps_per_s']:.3e} "
        f"particle-steps/s, SOR its max {out['sor_iterations_last_step']['max']}")
    ens.close()
    if not args.no_worlds:
        # the same boxes as separate worlds through the per-world entry points
        worlds = []
        for i, b in enumerate(bx):
            L = b["L"]
            w = P.World(b["n"], b["n"], b["n"])
            w.set_extents((-L / 2,) * 3, (L / 2,) * 3)
            sps = []
            for s, (m, q) in enumerate(spec):
                lo = (-L / 2,) * 3 if s == 0 else (0.0,) * 3
                vol = L ** 3 if s == 0 else (L / 2) ** 3
                sp = P.Species(w, m, q, b["den"] * vol / counts[i][s], P.MIXED,
                               capacity=int(counts[i][s] * 1.05) + 1024)
                sp.load_box_device(lo, (L / 2,) * 3, b["den"], counts[i][s], seed=1)
                sps.append(sp)
            solver = P.PotentialSolver(w, P.SOR_LINEAR, 10000, ENS_TOL)
            for sp in sps:
                sp.compute_number_density()
            w.compute_charge_density(sps)
            solver.solve()
            solver.compute_ef()
            worlds.append((w, sps, solver))
        log("[bench] worlds: loaded")

        def world_steps(k):
            for _ in range(k):
                for w, sps, solver in worlds:
                    for sp in sps:
                        sp.advance(P.REFLECT, ENS_DT)
                    for sp in sps:
                        sp.compute_number_density()
                    w.compute_charge_density(sps)
                    solver.solve()
                    solver.compute_ef()
            for w, _, _ in worlds:
                w.sync()
        world_steps(args.warmup)
        k = max(2, args.ensemble_steps // 4)
        t0 = time.perf_counter()
        world_steps(k)
        wms = 1e3 * (time.perf_counter() - t0) / k
        out["separate_worlds"] = {"ms_per_step": wms, "particle_steps_per_s": nparts / (wms * 1e-3),
                                  "ensemble_speedup": wms / st["ms_per_step"]}
        log(f"[bench] separate worlds: {wms:.2f} ms/step ({wms / st['ms_per_step']:.2f}x the ensemble)")
        for w, sps, solver in worlds:
            solver.close()
            for sp in sps:
                sp.close()
            w.close()
    return out


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, grid, ef, T_ion):
    """Times the reference's own Species::advance + computeNumberDensity (oracle/_ref,
    built from the reference sources) on a bounded sample of the same workload: the same
    grid, sphere and E field, cpu-particles particles of the same distribution."""
    from oracle import ref as R
    if not (R.available("ch3") and R.available("mt")):
        return None
    ni, nj, nk = grid
    n = int(args.cpu_particles)
    rng = np.random.default_rng(7)
    lo, hi = np.array(Case.x0), np.array(Case.xm)
    pos = lo + rng.random((n, 3)) * (hi - lo)
    c = np.array(Case.sphere_c)
    pos = pos[np.sum((pos - c) ** 2, axis=1) > Case.sphere_r ** 2]
    n = len(pos)
    vth = np.sqrt(2 * K_B * T_ion / (16 * AMU))
    vel = rng.normal(0, vth / np.sqrt(2), (n, 3))
    vel[:, 2] += Case.v_drift
    vol = np.prod(hi - lo)
    mpw = np.full(n, Case.n0 * vol / args.particles)
    res = {}
    for kind, ch, threads in (("serial", "ch3", 1), ("mt", "mt", args.cpu_threads)):
        sim = R.RefSim(ch, ni, nj, nk, Case.x0, Case.xm, Case.dt, 100)
        sim.add_sphere(Case.sphere_c, Case.sphere_r, Case.phi_s)
        sim.add_inlet()
        sp = sim.add_species("O+", 16 * AMU, QE, float(mpw[0]))
        if ch == "mt":
            sim.set_num_threads(threads)
        sim.set_field(R.EF, ef)
        sim.set_particles(sp, pos, vel, mpw)
        t, ps = sim.time_particle_steps(args.cpu_steps)
        res[kind] = dict(value=ps / t, seconds=t, particle_steps=ps, threads=threads)
        sim.close()
        log(f"[bench] cpu {kind}: {ps / t:.3e} particle-steps/s ({t:.1f}s, {threads} threads)")
    return res, n


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus "
              f"{args.gpus}. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    copy_gbps = None if args.push_only else copy_bandwidth()
    import pic_amd as P
    if P.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    grid = tuple(int(x) for x in args.grid.split(","))
    if args.push_only:
        push_only(P, args, grid)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_case(P, args)}), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_case(P, args)}), flush=True)
        return
    horizon = args.horizon if args.horizon is not None else (0 if args.no_fp64 else 100)
    mixed = run_gpu(P, grid, args.particles, P.MIXED, args.steps, args.warmup, args.T_ion,
                    want_fields=True, horizon=horizon)
    log(f"[bench] mixed: {mixed['particle_steps_per_s']:.4e} particle-steps/s, "
        f"push {mixed['push_avg_ms']:.3f} ms/launch, {mixed['ms_per_step']:.1f} ms/step")
    fp64 = None
    if not args.no_fp64:
        # the same seeded particles and the same initial potential: the field differences
        # after warmup+steps steps are those of the particle arithmetic alone
        fp64 = run_gpu(P, grid, args.particles, P.FP64, args.steps, args.warmup, args.T_ion,
                       want_fields=True, phi=mixed["phi_init"], horizon=horizon)
        log(f"[bench] fp64: {fp64['particle_steps_per_s']:.4e} particle-steps/s, "
            f"push {fp64['push_avg_ms']:.3f} ms/launch")

    # roofline of the dominant kernel (the fused push): algorithmic bytes per launch over
    # the kernel's average HIP-event duration (events around the k_push launches alone: the
    # two-cell-rim instantiation and the wide-rim one for the tiles the plan flagged)
    steps = args.steps
    np_avg = mixed["particle_steps"] / steps
    bytes_launch = np_avg * BYTES_PER_PARTICLE_STEP + mixed["nn"] * BYTES_PER_NODE_E
    achieved = bytes_launch / (mixed["push_kernel_avg_ms"] * 1e-3)
    # measured HBM bytes of the push (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated; see
    # profiles/push_sphere257_traffic.md) scaled to this run's particle count
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "push_mixed_sphere257_traffic.json")
    if os.path.exists(tfile) and tuple(grid) == (257, 257, 513):
        try:
            traffic = json.load(open(tfile))["hbm_bytes_per_particle"] * np_avg
        except (OSError, ValueError, KeyError):
            traffic = None
    traffic_src = None
    if traffic is not None:
        tj = json.load(open(tfile))
        traffic_src = (f"stored counter measurement: {tj['hbm_bytes_per_particle']:.2f} B/particle from "
                       f"rocprofv3 FETCH_SIZE + WRITE_SIZE of bench.py --push-only, library of commit "
                       f"{tj.get('measured_with_commit', '?')} (profiles/push_mixed_sphere257_traffic.json), "
                       f"times this run's particles")
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "k_push<PIC_MIXED, HALO> + k_far_arrivals<PIC_MIXED>", "bytes_per_launch": bytes_launch,
                "kernel_ms": mixed["push_kernel_avg_ms"],
                "copy_kernel_gbps": copy_gbps}

    cpu = None
    if not args.no_cpu:
        r = cpu_baseline(args, grid, mixed["ef"], args.T_ion)
        if r is not None:
            res, n = r
            cpu = {"value": res["serial"]["value"], "unit": "particle-steps/s", "cores": 1,
                   "kind": "reference", "cpu_model": cpu_model(),
                   "sample": f"reference ch3/ver2 Species::advance+computeNumberDensity, "
                             f"{n} particles x {args.cpu_steps} steps on the same "
                             f"{grid[0]}x{grid[1]}x{grid[2]} grid, sphere and E field",
                   "threaded": {"value": res["mt"]["value"], "cores": res["mt"]["threads"],
                                "kind": "reference", "code": "ch9/MT"}}

    line = {
        "metric": "particle-steps/s (push+gather+deposit) + Poisson ms/step, mixed vs FP64",
        "value": mixed["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"sphere {grid[0]}x{grid[1]}x{grid[2]}, {args.particles:.3g} O+ "
                        f"mixed-precision particles, Boltzmann-electron Newton-MGPCG",
            "particles": args.particles, "grid": list(grid), "dt": Case.dt,
            "solver": "Newton + multigrid-preconditioned CG on the Neumann-reduced system "
                      "(PIC_SOLVER_MG), FP64, tol 1e-4, NR tol 1e-3",
            "particles_final": mixed["np_final"],
        },
        "poisson_ms_per_step": mixed["solve_ms"] / steps,
        "cg_iters_per_step": float(np.mean(mixed["pcg_iters"])),
        "phase_ms_per_step": {k: mixed[k + "_ms"] / steps for k in
                              ("push", "deposit", "inject", "solve", "ef", "rho")},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        line["fp64"] = {"value": fp64["particle_steps_per_s"], "unit": "particle-steps/s",
                        "ms_per_step": fp64["ms_per_step"],
                        "poisson_ms_per_step": fp64["solve_ms"] / steps,
                        "push_avg_ms": fp64["push_avg_ms"],
                        "push_kernel_avg_ms": fp64["push_kernel_avg_ms"],
                        "mixed_speedup": mixed["particle_steps_per_s"] / fp64["particle_steps_per_s"]}

        def rel_l2(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        line["mixed_vs_fp64_rel_l2"] = {
            "den": rel_l2(mixed["den"], fp64["den"]), "phi": rel_l2(mixed["phi"], fp64["phi"]),
            "ef": rel_l2(mixed["ef"], fp64["ef"]), "after_steps": mixed["steps_total"],
            "start": "same seeded particles, same initial phi"}
        hm, hf = mixed["history"], fp64["history"]
        if len(hm) and len(hm) == len(hf):
            # largest relative difference over the history of each diagnostic
            def hdiff(c, scale=None):
                ref = np.abs(hf[:, c]) if scale is None else scale
                return float(np.max(np.abs(hm[:, c] - hf[:, c]) / np.maximum(ref, 1e-300)))
            line["mixed_vs_fp64_history"] = {
                "steps": f"{args.warmup + steps + 1}-{mixed['steps_total']}",
                "particle_count": hdiff(0), "real_count": hdiff(1), "kinetic_energy": hdiff(3),
                "momentum_z": hdiff(2), "potential_energy": hdiff(4)}
            # push time late in the run, once the wake behind the sphere has formed
            line["late_push_ms"] = {
                "steps": f"{mixed['steps_total'] - 19}-{mixed['steps_total']}",
                "mixed": float(np.mean(mixed["late_push_ms"][-20:])),
                "fp64": float(np.mean(fp64["late_push_ms"][-20:]))}
            # the same roofline as above over the last 20 steps of the continuation
            np_late = float(np.mean(mixed["late_np"][-20:]))
            ms_late = float(np.mean(mixed["late_kernel_ms"][-20:]))
            b_late = np_late * BYTES_PER_PARTICLE_STEP + mixed["nn"] * BYTES_PER_NODE_E
            a_late = b_late / (ms_late * 1e-3)
            line["late_roofline"] = {
                "steps": line["late_push_ms"]["steps"], "bound": "hbm", "achieved": a_late / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": a_late / HBM_PEAK,
                "bytes_per_launch": b_late, "kernel_ms": ms_late, "particles": np_late,
                "particle_steps_per_s": np_late / (ms_late * 1e-3)}
    if not args.no_adjoint:
        line["adjoint"] = adjoint_case(P, args)
    if not args.no_ensemble:
        line["ensemble"] = ensemble_case(P, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
