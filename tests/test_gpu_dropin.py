"""The C++ drop-in classes run the reference's main-loop call sequence on the GPU (the grounded box
and the ch3 sphere) and agree with the oracle replaying the same seeded run.

With the solver tolerance tight (1e-8 here) the red-black and lexicographic SOR iterates
coincide to ~1e-9, so the FP64 run matches the oracle to the 6 significant digits that
runtime_diags.csv carries (the reference's default stream precision) and the
mixed-precision run to 1e-4.
"""
import csv
import os
import select
import subprocess
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "plasma-simulations-by-example_amd")


def run_example(tmp_path, name, *args):
    exe = os.path.join(PKG, "examples", name)
    assert os.path.exists(exe), f"{exe} not built"
    os.makedirs(tmp_path / "results", exist_ok=True)
    print(f"running {name} {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([exe, *map(str, args)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = list(csv.DictReader(open(tmp_path / "runtime_diags.csv")))
    return r.stdout, rows


def oracle_grounded_box(num_ts, tol):
    x0, xm, dt = (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), 2e-10
    g = O.OGrid(21, 21, 21, x0, xm)
    nv = g.node_volumes()
    rng = O.Rng(1)
    ef = g.zeros(3)
    Pi, Pe = O.Particles(80000), O.Particles(10000)
    O.load_box(g, Pi, rng, x0, xm, 1e11, 80000, ef, dt, O.QE, 16 * O.AMU)
    O.load_box(g, Pe, rng, (0, 0, 0), xm, 1e11, 10000, ef, dt, -O.QE, O.ME)
    phi = g.zeros()
    O.solve_sor_linear(g, phi, g.zeros(), 10000, tol)
    ef = O.compute_ef(g, phi)
    out = {}
    for ts in range(num_ts + 1):
        O.push_reflect(g, Pi, ef, dt, O.QE, 16 * O.AMU)
        O.push_reflect(g, Pe, ef, dt, -O.QE, O.ME)
        di = O.number_density(g, Pi.pos[:Pi.n], Pi.mpw[:Pi.n], nv)
        de = O.number_density(g, Pe.pos[:Pe.n], Pe.mpw[:Pe.n], nv)
        rho = O.charge_density([O.QE, -O.QE], [di, de])
        O.solve_sor_linear(g, phi, rho, 10000, tol)
        ef = O.compute_ef(g, phi)
        out[ts] = dict(KE_i=O.species_diag(Pi.vel[:Pi.n], Pi.mpw[:Pi.n], 16 * O.AMU)[4],
                       KE_e=O.species_diag(Pe.vel[:Pe.n], Pe.mpw[:Pe.n], O.ME)[4],
                       PE=O.get_pe(g, ef, nv))
    return out


@pytest.mark.parametrize("fp64,rtol", [(1, 5e-6), (0, 1e-4)])
def test_ch2_box_dropin_matches_oracle(tmp_path, fp64, rtol):
    num_ts = 10
    stdout, rows = run_example(tmp_path, "ch2_box", num_ts, 1, fp64, 1e-8)
    assert "Simulation took" in stdout
    ref = oracle_grounded_box(num_ts, 1e-8)
    last = rows[-1]
    ts = int(last["ts"])
    assert ts == num_ts
    assert float(last["mp_count.O+"]) == 80000 and float(last["mp_count.e-"]) == 10000
    np.testing.assert_allclose(float(last["KE.e-"]), ref[ts]["KE_e"], rtol=rtol)
    np.testing.assert_allclose(float(last["KE.O+"]), ref[ts]["KE_i"], rtol=rtol)
    np.testing.assert_allclose(float(last["PE"]), ref[ts]["PE"], rtol=rtol)
    # written after the loop, at ts = num_ts + 1 (advanceTime pre-increments)
    assert os.path.exists(tmp_path / "results" / f"fields_{num_ts + 1:05d}.vti")


def test_ch3_sphere_dropin_runs_and_conserves(tmp_path):
    """ch3 main loop with cold-beam injection and NR-PCG: the particle count follows
    injection minus absorption, and the run is reproducible for a given seed."""
    stdout, rows = run_example(tmp_path, "ch3_sphere", 30, "PCG", -100, 0, 3)
    np_ = [float(r["mp_count.O+"]) for r in rows]
    assert np_[0] > 4000 and all(b > a for a, b in zip(np_[:10], np_[1:11]))  # filling up
    stdout2, rows2 = run_example(tmp_path / "again", "ch3_sphere", 30, "PCG", -100, 0, 3)
    assert [r["mp_count.O+"] for r in rows2] == [r["mp_count.O+"] for r in rows]
    # absorbing walls: the real count equals mp_count * mpw0
    for r in rows:
        assert abs(float(r["real_count.O+"]) - 50.0 * float(r["mp_count.O+"])) < 1e-6 * float(
            r["real_count.O+"])


def test_ch3_sphere_dropin_mg_matches_pcg(tmp_path):
    """The same ch3 main with SolverType MG (multigrid-preconditioned CG on the reduced
    system) follows the PCG run: both stop on the reference's tolerances, so the potentials
    differ by at most the solver tolerance and the diagnostics histories agree closely."""
    _, rows_p = run_example(tmp_path / "pcg", "ch3_sphere", 30, "PCG", -100, 0, 3)
    _, rows_m = run_example(tmp_path / "mg", "ch3_sphere", 30, "MG", -100, 0, 3)
    assert len(rows_p) == len(rows_m)
    for rp, rm in zip(rows_p, rows_m):
        assert abs(float(rm["mp_count.O+"]) - float(rp["mp_count.O+"])) <= 2e-3 * float(rp["mp_count.O+"])
        assert abs(float(rm["PE"]) - float(rp["PE"])) <= 2e-3 * abs(float(rp["PE"]))


# ---- the drop-in classes against the reference's own outputs ------------------------------

REPO = os.path.dirname(PKG)
GOLD = os.path.join(REPO, "tests", "golden")
REF_DIR = os.path.join(REPO, "oracle", "_ref")


def run_streamed(cmd, cwd, timeout, env=None, echo_every=1):
    """Runs cmd with its output echoed as it comes (every echo_every-th line), so that a slow
    run shows where it is; kills it after timeout seconds.  Returns (returncode, output)."""
    print(f"running {' '.join(map(str, cmd))}", flush=True)
    p = subprocess.Popen(list(map(str, cmd)), cwd=cwd, env=env, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, bufsize=1)
    lines, deadline = [], time.monotonic() + timeout
    while True:
        ready, _, _ = select.select([p.stdout], [], [], 1.0)
        if ready:
            line = p.stdout.readline()
            if line == "":
                break
            lines.append(line)
            if (len(lines) - 1) % echo_every == 0:
                print("  | " + line.rstrip(), flush=True)
        if time.monotonic() > deadline:
            p.kill()
            p.wait()
            raise AssertionError(f"{cmd[0]} killed after {timeout} s:\n" + "".join(lines[-20:]))
    return p.wait(), "".join(lines)


def run_check(cwd, *args):
    """examples/dropin_check (the reference's Main.cpp sequences on the drop-in classes)."""
    exe = os.path.join(PKG, "examples", "dropin_check")
    assert os.path.exists(exe), f"{exe} not built"
    os.makedirs(os.path.join(cwd, "results"), exist_ok=True)
    rc, out = run_streamed([exe, *args], cwd, 150)
    assert rc == 0, out[-2000:]
    return out


def first_difference(a: bytes, b: bytes):
    """A short account of where two files differ (pytest's own diff of megabyte strings takes
    minutes)."""
    la, lb = a.split(b"\n"), b.split(b"\n")
    for n, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            fx, fy = x.split(b" "), y.split(b" ")
            for m, (u, v) in enumerate(zip(fx, fy)):
                if u != v:
                    return f"line {n} field {m}: ours {u[:60]!r} reference {v[:60]!r}"
            return f"line {n}: {len(fx)} fields against {len(fy)}"
    return f"{len(la)} lines against {len(lb)}"


def vti_sections(raw: bytes):
    """(header line, body lines) of a .vti file: the lines before the first DataArray, then
    each DataArray, then the closing lines."""
    out, head, body = [], b"preamble", []
    for line in raw.split(b"\n"):
        if line.startswith(b"<DataArray") or line.startswith(b"</PointData"):
            out.append((head, body))
            head, body = line, []
        elif line == b"</DataArray>":
            continue
        else:
            body.append(line)
    out.append((head, body))
    return out


def read_dump(path):
    """The check program's state dump (format in examples/dropin_check.cpp)."""
    raw = open(path, "rb").read()
    nn, np_, ts = np.frombuffer(raw[:24], dtype="<i8")
    vals = np.frombuffer(raw[32:], dtype="<f8")
    sizes = [("phi", nn), ("rho", nn), ("ef", 3 * nn), ("den", nn), ("den_ave", nn),
             ("pos", 3 * np_), ("vel", 3 * np_), ("mpw", np_)]
    out, o = dict(ts=int(ts), time=float(np.frombuffer(raw[24:32], dtype="<f8")[0])), 0
    for name, k in sizes:
        out[name] = vals[o:o + k]
        o += k
    assert o == len(vals)
    out["pos"] = out["pos"].reshape(-1, 3)
    out["vel"] = out["vel"].reshape(-1, 3)
    return out


@pytest.mark.parametrize("fp64", [True, False], ids=["fp64", "mixed"])
def test_qs_loader_through_dropin_matches_reference(tmp_path, fp64):
    """Species::loadParticlesBoxQS of the drop-in (ch2/Species.cpp:101-141) on the golden
    case: the same lattice points and weights as the compiled reference (exactly in FP64;
    to the mixed storage's FP32 offset otherwise), zero velocities."""
    G = np.load(os.path.join(GOLD, "qs_ch2.npz"))
    run_check(tmp_path, "qs", "qs.bin", *(["fp64"] if fp64 else []))
    d = read_dump(tmp_path / "qs.bin")
    assert len(d["mpw"]) == len(G["mpw"]) == 11 ** 3

    def rows(pos, mpw):  # device order is not the load order: compare sorted rows
        a = np.column_stack([pos, mpw])
        return a[np.lexsort(a.T[::-1])]

    a, b = rows(d["pos"], d["mpw"]), rows(G["pos"], G["mpw"])
    if fp64:
        assert np.array_equal(a, b)
    else:
        np.testing.assert_allclose(a[:, :3], b[:, :3], rtol=0, atol=1e-9)
        np.testing.assert_allclose(a[:, 3], b[:, 3], rtol=1e-7)
    assert np.all(d["vel"] == 0)


@pytest.fixture(scope="module")
def sphere_dropin(tmp_path_factory):
    """FP64 sphere runs of the check program (the golden history's case), by step count."""
    runs = {}

    def get(steps):
        if steps not in runs:
            tmp = tmp_path_factory.mktemp(f"sphere{steps}")
            out = run_check(tmp, "sphere", steps, "sphere.bin", "fp64")
            runs[steps] = (tmp, out, read_dump(tmp / "sphere.bin"))
        return runs[steps]

    return get


def test_sphere_dropin_steady_state_averages_and_diags(sphere_dropin):
    """World::steadyState and Species::updateAverages of the drop-in over the golden sphere
    history (tests/golden/hist_ch3.npz, written by the compiled reference): the same
    steady-state step, reported once; den_ave and phi to rounding; runtime_diags.csv with
    the reference's particle count on every row and its diagnostics to the file's 6 digits."""
    tmp, out, d = sphere_dropin(400)
    G = np.load(os.path.join(GOLD, "hist_ch3.npz"))
    step = int(np.argmax(G["steady"]))
    assert out.count("Steady state reached") == 1
    assert f"Steady state reached at time step {step}\n" in out
    assert len(d["mpw"]) == G["np"][-1]
    for k in ("den_ave", "phi"):
        assert np.linalg.norm(d[k] - G[k]) <= 1e-12 * np.linalg.norm(G[k]), k
    rows = list(csv.DictReader(open(tmp / "runtime_diags.csv")))
    assert [int(r["ts"]) for r in rows] == list(range(len(G["np"])))
    assert [int(r["mp_count.O+"]) for r in rows] == list(G["np"])
    cols = ["real_count.O+", "px.O+", "py.O+", "pz.O+", "KE.O+"]
    csv_diag = np.array([[float(r[c]) for c in cols] for r in rows])
    scale = np.abs(G["diag"][:, 1:]).max(axis=0)
    scale[1:4] = scale[3]  # momentum columns on the z momentum's scale
    assert (np.abs(csv_diag - G["diag"][:, 1:]) / scale).max() < 1e-5
    pe = np.array([float(r["PE"]) for r in rows])
    assert (np.abs(pe - G["pe"]) / np.abs(G["pe"]).max()).max() < 1e-5


def test_output_files_match_reference_writer_byte_for_byte(sphere_dropin, tmp_path):
    """host/Output.cpp against the reference's own Output.cpp (ch3/ver2/Output.cpp:12-132,
    compiled into oracle/_ref) on the same state: the drop-in's fields and particles after
    100 steps are loaded into the reference's World and Species, whose writers then produce
    results/fields_00100.vti and a runtime_diags.csv row.  The .vti files are identical
    byte for byte; so are the CSV header and the last row except its wall-clock column (the
    row's sums are each side's own reductions of the same particles, printed to 6 digits)."""
    from oracle import ref as R
    if not R.available("ch3"):
        pytest.skip("oracle/_ref not built (make -C oracle ref)")
    steps = 100  # a multiple of 50: the check program writes the fields of its last step
    tmp, _, d = sphere_dropin(steps)
    sim = R.RefSim("ch3", 21, 21, 41, (-0.1, -0.1, 0.0), (0.1, 0.1, 0.4), 2e-7, steps)
    sim.add_sphere((0, 0, 0.15), 0.05, -100.0)
    sim.add_inlet()
    ion = sim.add_species("O+", 16 * O.AMU, O.QE, 200.0)
    sim.set_particles(ion, d["pos"], d["vel"], d["mpw"])
    sim.set_field(R.PHI, d["phi"])
    sim.set_field(R.RHO, d["rho"])
    sim.set_field(R.EF, d["ef"])
    sim.set_field(100 + ion, d["den"])
    sim.set_field(200 + ion, d["den_ave"])
    for _ in range(steps + 1):  # ts -1 -> 400, the time summed as the check program's
        sim.advance_time()
    os.makedirs(tmp_path / "results")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        sim.output_fields()
        sim.output_diag()
    finally:
        os.chdir(cwd)
    name = f"results/fields_{steps:05d}.vti"
    ours, theirs = (tmp / name).read_bytes(), (tmp_path / name).read_bytes()
    assert len(theirs) > 100000
    # The reference allocates node_vol as (ni, nk, nk) (ch3/ver2/World.cpp:17), so on this
    # 21x21x41 grid its NodeVol array holds 21*41 values per k plane, the j >= nj ones zero,
    # against the 21*21 of the declared extent.  The drop-in writes the declared extent:
    # NodeVol is compared on the nodes that exist, every other byte exactly.
    a, b = vti_sections(ours), vti_sections(theirs)
    assert [h for h, _ in a] == [h for h, _ in b]
    ni, nj, nk = 21, 21, 41
    for (head, la), (_, lb) in zip(a, b):
        if b'Name="NodeVol"' in head:
            assert len(la) == len(lb) == nk
            for x, y in zip(la, lb):
                tx, ty = x.split(), y.split()
                assert len(tx) == ni * nj and len(ty) == ni * nk
                assert tx == ty[:ni * nj] and set(ty[ni * nj:]) == {b"0"}
        else:
            same = la == lb  # outside the assert: pytest would diff megabytes
            assert same, (head, first_difference(b"\n".join(la), b"\n".join(lb)))
    mine = (tmp / "runtime_diags.csv").read_text().splitlines()
    ref = (tmp_path / "runtime_diags.csv").read_text().splitlines()
    assert len(ref) == 2 and mine[0] == ref[0]

    def masked(row):
        f = row.split(",")
        return f[:2] + f[3:]  # ts, time, (wall_time), ...

    assert masked(mine[-1]) == masked(ref[1]), (mine[-1], ref[1])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_unchanged_reference_mains_run_on_dropin(tmp_path, prec):
    """The reference's ch2/Main.cpp and ch3/ver2/Main.cpp, compiled unchanged against the
    drop-in classes (oracle/Makefile `mains`), run to completion on the GPU with the
    precision PICAMD_PRECISION selects: ch2 keeps every QS-loaded particle (81^3 ions, 41^3
    electrons) for its 10000 steps; ch3 fills the domain from its beam source with the real
    count mpw0 times the particle count."""
    env = dict(os.environ, PICAMD_PRECISION=prec)
    for exe, steps in (("ch2_main", 10000), ("ch3_main", 400)):
        path = os.path.join(REF_DIR, exe)
        if not os.path.exists(path):
            pytest.skip(f"{path} not built (make -C oracle mains)")
        cwd = tmp_path / exe
        os.makedirs(cwd / "results")
        rc, out = run_streamed([path], cwd, 280, env=env, echo_every=500)
        assert rc == 0, out[-2000:]
        assert "Simulation took" in out
        rows = list(csv.DictReader(open(cwd / "runtime_diags.csv")))
        assert [int(x["ts"]) for x in rows] == list(range(steps + 1))
        if exe == "ch2_main":
            assert {(x["mp_count.O+"], x["mp_count.e-"]) for x in rows} == {("531441", "68921")}
            assert os.path.exists(cwd / "results" / "fields_10000.vti")
        else:
            nps = [int(x["mp_count.O+"]) for x in rows]
            assert nps[0] > 0 and nps[-1] > 50 * nps[0]
            for x in rows:
                assert abs(float(x["real_count.O+"]) - 50 * int(x["mp_count.O+"])) <= \
                    1e-5 * float(x["real_count.O+"])
            assert os.path.exists(cwd / "results" / "fields_00400.vti")
