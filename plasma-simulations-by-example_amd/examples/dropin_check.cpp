// Check program for the GPU tests of the drop-in classes (not an example of use): drives
// World / Species / PotentialSolver / Output the way the reference's Main.cpp does and dumps
// the state at full precision for comparison with the reference's own outputs.
//
//   dropin_check qs <dump> [fp64]
//       Species::loadParticlesBoxQS on an 11^3 lattice over [-0.1,-0.1,0]..[0.1,0.1,0.2]
//       (the golden case of ch2/Species.cpp:101-141), then the particles.
//   dropin_check sphere <steps> <dump> [fp64]
//       ch3/ver2/Main.cpp's loop (:48-80) on its 21x21x41 sphere world with the QN solver, at
//       dt 2e-7 and mpw0 200: each step injects the cold beam with the draws of
//       ColdBeamSource::sample (ch3/ver2/Source.cpp:4-27) from rnd seeded with 7 and passed
//       through addParticle, then advance, density, rho, solve, computeEF, steadyState and
//       updateAverages, Output::diagOutput, and Output::fields every 50 steps and on the
//       step isLastTimeStep() names; the dump after the loop.
//
// Dump (little-endian): int64 nn, np, ts; double time; then phi, rho (nn each), ef (3 nn),
// den, den_ave (nn each) in the flat order U = k ni nj + j ni + i, then pos (3 np), vel (3 np),
// mpw (np).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "Output.h"
#include "PotentialSolver.h"
#include "Species.h"
#include "World.h"

using namespace std;
using namespace Const;

// progress lines (stdout, flushed): the caller watches them
static void progress(const string &what) {
    static const auto t0 = chrono::steady_clock::now();
    double s = chrono::duration<double>(chrono::steady_clock::now() - t0).count();
    cout << "[dropin_check " << s << " s] " << what << endl;
}

static void dump(const char *path, World &world, Species &sp) {
    world.updateHost();
    sp.syncHost(true);
    FILE *f = fopen(path, "wb");
    if (!f) {
        perror(path);
        exit(2);
    }
    int64_t nn = (int64_t)world.ni * world.nj * world.nk, np = (int64_t)sp.particles.size();
    int64_t ts = world.getTs();
    double time = world.getTime();
    fwrite(&nn, 8, 1, f);
    fwrite(&np, 8, 1, f);
    fwrite(&ts, 8, 1, f);
    fwrite(&time, 8, 1, f);
    vector<double> flat(nn);
    for (Field *fld : {&world.phi, &world.rho}) {
        fld->deflate(flat.data());
        fwrite(flat.data(), 8, nn, f);
    }
    vector<double3> e3(nn);
    world.ef.deflate(e3.data());
    for (const double3 &e : e3) fwrite(&e, 8, 3, f);
    for (Field *fld : {&sp.den, &sp.den_ave}) {
        fld->deflate(flat.data());
        fwrite(flat.data(), 8, nn, f);
    }
    for (Particle &p : sp.particles) fwrite(&p.pos, 8, 3, f);
    for (Particle &p : sp.particles) fwrite(&p.vel, 8, 3, f);
    for (Particle &p : sp.particles) fwrite(&p.mpw, 8, 1, f);
    fclose(f);
}

static int run_qs(const char *out, bool fp64) {
    progress("start");
    World world(21, 21, 21);
    progress("world created");
    world.setExtents({-0.1, -0.1, 0}, {0.1, 0.1, 0.2});
    world.setTime(2e-10, 1);
    world.setPrecision(fp64 ? PIC_FP64 : PIC_MIXED);
    vector<Species> species;
    species.push_back(Species("O+", 16 * AMU, QE, world));
    species[0].loadParticlesBoxQS(world.getX0(), world.getXm(), 1e11, {11, 11, 11});
    species[0].getNp();  // flushes the queued additions
    dump(out, world, species[0]);
    return 0;
}

static int run_sphere(int steps, const char *out, bool fp64) {
    const double dt = 2e-7, mpw0 = 200.0, v_drift = 7000.0, ndi = 1e10;
    progress("start");
    World world(21, 21, 41);
    progress("world created");
    world.setExtents({-0.1, -0.1, 0}, {0.1, 0.1, 0.4});
    world.setTime(dt, steps);
    world.setPrecision(fp64 ? PIC_FP64 : PIC_MIXED);
    world.addSphere({0, 0, 0.15}, 0.05, -100.0);
    world.addInlet();
    vector<Species> species;
    species.push_back(Species("O+", 16 * AMU, QE, mpw0, world));
    Species &ions = species[0];
    if (!fp64) {  // as a ColdBeamSource on this species would set it
        const double u[3] = {0, 0, v_drift};
        pic_check(pic_species_set_frame_velocity(ions.dev(), u), "frame velocity");
    }
    PotentialSolver solver(world, QN, 20000, 1e-4);
    solver.setReferenceValues(0, 1.5, ndi);
    solver.solve();
    solver.computeEF();
    rnd.seed(7);
    double3 dh = world.getDh(), x0 = world.getX0();
    const double Lx = dh[0] * (world.ni - 1), Ly = dh[1] * (world.nj - 1);
    while (world.advanceTime()) {
        // the draws of ColdBeamSource::sample, in its order and arithmetic
        const double area = Lx * Ly;
        int num = (int)(ndi * v_drift * area * world.getDt() / mpw0 + rnd());
        for (int q = 0; q < num; q++) {
            double x = x0[0] + rnd() * Lx;
            double y = x0[1] + rnd() * Ly;
            ions.addParticle({x, y, x0[2]}, {0, 0, v_drift}, mpw0);
        }
        ions.advance();
        ions.computeNumberDensity();
        world.computeChargeDensity(species);
        solver.solve();
        solver.computeEF();
        if (world.steadyState(species)) ions.updateAverages();
        Output::diagOutput(world, species);
        if (world.getTs() % 50 == 0 || world.isLastTimeStep()) {
            Output::fields(world, species);
            progress("step " + to_string(world.getTs()) + ", " + to_string(ions.getNp()) + " particles");
        }
    }
    dump(out, world, ions);
    progress("dumped");
    return 0;
}

int main(int argc, char *argv[]) {
    string mode = argc > 1 ? argv[1] : "";
    if (mode == "qs" && argc >= 3) return run_qs(argv[2], argc > 3 && strcmp(argv[3], "fp64") == 0);
    if (mode == "sphere" && argc >= 4)
        return run_sphere(atoi(argv[2]), argv[3], argc > 4 && strcmp(argv[4], "fp64") == 0);
    cerr << "usage: dropin_check qs <dump> [fp64] | sphere <steps> <dump> [fp64]" << endl;
    return 2;
}
