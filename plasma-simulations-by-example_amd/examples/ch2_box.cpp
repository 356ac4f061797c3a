// Grounded box (21^3 nodes, 80k O+ and 10k e-): O+ ions over the box and electrons over one octant, reflecting
// walls, linear SOR.  The call sequence is that of ch2/Main.cpp:14-75, written against the
// drop-in classes; the particle load draws from a seeded rnd, so the initial state is the
// reference's for the same seed.
//   usage: ch2_box [num_ts=10000] [seed=1] [fp64=0] [tol=1e-2]
#include <cstdlib>
#include <iostream>
#include <vector>

#include "Output.h"
#include "PotentialSolver.h"
#include "Species.h"
#include "World.h"

using namespace std;
using namespace Const;

int main(int argc, char *argv[]) {
    int num_ts = argc > 1 ? atoi(argv[1]) : 10000;
    unsigned seed = argc > 2 ? (unsigned)atoi(argv[2]) : 1u;
    bool fp64 = argc > 3 && atoi(argv[3]) != 0;
    double tol = argc > 4 ? atof(argv[4]) : 1e-2;

    World world(21, 21, 21);
    world.setExtents({-0.1, -0.1, -0.1}, {0.1, 0.1, 0.1});
    world.setTime(2e-10, num_ts);
    world.setPrecision(fp64 ? PIC_FP64 : PIC_MIXED);

    vector<Species> species;
    species.reserve(2);
    species.push_back(Species("O+", 16 * AMU, QE, world));
    species.push_back(Species("e-", ME, -1 * QE, world));

    rnd.seed(seed);
    species[0].loadParticlesBox(world.getX0(), world.getXm(), 1e11, 80000);
    species[1].loadParticlesBox({0, 0, 0}, world.getXm(), 1e11, 10000);

    PotentialSolver solver(world, 10000, tol);
    solver.solve();
    solver.computeEF();

    while (world.advanceTime()) {
        for (Species &sp : species) {
            sp.advance();
            sp.computeNumberDensity();
        }
        world.computeChargeDensity(species);
        solver.solve();
        solver.computeEF();
        if (world.getTs() % 100 == 0 || world.getTs() == num_ts) {
            Output::screenOutput(world, species);
            Output::diagOutput(world, species);
        }
    }
    Output::fields(world, species);
    cout << "Simulation took " << world.getWallTime() << " seconds" << endl;
    return 0;
}
