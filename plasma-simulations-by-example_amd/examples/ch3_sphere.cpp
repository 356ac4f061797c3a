// Flow past a sphere with Boltzmann electrons: cold (or warm) O+ beam injected at z=0,
// absorbing sphere and walls.  The call sequence is that of ch3/ver2/Main.cpp:16-91,
// written against the drop-in classes.
//   usage: ch3_sphere [num_ts=400] [solver=PCG|GS|QN|PCGS|MG] [phi_sphere=-100] [T_ion=0] [seed=1]
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <vector>

#include "Output.h"
#include "PotentialSolver.h"
#include "Source.h"
#include "Species.h"
#include "World.h"

using namespace std;
using namespace Const;

int main(int argc, char *argv[]) {
    int num_ts = argc > 1 ? atoi(argv[1]) : 400;
    SolverType st = PCG;
    if (argc > 2 && !strcmp(argv[2], "GS")) st = GS;
    if (argc > 2 && !strcmp(argv[2], "QN")) st = QN;
    if (argc > 2 && !strcmp(argv[2], "PCGS")) st = PCGS;
    if (argc > 2 && !strcmp(argv[2], "MG")) st = MG;
    double phi_sphere = argc > 3 ? atof(argv[3]) : -100;
    double T_ion = argc > 4 ? atof(argv[4]) : 0.0;
    rnd.seed(argc > 5 ? (unsigned)atoi(argv[5]) : 1u);

    World world(21, 21, 41);
    world.setExtents({-0.1, -0.1, 0}, {0.1, 0.1, 0.4});
    world.setTime(1e-7, num_ts);
    cout << "Sphere potential: " << phi_sphere << " V" << endl;
    world.addSphere({0, 0, 0.15}, 0.05, phi_sphere);
    world.addInlet();

    vector<Species> species;
    species.push_back(Species("O+", 16 * AMU, QE, 5e1, world));

    const double ndi = 1e10;
    vector<unique_ptr<Source>> sources;
    if (T_ion > 0)
        sources.emplace_back(new WarmBeamSource(species[0], world, 7000, ndi, T_ion));
    else
        sources.emplace_back(new ColdBeamSource(species[0], world, 7000, ndi));

    PotentialSolver solver(world, st, st == GS ? 20000 : 10000, 1e-4);
    solver.setReferenceValues(0, 1.5, ndi);
    solver.solve();
    solver.computeEF();

    while (world.advanceTime()) {
        for (auto &src : sources) src->sample();
        for (Species &sp : species) {
            sp.advance();
            sp.computeNumberDensity();
        }
        world.computeChargeDensity(species);
        solver.solve();
        solver.computeEF();
        if (world.steadyState(species))
            for (Species &sp : species) sp.updateAverages();
        Output::screenOutput(world, species);
        Output::diagOutput(world, species);
        if (world.getTs() % 50 == 0 || world.isLastTimeStep()) Output::fields(world, species);
    }
    cout << "Simulation took " << world.getWallTime() << " seconds" << endl;
    return 0;
}
