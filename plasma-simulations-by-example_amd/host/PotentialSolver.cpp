// PotentialSolver over libpicamd (see PotentialSolver.h).
#include "PotentialSolver.h"

#include <iostream>

#include "Species.h"

using namespace std;

PotentialSolver::PotentialSolver(World &world, SolverType type, int max_it, double tol)
    : world(world),
      kind(type == GS     ? PIC_SOLVER_GS
           : type == PCG  ? PIC_SOLVER_PCG
           : type == PCGS ? PIC_SOLVER_PCG_SYM
           : type == MG   ? PIC_SOLVER_MG
                          : PIC_SOLVER_QN) {
    pic_check(pic_solver_create(world.dev(), kind, max_it, tol, &handle), "pic_solver_create");
    buildMatrix();
}

PotentialSolver::PotentialSolver(World &world, int max_it, double tol)
    : world(world), kind(PIC_SOLVER_SOR_LINEAR) {
    pic_check(pic_solver_create(world.dev(), kind, max_it, tol, &handle), "pic_solver_create");
}

PotentialSolver::~PotentialSolver() { pic_solver_destroy(handle); }

void PotentialSolver::setReferenceValues(double phi0, double Te0, double n0) {
    pic_check(pic_solver_set_reference_values(handle, phi0, Te0, n0), "setReferenceValues");
}

// the reference constructor ends buildMatrix with solveQN() on the reference values it
// has at that point (ch3/ver2/PotentialSolver.cpp:200)
void PotentialSolver::buildMatrix() { pic_check(pic_solver_solve_qn(handle), "solveQN"); }

bool PotentialSolver::solve() {
    int rc = pic_solver_solve(handle, &last_res, &last_its);
    if (rc == PIC_NOT_CONVERGED) {
        if (kind == PIC_SOLVER_PCG || kind == PIC_SOLVER_PCG_SYM || kind == PIC_SOLVER_MG)
            cout << "NR+PCG failed to converge, norm = " << last_res << endl;
        else
            cerr << "GS failed to converge, L2=" << last_res << endl;
        return false;
    }
    pic_check(rc, "solve");
    return true;
}

void PotentialSolver::computeEF() {
    // queued particles are rewound with the E they were added under
    Species::flushAll(world);
    pic_check(pic_solver_compute_ef(handle), "computeEF");
}
