// PotentialSolver over libpicamd: the ch3/ver2 interface (PotentialSolver.h:40-119) with
// SolverType {GS, PCG, QN}, plus the ch2 constructor (world, max_it, tol) selecting the
// linear SOR of ch2/PotentialSolver.cpp.  The 7-point matrix of the reference is implicit
// (matrix-free kernels); buildMatrix() keeps its one observable effect, the solveQN
// initial guess with the default reference values.
#ifndef PICAMD_HOST_POTENTIALSOLVER_H
#define PICAMD_HOST_POTENTIALSOLVER_H

#include "World.h"

// PCGS: Newton + PCG on the Neumann-reduced (symmetric) Jacobian, PIC_SOLVER_PCG_SYM;
// MG: the same Newton iteration with multigrid-preconditioned CG, PIC_SOLVER_MG
enum SolverType { GS, PCG, QN, PCGS, MG };

class PotentialSolver {
public:
    PotentialSolver(World &world, SolverType type, int max_it, double tol);
    PotentialSolver(World &world, int max_it, double tol);  // ch2: linear SOR, w = 1.4
    ~PotentialSolver();
    PotentialSolver(const PotentialSolver &) = delete;
    PotentialSolver &operator=(const PotentialSolver &) = delete;

    void setReferenceValues(double phi0, double Te0, double n0);
    void computeEF();
    void buildMatrix();
    bool solve();

    // residual norm and iteration count of the last solve()
    double lastResidual() const { return last_res; }
    int lastIterations() const { return last_its; }

protected:
    World &world;
    int kind;
    pic_solver *handle = nullptr;
    double last_res = 0;
    int last_its = 0;
};

#endif
