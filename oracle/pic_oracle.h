/*
 * pic_oracle.h -- FP64 CPU restatement of the reference ES-PIC inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared against;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path (libpicamd.so) never calls into it.
 *
 * Every function restates one reference function with the same loop order and the same
 * floating-point operation order, so that it is bit-exact against the reference built
 * from source (oracle/_ref) under the same compiler flags.  Node arrays are flat with
 * U = k*ni*nj + j*ni + i (ch3/ver2/Field.h:161); 3-vectors are interleaved (3 per node /
 * particle).  object_id: 0 open node, >0 fixed (Dirichlet) node.
 *
 * Pinning (what checks each function, beyond the reference build):
 *   see the module docstring of oracle/oracle.py.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* physical constants, ch2/World.h:11-20 */
#define ORC_EPS_0 8.85418782e-12
#define ORC_QE 1.602176565e-19
#define ORC_AMU 1.660538921e-27
#define ORC_ME 9.10938215e-31
#define ORC_K 1.380648e-23
#define ORC_PI 3.141592653

typedef struct {
    int ni, nj, nk;
    double x0[3], xm[3], dh[3], xc[3];
} orc_grid;

typedef struct {
    uint32_t mt[624];
    int idx;
} orc_mt19937;

/* World::setExtents (ch2/World.cpp:21-37) */
void orc_grid_init(orc_grid *g, int ni, int nj, int nk, const double *x0, const double *xm);
/* World::computeNodeVolumes (ch2/World.cpp:58-69) */
void orc_node_volumes(const orc_grid *g, double *nv);
/* World::XtoL (ch2/World.h:72-78) */
void orc_xtol(const orc_grid *g, const double *x, double *lc);
/* World::inBounds (ch2/World.h:59-63) */
int orc_in_bounds(const orc_grid *g, const double *x);
/* World::inSphere (ch3/ver2/World.cpp:119-126) */
int orc_in_sphere(const double *x, const double *c, double r2);
/* Field_<double3>::gather (ch2/Field.h:186-208) */
void orc_gather3(const orc_grid *g, const double *ef3, const double *lc, double *out);
/* Field_<double>::gather */
double orc_gather1(const orc_grid *g, const double *f, const double *lc);
/* Field_<double>::scatter (ch2/Field.h:164-183) */
void orc_scatter(const orc_grid *g, double *f, const double *lc, double v);

/* Species::advance, reflecting walls (ch2/Species.cpp:7-38) */
void orc_push_reflect(const orc_grid *g, long np, double *pos, double *vel, const double *ef3,
                      double dt, double charge, double mass);
/* Species::advance, absorbing sphere + walls, swap-with-last compaction
 * (ch3/ver2/Species.cpp:7-48).  ids (nullable) is permuted with the particles.
 * Returns the new particle count. */
long orc_push_absorb(const orc_grid *g, long np, double *pos, double *vel, double *mpw, long *ids,
                     const double *ef3, double dt, double charge, double mass,
                     const double *sphere_c, double sphere_r2);
/* Species::computeNumberDensity (ch2/Species.cpp:41-52): clear, scatter, /= node_vol */
void orc_number_density(const orc_grid *g, long np, const double *pos, const double *mpw,
                        const double *nv, double *den);
/* World::computeChargeDensity (ch2/World.cpp:46-54) */
void orc_charge_density(long nn, int nsp, const double *charges, const double *const *dens,
                        double *rho);

/* Species::addParticle (ch2/Species.cpp:55-71): returns 1 if appended at *np. */
int orc_add_particle(const orc_grid *g, const double *ef3, double dt, double charge, double mass,
                     const double *x, const double *v, double w, long *np, double *pos,
                     double *vel, double *mpw);
/* Species::loadParticlesBox (ch2/Species.cpp:74-97) */
void orc_load_box(const orc_grid *g, orc_mt19937 *rng, const double *x1, const double *x2,
                  double num_den, int num_mp, const double *ef3, double dt, double charge,
                  double mass, long *np, double *pos, double *vel, double *mpw);
/* Species::loadParticlesBoxQS (ch2/Species.cpp:101-141) */
void orc_load_box_qs(const orc_grid *g, const double *x1, const double *x2, double num_den,
                     const int *num_mp, const double *ef3, double dt, double charge, double mass,
                     long *np, double *pos, double *vel, double *mpw);
/* ColdBeamSource::sample (ch3/ver2/Source.cpp:4-27); appends at *np (capacity cap) */
int orc_cold_beam_sample(const orc_grid *g, orc_mt19937 *rng, double v_drift, double den,
                         double mpw0, const double *ef3, double dt, double charge, double mass,
                         long cap, long *np, double *pos, double *vel, double *mpw);
/* WarmBeamSource::sample (ch4/Source.cpp:31-56) with Species::sampleVth /
 * sampleIsotropicVel (ch4/Species.cpp:149-174) */
int orc_warm_beam_sample(const orc_grid *g, orc_mt19937 *rng, double v_drift, double den,
                         double T, double mpw0, const double *ef3, double dt, double charge,
                         double mass, long cap, long *np, double *pos, double *vel, double *mpw);

/* PotentialSolver::solve, linear GS-SOR w=1.4 (ch2/PotentialSolver.cpp:11-67).
 * Returns 1 if converged.  *l2 last residual norm, *its iterations run. */
int orc_solve_sor_linear(const orc_grid *g, double *phi, const double *rho, int max_it, double tol,
                         double *l2, int *its);
/* PotentialSolver::solveGS, nonlinear Boltzmann GS-SOR (ch3/ver2/PotentialSolver.cpp:334-430) */
int orc_solve_gs_boltzmann(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                           double phi0, double Te0, double n0, int max_it, double tol, double *l2,
                           int *its);
/* PotentialSolver::solveQN (ch3/ver2/PotentialSolver.cpp:204-222) */
void orc_solve_qn(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                  double phi0, double Te0, double n0);
/* PotentialSolver::solveNRPCG + solvePCGLinear + solveGSLinear, matrix-free restatement of
 * buildMatrix / Matrix ops (ch3/ver2/PotentialSolver.cpp:24-76,146-201,225-331,433-461).
 * stats (nullable, 4 doubles): NR iterations, total PCG iterations, GS fallbacks, final norm */
int orc_solve_nrpcg(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                    double phi0, double Te0, double n0, int max_it, double tol, double *stats);
/* PotentialSolver::computeEF (ch3/ver2/PotentialSolver.cpp:465-504) */
void orc_compute_ef(const orc_grid *g, const double *phi, double *ef3);

/* World::addSphere (ch3/ver2/World.cpp:88-106), World::addInlet (:109-116) */
void orc_add_sphere(const orc_grid *g, int *object_id, double *phi, const double *c, double r,
                    double phi_s);
void orc_add_inlet(const orc_grid *g, int *object_id, double *phi);

/* World::getPE (ch2/World.cpp:72-84) */
double orc_get_pe(const orc_grid *g, const double *ef3, const double *nv);
/* Species::getRealCount / getMomentum / getKE (ch2/Species.cpp:144-168).
 * out: real count, px, py, pz, KE */
void orc_species_diag(long np, const double *vel, const double *mpw, double mass, double *out);
/* Field_::updateAverage (ch3/ver2/Field.h:214-220) */
void orc_update_average(long nn, double *ave, const double *inst, int *samples);

/* std::mt19937 + libstdc++ std::uniform_real_distribution<double>(0,1) (World.h:24-33) */
void orc_mt_seed(orc_mt19937 *r, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937 *r);
double orc_mt_uniform(orc_mt19937 *r);

/* ---- differentiable flow past a sphere (pic_oracle_diff.c) --------------------------------
 * The checker of the GPU's pic_adjoint_* (no reference counterpart): the same loop, soft
 * absorption, injection and Philox draws, in FP64 on the CPU.  Field order as
 * pic_adjoint_config. */
typedef struct {
    int ni, nj, nk;
    double x0[3], xm[3];
    double sphere_c[3], sphere_r;
    double dt;
    int steps, checkpoint_every;
    double mass, charge, n0, T_ion;
    long n_init, n_inject;
    double soft_width, cg_tol, nr_tol, adj_rel_tol;
    uint64_t seed;
} orc_dsph_config;

/* Philox4x32-10 (Salmon et al., SC'11) */
void orc_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]);
/* Newton on the Neumann-reduced system (the fixed point of solveNRPCG,
 * ch3/ver2/PotentialSolver.cpp:225-296) with Jacobi-preconditioned CG from y = 0 until
 * RMS(J y - F) < cg_tol, until RMS(y) < nr_tol.  stats: Newton its, CG its, 0, last norm. */
int orc_solve_nr_sym(const orc_grid *g, double *phi, const double *rho, const int *object_id,
                     double phi0, double Te0, double n0, double cg_tol, double nr_tol, int max_cg,
                     double *stats);
/* K steps at theta = (drift, sphere potential, Te).  current: the ion current to the sphere;
 * den_out, phi_out (nullable, nn): density and potential after the last step; stats
 * (nullable, 4): solves, CG its, live weight at the end, weight taken by the sphere.
 * Returns 1 if every solve converged. */
int orc_dsph_run(const orc_dsph_config *c, const double theta[3], double *current, double *den_out,
                 double *phi_out, double *stats);

#ifdef __cplusplus
}
#endif
#endif
