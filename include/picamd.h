/*
 * picamd.h -- C ABI of the MI355X (gfx950) ES-PIC inner loop.
 *
 * The reference has no FFI: its programs compile World / Species / PotentialSolver
 * straight into Main.cpp.  This ABI is the boundary those classes call through, one
 * entry point per reference operation on the hot path (the reference function each one
 * replaces is cited beside it).  The C++ drop-in classes in
 * plasma-simulations-by-example_amd/host/ and the Python binding pic_amd.py both bind
 * these symbols; nothing else is exported.
 *
 * Conventions
 *  - Plain C types only; device memory is owned by the library behind opaque handles.
 *    Host arrays are passed as pointers + sizes and copied synchronously.
 *  - Node arrays use the reference's flat index U = k*ni*nj + j*ni + i
 *    (ch3/ver2/Field.h:161); 3-vectors are interleaved (x,y,z per node / particle).
 *  - Every call returns an int status: PIC_OK (0), PIC_NOT_CONVERGED (1, solver only,
 *    mirrors `bool solve()` returning false), or a negative error.  Nothing throws
 *    across the ABI; pic_last_error() holds a message for the calling thread.
 *  - All work of one world is ordered on that world's own HIP stream (default
 *    priority).  Calls that return host data synchronise that stream first.
 */
#ifndef PICAMD_H
#define PICAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: PIC_T_COUNT grew to 8 (PIC_T_PUSH_KERNEL); 3: pic_solver_get_info; 4: pic_adjoint_*;
 * 5: pic_ensemble_*; 6: tuning table, pic_species_set_launch, pic_solver_set_mg;
 * 7: pic_species_set_frame_velocity */
#define PICAMD_ABI_VERSION 7

/* status codes */
#define PIC_OK 0
#define PIC_NOT_CONVERGED 1
#define PIC_ERR_ARG (-1)
#define PIC_ERR_HIP (-2)
#define PIC_ERR_CAPACITY (-3)
#define PIC_ERR_STATE (-4)
#define PIC_ERR_NODEVICE (-5)

/* particle storage / arithmetic of a species */
#define PIC_MIXED 0 /* int32 packed cell + FP32 in-cell offset, FP32 velocity and weight */
#define PIC_FP64 1  /* FP64 physical position, velocity and weight (the reference's form) */

/* wall treatment in advance() */
#define PIC_BC_REFLECT 0 /* specular walls, ch2/Species.cpp:32-36 */
#define PIC_BC_ABSORB 1  /* absorbing walls and sphere, ch3/ver2/Species.cpp:28-32 */

/* solver types: GS, PCG, QN as ch3/ver2/PotentialSolver.h:40; SOR_LINEAR is the ch2 solver.
 * PCG_SYM is the Newton + Jacobi-PCG of PCG on the system with the Neumann face nodes
 * eliminated (each face node equals the interior node its copy chain reaches, as the
 * reference's Neumann rows require at convergence): the same discrete solution, but a
 * symmetric negative-definite Jacobian, so CG converges where the reference's one-sided
 * Neumann rows (ch3/ver2/PotentialSolver.cpp:180-187) make it stall on large grids. */
#define PIC_SOLVER_GS 0
#define PIC_SOLVER_PCG 1
#define PIC_SOLVER_QN 2
#define PIC_SOLVER_SOR_LINEAR 3
#define PIC_SOLVER_PCG_SYM 4
/* MG: the Newton iteration and stopping tests of PCG_SYM with CG preconditioned by one
 * geometric-multigrid V-cycle instead of the diagonal (Matrix::invDiagonal,
 * ch3/ver2/PotentialSolver.cpp:52-57): the same discrete solution in a number of CG
 * iterations that barely grows with the grid. */
#define PIC_SOLVER_MG 5

/* world node fields for pic_world_get_field / pic_world_set_field */
#define PIC_FIELD_PHI 0       /* World::phi */
#define PIC_FIELD_RHO 1       /* World::rho */
#define PIC_FIELD_NODE_VOL 2  /* World::node_vol */
#define PIC_FIELD_EF 3        /* World::ef (3 doubles per node) */
#define PIC_FIELD_OBJECT_ID 4 /* World::object_id (as double) */

/* phase timers (pic_world_get_timers) */
#define PIC_T_PUSH 0    /* fused push + gather + deposit + rebin */
#define PIC_T_DEPOSIT 1 /* density finish (window reduction, /node_vol) */
#define PIC_T_SORT 2    /* full counting sort */
#define PIC_T_SOLVE 3   /* Poisson solve */
#define PIC_T_EF 4      /* computeEF */
#define PIC_T_RHO 5     /* computeChargeDensity */
#define PIC_T_INJECT 6  /* device injection */
#define PIC_T_PUSH_KERNEL 7 /* the fused push kernel alone (PIC_T_PUSH less the segment plan) */
#define PIC_T_COUNT 8

typedef struct pic_world pic_world;
typedef struct pic_species pic_species;
typedef struct pic_solver pic_solver;

/* ---- library ---------------------------------------------------------------------- */
int pic_abi_version(void);
const char *pic_last_error(void);
/* number of visible HIP devices (0 without a GPU; no other call needs one to be absent) */
int pic_device_count(int *n);

/* ---- World (ch3/ver2/World.h:38-147) ---------------------------------------------- */
/* World::World(ni,nj,nk) (ch3/ver2/World.cpp:14-20): allocates the node fields on the
 * current device and creates the world's stream. */
int pic_world_create(int ni, int nj, int nk, pic_world **out);
int pic_world_destroy(pic_world *w);
/* World::setExtents (ch3/ver2/World.cpp:23-37): dh = (xm-x0)/(nn-1); node volumes. */
int pic_world_set_extents(pic_world *w, const double x0[3], const double xm[3]);
/* World::addSphere (ch3/ver2/World.cpp:88-106): object_id=1, phi=phi_s on nodes with
 * |x-c|^2 <= r^2; the sphere also absorbs particles in advance(PIC_BC_ABSORB). */
int pic_world_add_sphere(pic_world *w, const double c[3], double r, double phi_s);
/* World::addInlet (ch3/ver2/World.cpp:109-116): k=0 plane object_id=2, phi=0. */
int pic_world_add_inlet(pic_world *w);
/* Host copy of a node field (the updateHostPhi() boundary of ch9/CUDA/PotentialSolver.cu:
 * 323-329).  out holds nn (or 3*nn for PIC_FIELD_EF) doubles. */
int pic_world_get_field(pic_world *w, int which, double *out);
/* Host write of PIC_FIELD_PHI, PIC_FIELD_RHO or PIC_FIELD_EF. */
int pic_world_set_field(pic_world *w, int which, const double *in);
/* World::computeChargeDensity (ch3/ver2/World.cpp:47-55): rho = sum q_s*den_s over
 * charged species (den_s must be current: call pic_species_compute_number_density). */
int pic_world_compute_charge_density(pic_world *w, pic_species *const *sp, int nsp);
/* World::getPE (ch3/ver2/World.cpp:73-85): 0.5*eps0*sum |E|^2 * node_vol (FP64). */
int pic_world_get_pe(pic_world *w, double *pe);
/* Waits for all queued work of the world. */
int pic_world_sync(pic_world *w);
/* Raw hipStream_t of the world (for callers that time or overlap work on it). */
int pic_world_stream(pic_world *w, void **stream);
/* Phase timers (HIP events around each launch; off by default).  ms[PIC_T_COUNT] receives
 * accumulated milliseconds, calls[PIC_T_COUNT] launch counts; reset clears them. */
int pic_world_set_timing(pic_world *w, int enable);
int pic_world_get_timers(pic_world *w, double *ms, long *calls, int reset);

/* ---- Species (ch3/ver2/Species.h:11-72) ------------------------------------------- */
/* Species(name, mass, charge, mpw0, world).  precision: PIC_MIXED or PIC_FP64.
 * capacity: particles the species may hold (grown on demand by the host calls that add
 * particles).  track_ids: carry a uint32 id per particle (for per-particle parity). */
int pic_species_create(pic_world *w, double mass, double charge, double mpw0, int precision,
                       long capacity, int track_ids, pic_species **out);
int pic_species_destroy(pic_species *s);
/* Species::addParticle (ch3/ver2/Species.cpp:65-81) for n particles: rejects positions
 * outside [x0,xm), rewinds v by (q/m)*E(x)*dt/2 with the current E, appends.
 * *n_added (nullable) receives the number appended.  ids are assigned in call order. */
int pic_species_add_particles(pic_species *s, long n, const double *pos, const double *vel,
                              const double *mpw, double dt, long *n_added);
/* Raw replacement of all particles (no rewind, no bounds rejection); ids nullable. */
int pic_species_set_particles(pic_species *s, long n, const double *pos, const double *vel,
                              const double *mpw, const uint32_t *ids);
/* Host copy of all particles in device order (pos/vel: 3*np doubles). ids nullable. */
int pic_species_get_particles(pic_species *s, double *pos, double *vel, double *mpw,
                              uint32_t *ids);
/* Species::getNp (ch3/ver2/Species.h:29) */
int pic_species_get_np(pic_species *s, long *np);
/* Species::advance: fused leapfrog push + CIC gather (+ CIC deposit of the new positions,
 * consumed by computeNumberDensity) with reflecting (ch2/Species.cpp:7-38) or absorbing
 * walls/sphere and compaction (ch3/ver2/Species.cpp:7-48).  A push that had to drop
 * particles (its staging list full) is reported one call later, without a stream sync: the
 * next advance, computeNumberDensity, inject or getNp returns PIC_ERR_CAPACITY. */
int pic_species_advance(pic_species *s, int bc, double dt);
/* Species::computeNumberDensity (ch3/ver2/Species.cpp:51-62): den = scatter(mpw)/node_vol.
 * Reuses the deposit fused into advance when nothing changed since; deposits afresh
 * otherwise. */
int pic_species_compute_number_density(pic_species *s);
/* Host copy of Species::den / den_ave (nn doubles). */
int pic_species_get_den(pic_species *s, double *out);
int pic_species_get_den_ave(pic_species *s, double *out);
/* Species::updateAverages -> Field_::updateAverage (ch3/ver2/Field.h:214-220) */
int pic_species_update_averages(pic_species *s);
/* Species::getRealCount / getMomentum / getKE (ch3/ver2/Species.cpp:84-108), FP64
 * reductions.  out[5] = real count, px, py, pz, KE. */
int pic_species_diag(pic_species *s, double out[5]);
/* No counterpart in the reference.  For callers that cannot allocate device or pinned memory
 * (HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * outputs as pic_species_diag's full reduction over all particles, with the partial sums in a
 * static device array of the library (6.3 MB per GPU, held until the process ends) instead
 * of the species' own buffers.  It allocates nothing.  Grids of more than 131008 tiles
 * (8x8x8 cells each) return PIC_ERR_CAPACITY.  Calls on one GPU share that array: calls
 * that overlap on one GPU (several host threads or streams at once) race. */
int pic_species_diag_noalloc(pic_species *s, double out[5]);
/* Counting sort of all particles into tile bins (tile-major cell order). */
int pic_species_sort(pic_species *s);
/* No counterpart in the reference (launch shape of the push, the autotuner's parameters):
 * particles per push workgroup and push-map entries per XCD stripe.  Species take both from
 * the tuning table at creation (else 32768 and 32). */
int pic_species_set_launch(pic_species *s, long push_chunk, int xcd_stripe);
/* No counterpart in the reference (a storage choice of the mixed form): a PIC_MIXED species
 * stores each velocity as v - u in FP32, u its frame velocity (default 0).  A drifting
 * species (a beam at v_drift along z) keeps FP32 rounding at the scale of its deviations
 * from the drift instead of the drift itself.  Every interface still takes and returns v;
 * stored particles are re-expressed in the new frame.  PIC_FP64 species store v and ignore
 * the call.  The C++ Source classes set u = v_drift z on a species they feed. */
int pic_species_set_frame_velocity(pic_species *s, const double u[3]);
/* ColdBeamSource::sample (ch3/ver2/Source.cpp:4-27) when T == 0, WarmBeamSource::sample
 * (ch4/Source.cpp:31-56) when T > 0 [K], on the device: num = (int)(den*v*A*dt/mpw0 + U)
 * particles uniform on the k=0 face, v = v_drift z (+ isotropic thermal speed), then the
 * addParticle rewind.  Random numbers: Philox4x32-10 keyed by (seed, counter), so a run
 * is reproducible; *n_added (nullable) receives the count. */
int pic_species_inject_beam(pic_species *s, double v_drift, double den, double T, double dt,
                            uint64_t seed, uint64_t counter, long *n_added);
/* Device bulk load of num_mp particles uniform in [x1,x2) with zero velocity and weight
 * num_den*V/num_mp (Species::loadParticlesBox, ch2/Species.cpp:74-97), Philox random
 * numbers; for the 1e8-1e9 particle workloads that a host load cannot feed. Optional
 * drift/thermal velocity as inject_beam. */
int pic_species_load_box_device(pic_species *s, const double x1[3], const double x2[3],
                                double num_den, long num_mp, double v_drift, double T,
                                uint64_t seed);

/* ---- PotentialSolver (ch3/ver2/PotentialSolver.h:40-119; ch2/PotentialSolver.h) ---- */
/* PotentialSolver(world, type, max_it, tol).  Unlike the reference constructor this does
 * not run solveQN(); the drop-in class calls pic_solver_solve_qn to reproduce it. */
int pic_solver_create(pic_world *w, int type, int max_it, double tol, pic_solver **out);
int pic_solver_destroy(pic_solver *s);
/* setReferenceValues(phi0, Te0, n0) -- note Te0 second, as the reference. */
int pic_solver_set_reference_values(pic_solver *s, double phi0, double Te0, double n0);
/* solve(): PIC_OK when converged, PIC_NOT_CONVERGED otherwise.  residual (nullable) gets
 * the last convergence norm, iters (nullable) the iterations (PCG: total inner PCG
 * iterations over the Newton steps). */
int pic_solver_solve(pic_solver *s, double *residual, int *iters);
/* What the last solve() did, for callers that check which path ran: info[0] Newton
 * iterations (NR solvers; 0 otherwise), info[1] the iterations solve() reported (SOR sweeps,
 * or inner PCG iterations over all Newton steps), info[2] how many Newton steps fell back
 * from PCG to solveGSLinear (ch3/ver2/PotentialSolver.cpp:270-271), info[3] 1 when it
 * converged. */
int pic_solver_get_info(pic_solver *s, int info[4]);
/* No counterpart in the reference (launch shapes of PIC_SOLVER_MG, the autotuner's
 * parameters): smoothing sweep pairs on the 7-point levels, first level with Galerkin
 * operators, largest level of the single-workgroup coarse cycle (nodes); 0 keeps the tuning
 * table's value (else the built-in default).  The hierarchy is rebuilt by the next solve. */
int pic_solver_set_mg(pic_solver *s, int sweeps, int gal_from, long fuse_max);
/* solveQN() with the solver's reference values (ch3/ver2/PotentialSolver.cpp:204-222). */
int pic_solver_solve_qn(pic_solver *s);
/* computeEF() (ch3/ver2/PotentialSolver.cpp:465-504). */
int pic_solver_compute_ef(pic_solver *s);

/* ---- Differentiable flow past a sphere and its adjoint ----------------------------------
 * No counterpart in the reference: K steps of the ch3/ver2 loop (ch3/ver2/Main.cpp:50-70:
 * inject with the addParticle rewind, advance, computeNumberDensity, computeChargeDensity,
 * the Newton solve of the Boltzmann-electron Poisson equation, computeEF) in FP64, as a
 * differentiable function of theta = (drift speed [m/s], sphere potential [V], Te [eV]), with
 * a hand-written adjoint (reverse leapfrog, transposes of the CIC gather and scatter and of
 * computeEF, one linear solve per step with the symmetric Newton Jacobian of the reduced
 * system) over checkpoints stored every checkpoint_every steps.  Differences from the
 * reference loop that make it differentiable: absorption at the sphere and the five open
 * walls multiplies a particle's weight by clamp(distance/(soft_width*dh), 0, 1) (0 inside,
 * as inSphere || !inBounds kills it) instead of killing it outright; each step injects
 * n_inject particles whose weight n0*v_drift*A*dt/n_inject carries the flux; random numbers
 * are Philox keyed by (seed, slot, step).  The world is config's sphere domain with an inlet
 * at k = 0; the solve is Newton + multigrid-preconditioned CG (PIC_SOLVER_MG). */
typedef struct pic_adjoint pic_adjoint;
typedef struct {
    int ni, nj, nk;
    double x0[3], xm[3];
    double sphere_c[3], sphere_r;
    double dt;
    int steps;             /* K */
    int checkpoint_every;  /* M */
    double mass, charge;   /* the ion species */
    double n0;             /* ion density of the initial fill and the beam; Boltzmann n0 */
    double T_ion;          /* ion temperature [K] of the fill and the beam (0: cold) */
    long n_init;           /* macro-particles of the initial uniform fill */
    long n_inject;         /* macro-particles injected per step */
    double soft_width;     /* width of the soft absorbing layers, in cells */
    double cg_tol;         /* CG tolerance of each Newton step (<= 0: 1e-6) */
    double nr_tol;         /* Newton stopping test, RMS of the correction [V] (<= 0: 1e-8) */
    double adj_rel_tol;    /* adjoint linear solves, relative to the RHS RMS (<= 0: 1e-8) */
    uint64_t seed;
} pic_adjoint_config;
int pic_adjoint_create(const pic_adjoint_config *cfg, pic_adjoint **out);
int pic_adjoint_destroy(pic_adjoint *a);
/* Runs the K steps at theta.  current (nullable): the ion current to the sphere,
 * charge * sum over steps of the weight the sphere took / (K dt) [A].  den_out, phi_out
 * (nullable): DEVICE pointers, nn doubles each: the ion number density and the potential
 * after step K. */
int pic_adjoint_forward(pic_adjoint *a, const double theta[3], double *current, double *den_out,
                        double *phi_out);
/* Gradient of L at the last forward's theta, for dL/dcurrent = dcurrent, dL/dden = dden and
 * dL/dphi = dphi (DEVICE pointers, nn doubles each; null for 0).  grad[3] = dL/dtheta. */
int pic_adjoint_backward(pic_adjoint *a, double dcurrent, const double *dden, const double *dphi,
                         double grad[3]);
/* out[4]: last forward ms, last backward ms, Newton solves of the last forward, CG iterations
 * of the last backward's linear solves. */
int pic_adjoint_stats(pic_adjoint *a, double out[4]);

/* ---- Tuning table --------------------------------------------------------------------------
 * No counterpart in the reference.  Launch shapes measured by the autotuner
 * (plasma-simulations-by-example_amd/autotune.py), one entry per line:
 * "<kind> <nodes> <particles> key=value ...", kinds push, mg and ensemble.  Creation of a
 * species, a multigrid hierarchy or an ensemble takes the entry of its kind nearest to its
 * size; without one the built-in defaults apply.  The table is read from PICAMD_TUNING, else
 * tuning/table.txt beside libpicamd.so, on first use. */
/* (Re)loads the table from path (null: the default location). */
int pic_tuning_load(const char *path);
/* The value of key for kind at (nodes, particles), dflt when the table has none. */
int pic_tuning_lookup(const char *kind, double nodes, double particles, const char *key,
                      double dflt, double *value);

/* ---- Batched ensemble of grounded boxes ----------------------------------------------------
 * No counterpart in the reference: many independent ch2 worlds (ch2/Main.cpp:14-75) of uneven
 * size stepped together.  Each box has specular walls (ch2/Species.cpp:32-36), linear SOR
 * w = 1.4 with the residual every 25 iterations (ch2/PotentialSolver.cpp:11-67) and
 * computeEF (:70-110).  One launch per phase covers every box through a box-descriptor
 * table, and the step's launches are captured once in a hipGraph and replayed.  Particle
 * counts per (box, species) are fixed at creation (the walls reflect: none is lost). */
typedef struct pic_ensemble pic_ensemble;
typedef struct {
    int ni, nj, nk;       /* nodes */
    double x0[3], xm[3];  /* World::setExtents */
} pic_box;
/* nbox boxes, nsp species (mass, charge per species), np[b*nsp + s] particles of species s
 * in box b; precision PIC_MIXED or PIC_FP64; track_ids: carry a uint32 id per particle. */
int pic_ensemble_create(int nbox, const pic_box *boxes, int nsp, const double *mass,
                        const double *charge, const long *np, int precision, int track_ids,
                        pic_ensemble **out);
int pic_ensemble_destroy(pic_ensemble *e);
/* Host upload of all n particles of (box b, species s) (n must equal the count at creation;
 * FP64 interleaved positions and velocities, no rewind); ids become 0..n-1. */
int pic_ensemble_set_particles(pic_ensemble *e, int b, int s, long n, const double *pos,
                               const double *vel, const double *mpw);
/* Species::loadParticlesBox (ch2/Species.cpp:74-97) on the device: the range's particles
 * uniform in [x1, x2) at rest, weight num_den*V/n; Philox draws keyed by (seed, b, s). */
int pic_ensemble_load_box(pic_ensemble *e, int b, int s, const double x1[3], const double x2[3],
                          double num_den, uint64_t seed);
int pic_ensemble_get_particles(pic_ensemble *e, int b, int s, double *pos, double *vel,
                               double *mpw, uint32_t *ids);
/* Counting sort of every (box, species) range by cell. */
int pic_ensemble_sort(pic_ensemble *e);
/* Tile and launch shapes (the autotuner's parameters): particles per push workgroup, LDS
 * deposit window in nodes (0: global atomics only; at most 8192), threads per SOR
 * workgroup (256, 512, 1024), steps between counting sorts (0: never); use_graph 0 launches
 * each step's kernels eagerly instead of replaying the captured graph. */
int pic_ensemble_set_tuning(pic_ensemble *e, int chunk_particles, int window_nodes,
                            int sor_threads, int sort_every, int use_graph);
/* The initial solve of every box before the loop (ch2/Main.cpp:38-42): density, rho, SOR,
 * computeEF with the particles where they stand. */
int pic_ensemble_solve_initial(pic_ensemble *e, int max_it, double tol);
/* nsteps steps of advance -> computeNumberDensity -> computeChargeDensity -> solve ->
 * computeEF for every box (ch2/Main.cpp:45-61), one hipGraph replay per step. */
int pic_ensemble_step(pic_ensemble *e, int nsteps, double dt, int max_it, double tol);
/* Host copy of box b's field: 0 phi, 1 rho, 2 E (3 per node), 10+s density of species s. */
int pic_ensemble_get_field(pic_ensemble *e, int b, int which, double *out);
/* Per box: SOR iterations of the last solve (negative when it did not converge) and the last
 * residual norm (nullable arrays of nbox). */
int pic_ensemble_solver_status(pic_ensemble *e, int *iters, double *residual);
/* out[5*(b*nsp+s) + q]: sum w, sum w vx, sum w vy, sum w vz, sum w |v|^2 of each range. */
int pic_ensemble_diag(pic_ensemble *e, double *out);
/* out[5]: ms per step of the last pic_ensemble_step, kernels in the step's graph, particles,
 * nodes, push workgroups. */
int pic_ensemble_stats(pic_ensemble *e, double out[5]);

#ifdef __cplusplus
}
#endif
#endif
