// Test-infrastructure driver for the reference C++ classes (NOT product code).
//
// This file is compiled together with the reference's own, unchanged source files
// (World.cpp, Species.cpp, PotentialSolver.cpp and, for ch3/ch9, Source.cpp) straight
// from where they lie under the reference tree, into oracle/_ref/libref_<ch>.so.  It
// exposes the reference's class API through a flat extern "C" surface so that Python
// tests and the CPU-baseline leg of bench.py can drive the reference exactly the way a
// Main.cpp would (ch2/Main.cpp:14-75, ch3/ver2/Main.cpp:16-91, ch9/MT/Main.cpp).
//
// REF_CH selects the chapter API:
//   2  -> ch2       (Species(name,m,q,world), PotentialSolver(world,max_it,tol))
//   3  -> ch3/ver2  (Species(name,m,q,mpw0,world), PotentialSolver(world,type,max_it,tol))
//   9  -> ch9/MT    (as ch3 + World::setNumThreads)
//
// Field exchange uses the flat node index U = k*ni*nj + j*ni + i (ch3/ver2/Field.h:161).
#include <chrono>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "World.h"
#include "Species.h"
#include "PotentialSolver.h"
#if REF_CH != 2
#include "Source.h"
#endif
#if REF_CH == 3
#include "Output.h"
#endif

#define REF_API extern "C" __attribute__((visibility("default")))

// The reference seeds its global generator from std::random_device (World.h:24-33,
// World.cpp:10).  Goldens need a reproducible stream, so this accessor reseeds the very
// same std::mt19937 object.  It adds no members, so the object layout is Rnd's own.
struct RndReseed : public Rnd {
    void reseed(unsigned s) { mt_gen.seed(s); }
};

namespace {

struct RefSim {
    World world;
    std::vector<Species> species;
    std::unique_ptr<PotentialSolver> solver;
#if REF_CH != 2
    std::vector<ColdBeamSource> sources;
#endif
    RefSim(int ni, int nj, int nk) : world(ni, nj, nk) { species.reserve(16); }
};

inline int U(const World& w, int i, int j, int k) { return k * w.ni * w.nj + j * w.ni + i; }

Field* scalar_field(RefSim* s, int which) {
    if (which == 0) return &s->world.phi;
    if (which == 1) return &s->world.rho;
    if (which == 2) return &s->world.node_vol;
    if (which >= 100 && which < 100 + (int)s->species.size()) return &s->species[which - 100].den;
#if REF_CH != 2
    if (which >= 200 && which < 200 + (int)s->species.size()) return &s->species[which - 200].den_ave;
#endif
    return nullptr;
}

}  // namespace

REF_API int ref_chapter() { return REF_CH; }

REF_API void ref_reseed(unsigned seed) { static_cast<RndReseed&>(rnd).reseed(seed); }

REF_API double ref_rnd() { return rnd(); }

REF_API void* ref_create(int ni, int nj, int nk, const double* x0, const double* xm, double dt,
                         int num_ts) {
    RefSim* s = new RefSim(ni, nj, nk);
    s->world.setExtents(double3(x0), double3(xm));
    s->world.setTime(dt, num_ts);
    return s;
}

REF_API void ref_destroy(void* h) { delete static_cast<RefSim*>(h); }

REF_API int ref_add_species(void* h, const char* name, double mass, double charge, double mpw0) {
    RefSim* s = static_cast<RefSim*>(h);
#if REF_CH == 2
    (void)mpw0;
    s->species.push_back(Species(name, mass, charge, s->world));
#else
    s->species.push_back(Species(name, mass, charge, mpw0, s->world));
#endif
    return (int)s->species.size() - 1;
}

#if REF_CH == 2
// loadParticlesBox / loadParticlesBoxQS are defined only in ch2 (ch2/Species.cpp:74-141).
REF_API void ref_load_box(void* h, int sp, const double* x1, const double* x2, double num_den,
                          int num_mp) {
    RefSim* s = static_cast<RefSim*>(h);
    s->species[sp].loadParticlesBox(double3(x1), double3(x2), num_den, num_mp);
}

REF_API void ref_load_box_qs(void* h, int sp, const double* x1, const double* x2, double num_den,
                             const int* nmp) {
    RefSim* s = static_cast<RefSim*>(h);
    s->species[sp].loadParticlesBoxQS(double3(x1), double3(x2), num_den, int3(nmp));
}
#endif

// addParticle semantics: rejects out-of-bounds, rewinds v by half a step with current ef.
REF_API void ref_add_particles(void* h, int sp, long n, const double* pos, const double* vel,
                               const double* mpw) {
    RefSim* s = static_cast<RefSim*>(h);
    for (long p = 0; p < n; p++)
        s->species[sp].addParticle(double3(pos + 3 * p), double3(vel + 3 * p), mpw[p]);
}

// Raw replacement of the particle array (no rewind, no bounds check).
REF_API void ref_set_particles(void* h, int sp, long n, const double* pos, const double* vel,
                               const double* mpw) {
    RefSim* s = static_cast<RefSim*>(h);
    auto& parts = s->species[sp].particles;
    parts.clear();
    parts.reserve(n);
    for (long p = 0; p < n; p++) parts.emplace_back(double3(pos + 3 * p), double3(vel + 3 * p), mpw[p]);
}

REF_API long ref_get_np(void* h, int sp) { return (long)static_cast<RefSim*>(h)->species[sp].getNp(); }

REF_API void ref_get_particles(void* h, int sp, double* pos, double* vel, double* mpw) {
    RefSim* s = static_cast<RefSim*>(h);
    long p = 0;
    for (Particle& part : s->species[sp].particles) {
        for (int c = 0; c < 3; c++) {
            pos[3 * p + c] = part.pos[c];
            vel[3 * p + c] = part.vel[c];
        }
        mpw[p] = part.mpw;
        p++;
    }
}

REF_API void ref_advance(void* h, int sp) { static_cast<RefSim*>(h)->species[sp].advance(); }

REF_API void ref_compute_number_density(void* h, int sp) {
    static_cast<RefSim*>(h)->species[sp].computeNumberDensity();
}

REF_API void ref_compute_charge_density(void* h) {
    RefSim* s = static_cast<RefSim*>(h);
    s->world.computeChargeDensity(s->species);
}

REF_API int ref_advance_time(void* h) { return static_cast<RefSim*>(h)->world.advanceTime() ? 1 : 0; }

REF_API int ref_get_ts(void* h) { return static_cast<RefSim*>(h)->world.getTs(); }

// which: 0 phi, 1 rho, 2 node_vol, 3 ef (3 doubles per node), 4 object_id,
//        100+s den of species s, 200+s den_ave of species s.
REF_API int ref_get_field(void* h, int which, double* out) {
    RefSim* s = static_cast<RefSim*>(h);
    World& w = s->world;
    for (int i = 0; i < w.ni; i++)
        for (int j = 0; j < w.nj; j++)
            for (int k = 0; k < w.nk; k++) {
                int u = U(w, i, j, k);
                if (which == 3) {
                    for (int c = 0; c < 3; c++) out[3 * u + c] = w.ef[i][j][k][c];
                } else if (which == 4) {
#if REF_CH == 2
                    out[u] = 0;
#else
                    out[u] = w.object_id[i][j][k];
#endif
                } else {
                    Field* f = scalar_field(s, which);
                    if (!f) return -1;
                    out[u] = (*f)[i][j][k];
                }
            }
    return 0;
}

REF_API int ref_set_field(void* h, int which, const double* in) {
    RefSim* s = static_cast<RefSim*>(h);
    World& w = s->world;
    for (int i = 0; i < w.ni; i++)
        for (int j = 0; j < w.nj; j++)
            for (int k = 0; k < w.nk; k++) {
                int u = U(w, i, j, k);
                if (which == 3) {
                    for (int c = 0; c < 3; c++) w.ef[i][j][k][c] = in[3 * u + c];
                } else {
                    Field* f = scalar_field(s, which);
                    if (!f) return -1;
                    (*f)[i][j][k] = in[u];
                }
            }
    return 0;
}

// type: 0 GS, 1 PCG, 2 QN (ch3/ver2/PotentialSolver.h:40); ignored by ch2.
REF_API void ref_make_solver(void* h, int type, int max_it, double tol) {
    RefSim* s = static_cast<RefSim*>(h);
#if REF_CH == 2
    (void)type;
    s->solver.reset(new PotentialSolver(s->world, max_it, tol));
#else
    s->solver.reset(new PotentialSolver(s->world, (SolverType)type, max_it, tol));
#endif
}

REF_API void ref_set_reference_values(void* h, double phi0, double Te0, double n0) {
#if REF_CH == 2
    (void)h; (void)phi0; (void)Te0; (void)n0;
#else
    static_cast<RefSim*>(h)->solver->setReferenceValues(phi0, Te0, n0);
#endif
}

REF_API int ref_solve(void* h) { return static_cast<RefSim*>(h)->solver->solve() ? 1 : 0; }

REF_API void ref_compute_ef(void* h) { static_cast<RefSim*>(h)->solver->computeEF(); }

REF_API double ref_get_pe(void* h) { return static_cast<RefSim*>(h)->world.getPE(); }

// out: np, real count, px, py, pz, KE
REF_API void ref_species_diag(void* h, int sp, double* out) {
    Species& s = static_cast<RefSim*>(h)->species[sp];
    out[0] = (double)s.getNp();
    out[1] = s.getRealCount();
    double3 m = s.getMomentum();
    out[2] = m[0];
    out[3] = m[1];
    out[4] = m[2];
    out[5] = s.getKE();
}

#if REF_CH != 2
REF_API void ref_add_sphere(void* h, const double* c, double r, double phi) {
    static_cast<RefSim*>(h)->world.addSphere(double3(c), r, phi);
}

REF_API void ref_add_inlet(void* h) { static_cast<RefSim*>(h)->world.addInlet(); }

REF_API int ref_add_cold_source(void* h, int sp, double v_drift, double den) {
    RefSim* s = static_cast<RefSim*>(h);
    s->sources.emplace_back(s->species[sp], s->world, v_drift, den);
    return (int)s->sources.size() - 1;
}

REF_API void ref_sample_sources(void* h) {
    for (ColdBeamSource& src : static_cast<RefSim*>(h)->sources) src.sample();
}

REF_API int ref_steady_state(void* h) {
    RefSim* s = static_cast<RefSim*>(h);
    return s->world.steadyState(s->species) ? 1 : 0;
}

REF_API void ref_update_averages(void* h, int sp) { static_cast<RefSim*>(h)->species[sp].updateAverages(); }
#endif

#if REF_CH == 3
// The reference's writers (ch3/ver2/Output.cpp) on the current state, into the working
// directory: results/fields_<ts>.vti and one runtime_diags.csv row (the header on the first
// call of the process); the CSV stream is flushed so that the row can be read back.
namespace Output {
extern std::ofstream f_diag;
}
REF_API void ref_output_fields(void* h) {
    RefSim* s = static_cast<RefSim*>(h);
    Output::fields(s->world, s->species);
}
REF_API void ref_output_diag(void* h) {
    RefSim* s = static_cast<RefSim*>(h);
    Output::diagOutput(s->world, s->species);
    Output::f_diag.flush();
}
#endif

#if REF_CH == 9
REF_API void ref_set_num_threads(void* h, int n) { static_cast<RefSim*>(h)->world.setNumThreads(n); }
#endif

// Times nsteps of {advance, computeNumberDensity} over every species with the current ef
// (the particle phase of ch2/Main.cpp:48-52 and ch3/ver2/Main.cpp:56-60).  Returns the
// wall seconds; *particle_steps receives sum over steps of getNp() after advance.
REF_API double ref_time_particle_steps(void* h, int nsteps, long* particle_steps) {
    RefSim* s = static_cast<RefSim*>(h);
    long ps = 0;
    auto t0 = std::chrono::high_resolution_clock::now();
    for (int it = 0; it < nsteps; it++)
        for (Species& sp : s->species) {
            sp.advance();
            ps += (long)sp.getNp();
            sp.computeNumberDensity();
        }
    auto t1 = std::chrono::high_resolution_clock::now();
    *particle_steps = ps;
    return std::chrono::duration<double>(t1 - t0).count();
}

// Times nsteps of {solve, computeEF}; returns wall seconds.
REF_API double ref_time_field_steps(void* h, int nsteps) {
    RefSim* s = static_cast<RefSim*>(h);
    auto t0 = std::chrono::high_resolution_clock::now();
    for (int it = 0; it < nsteps; it++) {
        s->solver->solve();
        s->solver->computeEF();
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
