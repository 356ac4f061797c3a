// Species: the reference's particle container (ch3/ver2/Species.h:11-72; ch2 form without
// mpw0) over the GPU library.  Particles live on the device in tile bins; `particles`,
// `den` and `den_ave` are host mirrors refreshed by syncHost().  addParticle() queues on
// the host and the queue is flushed to the device (with addParticle's half-step rewind in
// the current E) before any device operation that reads particles or changes E.
#ifndef PICAMD_HOST_SPECIES_H
#define PICAMD_HOST_SPECIES_H

#include <memory>
#include <string>
#include <vector>

#include "Field.h"
#include "World.h"

struct Particle {
    double3 pos;
    double3 vel;
    double mpw;
    Particle(double3 x, double3 v, double mpw) : pos{x}, vel{v}, mpw{mpw} {}
};

class Species {
public:
    Species(std::string name, double mass, double charge, double mpw0, World &world);
    Species(std::string name, double mass, double charge, World &world);  // ch2 form
    Species(Species &&) noexcept = default;
    Species(const Species &) = delete;
    Species &operator=(const Species &) = delete;

    size_t getNp();
    double getRealCount();
    double3 getMomentum();
    double getKE();

    void advance();
    void computeNumberDensity();
    void addParticle(double3 pos, double3 vel, double mpw);
    void loadParticlesBox(double3 x1, double3 x2, double num_den, int num_mp);
    void loadParticlesBoxQS(double3 x1, double3 x2, double num_den, int3 num_mp);
    void updateAverages();

    // ---- GPU boundary ---------------------------------------------------------------
    void syncHost(bool with_particles = false);  // den, den_ave (+ particles) mirrors
    void flush();                                // queued addParticle calls -> device
    pic_species *dev();
    static void flushAll(World &w);              // every species of w (before E changes)

    const std::string name;
    const double mass;
    const double charge;
    const double mpw0;

    std::vector<Particle> particles;  // host mirror (syncHost(true))
    Field den;
    Field den_ave;

    struct Impl;

protected:
    World *world;
    std::shared_ptr<Impl> impl;
    double diag_cache[5];
    void diag(double out[5]);
};

#endif
