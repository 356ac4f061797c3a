// Injection sources on the device: ColdBeamSource (ch3/ver2/Source.h) and WarmBeamSource
// (ch4/Source.h).  Each sample() draws its random numbers from a counter-based generator
// keyed by a seed taken from the global rnd at construction, so a seeded run repeats.  A
// source sets its species' frame velocity to its drift, so that a mixed-precision species
// rounds its FP32 velocities at the scale of the deviations from the beam.
#ifndef PICAMD_HOST_SOURCE_H
#define PICAMD_HOST_SOURCE_H

#include "Species.h"
#include "World.h"

class Source {
public:
    virtual void sample() = 0;
    virtual ~Source() {}
};

class ColdBeamSource : public Source {
public:
    ColdBeamSource(Species &species, World &world, double v_drift, double den);
    void sample() override;

protected:
    Species *sp;
    World *world;
    double v_drift;
    double den;
    unsigned long long seed;
    unsigned long long counter = 0;
};

class WarmBeamSource : public Source {
public:
    WarmBeamSource(Species &species, World &world, double v_drift, double den, double T);
    void sample() override;

protected:
    Species *sp;
    World *world;
    double v_drift;
    double den;
    double T;
    unsigned long long seed;
    unsigned long long counter = 0;
};

#endif
